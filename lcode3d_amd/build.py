"""Builds liblcode3d_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m lcode3d_amd.build`` (or ``build()`` from Python).
Each source is compiled to an object in parallel (one code object per
source: every kernel is launched from the file that defines it), then the
objects are linked into the shared library.
"""

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ("particles.hip", "fields.hip", "dense.hip", "diag.hip", "unet.hip", "api.hip")
OUTPUT = os.path.join(HERE, "liblcode3d_hip.so")
ARCH = "gfx950"

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _run(cmds, verbose):
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                      stderr=subprocess.STDOUT, text=True))
    failed = False
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")


def build(verbose=False):
    """Compile every source into OUTPUT (always recompiles)."""
    with tempfile.TemporaryDirectory(prefix="lcode3d_build_") as tmpdir:
        objs, cmds = [], []
        for s in SOURCES:
            obj = os.path.join(tmpdir, s.replace(".hip", ".o"))
            objs.append(obj)
            cmds.append([hipcc(), *FLAGS, "-c", os.path.join(CSRC, s), "-o", obj])
        _run(cmds, verbose)
        tmp = OUTPUT + ".tmp"
        _run([[hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs,
               "-o", tmp]], verbose)
        os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
