"""Builds liblcode3d_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m lcode3d_amd.build`` (or ``build()`` from Python).
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ("particles.hip", "fields.hip", "dense.hip", "diag.hip", "predictor.hip", "api.hip")
OUTPUT = os.path.join(HERE, "liblcode3d_hip.so")
ARCH = "gfx950"

FLAGS = [
    "-O3", "-std=c++17", "-shared", "-fPIC",
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


def build(verbose=False):
    """Compile every source into OUTPUT (always recompiles)."""
    tmp = OUTPUT + ".tmp"
    cmd = [hipcc(), *FLAGS, *(os.path.join(CSRC, s) for s in SOURCES),
           "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
