"""Writes oracle_641_1500.npz: the oracle's plain FP64 run of the 641^2
production workload over 1500 slices, the fixture of
tests/test_gpu_configs.py::test_plain_641_1500_slices_against_oracle_fixture.

Workload: grid_steps 641, grid_step_size 0.025, xi_step_size 0.005,
plasma_coarseness 3, plasma_fineness 2, subtraction trick 1, Variant A,
paddings 5/10, cosine-Gaussian beam of amplitude 0.05 and sigma 1
(config_example.py:21-28, no shift).

Stored (all float64):
  ez_axis          Ez[N//2, N//2] after every slice (1500 values), the
                   quantity the reference's driver records (lcode.py:1045)
  checkpoints      slice counts 100, 200, ..., 1500
  <name>_row       centre row    a[n//2, :] of each of the 15 state arrays
  <name>_col       centre column a[:, n//2]   at every checkpoint

Run once in the build container (about an hour on 8 cores):
    python tests/golden/make_641_fixture.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import lcode_oracle as oracle  # noqa: E402

STATE = ("x_offt", "y_offt", "px", "py", "pz", "Ex", "Ey", "Ez", "Bx", "By",
         "Bz", "ro", "jx", "jy", "jz")
SLICES = 1500
EVERY = 100
OUT = os.path.join(HERE, "oracle_641_1500.npz")


def main():
    cfg = oracle.make_config(grid_steps=641, grid_step_size=.025,
                             xi_step_size=.005, coarseness=3, fineness=2)
    xs, ys, const, vp, st = oracle.init(cfg)
    c = cfg.grid_steps // 2
    ez = np.zeros(SLICES)
    rows = {f"{n}_{p}": [] for n in STATE for p in ("row", "col")}
    t0 = time.time()
    for xi_i in range(SLICES):
        st = oracle.step(cfg, const, vp, st,
                         oracle.beam(cfg, xi_i, xs, ys, amplitude=.05,
                                     sigma=1.0))
        ez[xi_i] = st.Ez[c, c]
        if (xi_i + 1) % EVERY == 0:
            for name in STATE:
                a = getattr(st, name)
                m = a.shape[0] // 2
                rows[f"{name}_row"].append(np.array(a[m, :]))
                rows[f"{name}_col"].append(np.array(a[:, m]))
            print(f"slice {xi_i + 1}: Ez(0,0) {ez[xi_i]:+.6e}, "
                  f"{time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(
        OUT, ez_axis=ez,
        checkpoints=np.arange(EVERY, SLICES + 1, EVERY),
        **{k: np.stack(v) for k, v in rows.items()})
    print("wrote", OUT, flush=True)


if __name__ == "__main__":
    main()
