"""lcode_deposit_noalloc (the deposition with its scratch in a fixed device
array, for callers that cannot allocate) against lcode_deposit, both through
the C ABI: two calls in a row, the second with a smaller input, each within
relative L2 1e-12 of lcode_deposit's result (the FP64 LDS atomics sum in a
run-dependent order).  The first call displaces particles by up to 7 cells,
beyond the tiles' 6-cell halo, so it takes the spill list in the fixed
array; the second (smaller, 2 cells) must find that list's count cleared.
Then an nf beyond the array's 2048 limit is rejected before any launch."""

import ctypes
import types

import numpy as np
import pytest
import torch

from gpu_helpers import dev, host, rel_l2, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import lcode3d_amd
    assert torch.cuda.is_available(), "no HIP device visible"
    return lcode3d_amd


def _cfg(N, c, f):
    return types.SimpleNamespace(
        grid_steps=N, grid_step_size=3.2 / N, xi_step_size=.05,
        field_solver_subtraction_trick=1, field_solver_variant_A=True,
        reflect_padding_steps=5, plasma_padding_steps=10,
        plasma_coarseness=c, plasma_fineness=f)


def _inputs(oracle, cfg, seed, cells):
    xs, ys, const, vp, st = oracle.init(cfg)
    rng = np.random.default_rng(seed)
    nc = const.m.shape[0]
    amp = cells * cfg.grid_step_size
    xo, yo = (amp * rng.uniform(-1, 1, (nc, nc)) for _ in range(2))
    px, py, pz = (.3 * rng.uniform(-1, 1, (nc, nc)) for _ in range(3))
    parts = [dev(a) for a in (xo, yo, const.m, const.q, px, py, pz)]
    dvp = types.SimpleNamespace(**{k: dev(getattr(vp, k)) for k in (
        "influence_prev", "influence_next", "fine_grid", "indices_prev",
        "indices_next")})
    return dev(const.ro_initial), parts, dvp


def _noalloc(lc, cfg, ro0, parts, dvp):
    from lcode3d_amd import _lib
    from lcode3d_amd import lcode as L
    lib = _lib.load()
    nc, nf = parts[0].shape[0], dvp.fine_grid.numel()
    cs = L._config_struct(cfg, nc, nf)
    vp = _lib.LcodeVirtParams(*(getattr(dvp, k).data_ptr() for k in (
        "influence_prev", "influence_next", "fine_grid", "indices_prev",
        "indices_next")))
    n = cfg.grid_steps
    outs = [torch.full((n, n), np.nan, dtype=torch.float64, device="cuda")
            for _ in range(4)]
    o_pp = (ctypes.c_void_p * 4)(*(o.data_ptr() for o in outs))
    rc = lib.lcode_deposit_noalloc(
        ctypes.byref(cs), ro0.data_ptr(), *[p.data_ptr() for p in parts],
        ctypes.byref(vp), o_pp, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, outs


def test_noalloc_deposit_matches_deposit_twice(lc, oracle):
    for N, c, f, cells in ((641, 3, 2, 7), (129, 2, 2, 2)):
        say(f"deposit_noalloc N={N} c={c} f={f} offsets<={cells} cells")
        cfg = _cfg(N, c, f)
        ro0, parts, dvp = _inputs(oracle, cfg, N, cells)
        rc, got = _noalloc(lc, cfg, ro0, parts, dvp)
        from lcode3d_amd import _lib
        assert rc == 0, _lib.load().lcode_last_error()
        want = lc.deposit(cfg, ro0, *parts, dvp)
        torch.cuda.synchronize()
        for name, g, w in zip(("ro", "jx", "jy", "jz"), got, want):
            err = rel_l2(host(g), host(w))
            say(f"  {name}: rel L2 {err:.1e}")
            assert err < 1e-12, name


def test_noalloc_rejects_too_many_fine_particles(lc, oracle):
    from lcode3d_amd import _lib
    from lcode3d_amd import lcode as L
    cfg = _cfg(129, 2, 2)
    ro0, parts, dvp = _inputs(oracle, cfg, 1, 1)
    cs = L._config_struct(cfg, parts[0].shape[0], 2049)
    vp = _lib.LcodeVirtParams(*(getattr(dvp, k).data_ptr() for k in (
        "influence_prev", "influence_next", "fine_grid", "indices_prev",
        "indices_next")))
    outs = (ctypes.c_void_p * 4)(*(ro0.data_ptr() for _ in range(4)))
    lib = _lib.load()
    rc = lib.lcode_deposit_noalloc(ctypes.byref(cs), ro0.data_ptr(),
                                   *[p.data_ptr() for p in parts],
                                   ctypes.byref(vp), outs, None)
    assert rc == -1
    assert b"2048" in lib.lcode_last_error()
