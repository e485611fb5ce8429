"""Throughput of the LCODE 3D xi step on one MI355X.

Workload: the 641^2 production configuration (grid_step_size 0.025,
xi_step_size 0.005, plasma_coarseness 3, plasma_fineness 2, Variant A,
subtraction trick 1, paddings 5/10) driven by the cosine-Gaussian beam of
amplitude 0.05, sigma 1, with the plain FP64 HIP step.  One "step" is one
xi slice: estimate, 3 pushes, 3 depositions, 2 six-field solves.  The beam
slices of the timed region are generated on the device beforehand, so the
timed loop reads only HBM-resident inputs.

Prints ONE JSON line on stdout (everything else goes to stderr).
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_GRID = 641
WORKLOAD = dict(grid_steps=641, grid_step_size=.025, xi_step_size=.005,
                plasma_coarseness=3, plasma_fineness=2,
                field_solver_subtraction_trick=1, field_solver_variant_A=True,
                reflect_padding_steps=5, plasma_padding_steps=10)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_config():
    import types
    return types.SimpleNamespace(**WORKLOAD)


def algorithmic_bytes_per_step(nc, n):
    """8 (79 nc^2 + 91 N^2): every stage reads its inputs once and writes
    its outputs once (estimate 10 + push 3x16 + deposit 3x7 coarse-array
    touches; push 3x6 + deposit 3x5 + field solve 2x29 grid touches)."""
    return 8 * (79 * nc * nc + 91 * n * n)


HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
#: the step's dominant kernel, timed with HIP events inside the timed region
ROOFLINE_KERNEL = "deposit_tile"
EVENT_EVERY = 4
#: the plain step's kernels, timed one name at a time after the timed region
KERNEL_NAMES = ("deposit_tile", "solve_pass_a", "solve_pass_b", "solve_pass_c",
                "push", "deposit_spill")


def kernel_bytes(name, nc, n):
    """Algorithmic HBM bytes of one launch of each kernel at this workload:
    every array the kernel must read is read once, every output written
    once (FP64, 8 B per value; c2 = nc^2 coarse, n2 = N^2 grid values).

    push           11 coarse inputs + 5 outputs, 6 gathered fields
    deposit_tile   7 coarse inputs, ro_initial, 4 grids out
    solve_pass_a   13 RHS source grids in, 6 transposed spectra out
    solve_pass_b   6 spectra + 6 eigenvalue tables in, 6 spectra out
    solve_pass_c   6 spectra + 6 prev fields in, 6 averages out, plus the
                   6 fields on the corrector pass (mean: 9 grids out)
    """
    c2, n2 = nc * nc, n * n
    table = {
        "push": 16 * c2 + 6 * n2,
        "deposit_tile": 7 * c2 + 5 * n2,
        "solve_pass_a": 19 * n2,
        "solve_pass_b": 18 * n2,
        "solve_pass_c": 21 * n2,
    }
    return 8 * table[name] if name in table else None


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload from the two
    rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE in KiB per dispatch)
    recorded in profiles/pmc_traffic.json, each scaled by the factor the
    same counter reads for a device copy of known size; None without it."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    return 1024 * (k["FETCH_SIZE_KiB"] * t["fetch_factor"]
                   + k["WRITE_SIZE_KiB"] * t["write_factor"])


def dense_flops_per_solve(n):
    """Algorithmic flops of one six-field solve in the dense-matrix form
    (forward and inverse transform along both axes, 2 flops per
    multiply-add): 8 N (N-1)(N-2) for each of the four transverse fields
    (an (N-2) x N grid, DST-I along one axis, DCT-I along the other),
    8 (N-2)^3 for Ez (DST-I both ways) and 8 N^3 for Bz (DCT-I both ways)."""
    return 32 * n * (n - 1) * (n - 2) + 8 * (n - 2) ** 3 + 8 * n ** 3


SOLVE_KERNELS = ("solve_pass_a", "solve_pass_b", "solve_pass_c", "dense_rhs",
                 "dense_gemm_1", "dense_gemm_2", "dense_gemm_3", "dense_gemm_4")
MFMA_PEAK_TFS = {"fp64": 78.6, "bf16x3": 2500.0}


def solver_legs(lc, torch, cfg, const, vp, state, beams, steps, prof_steps):
    """The same step with each field-solver back-end from the same developed
    state: whole-step throughput over `steps` slices and the six-field solve
    time from HIP events around each solve kernel's launches (one kernel
    name per pass of `prof_steps` slices)."""
    import copy
    legs = []
    for transform, precision in (("fft", "fp64"), ("fft", "fp32"), ("fft", "mixed"),
                                 ("dense", "bf16x3"), ("dense", "fp64")):
        c = copy.copy(cfg)
        c.field_solver_transform = transform
        c.field_solver_precision = precision
        st = state
        if precision == "mixed":   # the same state with FP32 grids
            st = lc.GPUArrays(**{k: (v.float() if v.shape == state.Ex.shape else v)
                                 for k, v in vars(state).items()})
        for k in range(10):
            st = lc.step(c, const, vp, st, beams[k])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            st = lc.step(c, const, vp, st, beams[k % len(beams)])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        # one solve kernel name at a time (see the headline's kernel table)
        kus, calls = {}, {}
        names = SOLVE_KERNELS[:3] if transform == "fft" else SOLVE_KERNELS[3:]
        for name in names:
            for k in range(2):
                st = lc.step(c, const, vp, st, beams[k % len(beams)])
            lc.profile_begin(16 * prof_steps, only=name)
            for k in range(prof_steps):
                st = lc.step(c, const, vp, st, beams[k % len(beams)])
            prof = lc.profile_end()
            if name in prof:
                kus[name] = prof[name][1] / prof[name][0] * 1e3
                calls[name] = prof[name][0] / (2 * prof_steps)
        solve_us = sum(kus[k] * calls[k] for k in kus)
        leg = {"transform": transform, "precision": precision,
               "xi_steps_per_s": steps / dt, "solve_us": solve_us,
               "kernels_us": kus}
        if transform == "dense":
            tf = dense_flops_per_solve(cfg.grid_steps) / (solve_us * 1e-6) / 1e12
            leg["algorithmic_TFLOPs"] = tf
            leg["mfma_peak_TFLOPs"] = MFMA_PEAK_TFS[precision]
            leg["frac"] = tf / MFMA_PEAK_TFS[precision]
        legs.append(leg)
        log(f"  solver {transform:5s} {precision:6s}: {steps / dt:7.1f} xi-steps/s, "
            f"six-field solve {solve_us:8.1f} us"
            + (f", {leg['algorithmic_TFLOPs']:.1f} TFLOP/s algorithmic "
               f"({leg['frac'] * 100:.1f}% of {precision} MFMA peak)"
               if transform == "dense" else ""))
    return legs


GUIDED_THRESHOLD = 2e-3   # above the one-pass corrector's own residual
GUIDED_BACKOFF = 8        # plain slices after a rejected guess (doubling to 64)
GUIDED_ANCHOR = 8         # a plain slice after this many guided ones in a row
LEARNED_REFRESH = 4       # network evaluations: every 4th guided slice
LEARNED_THRESHOLD = 2e-3  # corrector residual threshold of the learned legs
                          # (about 0.1 xi_step_size = 5.5e-4, p90 1.1e-3)
RUN_SLICES = 1500         # the 641^2 run of the headline workload


def run_beams(cfg, cosine_gaussian, slices, **beam):
    """Beam slices 0 .. slices-1 of a run, generated on the device before
    any clock starts (scalar 0 behind the bunch)."""
    return [cosine_gaussian(cfg, i, **beam) for i in range(slices)]


def continuous_run(lc, torch, cfg, beams, window, stepper=None, record=False):
    """One run from init over len(beams) slices, plain steps or
    `stepper(config, const, virt_params)`'s steps.  Times the slice window
    [w0, w1) (the run's own history leads into it: developed wake) and the
    whole run.  record: also the on-axis Ez of every slice and the six
    fields, densities and momenta every 100 slices (device copies)."""
    _, _, const, vp, st = lc.init(cfg)
    gs = stepper(cfg, const, vp) if stepper else None
    n, c = len(beams), cfg.grid_steps // 2
    w0, w1 = window
    ez = torch.zeros(n, dtype=torch.float64, device="cuda") if record else None
    checks = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    t_w = dt_w = 0.0
    for i in range(n):
        if i == w0:
            torch.cuda.synchronize()
            t_w = time.perf_counter()
        st = gs.step(st, beams[i]) if gs else lc.step(cfg, const, vp, st, beams[i])
        if i + 1 == w1:
            torch.cuda.synchronize()
            dt_w = time.perf_counter() - t_w
        if record:
            ez[i] = st.Ez[c, c]
            if (i + 1) % 100 == 0:
                checks.append({k: getattr(st, k).clone() for k in (
                    "Ex", "Ey", "Ez", "Bx", "By", "Bz", "ro", "px", "py", "pz")})
    torch.cuda.synchronize()
    dt_all = time.perf_counter() - t_all
    if gs is not None:
        gs.drain(wait=True)
    out = {"window_xi_steps_per_s": (w1 - w0) / dt_w, "window": [w0, w1],
           "run_xi_steps_per_s": n / dt_all, "run_slices": n,
           "faults": lc.fault_count(cfg, const.m.shape[0])}
    if gs is not None:
        out.update({"guided_slices": gs.guided, "accepted": gs.accepted,
                    "plain_slices": gs.plain})
    return out, ez, checks


def compare_runs(torch, ez_ref, checks_ref, ez, checks):
    """Ez(xi, 0, 0) relative L2 over the run and the worst relative L2 of
    each recorded array over the 100-slice checkpoints (Bz against |Bx|)."""
    axis = (torch.linalg.norm(ez - ez_ref) / torch.linalg.norm(ez_ref)).item()
    worst = {}
    for r, g in zip(checks_ref, checks):
        for k, v in r.items():
            den = torch.linalg.norm(r["Bx"] if k == "Bz" else v).item()
            e = torch.linalg.norm(g[k].double() - v.double()).item() / max(den, 1e-300)
            worst[k] = max(worst.get(k, 0.0), e)
    return {"ez_axis_rel_l2": axis, "checkpoint_rel_l2_max": worst}


def stepper_factory(lc, predictor=None, extrapolation="gap"):
    def make(cfg, const, vp):
        # the learned legs have a threshold of their own (measured at 641^2
        # with the network evaluated every 4th guided slice: 5e-3 and 1e-2
        # accept no larger share of guesses than 2e-3 but let the transverse
        # magnetic field drift from the plain step by 10-40 %)
        t = LEARNED_THRESHOLD if predictor is not None else GUIDED_THRESHOLD
        return lc.GuidedStepper(cfg, const, vp, t,
                                predictor=predictor, backoff=GUIDED_BACKOFF,
                                extrapolation=extrapolation, anchor=GUIDED_ANCHOR)
    return make


def mode_legs(lc, torch, cosine_gaussian, window, plain_ref, modes):
    """Each (name, precision, stepper) mode as a continuous 641^2 run
    against the plain FP64 run `plain_ref` = (ez, checks, stats)."""
    ez_p, ch_p, stats_p = plain_ref
    legs = {}
    for name, precision, stepper in modes:
        cfg = make_config()
        cfg.field_solver_precision = precision
        beams = run_beams(cfg, cosine_gaussian, RUN_SLICES)
        timing, _, _ = continuous_run(lc, torch, cfg, beams, window, stepper)
        acc, ez, checks = continuous_run(lc, torch, cfg, beams, window, stepper,
                                         record=True)
        del beams
        leg = {"precision": precision, **timing}
        for k in ("guided_slices", "accepted", "plain_slices"):
            if k in acc:
                leg[k] = acc[k]
        leg.update(compare_runs(torch, ez_p, ch_p, ez, checks))
        leg["window_vs_plain"] = leg["window_xi_steps_per_s"] / stats_p["window_xi_steps_per_s"]
        leg["run_vs_plain"] = leg["run_xi_steps_per_s"] / stats_p["run_xi_steps_per_s"]
        legs[name] = leg
        log(f"{name:14s}: window {leg['window_xi_steps_per_s']:7.1f} xi-steps/s "
            f"({leg['window_vs_plain']:.2f}x plain), run of {RUN_SLICES} "
            f"{leg['run_xi_steps_per_s']:7.1f} ({leg['run_vs_plain']:.2f}x); Ez axis rel L2 "
            f"{leg['ez_axis_rel_l2']:.1e}, checkpoints "
            + " ".join(f"{k} {v:.1e}" for k, v in leg["checkpoint_rel_l2_max"].items())
            + (f"; {leg['accepted']} of {leg['guided_slices']} guesses accepted, "
               f"{leg['plain_slices']} plain" if "guided_slices" in leg else ""))
    return legs


def train_learned(lc, cosine_gaussian, torch, runs=4, slices=RUN_SLICES, every=2,
                  epochs=4, cfg=None, beam_kw=None):
    """The U-Net predictor trained on the GPU on ~3000 slices of `runs` plain
    FP64 runs of `cfg` (default the 641^2 workload; amplitude ~ U[0.01,
    0.1], sigma ~ U[0.5, 2] from default_rng(0), other beam arguments from
    `beam_kw`), every `every`-th slice; target F1 - Q (the predictor pass's
    fields minus the quadratic extrapolation)."""
    import numpy as np
    from lcode3d_amd import unet
    cfg = cfg or make_config()
    rng = np.random.default_rng(0)
    params = [(float(rng.uniform(.01, .1)), float(rng.uniform(.5, 2))) for _ in range(runs)]
    net = unet.UNetPredictor(cfg.grid_steps, seed=0, lr=1e-3)
    net.refresh_every = LEARNED_REFRESH
    net.base = "alternating"     # correct the Variant-A-aware extrapolation
    beam_fn = None
    if beam_kw:
        def beam_fn(c, i, amplitude, sigma):
            return cosine_gaussian(c, i, amplitude=amplitude, sigma=sigma, **beam_kw)
    t0 = time.perf_counter()
    samples, sx, st = unet.collect_training_data(lc, net, cfg, params, slices, every=every,
                                                 beam_fn=beam_fn)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = unet.train(net, samples, (sx, st), epochs)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    n = len(samples)
    del samples
    log(f"learned predictor at {cfg.grid_steps}^2: {n} samples from {runs} runs "
        f"({t1 - t0:.1f} s), {epochs} epochs ({t2 - t1:.1f} s), "
        f"loss {hist[0]:.3g} -> {hist[-1]:.3g}")
    info = {"predictor": "U-Net, 4 levels x 32 channels, BF16 MFMA, target F1 - E "
                         "(E = 2 F(n-1) - 2 F(n-3) + F(n-4))",
            "refresh_every": LEARNED_REFRESH,
            "train_runs": params, "train_samples": n, "epochs": epochs,
            "collect_seconds": t1 - t0, "train_seconds": t2 - t1, "train_loss": hist}
    return net, info


def precision_table(lc, cosine_gaussian, torch, slices):
    """Error of the mixed-precision field solvers against the plain FP64
    step (itself within 1e-10 of the FP64 oracle) on the 641^2 workload for
    beam amplitudes 0.01, 0.05, 0.1 x sigma 0.5, 1, 2: relative L2 of each
    field and of the momenta, worst over checkpoints every 100 slices."""
    names = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "px", "py", "pz")
    modes = {"fft-fp32": ("fft", "fp32"), "dense-bf16x3": ("dense", "bf16x3")}
    table = []
    for amp in (0.01, 0.05, 0.1):
        for sig in (0.5, 1.0, 2.0):
            cfgs = {"fp64": make_config()}
            for m, (t, p) in modes.items():
                c = make_config()
                c.field_solver_transform, c.field_solver_precision = t, p
                cfgs[m] = c
            states, const, vp = {}, None, None
            for m, c in cfgs.items():
                _, _, const, vp, states[m] = lc.init(c)
            worst = {m: {n: 0.0 for n in names} for m in modes}
            for i in range(slices):
                beam = cosine_gaussian(cfgs["fp64"], i, amplitude=amp, sigma=sig)
                for m, c in cfgs.items():
                    states[m] = lc.step(c, const, vp, states[m], beam)
                if (i + 1) % 100 == 0 or i + 1 == slices:
                    ref = states["fp64"]
                    for m in modes:
                        for n in names:
                            r = getattr(ref, n)
                            floor = torch.linalg.norm(ref.Bx) if n == "Bz" else 0
                            den = max(torch.linalg.norm(r).item(), float(floor))
                            e = torch.linalg.norm(getattr(states[m], n) - r).item() / max(den, 1e-300)
                            worst[m][n] = max(worst[m][n], e)
            row = {"amplitude": amp, "sigma": sig, "slices": slices, "rel_l2_max": worst}
            table.append(row)
            log(f"  precision amp {amp:g} sigma {sig:g}: " + "; ".join(
                f"{m} " + " ".join(f"{n} {worst[m][n]:.1e}" for n in names) for m in modes))
    return table


SMALL_WORKLOAD = dict(grid_steps=129, grid_step_size=.05, xi_step_size=.05,
                      plasma_coarseness=2, plasma_fineness=2,
                      field_solver_subtraction_trick=1, field_solver_variant_A=True,
                      reflect_padding_steps=5, plasma_padding_steps=10)


LARGE_WORKLOAD = dict(grid_steps=1281, grid_step_size=.0125, xi_step_size=.0025,
                      plasma_coarseness=3, plasma_fineness=3)
LARGE_BEAM = dict(amplitude=.1, sigma=.7, shift_x=.5)


def large_workload_leg(lc, torch, cosine_gaussian, slices, learned=True):
    """The large non-axisymmetric case: 1281^2 (h 0.0125, xi step 0.0025,
    coarseness 3, fineness 3), beam amplitude 0.1, sigma 0.7, offset +0.5
    in x.  The plain FP64 run, the mixed-precision run, the mixed run with
    the guided step (quadratic extrapolation) and, with `learned`, with the
    U-Net guess (trained at 1281^2 on two plain runs of other beams with
    the same offset) over the same `slices` slices, each a continuous run
    timed whole and over slices [slices/3, slices/3 + 300); Ez(xi, 0, 0)
    and the fields every 100 slices compared with the plain run."""
    def cfg(precision):
        c = make_config()
        for k, v in LARGE_WORKLOAD.items():
            setattr(c, k, v)
        c.field_solver_precision = precision
        return c
    w0 = slices // 3
    window = (w0, min(w0 + 300, slices))
    out = {"workload": "1281^2, coarseness 3, fineness 3, beam 0.1 / 0.7 / +0.5 in x",
           "slices": slices}
    c = cfg("fp64")
    beams = run_beams(c, cosine_gaussian, slices, **LARGE_BEAM)
    plain, ez_p, ch_p = continuous_run(lc, torch, c, beams, window, record=True)
    out["plain"] = plain
    modes = [("mixed", "mixed", None), ("mixed_guided", "mixed", stepper_factory(lc)),
             ("mixed_guided_quad", "mixed", stepper_factory(lc, extrapolation="quadratic"))]
    if learned:
        shift = {"shift_x": LARGE_BEAM["shift_x"]}
        net, info = train_learned(lc, cosine_gaussian, torch, runs=2, slices=min(slices, 1500),
                                  every=1, cfg=c, beam_kw=shift)
        out["learned_training"] = info
        modes.append(("mixed_learned", "mixed", stepper_factory(lc, net.guess)))
    for name, prec, stepper in modes:
        r, ez, ch = continuous_run(lc, torch, cfg(prec), beams, window, stepper,
                                   record=True)
        r.update(compare_runs(torch, ez_p, ch_p, ez, ch))
        r["run_vs_plain"] = r["run_xi_steps_per_s"] / plain["run_xi_steps_per_s"]
        r["window_vs_plain"] = r["window_xi_steps_per_s"] / plain["window_xi_steps_per_s"]
        out[name] = r
        del ez, ch
    del beams, ch_p
    return out


def small_workload_leg(lc, cosine_gaussian, torch, with_cpu):
    """The 129^2 reference-runnable case in full (200 slices of the same
    beam shape): the HIP step's throughput and, when asked, the oracle's
    on the host cores, plus their on-axis Ez agreement."""
    import types
    cfg = types.SimpleNamespace(**SMALL_WORKLOAD)
    _, _, const, vp, st = lc.init(cfg)
    beams = [cosine_gaussian(cfg, i) for i in range(200)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ez = []
    for i in range(200):
        st = lc.step(cfg, const, vp, st, beams[i])
        ez.append(st.Ez[64, 64])
    torch.cuda.synchronize()
    gpu = 200 / (time.perf_counter() - t0)
    out = {"workload": "129^2, coarseness 2, fineness 2, 200 slices", "gpu_value": gpu,
           "unit": "xi-steps/s"}
    if with_cpu:
        import numpy as np
        from oracle import lcode_oracle as oracle
        cores = len(os.sched_getaffinity(0))
        oracle.set_fft_workers(cores)
        ocfg = types.SimpleNamespace(**SMALL_WORKLOAD)
        xs, ys, oconst, ovp, ost = oracle.init(ocfg)
        t0 = time.perf_counter()
        oez = []
        for i in range(200):
            ost = oracle.step(ocfg, oconst, ovp, ost, oracle.beam(ocfg, i, xs, ys))
            oez.append(ost.Ez[64, 64])
        out["cpu_value"] = 200 / (time.perf_counter() - t0)
        out["cpu_cores"] = cores
        g = torch.stack(ez).cpu().numpy()
        out["ez_axis_max_rel_diff"] = float(np.abs(g - np.array(oez)).max()
                                            / np.abs(np.array(oez)).max())
    return out


def cpu_baseline(steps):
    """The oracle (NumPy/SciPy restatement of the reference step) on the
    host cores, on a bounded sample of the same workload."""
    import numpy as np
    from oracle import lcode_oracle as oracle
    cores = len(os.sched_getaffinity(0))
    oracle.set_fft_workers(cores)
    cfg = oracle.make_config(**{k: v for k, v in dict(
        grid_steps=641, grid_step_size=.025, xi_step_size=.005,
        coarseness=3, fineness=2).items()})
    xs, ys, const, vp, st = oracle.init(cfg)
    start = 300  # a slice inside the bunch
    st = oracle.step(cfg, const, vp, st, oracle.beam(cfg, start, xs, ys))
    t0 = time.perf_counter()
    for k in range(steps):
        st = oracle.step(cfg, const, vp, st,
                         oracle.beam(cfg, start + 1 + k, xs, ys))
    dt = time.perf_counter() - t0
    assert np.isfinite(st.Ez).all()
    return {"value": steps / dt, "unit": "xi-steps/s", "cores": cores,
            "kind": "port",
            "sample": f"{steps} xi-steps at 641^2 after 1 warm-up step "
                      f"(slices {start + 1}-{start + steps}), scipy.fft "
                      f"workers={cores}, NumPy bincount deposition"}


def main():
    global GUIDED_THRESHOLD, GUIDED_BACKOFF, LEARNED_REFRESH, LEARNED_THRESHOLD, GUIDED_ANCHOR
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--solver-steps", type=int, default=50,
                    help="slices per field-solver back-end comparison leg (0: skip)")
    ap.add_argument("--settle", type=int, default=400,
                    help="untimed slices run before warm-up: the timed window "
                         "starts at slice settle + warmup, in the developed wake")
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="steps timed per kernel with HIP events")
    ap.add_argument("--no-modes", dest="modes", action="store_false",
                    help="skip the 1500-slice runs of the mixed / guided / learned "
                         "modes against the plain FP64 run")
    ap.add_argument("--guided-threshold", type=float, default=GUIDED_THRESHOLD,
                    help="corrector residual threshold of the guided legs")
    ap.add_argument("--precision-table", type=int, default=0, metavar="SLICES",
                    help="mixed-precision error table over SLICES slices (0: skip)")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip the learned-predictor leg")
    ap.add_argument("--learned-threshold", type=float, default=LEARNED_THRESHOLD,
                    help="corrector residual threshold of the learned legs")
    ap.add_argument("--learned-refresh", type=int, default=LEARNED_REFRESH,
                    help="evaluate the network on every k-th guided slice and reuse "
                         "its last correction in between")
    ap.add_argument("--no-small-workload", dest="small_workload", action="store_false",
                    help="skip the 129^2 leg (GPU and oracle)")
    ap.add_argument("--guided-anchor", type=int, default=GUIDED_ANCHOR,
                    help="a plain slice after this many guided slices in a row (0: never)")
    ap.add_argument("--guided-backoff", type=int, default=GUIDED_BACKOFF,
                    help="plain slices after a rejected guess (0: guess every slice)")
    ap.add_argument("--large", type=int, default=3000, metavar="SLICES",
                    help="the 1281^2 fineness-3 off-axis case, plain FP64 against mixed, "
                         "mixed + guided and mixed + learned, for SLICES slices (0: skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a device copy of this size (PMC calibration)")
    args = ap.parse_args()
    GUIDED_THRESHOLD = args.guided_threshold
    GUIDED_BACKOFF = args.guided_backoff
    LEARNED_REFRESH = args.learned_refresh
    GUIDED_ANCHOR = args.guided_anchor
    LEARNED_THRESHOLD = args.learned_threshold
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): "
              "the xi step is sequential in xi and this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    import lcode3d_amd as lc
    from lcode3d_amd.beams import cosine_gaussian

    torch.cuda.set_device(0)
    cfg = make_config()
    xs, ys, const, vp, state = lc.init(cfg)
    nc = const.m.shape[0]
    n = cfg.grid_steps
    total = args.warmup + args.steps
    w0 = args.settle + args.warmup
    # beam slices of the run, resident in HBM before the clock starts
    beams = run_beams(cfg, cosine_gaussian,
                      max(RUN_SLICES, w0 + args.steps
                          + len(KERNEL_NAMES) * (args.profile_steps + 2)))
    for k in range(args.settle):  # bring the plasma into the developed wake
        state = lc.step(cfg, const, vp, state, beams[k])
    torch.cuda.synchronize()
    log(f"workload 641^2, nc={nc}, nf={vp.fine_grid.numel()}, "
        f"{args.warmup} warm-up + {args.steps} timed steps (slices {w0}-{w0 + args.steps - 1})")
    for k in range(args.warmup):
        state = lc.step(cfg, const, vp, state, beams[args.settle + k])
    torch.cuda.synchronize()
    # the timed region carries a pair of HIP events on the roofline kernel's
    # launches in every EVENT_EVERY-th step: events attached to the
    # dispatch (hipExtLaunchKernelGGL, on the step's stream), which the
    # runtime stamps with the kernel's own start and end
    lc.profile_begin(4 * args.steps, only=ROOFLINE_KERNEL)
    t0 = time.perf_counter()
    for k in range(args.steps):
        lc.profile_pause(k % EVENT_EVERY != 0)
        state = lc.step(cfg, const, vp, state, beams[w0 + k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timed_prof = lc.profile_end()
    finite = bool(torch.isfinite(state.Ez).all().item())
    disp = max(state.x_offt.abs().max().item(), state.y_offt.abs().max().item())
    log(f"max coarse displacement {disp / cfg.grid_step_size:.2f} cells")
    faults = lc.fault_count(cfg, nc)
    if faults:
        log(f"WARNING: {faults} clamped stencil lookups")
    steps_per_s = args.steps / dt
    ms = dt / args.steps * 1e3
    step_bytes = algorithmic_bytes_per_step(nc, n)
    log(f"{steps_per_s:.1f} xi-steps/s, {ms:.3f} ms/step, "
        f"step-level {step_bytes / (dt / args.steps) / 1e9:.1f} GB/s algorithmic")

    # the other kernels' durations over the next slices, one kernel name
    # at a time (attached events), each pass led by two untimed steps
    roofline, kernels = None, {}
    calls, kms = timed_prof.get(ROOFLINE_KERNEL, (0, 0.0))
    if calls:
        avg_us = kms / calls * 1e3
        kernels[ROOFLINE_KERNEL] = {
            "launches_per_step": calls / len(range(0, args.steps, EVENT_EVERY)),
            "avg_us": avg_us,
            "window": f"timed region, every {EVENT_EVERY}th step ({calls} launches)"}
    if args.profile_steps:
        sl = w0 + args.steps
        for name in KERNEL_NAMES:
            if name == ROOFLINE_KERNEL and calls:
                continue
            for k in range(2):
                state = lc.step(cfg, const, vp, state, beams[sl])
                sl += 1
            lc.profile_begin(8 * args.profile_steps, only=name)
            for k in range(args.profile_steps):
                state = lc.step(cfg, const, vp, state, beams[sl])
                sl += 1
            prof = lc.profile_end()
            if name in prof:
                c, t = prof[name]
                kernels[name] = {"launches_per_step": c / args.profile_steps,
                                 "avg_us": t / c * 1e3,
                                 "window": f"slices {sl - args.profile_steps}-{sl - 1}"}
    if kernels:
        ksum = sum(v["launches_per_step"] * v["avg_us"] for v in kernels.values())
        for name, v in kernels.items():
            b = kernel_bytes(name, nc, n)
            v["share"] = v["launches_per_step"] * v["avg_us"] / ksum
            v["GBps"] = (b / (v["avg_us"] * 1e-6) / 1e9) if b else None
        kernels = dict(sorted(kernels.items(), key=lambda kv: -kv[1]["share"]))
        log(f"  kernel durations (HIP events; their sum {ksum:.1f} us per step "
            f"against {ms * 1e3:.1f} us measured):")
        for name, v in kernels.items():
            log(f"  {name:14s} {v['launches_per_step']:4.1f}/step "
                f"{v['avg_us']:8.2f} us  {v['share'] * 100:5.1f}%"
                + (f"  {v['GBps']:7.0f} GB/s" if v["GBps"] else "")
                + f"  ({v['window']})")
    if ROOFLINE_KERNEL in kernels:
        v = kernels[ROOFLINE_KERNEL]
        roofline = {"bound": "hbm", "kernel": ROOFLINE_KERNEL, "achieved": v["GBps"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": v["GBps"] / HBM_PEAK_GBS,
                    "traffic": measured_traffic(ROOFLINE_KERNEL),
                    "bytes_per_launch": kernel_bytes(ROOFLINE_KERNEL, nc, n),
                    "avg_us": v["avg_us"], "launches": calls}
    if "deposit_tile" in kernels:
        # the reference's deposition rate measure: 4 quantities x 9 stencil
        # cells = 36 atomic adds per fine particle (lcode.py:434-449,
        # :660-663); the tile kernel sums them in registers and LDS but
        # carries out the same 36 updates
        v = kernels["deposit_tile"]
        updates = 36 * vp.fine_grid.numel() ** 2
        v["fine_particle_updates_per_launch"] = updates
        v["fine_particle_updates_per_s"] = updates / (v["avg_us"] * 1e-6)
        log(f"  deposit_tile: {updates / 1e6:.1f} M fine-particle stencil updates per launch, "
            f"{v['fine_particle_updates_per_s'] / 1e12:.2f} T updates/s")
    if args.copy_probe_mib:
        src = torch.ones(args.copy_probe_mib << 17, dtype=torch.float64,
                         device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 3 x {args.copy_probe_mib} MiB device copies")

    solvers = None
    if args.solver_steps:
        solvers = solver_legs(lc, torch, cfg, const, vp, state, beams[w0:w0 + total],
                              args.solver_steps, min(args.profile_steps or 10, 10))
    del state

    # the whole 1500-slice run, plain FP64, and the other modes against it
    window = (w0, w0 + args.steps)
    run_beams_plain = beams[:RUN_SLICES]
    plain_run, _, _ = continuous_run(lc, torch, cfg, run_beams_plain, window)
    log(f"plain FP64 run of {RUN_SLICES} slices: {plain_run['run_xi_steps_per_s']:.1f} "
        f"xi-steps/s (window {plain_run['window_xi_steps_per_s']:.1f})")
    modes = None
    if args.modes:
        _, ez_p, ch_p = continuous_run(lc, torch, cfg, run_beams_plain, window, record=True)
        mode_list = [("mixed", "mixed", None),
                     ("mixed_guided", "mixed", stepper_factory(lc)),
                     ("mixed_guided_quad", "mixed", stepper_factory(lc, extrapolation="quadratic")),
                     ("mixed_guided_alt", "mixed", stepper_factory(lc, extrapolation="alternating")),
                     ("fp64_guided", "fp64", stepper_factory(lc))]
        learned_info = None
        if args.learned:
            net, learned_info = train_learned(lc, cosine_gaussian, torch)
            mode_list += [("mixed_learned", "mixed", stepper_factory(lc, net.guess)),
                          ("fp64_learned", "fp64", stepper_factory(lc, net.guess))]
        modes = mode_legs(lc, torch, cosine_gaussian, window, (ez_p, ch_p, plain_run),
                          mode_list)
        if learned_info:
            modes["mixed_learned"].update(learned_info)
        del ch_p
    del beams, run_beams_plain

    large = None
    if args.large:
        large = large_workload_leg(lc, torch, cosine_gaussian, args.large, args.learned)
        for k in [m for m in ("mixed", "mixed_guided", "mixed_guided_quad", "mixed_learned")
                  if m in large]:
            r = large[k]
            log(f"1281^2 x {args.large} {k}: {r['run_xi_steps_per_s']:.1f} xi-steps/s "
                f"({r['run_vs_plain']:.2f}x plain {large['plain']['run_xi_steps_per_s']:.1f}); "
                f"Ez axis rel L2 {r['ez_axis_rel_l2']:.1e}")

    small = None
    if args.small_workload:
        small = small_workload_leg(lc, cosine_gaussian, torch, not args.no_cpu_baseline)
        log(f"129^2 x 200 slices: {small['gpu_value']:.1f} xi-steps/s on the GPU"
            + (f", {small['cpu_value']:.2f} on {small['cpu_cores']} host cores (oracle), "
               f"Ez(xi,0,0) agreement {small['ez_axis_max_rel_diff']:.1e}"
               if "cpu_value" in small else ""))

    ptable = None
    if args.precision_table:
        ptable = precision_table(lc, cosine_gaussian, torch, args.precision_table)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_steps)
        log(f"cpu baseline {cpu['value']:.3f} xi-steps/s on {cpu['cores']} cores")

    out = {
        "metric": "xi-steps/s at 641^2 grid (plain FP64 HIP step)",
        "value": steps_per_s,
        "unit": "xi-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic cosine-Gaussian beam (amplitude 0.05, sigma 1); timed "
                f"slices {w0}-{w0 + args.steps - 1} of one run (developed wake)",
        "config": {"workload": "lcode3d xi step, 641^2 grid, coarseness 3, "
                               "fineness 2, plain FP64",
                   "grid_steps": n, "coarse_particles": nc * nc,
                   "fine_particles": int(vp.fine_grid.numel()) ** 2},
        "roofline": roofline,
        "kernels": kernels,
        "full_run": plain_run,
        "solver_backends": solvers,
        "modes": modes,
        "precision_table": ptable,
        "small_workload": small,
        "large_workload": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if not finite:
        log("ERROR: non-finite fields at the end of the timed run")
        sys.exit(3)


if __name__ == "__main__":
    main()
