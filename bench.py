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


def solver_legs(lc, torch, cfg, const, vp, state, beams, steps, prof_steps,
                bracket_us=0.0):
    """The same step with each field-solver back-end from the same developed
    state: whole-step throughput over `steps` slices and the six-field solve
    time from HIP events over `prof_steps` slices (less `bracket_us`, the
    event-pair overhead per launch)."""
    import copy
    legs = []
    for transform, precision in (("fft", "fp64"), ("fft", "fp32"),
                                 ("dense", "bf16x3"), ("dense", "fp64")):
        c = copy.copy(cfg)
        c.field_solver_transform = transform
        c.field_solver_precision = precision
        st = state
        for k in range(10):
            st = lc.step(c, const, vp, st, beams[k])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            st = lc.step(c, const, vp, st, beams[k % len(beams)])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        lc.profile_begin(64 * prof_steps)
        for k in range(prof_steps):
            st = lc.step(c, const, vp, st, beams[k % len(beams)])
        prof = lc.profile_end()
        kus = {k: v[1] / v[0] * 1e3 - bracket_us for k, v in prof.items()
               if k in SOLVE_KERNELS}
        calls = {k: v[0] / (2 * prof_steps) for k, v in prof.items() if k in SOLVE_KERNELS}
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
                          # (about 0.1 xi_step_size = 5.5e-4, p90 1.1e-3)


def _guided_runs(lc, cosine_gaussian, torch, slices, timed, precision,
                 predictor=None):
    """The guided step (GuidedStepper: the guess on the slices the backoff
    policy selects, GUIDED_THRESHOLD, GUIDED_BACKOFF) against the plain FP64
    step over `slices` slices from init: on-axis Ez(xi) relative L2 and the
    worst full-Ez relative L2 of the 100-slice checkpoints; then its
    throughput over `timed` slices of the bunch (beam slices 103 on), guess
    inference and the residual read-backs included."""
    cfg_p = make_config()
    cfg_g = make_config()
    cfg_g.field_solver_precision = precision
    _, _, const, vp, sp = lc.init(cfg_p)
    lc.init(cfg_g)
    sg = sp
    n = cfg_p.grid_steps
    gs = lc.GuidedStepper(cfg_g, const, vp, GUIDED_THRESHOLD, predictor=predictor,
                          backoff=GUIDED_BACKOFF)
    ez_p, ez_g, full = [], [], 0.0
    for i in range(slices):
        beam = cosine_gaussian(cfg_p, i)
        sp = lc.step(cfg_p, const, vp, sp, beam)
        sg = gs.step(sg, beam)
        ez_p.append(sp.Ez[n // 2, n // 2])
        ez_g.append(sg.Ez[n // 2, n // 2])
        if (i + 1) % 100 == 0:
            e = (torch.linalg.norm(sg.Ez - sp.Ez) / torch.linalg.norm(sp.Ez)).item()
            full = max(full, e)
    ez_p, ez_g = torch.stack(ez_p), torch.stack(ez_g)
    axis = (torch.linalg.norm(ez_g - ez_p) / torch.linalg.norm(ez_p)).item()
    out = {"precision": precision, "slices_compared": slices,
           "ez_axis_rel_l2": axis, "ez_checkpoint_rel_l2_max": full,
           "guided_slices": gs.guided, "accepted": gs.accepted, "plain_slices": gs.plain,
           "guess_accept_rate": gs.accepted / max(gs.guided, 1),
           "backoff": GUIDED_BACKOFF}
    beams = [cosine_gaussian(cfg_g, 100 + k) for k in range(timed + 3)]
    _, _, const, vp, st = lc.init(cfg_g)
    gs = lc.GuidedStepper(cfg_g, const, vp, GUIDED_THRESHOLD, predictor=predictor,
                          backoff=GUIDED_BACKOFF)
    for k in range(3):
        st = lc.step(cfg_g, const, vp, st, beams[k])
        gs.push(st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(timed):
        st = gs.step(st, beams[3 + k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out.update({"value": timed / dt, "unit": "xi-steps/s",
                "timed_guided_slices": gs.guided, "timed_accepted": gs.accepted,
                "timed_plain_slices": gs.plain})
    return out


def mixed_guided_leg(lc, cosine_gaussian, torch, slices, timed, precision):
    """Mixed precision (FP32 transforms) + the guided step with the
    quadratic extrapolation of the last three slices' fields as the guess."""
    out = _guided_runs(lc, cosine_gaussian, torch, slices, timed, precision)
    out["predictor"] = ("quadratic extrapolation of the last 3 slices' fields, "
                        f"corrector residual threshold {GUIDED_THRESHOLD:g}, "
                        f"backoff {GUIDED_BACKOFF} slices after a rejection")
    return out


def learned_leg(lc, cosine_gaussian, torch, slices, timed, precision,
                 train_runs=2, train_slices=220, epochs=3, hidden=(8, 8)):
    """Mixed precision + the learned predictor: a ConvPredictor trained on
    `train_runs` plain 641^2 runs (amplitude ~ U[0.01, 0.1], sigma ~
    U[0.5, 2] from default_rng(0)), its guess in place of the predictor
    pass with residual threshold GUIDED_THRESHOLD; accuracy against the
    plain FP64 step over `slices` slices of the bench beam, throughput over
    `timed` slices (guess inference included)."""
    import collections
    import numpy as np
    from lcode3d_amd.predictor import ConvPredictor
    cfg_p = make_config()
    rng = np.random.default_rng(0)
    params = [(rng.uniform(.01, .1), rng.uniform(.5, 2)) for _ in range(train_runs)]
    net = ConvPredictor(cfg_p.grid_steps, hidden=hidden, seed=0, lr=2e-3)
    t0 = time.perf_counter()

    def windows(amp, sig):
        _, _, const, vp, st = lc.init(cfg_p)
        last = collections.deque(maxlen=4)
        for i in range(train_slices):
            beam = cosine_gaussian(cfg_p, i, amplitude=amp, sigma=sig)
            st = lc.step(cfg_p, const, vp, st, beam)
            last.append(st)
            if i >= 23 and len(last) == 4:
                yield last[3], last[2], last[1], last[0], beam
    acc = net.scale_accumulator()
    for s0, s1, s2, s3, _ in windows(*params[0]):
        acc.add(s0, s1, s2, s3)
    acc.finish()
    samples = []
    for amp, sig in params:
        for s0, s1, s2, s3, beam in windows(amp, sig):
            samples.append((net.inputs(s1, s2, s3, beam).clone(),
                            net.targets(s1, s2, s3, s0).clone()))
    hist = net.fit(samples, epochs=epochs)
    del samples
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"learned predictor: {len(params)} runs x {train_slices} slices, {epochs} epochs, "
        f"loss {hist[0]:.3g} -> {hist[-1]:.3g}, {train_s:.1f} s")

    out = _guided_runs(lc, cosine_gaussian, torch, slices, timed, precision,
                       predictor=net.guess)
    out["predictor"] = (f"ConvPredictor hidden {list(hidden)}, trained on {len(params)} "
                        f"plain runs x {train_slices} slices, {epochs} epochs; corrector "
                        f"residual threshold {GUIDED_THRESHOLD:g}, backoff "
                        f"{GUIDED_BACKOFF} slices after a rejection")
    out.update({"train_seconds": train_s, "train_loss": hist})
    return out


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


def guided_trace_leg(lc, cosine_gaussian, torch, slices):
    """The guided step in the developed wake: `slices` plain FP32-transform
    steps (beam slices 103 on), then `slices` guided steps (quadratic
    extrapolation, threshold GUIDED_THRESHOLD), each block timed without a
    synchronisation inside; then `slices` more guided steps with the
    residual read back after each, for the acceptance count."""
    cfg = make_config()
    cfg.field_solver_precision = "fp32"
    _, _, const, vp, st = lc.init(cfg)
    nc, nf = const.m.shape[0], vp.fine_grid.numel()
    ex = lc.FieldExtrapolator(order=2)
    beams = [cosine_gaussian(cfg, 100 + k) for k in range(3 + 3 * slices)]
    for k in range(3):
        st = lc.step(cfg, const, vp, st, beams[k])
        ex.push(st)
    out = {"workload": "641^2, FP32 transforms, slices 103 on", "slices": slices,
           "threshold": GUIDED_THRESHOLD}
    for b, mode in enumerate(("plain", "guided")):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(slices):
            beam = beams[3 + b * slices + k]
            if mode == "plain":
                st = lc.step(cfg, const, vp, st, beam)
            else:
                st = lc.step(cfg, const, vp, st, beam, guess=ex.guess(),
                             threshold=GUIDED_THRESHOLD)
            ex.push(st)
        torch.cuda.synchronize()
        out[f"{mode}_us_per_slice"] = (time.perf_counter() - t0) / slices * 1e6
    accepted = 0
    for k in range(slices):
        st = lc.step(cfg, const, vp, st, beams[3 + 2 * slices + k],
                     guess=ex.guess(), threshold=GUIDED_THRESHOLD)
        ex.push(st)
        accepted += lc.guided_residual(cfg, nc, nf) <= GUIDED_THRESHOLD
    out["accepted_after"] = accepted
    return out


LARGE_WORKLOAD = dict(grid_steps=1281, grid_step_size=.0125, xi_step_size=.0025,
                      plasma_coarseness=3, plasma_fineness=3)
LARGE_BEAM = dict(amplitude=.1, sigma=.7, shift_x=.5)


def large_workload_leg(lc, cosine_gaussian, torch, slices):
    """The large non-axisymmetric case: 1281^2 (h 0.0125, xi step 0.0025,
    coarseness 3, fineness 3), beam amplitude 0.1, sigma 0.7, offset +0.5
    in x.  The plain FP64 run and the mixed run (FP32 transforms,
    GuidedStepper with the quadratic extrapolation) over the same `slices`
    slices, each timed on its own; Ez(xi, 0, 0) and the full Ez every 500
    slices compared between them."""
    def cfg(precision):
        c = make_config()
        for k, v in LARGE_WORKLOAD.items():
            setattr(c, k, v)
        c.field_solver_precision = precision
        return c
    out = {"workload": "1281^2, coarseness 3, fineness 3, beam 0.1 / 0.7 / +0.5 in x",
           "slices": slices}
    runs = {}
    for mode in ("plain", "mixed"):
        c = cfg("fp64" if mode == "plain" else "fp32")
        _, _, const, vp, st = lc.init(c)
        beams = [cosine_gaussian(c, i, **LARGE_BEAM) for i in range(slices)]
        gs = (lc.GuidedStepper(c, const, vp, GUIDED_THRESHOLD, backoff=GUIDED_BACKOFF)
              if mode == "mixed" else None)
        n = c.grid_steps
        axis, checks = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(slices):
            st = gs.step(st, beams[i]) if gs else lc.step(c, const, vp, st, beams[i])
            axis.append(st.Ez[n // 2, n // 2])
            if (i + 1) % 500 == 0:
                checks.append(st.Ez.clone())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        del beams
        runs[mode] = (torch.stack(axis), checks)
        out[f"{mode}_xi_steps_per_s"] = slices / dt
        out[f"{mode}_faults"] = lc.fault_count(c, const.m.shape[0])
        if gs:
            out.update({"guided_slices": gs.guided, "accepted": gs.accepted,
                        "plain_slices": gs.plain})
        else:
            out["max_offset_cells"] = max(st.x_offt.abs().max().item(),
                                          st.y_offt.abs().max().item()) / c.grid_step_size
        del st
    (ap, cp), (am, cm) = runs["plain"], runs["mixed"]
    out["mixed_vs_plain"] = out["mixed_xi_steps_per_s"] / out["plain_xi_steps_per_s"]
    out["ez_axis_rel_l2"] = (torch.linalg.norm(am - ap) / torch.linalg.norm(ap)).item()
    out["ez_checkpoint_rel_l2_max"] = max(
        [(torch.linalg.norm(m - p) / torch.linalg.norm(p)).item() for p, m in zip(cp, cm)],
        default=0.0)
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
    global GUIDED_THRESHOLD, GUIDED_BACKOFF
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64",
                    help="field-solve transform precision (fp32: mixed)")
    ap.add_argument("--solver-steps", type=int, default=50,
                    help="slices per field-solver back-end comparison leg (0: skip)")
    ap.add_argument("--settle", type=int, default=100,
                    help="untimed slices run before warm-up (plasma develops)")
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="steps timed per kernel with HIP events")
    ap.add_argument("--mixed-slices", type=int, default=1500,
                    help="slices of the mixed+guided vs plain comparison (0: skip)")
    ap.add_argument("--guided-threshold", type=float, default=GUIDED_THRESHOLD,
                    help="corrector residual threshold of the guided legs")
    ap.add_argument("--precision-table", type=int, default=0, metavar="SLICES",
                    help="mixed-precision error table over SLICES slices (0: skip)")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip the mixed+learned leg")
    ap.add_argument("--no-small-workload", dest="small_workload", action="store_false",
                    help="skip the 129^2 leg (GPU and oracle)")
    ap.add_argument("--guided-backoff", type=int, default=GUIDED_BACKOFF,
                    help="plain slices after a rejected guess (0: guess every slice)")
    ap.add_argument("--large", type=int, default=0, metavar="SLICES",
                    help="also run the 1281^2 fineness-3 off-axis case, plain FP64 and "
                         "mixed + guided, for SLICES slices (0: skip)")
    ap.add_argument("--guided-trace", type=int, default=0, metavar="SLICES",
                    help="only time plain against guided steps in the developed "
                         "wake (SLICES each) and count accepted guesses")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a device copy of this size (PMC calibration)")
    args = ap.parse_args()
    GUIDED_THRESHOLD = args.guided_threshold
    GUIDED_BACKOFF = args.guided_backoff
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): "
              "the xi step is sequential in xi and this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    import lcode3d_amd as lc
    from lcode3d_amd.beams import cosine_gaussian

    torch.cuda.set_device(0)
    if args.guided_trace:
        r = guided_trace_leg(lc, cosine_gaussian, torch, args.guided_trace)
        log(f"plain {r['plain_us_per_slice']:.1f} us/slice, guided "
            f"{r['guided_us_per_slice']:.1f} us/slice, accepted "
            f"{r['accepted_after']} of {r['slices']} after")
        print(json.dumps({"guided_trace": r}))
        return
    cfg = make_config()
    cfg.field_solver_precision = args.precision
    xs, ys, const, vp, state = lc.init(cfg)
    nc = const.m.shape[0]
    n = cfg.grid_steps
    total = args.warmup + args.steps
    # beam slices of the run, resident in HBM before the clock starts
    first = args.settle
    beams = [cosine_gaussian(cfg, first + k) for k in range(total)]
    for k in range(first):  # bring the plasma into a developed state
        state = lc.step(cfg, const, vp, state, cosine_gaussian(cfg, k))
    torch.cuda.synchronize()
    log(f"workload 641^2, nc={nc}, nf={vp.fine_grid.numel()}, "
        f"{args.warmup} warm-up + {args.steps} timed steps")
    for k in range(args.warmup):
        state = lc.step(cfg, const, vp, state, beams[k])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        state = lc.step(cfg, const, vp, state, beams[args.warmup + k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
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

    # per-kernel times over a separate region, HIP events on the step stream
    roofline, kernels, bracket_us = None, {}, 0.0
    if args.profile_steps:
        lc.profile_begin(64 * args.profile_steps)
        for k in range(args.profile_steps):
            state = lc.step(cfg, const, vp, state, beams[k % total])
        prof = lc.profile_end()
        for name, (calls, kms) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
            avg_us = kms / calls * 1e3
            kernels[name] = {"launches_per_step": calls / args.profile_steps,
                             "avg_us": avg_us,
                             "share": kms / sum(v[1] for v in prof.values())}
        # An event pair around a launch also times the launch's dispatch:
        # the per-launch excess of the bracketed kernel times over the
        # unbracketed timed loop is that overhead; subtracting it gives the
        # kernel's own duration (what a rocprofv3 kernel trace reports).
        launches = sum(v["launches_per_step"] for v in kernels.values())
        bracket_sum = sum(v["launches_per_step"] * v["avg_us"] for v in kernels.values())
        bracket_us = max(0.0, (bracket_sum - ms * 1e3) / launches)
        for name, v in kernels.items():
            v["kernel_us"] = v["avg_us"] - bracket_us
            b = kernel_bytes(name, nc, n)
            v["GBps"] = (b / (v["kernel_us"] * 1e-6) / 1e9) if b else None
        log(f"  event-pair overhead {bracket_us:.2f} us per launch; kernel durations:")
        for name, v in kernels.items():
            log(f"  {name:14s} {v['launches_per_step']:4.1f}/step "
                f"{v['kernel_us']:8.2f} us  {v['share'] * 100:5.1f}%"
                + (f"  {v['GBps']:7.0f} GB/s" if v["GBps"] else ""))
        dom = max((k for k in kernels if kernels[k]["GBps"]),
                  key=lambda k: kernels[k]["share"])
        ach = kernels[dom]["GBps"]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": ach,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(dom),
                    "bytes_per_launch": kernel_bytes(dom, nc, n),
                    "avg_us": kernels[dom]["kernel_us"],
                    "avg_us_with_event_pair": kernels[dom]["avg_us"],
                    "event_pair_overhead_us": bracket_us}
        if "deposit_tile" in kernels:
            # the reference's deposition rate measure: 4 quantities x 9
            # stencil cells = 36 atomic adds per fine particle
            # (lcode.py:434-449, :660-663); the tile kernel sums them in
            # registers and LDS but carries out the same 36 updates
            v = kernels["deposit_tile"]
            updates = 36 * vp.fine_grid.numel() ** 2
            v["fine_particle_updates_per_launch"] = updates
            v["fine_particle_updates_per_s"] = updates / (v["kernel_us"] * 1e-6)
            log(f"  deposit_tile: {updates / 1e6:.1f} M fine-particle stencil updates per launch, "
                f"{v['fine_particle_updates_per_s'] / 1e12:.2f} T updates/s "
                f"({vp.fine_grid.numel() ** 2 / (v['kernel_us'] * 1e-6) / 1e9:.1f} G fine particles/s)")
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
        solvers = solver_legs(lc, torch, cfg, const, vp, state, beams,
                              args.solver_steps, min(args.profile_steps or 10, 10),
                              bracket_us)

    mixed = None
    if args.mixed_slices:
        mixed = mixed_guided_leg(lc, cosine_gaussian, torch, args.mixed_slices,
                                 args.steps, "fp32")
        mixed["vs_plain"] = mixed["value"] / steps_per_s
        log(f"mixed+guided: {mixed['value']:.1f} xi-steps/s "
            f"({mixed['vs_plain']:.2f}x plain), Ez axis rel L2 "
            f"{mixed['ez_axis_rel_l2']:.2e}, checkpoints max "
            f"{mixed['ez_checkpoint_rel_l2_max']:.2e} over "
            f"{mixed['slices_compared']} slices ({mixed['accepted']} of "
            f"{mixed['guided_slices']} guesses accepted, {mixed['plain_slices']} plain "
            f"slices); timed: {mixed['timed_accepted']} of {mixed['timed_guided_slices']} "
            f"accepted, {mixed['timed_plain_slices']} plain")

    learned = None
    if args.mixed_slices and args.learned:
        learned = learned_leg(lc, cosine_gaussian, torch, args.mixed_slices,
                              args.steps, "fp32")
        learned["vs_plain"] = learned["value"] / steps_per_s
        log(f"mixed+learned: {learned['value']:.1f} xi-steps/s "
            f"({learned['vs_plain']:.2f}x plain), Ez axis rel L2 "
            f"{learned['ez_axis_rel_l2']:.2e}, checkpoints max "
            f"{learned['ez_checkpoint_rel_l2_max']:.2e} over "
            f"{learned['slices_compared']} slices ({learned['accepted']} of "
            f"{learned['guided_slices']} guesses accepted, {learned['plain_slices']} plain "
            f"slices); timed: {learned['timed_accepted']} of {learned['timed_guided_slices']} "
            f"accepted, {learned['timed_plain_slices']} plain")

    large = None
    if args.large:
        large = large_workload_leg(lc, cosine_gaussian, torch, args.large)
        log(f"1281^2 x {large['slices']} slices: plain {large['plain_xi_steps_per_s']:.1f}, "
            f"mixed+guided {large['mixed_xi_steps_per_s']:.1f} xi-steps/s "
            f"({large['mixed_vs_plain']:.2f}x); Ez axis rel L2 {large['ez_axis_rel_l2']:.2e}, "
            f"checkpoints max {large['ez_checkpoint_rel_l2_max']:.2e}; "
            f"{large['accepted']} of {large['guided_slices']} guesses accepted; max offset "
            f"{large['max_offset_cells']:.2f} cells; faults {large['plain_faults']}/"
            f"{large['mixed_faults']}")

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
        "metric": "xi-steps/s at 641^2 grid (plain FP64 HIP step)"
                  if args.precision == "fp64" else
                  "xi-steps/s at 641^2 grid (FP32-transform mixed HIP step)",
        "value": steps_per_s,
        "unit": "xi-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32 transforms",
        "data": "synthetic cosine-Gaussian beam (amplitude 0.05, sigma 1), "
                f"slices {args.settle}+ of the bunch",
        "config": {"workload": "lcode3d xi step, 641^2 grid, coarseness 3, "
                               f"fineness 2, {'plain FP64' if args.precision == 'fp64' else 'FP32 transforms'}",
                   "grid_steps": n, "coarse_particles": nc * nc,
                   "fine_particles": int(vp.fine_grid.numel()) ** 2},
        "roofline": roofline,
        "kernels": kernels,
        "solver_backends": solvers,
        "mixed_guided": mixed,
        "mixed_learned": learned,
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
