#!/usr/bin/env python3
"""Benchmark of the IBPM time step on one MI355X.

Workload: stationary circular cylinder, 1024 x 1024 cells on [-2,6] x [-4,4] (h = 1/128),
5 multi-domain levels, Re 200, 402 panels, dt 0.001 (default_scheme with CFL 0.2),
impulsive start, synthetic (no data files). One "step" is one advance! of the solver.

Prints one JSON line: steps/s of the fastest path whose forces stay within 1e-6 of the
FP64 path (headline), the FP64 and mixed paths' steps/s and the force error between them,
CG iterations per step, the roofline of the dominant kernel (the sine-contraction pass)
measured with HIP events on the solver's stream, the FP64 CPU oracle timed on the host
cores, CG iterations per step of a pitching plate with the learned warm start against the
reference's guess, batched ensembles, and plunging ellipses (moving bodies) at 2048 x 1024.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "immersedbodies.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

PEAK = {"mixed": 2500.0, "fp64": 78.6}  # dense MFMA TFLOP/s: bf16, fp64


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
HBM_PEAK = 8000.0  # GB/s


def time_solver(prob, precision, refine, steps, warmup):
    import immersedbodies as ib

    t0 = time.perf_counter()
    solver = ib.create_solver(prob, precision=precision, refine_sweeps=refine)
    setup_s = time.perf_counter() - t0
    wstats = solver.advance(warmup) if warmup else []
    fw = solver.forces(wstats) if warmup else np.zeros((0, 2))
    from immersedbodies import _lib

    _lib.check(_lib.load().ibpm_synchronize(solver._h))
    t0 = time.perf_counter()
    stats = solver.advance(steps)  # returns after the stream is synchronised
    dt = time.perf_counter() - t0
    forces = np.concatenate([fw, solver.forces(stats)])
    iters = solver.cg_iters(stats)
    return solver, dt, setup_s, forces, iters


def kernel_roofline(solver, kernel, reps, peak_tflops):
    import ctypes as C

    from immersedbodies import _lib

    ms, fl, by = C.c_double(), C.c_double(), C.c_double()
    _lib.check(_lib.load().ibpm_time_kernel(solver._h, kernel, reps, C.byref(ms), C.byref(fl),
                                            C.byref(by)))
    sec = ms.value * 1e-3
    if kernel == 2:
        return round(fl.value / sec / 1e12, 2)
    if kernel == 0:
        ach = fl.value / sec / 1e12
        g = solver.state.prob.fluid.grid
        nk, nb = g.base.n[0] - 1, g.base.n[1] - 1
        dense = g.nlevel * 2.0 * nk * nk * nb  # the unfolded sine-matrix product
        return {"bound": "mfma", "achieved": round(ach, 3), "peak": peak_tflops,
                "unit": "TFLOP/s", "frac": round(ach / peak_tflops, 4), "traffic": None,
                "avg_launch_ms": round(ms.value, 5), "flops_per_launch": fl.value,
                "flops_basis": "matrix-core flops of the folded contraction the launch runs",
                "dense_flops_per_launch": dense,
                "dense_equivalent_tflops": round(dense / sec / 1e12, 2),
                "min_bytes_per_launch": by.value,
                "hbm_GBps_at_min_bytes": round(by.value / sec / 1e9, 1)}
    ach = by.value / sec / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK, 4), "traffic": None,
            "avg_launch_ms": round(ms.value, 5), "bytes_per_launch": by.value}


def measured_traffic(precision, tiles, levels):
    """HBM bytes per launch of the sine pass timed by kernel_roofline (the 5-plane launch of
    the 1024^2 workload, tile variant ``tiles``), from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes summarised in profiles/ (scaled by the same counters on a copy of known size);
    None when not measured."""
    path = os.path.join(ROOT, "profiles", f"dst_pass_{precision}_cylinder1024.md")
    if levels != 5 or not os.path.exists(path):
        return None
    key = "traffic_bytes_per_launch[{},{},{},{}]:".format(*tiles)  # TB, TM, PF, fold
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(prob, steps):
    """FP64 CPU oracle on the host cores for a bounded number of steps of the same
    workload (B^-1 replaced by the identity: the dense assembly is setup, the per-step
    cost is the same GEMV)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cases
    import oracle as O

    threads = max(1, min(16, os.cpu_count() or 1))
    O.set_threads(threads)
    orc = cases.oracle_for(prob, timing_only=True)
    cases.oracle_advance(orc, prob, 1)
    t0 = time.perf_counter()
    cases.oracle_advance(orc, prob, steps)
    dt = time.perf_counter() - t0
    # setup: the dense B^-1 needs 2n applications of B; 4 are timed and scaled
    m = 2 * prob.bodies.npanel
    z = np.zeros(m)
    t0 = time.perf_counter()
    for k in range(4):
        z[:] = 0.0
        z[k] = 1.0
        orc.apply_B(z)
    t_b = (time.perf_counter() - t0) / 4
    orc.close()
    return {"value": round(steps / dt, 4), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same 1024x1024 5-level cylinder step after 1 "
                      "warm-up step, FP64 C/OpenMP restatement (FFT-based DST-I), B^-1 "
                      "replaced by the identity (same per-step GEMV cost)",
            "setup_s_extrapolated": round(m * t_b, 1),
            "setup_note": f"2n = {m} applications of B at {t_b * 1e3:.1f} ms each (4 timed), "
                          "LU of B not included"}


def warm_start_bench(args):
    """Learned CG warm start on a pitching plate (512^2, 4 levels, 33 points, CG to a
    1e-8 ||b||-relative residual): record samples on four reduced frequencies, train on the
    GPU, and compare CG iterations per step on a fifth, held-out frequency against CG from
    the previous step's F~b (the reference) and from the extrapolated guess alone."""
    import immersedbodies as ib
    from immersedbodies import _lib
    from immersedbodies.warmstart import WarmStartNet
    import cases

    ks_train = (0.2, 0.45, 0.7, 0.95)
    k_test = 0.6
    per = max(1, args.ws_samples // len(ks_train))
    cg = dict(maxiter=5000, reltol=1e-8)

    def make(k, precond=False):
        return ib.create_solver(cases.pitching_plate_512(k), precision=args.ws_precision,
                                refine_sweeps=args.refine, cg=cg, cg_tol_mode=1,
                                cg_precond=precond)

    net = WarmStartNet(channels=128, layers=4, kernel=5, max_points=64,
                       capacity=per * len(ks_train), max_batch=64, seed=1)
    t0 = time.perf_counter()
    for k in ks_train:
        s = make(k)
        s.set_warm_start("extrapolate", net=net, record=True)
        s.advance(per)
        s.close()
        log(f"warm start: recorded {per} steps at k={k}")
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = net.train(args.ws_iters, batch=64, lr=1e-3)
    t_train = time.perf_counter() - t0
    log(f"warm start: trained {args.ws_iters} iterations in {t_train:.1f} s")
    modes = {}
    forces = {}
    for key in ("previous", "extrapolate", "learned", "previous+precond",
                "extrapolate+precond", "learned+precond"):
        mode, precond = key.split("+")[0], key.endswith("+precond")
        s = make(k_test, precond)
        s.set_warm_start("extrapolate")
        s.advance(args.ws_warm)  # for +precond this includes the assembly (first step)
        s.set_warm_start(mode, net=net if mode == "learned" else None)
        _lib.check(_lib.load().ibpm_synchronize(s._h))
        t0 = time.perf_counter()
        st = s.advance(args.ws_eval_steps)
        dt = time.perf_counter() - t0
        modes[key] = {"cg_iters_per_step": round(float(s.cg_iters(st).mean()), 3),
                      "b_applies_per_step": round(float(s.b_applies(st).mean()), 3),
                      "steps_per_s": round(args.ws_eval_steps / dt, 2)}
        if mode == "learned":
            modes[key]["learned_guess_kept"] = float(np.mean([x.ws_used for x in st]))
        forces[key] = s.forces(st)
        s.close()
        log(f"warm start: {key} {modes[key]}")
    ref = forces["previous"]
    err = {m: float(np.max(np.linalg.norm(forces[m] - ref, axis=1) / np.linalg.norm(ref, axis=1)))
           for m in forces if m != "previous"}
    return {
        "workload": "pitching plate, 512x512 cells, 4 levels, Re 200, dt 0.002, 33 points "
                    "(spacing 2h), alpha = 20+10 sin(2 pi f t) deg; train k = "
                    f"{list(ks_train)}, test k = {k_test}; CG reltol 1e-8 of ||b||; "
                    "+precond: CG preconditioned by B^-1 of the first step, rotated with the plate",
        "train_samples": per * len(ks_train), "record_s": round(t_rec, 2),
        "train_iters": args.ws_iters, "train_s": round(t_train, 2),
        "train_loss_first_last": [float(loss[:50].mean()), float(loss[-50:].mean())],
        "eval_steps": args.ws_eval_steps, "precision": args.ws_precision, "modes": modes,
        "cg_iter_reduction_vs_previous": round(modes["previous"]["cg_iters_per_step"] /
                                               max(modes["learned"]["cg_iters_per_step"], 1e-9), 3),
        "cg_iter_reduction_learned_precond_vs_previous": round(
            modes["previous"]["cg_iters_per_step"] /
            max(modes["learned+precond"]["cg_iters_per_step"], 1e-9), 3),
        "b_apply_reduction_vs_previous": round(modes["previous"]["b_applies_per_step"] /
                                               modes["learned"]["b_applies_per_step"], 3),
        "force_rel_err_vs_previous_max": err,
    }


def batch_bench(args):
    """Batched ensemble: independent stationary-ellipse problems of uneven sizes
    (256^2..1024^2, 3-5 levels, Re 50-300, 68-400 points), advanced one graph replay per
    step, against the same solvers advanced one after another."""
    import immersedbodies as ib
    import cases

    probs = cases.ensemble(args.batch_problems, seed=2024)
    t0 = time.perf_counter()
    sol = [ib.create_solver(p, precision=args.batch_prec, refine_sweeps=args.refine)
           for p in probs]
    setup = time.perf_counter() - t0
    log(f"batch: {len(sol)} solvers set up in {setup:.1f} s")
    for s in sol:
        s.advance(2)
    t0 = time.perf_counter()
    for s in sol:
        s.advance(args.batch_steps)
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch = ib.Batch(sol)
    t_cap = time.perf_counter() - t0
    batch.advance(2)
    t0 = time.perf_counter()
    forces = batch.advance(args.batch_steps)
    t_b = time.perf_counter() - t0
    batch.close()
    n = len(sol)
    sizes = [p.fluid.grid.base.n[0] for p in probs]
    cells = sum(p.fluid.grid.base.n[0] * p.fluid.grid.base.n[1] * p.fluid.grid.nlevel
                for p in probs)
    for s in sol:
        s.close()
    res = {"workload": f"{n} stationary ellipses (a 0.5, b U[0.1,0.5], AoA U[0,30] deg, points "
                       "spaced h), N x N on [-2,6]x[-4,4] with N from {256,512,768,1024}, 3-5 "
                       "levels, Re U[50,300], dt = h/5, numpy default_rng(2024)",
           "problems": n, "grid_sizes": {str(k): sizes.count(k) for k in sorted(set(sizes))},
           "panels_min_max": [min(p.bodies.npanel for p in probs), max(p.bodies.npanel for p in probs)],
           "steps": args.batch_steps, "precision": args.batch_prec,
           "setup_s": round(setup, 2), "capture_s": round(t_cap, 3),
           "problem_steps_per_s": round(n * args.batch_steps / t_b, 2),
           "ensemble_steps_per_s": round(args.batch_steps / t_b, 3),
           "sequential_problem_steps_per_s": round(n * args.batch_steps / t_seq, 2),
           "speedup_vs_sequential": round(t_seq / t_b, 3),
           "level_cell_updates_per_s": round(cells * args.batch_steps / t_b, 1),
           "forces_finite": bool(np.all(np.isfinite(forces)))}
    log(f"batch {args.batch_prec}: {res['problem_steps_per_s']} problem-steps/s, "
        f"x{res['speedup_vs_sequential']} vs sequential")
    return res


def moving_bench(args):
    """Moving bodies at scale: three plunging ellipses on 2048 x 1024 cells of [-4,12] x
    [-4,4], 6 levels, 426 points, FP64, CG to a 1e-8 ||b||-relative residual; the
    reference's initial guess (the previous step's solution) against its quadratic
    extrapolation, each with the reference's unpreconditioned CG and with CG
    preconditioned by B^-1 of the first step (assembly timed in first_step_s)."""
    import immersedbodies as ib
    import cases

    prob = cases.plunging_ellipses()
    out = {"workload": "3 ellipses (a 0.5, b U[0.1,0.3], 3 apart) plunging y = 0.2 sin(2 pi f t "
                       "+ phase), f U[0.1,0.5], 2048x1024 cells on [-4,12]x[-4,4], 6 levels, "
                       "Re 200, dt 0.001, 426 points, impulsive start, numpy default_rng(7)",
           "precision": "fp64", "cg": {"reltol": 1e-8, "tol_mode": "||b||-relative"},
           "warmup_steps": args.moving_warmup, "steps": args.moving_steps, "modes": {}}
    forces = {}
    for key in ("previous", "extrapolate", "previous+precond", "extrapolate+precond"):
        mode, precond = key.split("+")[0], key.endswith("+precond")
        s = ib.create_solver(prob, cg=dict(maxiter=2000, reltol=1e-8), cg_tol_mode=1,
                             cg_precond=precond)
        s.set_warm_start(mode)
        t0 = time.perf_counter()
        s.advance(1)  # with +precond: assembles B (2n applications) and inverts it
        first = time.perf_counter() - t0
        s.advance(args.moving_warmup - 1)
        t0 = time.perf_counter()
        st = s.advance(args.moving_steps)
        dt = time.perf_counter() - t0
        it = s.cg_iters(st)
        ba = s.b_applies(st)
        out["modes"][key] = {"steps_per_s": round(args.moving_steps / dt, 3),
                             "ms_per_step": round(1e3 * dt / args.moving_steps, 3),
                             "cg_iters_per_step": round(float(it.mean()), 2),
                             "b_applies_per_step": round(float(ba.mean()), 2),
                             "ms_per_b_apply": round(1e3 * dt / max(1, ba.sum()), 4),
                             "first_step_s": round(first, 3),
                             "all_converged": bool(all(x.cg_converged for x in st))}
        forces[key] = s.forces(st)
        s.close()
        log(f"moving {key}: {out['modes'][key]}")
    ref = forces["previous"]
    out["force_rel_err_vs_previous_max"] = {
        k: float(np.max(np.linalg.norm(f - ref, axis=1) / np.linalg.norm(ref, axis=1)))
        for k, f in forces.items() if k != "previous"}
    out["speedup_precond_vs_previous"] = round(out["modes"]["previous"]["ms_per_step"] /
                                               out["modes"]["extrapolate+precond"]["ms_per_step"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="auto", choices=["auto", "mixed", "fp64"],
                    help="auto: time both paths, headline the fastest within the force tolerance")
    ap.add_argument("--roofline-reps", type=int, default=50)
    ap.add_argument("--refine", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--no-ws", action="store_true", help="skip the learned warm-start section")
    ap.add_argument("--ws-samples", type=int, default=4000)
    ap.add_argument("--ws-precision", default="fp64", choices=["mixed", "fp64"])
    ap.add_argument("--ws-iters", type=int, default=3000)
    ap.add_argument("--ws-warm", type=int, default=200)
    ap.add_argument("--ws-eval-steps", type=int, default=300)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble section")
    ap.add_argument("--batch-problems", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-precision", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--no-moving", action="store_true", help="skip the plunging-ellipse section")
    ap.add_argument("--moving-warmup", type=int, default=5)
    ap.add_argument("--moving-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every problem fits one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import cases

    prob = cases.cylinder_1024(nlevel=args.levels)
    steps, warm = args.steps, args.warmup
    paths = ["fp64", "mixed"] if args.precision == "auto" else [args.precision]
    runs = {}
    for prec in paths:
        sv, dt, setup, f, it = time_solver(prob, prec, args.refine, steps, warm)
        roof = kernel_roofline(sv, 0, args.roofline_reps, PEAK[prec])
        roof["tiles_TB_TM_PF_fold"] = list(sv.dst_tiles(args.levels))
        ceil = kernel_roofline(sv, 2, 5, None)  # sustained MFMA rate, register-only loop
        roof["mfma_sustained_tflops"] = ceil
        roof["frac_of_sustained"] = round(roof["achieved"] / ceil, 4) if ceil else None
        roof["traffic"] = measured_traffic(prec, roof["tiles_TB_TM_PF_fold"], args.levels)
        roof_nl = kernel_roofline(sv, 1, args.roofline_reps, None)
        sv.close()
        runs[prec] = dict(dt=dt, setup=setup, f=f, it=it, roof=roof, roof_nl=roof_nl)
        log(f"{prec}: {steps / dt:.2f} steps/s")
    tol = 1e-6
    err = None
    if "fp64" in runs and "mixed" in runs:
        fm, fr = runs["mixed"]["f"], runs["fp64"]["f"]
        n = min(len(fm), len(fr))
        err = np.linalg.norm(fm[:n] - fr[:n], axis=1) / np.linalg.norm(fr[:n], axis=1)
    # headline: the fastest path whose forces meet the 1e-6 tolerance (FP64 always does)
    ok = {p: (p == "fp64" or err is None or float(err.max()) <= tol) for p in runs}
    head = max((p for p in runs if ok[p]), key=lambda p: steps / runs[p]["dt"])
    h = runs[head]
    res = {
        "metric": "IBPM time steps/s at 1024^2 (mixed/batched vs FP64); CG iters/step; force rel err",
        "value": round(steps / h["dt"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(1e3 * h["dt"] / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "bf16x3 split / fp32 accumulate + fp64 refinement",
        "data": "synthetic",
        "config": {"workload": f"stationary cylinder, 1024x1024 cells, {args.levels} levels, Re 200, "
                               "402 panels, dt 0.001, impulsive start",
                   "precision": head, "refine_sweeps": args.refine if head == "mixed" else 0,
                   "selection": "fastest path with forces within 1e-6 of the FP64 path"},
        "cg_iters_per_step": float(np.mean(h["it"])) if len(h["it"]) else 0.0,
        "setup_s": round(h["setup"], 3),
        "final_force": [float(h["f"][-1, 0]), float(h["f"][-1, 1])],
        "roofline": h["roof"],
        "roofline_nonlinear_kernel": h["roof_nl"],
        "paths": {p: {"steps_per_s": round(steps / r["dt"], 3),
                      "ms_per_step": round(1e3 * r["dt"] / steps, 4),
                      "roofline_sine_pass": r["roof"]} for p, r in runs.items()},
    }
    if err is not None:
        res["mixed_vs_fp64"] = {"speedup": round(runs["fp64"]["dt"] / runs["mixed"]["dt"], 3),
                                "force_rel_err_max": float(err.max()),
                                "force_rel_err_final": float(err[-1]),
                                "meets_1e-6": bool(err.max() <= tol),
                                "refine_sweeps": args.refine}
    if not args.no_ws:
        res["warm_start"] = warm_start_bench(args)
    if not args.no_batch:
        precs = ["mixed", "fp64"] if args.batch_precision == "both" else [args.batch_precision]
        res["batch"] = {}
        for prec in precs:
            args.batch_prec = prec
            res["batch"][prec] = batch_bench(args)
    if not args.no_moving:
        res["moving_bodies"] = moving_bench(args)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(prob, args.cpu_steps)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
