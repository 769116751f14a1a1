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


def step_roofline(solver, ms_per_step, pass_flops):
    """Whole-step roofline of the FP64 step as it executes (folded sine passes, box-update
    products, Gauss-Seidel Helmholtz solves, spectral ring chains): per kernel class, the algorithmic matrix-core flops and minimum HBM bytes
    (each input read once, each output written once), its floor max(flops / 78.6 TFLOP/s,
    bytes / 8 TB/s), and the step's floor as the sum of those (the classes run one after
    another). pass_flops: matrix-core flops of one 5-plane sine pass (ibpm_time_kernel)."""
    g = solver.state.prob.fluid.grid
    nx, ny = g.base.n
    L = g.nlevel
    Nx, Ny = nx - 1, ny - 1
    P = Nx * Ny                       # plane values
    n = solver.state.prob.bodies.npanel
    B8 = 8.0
    comp = []

    def add(name, launches, flops, nbytes):
        comp.append({"kernel": name, "launches": launches, "flops": float(flops),
                     "bytes": float(nbytes)})

    boxes = [solver.box_update(l) for l in range(1, L + 1)]
    pbox = solver.box_update(0)
    on = boxes[0] is not None
    # the fluxes are formed from psi inside the kernel (curl!), so psi is read instead of q
    add("nonlin_rhs (curl, avg_flux, direct_product, rot, CNAB rhs; all levels)", 1, 0,
        L * (P + 4 * P) * B8)
    add("Helmholtz solves, Gauss-Seidel (all levels)", 1, 0, 2 * L * P * B8)
    # two vort2flux per step; coarsify reads the fine plane and writes the coarse centre
    # (trial); the final one re-evaluates each coarser level's box only (the fine values
    # under it read, the trial values saved, the new ones written), or the whole centre
    # without box updates
    add("coarsify (trial)", L - 1, 0, (L - 1) * (P + P / 4) * B8)
    if on:
        cells = sum(b[1] * b[3] for b in boxes[1:] if b is not None)
        add("coarsify on the boxes (final)", L - 1, 0, 7 * cells * B8)
    else:
        add("coarsify (final)", L - 1, 0, (L - 1) * (P + P / 4) * B8)
    # sine passes: forward pair of the trial, inverse pair of the final (box updates),
    # 5 planes in and out per launch
    npass = 4 if on else 8
    add("sine passes (folded, FP64 MFMA)", npass, npass * pass_flops, npass * 2 * L * P * B8)
    # ring chains: every level's spectral plane read, E written where psi is formed, the
    # ring operators M_x, M_y read per level step
    m_bytes = (Nx * Nx + Ny * Ny) * B8
    add("ring chain (sweep, reduce, factor) x 2", 2 * (3 * L - 2), 0,
        2 * L * P * B8 + (1 + L) * P * B8 + 2 * (L - 1) * m_bytes)
    if on:
        fl = by = 0.0
        for b in boxes:
            if b is None or b[1] == 0:
                continue
            bw, bh = b[1], b[3]
            # x first: the second product folded over the sine rows' parity (columns n and
            # Ny-1-n from half the products)
            fl += min(2.0 * Nx * bw * bh + 2.0 * Nx * bh * ((Ny + 1) // 2),
                      2.0 * bw * bh * Ny + 2.0 * Nx * bw * Ny)
            by += 2 * P * B8 + (bw * bh + Nx * bw + Ny * bh) * B8
        add("box update products (FP64 MFMA)", 2, fl, by)
    if pbox is not None:
        bw, bh = pbox[1], pbox[3]
        # the first product folded over its contraction (half the sine rows)
        add("trial psi on the panels' box (FP64 MFMA, K-split)", 4,
            2.0 * Nx * ((Ny + 1) // 2) * bh + 2.0 * bw * Nx * bh,
            (P + Ny * bh + bw * Nx + bw * bh) * B8)
    add("couple_surface (gather, dense B^-1 product, forces)", 3, 2.0 * (2 * n) ** 2,
        ((2 * n) ** 2 + 98 * 2 * n) * B8)
    if on:
        # the trial Gamma_1 saved on the box, rot E^T there, the box solve (b and the
        # saved values read, Gamma_1 written)
        add("project_circ (box save, rot E^T, Gauss-Seidel on the box)", 2, 0,
            6 * boxes[0][1] * boxes[0][3] * B8)
    else:
        add("project_circ (level-1 save, rot E^T, solve)", 3, 0, 5 * P * B8)
    floor = 0.0
    for c in comp:
        tf = c["flops"] / (PEAK["fp64"] * 1e12)
        tb = c["bytes"] / (HBM_PEAK * 1e9)
        c["bound"] = "mfma" if tf > tb else "hbm"
        c["floor_us"] = round(1e6 * max(tf, tb), 2)
        floor += max(tf, tb)
    return {"bound": "sum over kernel classes of max(flops / MFMA peak, bytes / HBM peak)",
            "floor_ms": round(1e3 * floor, 4), "measured_ms": round(ms_per_step, 4),
            "frac": round(1e3 * floor / ms_per_step, 4),
            "mfma_flops_per_step": sum(c["flops"] for c in comp),
            "hbm_bytes_per_step": sum(c["bytes"] for c in comp),
            "peaks": {"fp64_mfma_tflops": PEAK["fp64"], "hbm_GBps": HBM_PEAK},
            "components": comp}


def measured_traffic(precision, tiles, levels):
    """HBM bytes per launch of the sine pass timed by kernel_roofline (the 5-plane launch of
    the 1024^2 workload, tile variant ``tiles``), from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes summarised in profiles/ (scaled by the same counters on a copy of known size);
    None when not measured."""
    path = os.path.join(ROOT, "profiles", f"dst_pass_{precision}_cylinder1024.md")
    if levels != 5 or not os.path.exists(path):
        return None
    key = "traffic_bytes_per_launch[{},{},{},{},{}]:".format(*tiles)  # TB, TM, PF, fold, KC
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
    # the plate as first stated: 200 points (spacing 0.32 h), plain and preconditioned CG
    stated = {}
    for key, precond in (("previous", False), ("previous+precond", True)):
        s = ib.create_solver(cases.pitching_plate_512(k_test, npts=200),
                             precision=args.ws_precision, refine_sweeps=args.refine, cg=cg,
                             cg_tol_mode=1, cg_precond=precond)
        t0 = time.perf_counter()
        st = s.advance(3)
        dt = time.perf_counter() - t0
        stated[key] = {"cg_iters": [int(x.cg_iters) for x in st],
                       "converged": [bool(x.cg_converged) for x in st],
                       "cg_residual": [float(x.cg_residual) for x in st],
                       "s_per_step": round(dt / 3, 3)}
        s.close()
        log(f"warm start: 200-point plate {key} {stated[key]}")
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
        "stated_200_points": {"workload": "the same plate with 200 points (spacing 0.32 h), "
                                          "k = 0.6, first 3 steps, CG maxiter 5000",
                              "modes": stated},
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


def multibody_bench(args):
    """Combined workload: problems of 2-4 plunging ellipses each (cases.multibody_ensemble:
    2048 x 1024 cells on [-4,12] x [-4,4], 6 levels, Re U[100,1000]), all advanced together
    by two graph replays per step (up to the CG, and the rest). CG to a 1e-8 ||b||-relative
    residual, preconditioned by B^-1 of the first step, starting from the learned guess
    (a network trained on steps the solver recorded on two of the problems). Forces of one
    problem are checked against the FP64 path with the reference's plain CG."""
    import immersedbodies as ib
    from immersedbodies.warmstart import WarmStartNet
    import cases

    n_max = max(args.mb_problems, args.mb_mixed_problems)
    probs = cases.multibody_ensemble(n_max)
    cg = dict(maxiter=5000, reltol=1e-8)
    warm = 3
    npts = max(p.bodies.npanel for p in probs)
    net = WarmStartNet(channels=128, layers=4, kernel=5, max_points=1024,
                       capacity=2 * args.mb_record, max_batch=16, seed=2)
    t0 = time.perf_counter()
    for p in probs[:2]:
        s = ib.create_solver(p, box_update="off", cg=cg, cg_tol_mode=1, cg_precond=True)
        s.set_warm_start("extrapolate", net=net, record=True)
        s.advance(args.mb_record)
        s.close()
    loss = net.train(args.mb_train_iters, batch=16, lr=1e-3)
    t_net = time.perf_counter() - t0
    log(f"multibody: network recorded and trained in {t_net:.1f} s")
    out = {"workload": "problems of 2-4 ellipses (semi-axes U[0.1,0.5], centres U[0,4]x"
                       "[-1.5,1.5], no overlap) plunging y = 0.2 sin(2 pi f t), f U[0.1,0.5], "
                       "2048x1024 cells on [-4,12]x[-4,4], 6 levels, Re U[100,1000], dt 0.001, "
                       "panels spaced 2h, numpy default_rng(2025)",
           "cg": {"reltol": 1e-8, "tol_mode": "||b||-relative", "maxiter": 5000,
                  "preconditioner": "B^-1 of the first step, rotated with each body",
                  "initial_guess": "learned (network trained on "
                                   f"{2 * args.mb_record} recorded steps)"},
           "net_record_train_s": round(t_net, 2),
           "train_loss_first_last": [float(loss[:20].mean()), float(loss[-20:].mean())],
           "panels_max": npts, "runs": {}}
    f0 = None
    for prec, nprob in (("fp64", args.mb_problems), ("mixed", args.mb_mixed_problems)):
        if nprob <= 0:
            continue
        t0 = time.perf_counter()
        sol = []
        for p in probs[:nprob]:
            s = ib.create_solver(p, precision=prec, refine_sweeps=args.refine, box_update="off",
                                 cg=cg, cg_tol_mode=1, cg_precond=True)
            s.set_warm_start("extrapolate")
            s.advance(warm)  # the preconditioner's assembly and the guess's history
            s.set_warm_start("learned", net=net)
            sol.append(s)
        setup = time.perf_counter() - t0
        batch = ib.Batch(sol)
        fw = batch.advance(1)
        t0 = time.perf_counter()
        f = batch.advance(args.mb_steps)
        dt = time.perf_counter() - t0
        st = batch.last_stats
        it = batch.cg_iters()
        kept = np.mean([[x.ws_used for x in row] for row in st])
        conv = all(x.cg_converged for row in st for x in row)
        batch.close()
        for s in sol:
            s.close()
        run = {"problems": nprob, "precision": prec, "steps": args.mb_steps,
               "setup_s": round(setup, 2),
               "problem_steps_per_s": round(nprob * args.mb_steps / dt, 3),
               "ensemble_steps_per_s": round(args.mb_steps / dt, 4),
               "ms_per_problem_step": round(1e3 * dt / (nprob * args.mb_steps), 3),
               "cg_iters_per_step": round(float(it.mean()), 3),
               "learned_guess_kept": round(float(kept), 3), "all_converged": bool(conv),
               "bodies_per_problem": [len(p.bodies.sections) for p in probs[:nprob]]}
        out["runs"][prec] = run
        if prec == args.mb_check_precision or f0 is None:
            f0 = (prec, np.concatenate([fw[:, 0], f[:, 0]]))
        log(f"multibody {prec}: {run['problem_steps_per_s']} problem-steps/s, "
            f"{run['cg_iters_per_step']} CG iterations/step")
    # problem 0 with the FP64 path and the reference's plain CG from the previous solution
    ref = ib.create_solver(probs[0], cg=dict(maxiter=5000, reltol=1e-10), cg_tol_mode=1)
    fr = ref.forces(ref.advance(warm + 1 + args.mb_steps))[warm:]
    ref.close()
    prec, fb = f0
    e = np.linalg.norm(fb - fr, axis=1) / np.linalg.norm(fr, axis=1)
    out["force_check"] = {"problem": 0, "path": prec, "against": "FP64, plain CG to 1e-10",
                          "rel_err_max": float(e.max()), "meets_1e-6": bool(e.max() <= 1e-6)}
    return out


# ---- measurement helpers for the rocprofv3 summaries under profiles/ ----
# (bench.py --kernel-workload K: the program to profile; bench.py --summarize MODE PATH:
# the tables the summaries quote)

import sqlite3  # noqa: E402
from collections import defaultdict  # noqa: E402


def step_breakdown(db_path: str, steps: int = 20, marker: str = "nonlin_rhs_kernel"):
    db = sqlite3.connect(db_path)
    rows = list(db.execute("select name, start, end from kernels order by start"))
    # a step's marker is followed by the step's first Helmholtz-solve kernel, a sine
    # contraction pass or a Gauss-Seidel launch (isolated marker launches, such as a
    # kernel-timing loop, are not step boundaries)
    marks = [i for i, r in enumerate(rows)
             if marker in r[0] and i + 1 < len(rows)
             and ("dst_pass" in rows[i + 1][0] or "gs_helmholtz" in rows[i + 1][0])]
    if len(marks) < 2:
        raise ValueError("fewer than two marker dispatches")
    sel = marks[-(steps + 1):]
    n = len(sel) - 1
    per = defaultdict(lambda: [0, 0.0])
    busy = 0.0
    for i in range(sel[0], sel[-1]):
        name, s, e = rows[i]
        d = (e - s) * 1e-3  # ns -> us
        short = name.replace("(anonymous namespace)::", "").split("(")[0]
        short = short.replace("void ", "").replace("ibpm::", "")
        per[short][0] += 1
        per[short][1] += d
        busy += d
    wall = (rows[sel[-1]][1] - rows[sel[0]][1]) * 1e-3
    table = sorted(((k, c / n, t / n) for k, (c, t) in per.items()), key=lambda x: -x[2])
    return n, wall / n, busy / n, table


def workload(kernel: int, precision: str, reps: int, copy_mib: int = 2048, levels: int = 5,
             tiles: int = -1):
    """Time one kernel class of the 1024^2 cylinder step (ibpm_time_kernel: kernels 0 and 1
    inside ``reps`` real steps, kernel 2 in a loop), then one plain device-to-device copy of
    ``copy_mib`` MiB whose byte count is known: the reference for scaling the FETCH_SIZE /
    WRITE_SIZE counters."""
    import ctypes as C

    import cases  # workload builders (synthetic inputs)
    import immersedbodies as ib
    from immersedbodies import _lib

    s = ib.create_solver(cases.cylinder_1024(nlevel=levels), precision=precision, dst_tiles=tiles)
    s.advance(2)
    ms, fl, by = C.c_double(), C.c_double(), C.c_double()
    _lib.check(_lib.load().ibpm_time_kernel(s._h, kernel, reps, C.byref(ms), C.byref(fl),
                                            C.byref(by)))
    print(f"kernel {kernel} {precision}: avg {ms.value:.5f} ms, flops {fl.value:.6g}, "
          f"min bytes {by.value:.6g}", flush=True)
    s.close()
    # reference copy: one device-to-device hipMemcpy (the runtime's copy kernel) of a source
    # written before a 1 GiB flush buffer, so that neither L2 nor the 256 MB last-level
    # cache still holds it
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(copy_mib * (1 << 20))
    fbytes = C.c_size_t(1 << 30)
    a, b, f = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for ptr, nb in ((a, nbytes), (b, nbytes), (f, fbytes)):
        if hip.hipMalloc(C.byref(ptr), nb) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, nbytes)
    hip.hipMemset(f, 2, fbytes)
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    for ptr in (a, b, f):
        hip.hipFree(ptr)
    print(f"copy: {copy_mib} MiB read, {copy_mib} MiB written", flush=True)


def pmc_summary(db_path: str, match: str, last: int = 40, copy_match: str = "copyBuffer"):
    """From a rocprofv3 --pmc database: per kernel whose name contains ``match``, the mean
    counter value (KiB) over its last ``last`` dispatches with the grid of its final
    dispatch (the launches inside the timed steps, not the setup's), and the largest value
    on a dispatch of the reference copy kernel."""
    db = sqlite3.connect(db_path)
    rows = list(db.execute("select dispatch_id, kernel_name, grid_size, counter_name, value "
                           "from counters_collection order by dispatch_id"))
    per = defaultdict(lambda: defaultdict(float))  # counter -> dispatch -> value
    meta = {}
    for d, name, grid, counter, value in rows:
        per[counter][d] += value
        meta[d] = (name, grid)
    out = []
    for counter, disp in per.items():
        cv = [disp[d] for d in disp if copy_match in meta[d][0]]
        names = sorted({meta[d][0] for d in disp if match in meta[d][0]})
        for name in names:
            ks = [d for d in sorted(disp) if meta[d][0] == name]
            g = meta[ks[-1]][1]
            ks = [d for d in ks if meta[d][1] == g][-last:]
            out.append((counter, name, sum(disp[d] for d in ks) / len(ks), len(ks),
                        max(cv) if cv else None))
    return out


def trace_tail(csv_path: str, match: str, last: int = 200):
    """From a rocprofv3 --kernel-trace CSV: mean and spread of the durations (us) of the last
    ``last`` dispatches of the kernel whose name contains ``match`` that share the grid of
    its final dispatch (the timing loop that follows the steps in bench.py)."""
    import csv

    rows = []
    with open(csv_path) as fh:
        for r in csv.DictReader(fh):
            if match in r["Kernel_Name"]:
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], int(r["Grid_Size_X"]),
                             (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3))
    rows.sort()
    g, name = rows[-1][2], rows[-1][1]
    d = [r[3] for r in rows if r[2] == g and r[1] == name][-last:]
    mean = sum(d) / len(d)
    return name, g, len(d), mean, min(d), max(d)


def summarize(mode, path, match, last, steps):
    if mode == "trace":
        name, g, n, mean, lo, hi = trace_tail(path, match, last)
        print(f"`{name}` grid {g}: last {n} dispatches mean {mean:.2f} us "
              f"(min {lo:.2f}, max {hi:.2f})")
    elif mode == "pmc":
        for counter, kname, mean, n, copyv in pmc_summary(path, match, last):
            scale = 2048 * 1024.0 / copyv if copyv else None
            print(f"{counter}: `{kname}`: {mean:.1f} KiB per dispatch (mean of {n}); copy of "
                  f"2048 MiB: {copyv:.1f} KiB, scale {scale:.4f}; scaled "
                  f"{mean * scale / 1024:.2f} MiB per dispatch")
    else:
        n, wall, busy, table = step_breakdown(path, steps)
        print(f"steps aggregated: {n}; wall per step {wall:.1f} us; kernel time per step "
              f"{busy:.1f} us\n")
        print("| kernel | calls/step | us/step | % of kernel time |\n|---|---|---|---|")
        for k, c, t in table:
            print(f"| `{k}` | {c:.1f} | {t:.1f} | {100 * t / busy:.1f} |")


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
    ap.add_argument("--no-multibody", action="store_true",
                    help="skip the batched multi-body plunging section")
    ap.add_argument("--mb-problems", type=int, default=32)
    ap.add_argument("--mb-mixed-problems", type=int, default=8)
    ap.add_argument("--mb-steps", type=int, default=5)
    ap.add_argument("--mb-record", type=int, default=150)
    ap.add_argument("--mb-train-iters", type=int, default=1500)
    ap.add_argument("--mb-check-precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--kernel-workload", type=int, default=None,
                    help="only time kernel K of the 1024^2 step (ibpm_time_kernel; K = 0, 1 inside "
                         "--reps steps) and a 2048 MiB reference copy: the program rocprofv3 profiles")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--tiles", type=int, default=-1)
    ap.add_argument("--summarize", nargs=2, metavar=("MODE", "PATH"), default=None,
                    help="MODE breakdown (rocprofv3 results db), pmc (--pmc results db) or "
                         "trace (kernel-trace csv): print the table a profile summary quotes")
    ap.add_argument("--match", default="dst_pass")
    ap.add_argument("--last", type=int, default=200)
    args = ap.parse_args()
    if args.summarize:
        summarize(args.summarize[0], args.summarize[1], args.match, args.last, args.steps)
        return
    if args.kernel_workload is not None:
        workload(args.kernel_workload, "mixed" if args.precision == "mixed" else "fp64",
                 args.reps, levels=args.levels, tiles=args.tiles)
        return
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
        ceil = kernel_roofline(sv, 2, 5, None)  # sustained MFMA rate, register-only loop
        roof = kernel_roofline(sv, 0, args.roofline_reps, PEAK[prec])
        roof["tiles_TB_TM_PF_fold_KC"] = list(sv.dst_tiles(args.levels))
        roof["mfma_sustained_tflops"] = ceil
        roof["frac_of_sustained"] = round(roof["achieved"] / ceil, 4) if ceil else None
        roof["traffic"] = measured_traffic(prec, roof["tiles_TB_TM_PF_fold_KC"], args.levels)
        roof_nl = kernel_roofline(sv, 1, args.roofline_reps, None)
        step_roof = step_roofline(sv, 1e3 * dt / steps, roof["flops_per_launch"]) \
            if prec == "fp64" else None
        sv.close()
        runs[prec] = dict(dt=dt, setup=setup, f=f, it=it, roof=roof, roof_nl=roof_nl,
                          step_roof=step_roof)
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
        "roofline_step": h["step_roof"],
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
    if not args.no_multibody:
        res["multibody_batch"] = multibody_bench(args)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(prob, args.cpu_steps)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
