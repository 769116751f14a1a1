"""GPU parity: libibpm (through its C ABI) against the FP64 CPU oracle on identical inputs.

Tolerances: the FP64 path must match fields within 1e-10 and forces within 1e-9
(relative, 2-norm over the array); the mixed path (3xBF16 products, FP32 accumulation,
FP64 refinement) must match forces within 1e-6.
"""
import numpy as np
import pytest

import cases
import immersedbodies as ib

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
FORCE_TOL = 1e-9


def _compare_state(solver, orc, field_tol=FIELD_TOL, force_tol=FORCE_TOL, fields=None):
    fields = fields or ["Gamma", "psi", "q", "nonlin1", "nonlin2", "Ftilde", "fb", "xb", "ub"]
    errs = {}
    for name in fields:
        a, b = solver.read(name), orc.field(name)
        errs[name] = cases.relerr(a, b)
    for name, e in errs.items():
        tol = force_tol if name in ("Ftilde", "fb") else field_tol
        assert e <= tol, f"{name}: rel err {e:.3e} > {tol:.1e} ({errs})"
    return errs


@pytest.mark.parametrize("box", ["auto", "force"])
@pytest.mark.parametrize("builder", [cases.cylinder, cases.two_lines])
def test_static_fp64_matches_oracle(builder, box):
    """box "force": the box updates of the final vort2flux and the trial psi on the panels'
    box run on these small grids too (where they would not pay)."""
    prob = builder()
    print("case", builder.__name__, box, flush=True)
    solver = ib.create_solver(prob, box_update=box)
    if box == "force":
        assert solver.box_update(1) is not None
    orc = cases.oracle_for(prob)
    nsteps = 6
    stats = solver.advance(nsteps)
    f_ref = cases.oracle_advance(orc, prob, nsteps)
    f_gpu = solver.forces(stats)
    assert cases.relerr(f_gpu, f_ref) <= FORCE_TOL, (f_gpu, f_ref)
    _compare_state(solver, orc)
    assert solver.state.index == nsteps
    assert abs(solver.state.t - 5 * prob.scheme.dt) < 1e-15


@pytest.mark.parametrize("y0", [-1.5, -1.55])
def test_static_fp64_even_ny_matches_oracle(y0):
    """ny = 61: the interior length Ny = 60 is even, so the folded first product of the
    trial psi on the panels' box pairs every sine row with its mirror (no middle row); the
    two origins, one cell apart, start the box on rows of both parities."""
    prob = cases.cylinder(nx=60, ny=61, x0=(-1.0, y0))
    print("case even Ny", y0, flush=True)
    solver = ib.create_solver(prob, box_update="force")
    assert solver.box_update(0) is not None
    orc = cases.oracle_for(prob)
    nsteps = 6
    stats = solver.advance(nsteps)
    f_ref = cases.oracle_advance(orc, prob, nsteps)
    f_gpu = solver.forces(stats)
    assert cases.relerr(f_gpu, f_ref) <= FORCE_TOL, (f_gpu, f_ref)
    _compare_state(solver, orc)


def test_no_body_uniform_flow_stays_irrotational():
    prob = cases.no_body()
    solver = ib.create_solver(prob)
    orc = cases.oracle_for(prob)
    solver.advance(5)
    cases.oracle_advance(orc, prob, 5)
    G = solver.read("Gamma")
    assert np.abs(G).max() < 1e-12
    _compare_state(solver, orc, fields=["q", "Gamma", "psi"])


@pytest.mark.parametrize("box", ["auto", "force"])
def test_moving_cylinder_cg_matches_oracle(box):
    """box "force": B and the step by the box updates, the boxes following the body."""
    prob = cases.moving_cylinder()
    print("case moving_cylinder", box, flush=True)
    solver = ib.create_solver(prob, box_update=box)
    orc = cases.oracle_for(prob)
    nsteps = 4
    stats = solver.advance(nsteps)
    if box == "force":
        assert solver.box_update(1) is not None  # the last step's boxes
    f_ref = cases.oracle_advance(orc, prob, nsteps)
    assert cases.relerr(solver.forces(stats), f_ref) <= 1e-8
    iters = solver.cg_iters(stats)
    assert (iters > 0).all()
    _compare_state(solver, orc, field_tol=1e-9, force_tol=1e-8)


def test_mixed_precision_forces():
    prob = cases.cylinder()
    solver = ib.create_solver(prob, precision="mixed", refine_sweeps=2)
    orc = cases.oracle_for(prob)
    nsteps = 6
    stats = solver.advance(nsteps)
    f_ref = cases.oracle_advance(orc, prob, nsteps)
    e = cases.relerr(solver.forces(stats), f_ref)
    print("mixed force rel err", e)
    assert e <= 1e-6


def test_outside_region_error():
    # a cylinder touching the grid edge (test/solver.jl:66-79)
    prob = cases.cylinder(nx=16, ny=20, nlevel=3, h=0.1, x0=(-1.0, -1.0), Re=20.0, dt=0.01)
    with pytest.raises(Exception, match="outside allowed region"):
        ib.create_solver(prob)


def test_cylinder_200_full_run_matches_oracle():
    """The reference's CPU-sized workload end to end: 200 x 200 cells on [-1,3] x [-2,2],
    3 levels, Re 100, dt 0.004, 157 points, impulsive start, 500 steps. FP64 forces and
    fields stay within the FP64 tolerances of the oracle over the whole run; the mixed
    path's lift/drag history stays within 1e-6. Symmetry keeps the lift at round-off in
    the first steps."""
    import oracle as O

    O.set_threads(16)
    prob = cases.cylinder_200()
    orc = cases.oracle_for(prob)
    nsteps = 500
    f_ref = cases.oracle_advance(orc, prob, nsteps)
    fp64 = ib.create_solver(prob)
    f64 = fp64.forces(fp64.advance(nsteps))
    err = np.linalg.norm(f64 - f_ref, axis=1) / np.linalg.norm(f_ref, axis=1)
    print("fp64 force rel err max", err.max(), "final drag", f64[-1, 0], flush=True)
    assert err.max() <= FORCE_TOL
    _compare_state(fp64, orc, fields=["Gamma", "psi", "q", "Ftilde", "fb"])
    # the box updates forced on (they do not pay at this size): the same history
    boxed = ib.create_solver(prob, box_update="force")
    assert boxed.box_update(1) is not None and boxed.box_update(0) is not None
    fbx = boxed.forces(boxed.advance(nsteps))
    errb = np.linalg.norm(fbx - f_ref, axis=1) / np.linalg.norm(f_ref, axis=1)
    print("box-update force rel err max", errb.max(), flush=True)
    assert errb.max() <= FORCE_TOL
    _compare_state(boxed, orc, fields=["Gamma", "psi", "q", "Ftilde", "fb"])
    mixed = ib.create_solver(prob, precision="mixed")
    fm = mixed.forces(mixed.advance(nsteps))
    errm = np.linalg.norm(fm - f_ref, axis=1) / np.linalg.norm(f_ref, axis=1)
    print("mixed force rel err max", errm.max(), flush=True)
    assert errm.max() <= 1e-6
    # y-symmetric body and flow: lift stays at round-off early on
    assert np.abs(f64[:20, 1]).max() <= 1e-10 * np.abs(f64[:20, 0]).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_moving_grid_matches_oracle(precision):
    """MovingGrid (translating and turning frame, static ellipse in it): fields, forces and
    the base flux against the oracle."""
    prob = cases.rotating_frame()
    print("case rotating_frame", precision, flush=True)
    # FP64: the box updates forced (a static ellipse in the moving frame)
    solver = ib.create_solver(prob, precision=precision,
                              box_update="force" if precision == "fp64" else "auto")
    orc = cases.oracle_for(prob)
    stats = solver.advance(8)
    f_ref = cases.oracle_advance(orc, prob, 8)
    tol = FORCE_TOL if precision == "fp64" else 1e-6
    assert cases.relerr(solver.forces(stats), f_ref) <= tol
    assert cases.relerr(solver.read("q0"), orc.field("q0")) <= 1e-14
    if precision == "fp64":
        _compare_state(solver, orc, fields=["Gamma", "psi", "q", "nonlin1", "Ftilde", "fb"])


@pytest.mark.parametrize("box", ["auto", "force"])
def test_naca_airfoil_matches_oracle(box):
    """NACA 0012 at 20 degrees built by the Curves module (the reference's NACA example):
    fields and forces against the oracle, with and without the forced box updates."""
    prob = cases.naca_airfoil()
    print("case naca_airfoil 0012 20deg", box, flush=True)
    solver = ib.create_solver(prob, box_update=box)
    orc = cases.oracle_for(prob)
    stats = solver.advance(5)
    f_ref = cases.oracle_advance(orc, prob, 5)
    assert cases.relerr(solver.forces(stats), f_ref) <= FORCE_TOL
    _compare_state(solver, orc, fields=["Gamma", "psi", "q", "Ftilde", "fb"])


@pytest.mark.parametrize("box", ["auto", "force"])
def test_plunging_ellipses_matches_oracle(box):
    """Three plunging ellipses (moving bodies, CG) on 256 x 128 cells of [-4,12] x [-4,4],
    3 levels: fields and forces against the oracle, with and without the box updates."""
    prob = cases.plunging_ellipses(nx=256, ny=128, nlevel=3, h=1.0 / 16)
    print("case plunging_ellipses 256x128 L3", box, flush=True)
    solver = ib.create_solver(prob, box_update=box)
    orc = cases.oracle_for(prob)
    stats = solver.advance(5)
    f_ref = cases.oracle_advance(orc, prob, 5)
    assert cases.relerr(solver.forces(stats), f_ref) <= 1e-8
    assert all(s.cg_converged for s in stats)
    _compare_state(solver, orc, field_tol=1e-9, force_tol=1e-8, fields=["q", "Gamma", "psi", "xb"])


def test_plunging_ellipses_full_size_warm_start_invariance():
    """2048 x 1024 cells, 6 levels, three plunging ellipses (426 points): CG converges
    every step, and the forces do not depend on CG's initial guess (the previous step's
    solution or its extrapolation) beyond the CG tolerance."""
    prob = cases.plunging_ellipses()
    print("case plunging_ellipses 2048x1024 L6", flush=True)
    forces, iters = {}, {}
    for mode in ("previous", "extrapolate"):
        s = ib.create_solver(prob, cg=dict(maxiter=500, reltol=1e-10), cg_tol_mode=1)
        s.set_warm_start(mode)
        st = s.advance(6)
        assert all(x.cg_converged for x in st), mode
        forces[mode] = s.forces(st)
        iters[mode] = s.cg_iters(st)
        assert np.isfinite(forces[mode]).all()
        s.close()
    print("cg iterations", iters, flush=True)
    assert cases.relerr(forces["extrapolate"], forces["previous"]) <= 1e-7
    # the box updates (on here: the boxes follow the bodies) against the full transforms
    s = ib.create_solver(prob, cg=dict(maxiter=500, reltol=1e-10), cg_tol_mode=1, box_update="off")
    s.set_warm_start("previous")
    st = s.advance(6)
    assert all(x.cg_converged for x in st)
    off = s.forces(st)
    s.close()
    print("box on/off force rel err", cases.relerr(forces["previous"], off), flush=True)
    assert cases.relerr(forces["previous"], off) <= 1e-7


def test_staged_body_motion_matches_motion_array():
    """ibpm_set_body_motion + ibpm_advance(motion = NULL) equals the motion array, and
    the staging is checked: the time, a static body, a missing body, several steps."""
    prob = cases.moving_cylinder()
    a = ib.create_solver(prob)
    b = ib.create_solver(prob)
    dt = prob.scheme.dt
    body = prob.bodies["cyl"]
    for k in range(3):
        a.advance(1)
        t = b.state.t0 + dt * b.state.index
        b.set_body_motion(0, t, body.motion(t))
        b.advance_staged()
    for name in ("Gamma", "psi", "Ftilde", "xb", "ub"):
        assert np.array_equal(a.read(name), b.read(name)), name
    with pytest.raises(RuntimeError, match="has no motion"):
        b.advance_staged()
    t = b.state.t0 + dt * b.state.index
    b.set_body_motion(0, t + dt, body.motion(t + dt))
    with pytest.raises(RuntimeError, match="evaluated at t"):
        b.advance_staged()
    with pytest.raises(RuntimeError, match="out of range"):
        b.set_body_motion(1, t, body.motion(t))
    s = ib.create_solver(cases.cylinder())
    with pytest.raises(RuntimeError, match="is static"):
        s.set_body_motion(0, 0.0, body.motion(0.0))


def test_cylinder_1024_matches_oracle_fixture():
    """The full-size 1024 x 1024, 5-level cylinder for 5 steps against the oracle frozen in
    tests/golden/oracle_cylinder1024.npz: forces and F~b within 1e-9; per-level norms of
    Gamma, psi and q within 1e-10, and the error over 4096 sampled values within 1e-10 of
    the field's norm (what a full-field relative error of 1e-10 implies)."""
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "oracle_cylinder1024.npz")
    ref = np.load(path)
    prob = cases.cylinder_1024()
    print("case cylinder_1024 fixture", flush=True)
    solver = ib.create_solver(prob)
    stats = solver.advance(len(ref["forces"]))
    assert cases.relerr(solver.forces(stats), ref["forces"]) <= FORCE_TOL
    assert cases.relerr(solver.read("Ftilde"), ref["Ftilde"]) <= FORCE_TOL
    L = prob.fluid.grid.nlevel
    for name in ("Gamma", "psi", "q"):
        f = solver.read(name)
        norms = np.linalg.norm(f.reshape(L, -1), axis=1)
        assert np.allclose(norms, ref[f"{name}/norms"], rtol=1e-10, atol=0), name
        err = np.linalg.norm(f[ref[f"{name}/idx"]] - ref[f"{name}/values"])
        assert err <= 1e-10 * np.linalg.norm(ref[f"{name}/norms"]), (name, err)


def test_cylinder_1024_mixed_matches_oracle_fixture():
    """The mixed path (3xBF16 sine contractions with FP32 accumulation, two FP64 refinement
    sweeps) on the full-size 1024 x 1024, 5-level cylinder against the FP64 oracle frozen in
    tests/golden/oracle_cylinder1024.npz: the force within 1e-6 at every step (the mixed
    path's target), F~b within 1e-6, per-level norms within 1e-8."""
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "oracle_cylinder1024.npz")
    ref = np.load(path)
    prob = cases.cylinder_1024()
    print("case cylinder_1024 fixture, mixed", flush=True)
    solver = ib.create_solver(prob, precision="mixed", refine_sweeps=2)
    f = solver.forces(solver.advance(len(ref["forces"])))
    # per step, relative to the force vector (the lift of this symmetric flow is round-off)
    e = np.linalg.norm(f - ref["forces"], axis=1) / np.linalg.norm(ref["forces"], axis=1)
    print("force rel err per step", e, flush=True)
    assert (e <= 1e-6).all(), e
    assert cases.relerr(solver.read("Ftilde"), ref["Ftilde"]) <= 1e-6
    L = prob.fluid.grid.nlevel
    for name in ("Gamma", "psi", "q"):
        norms = np.linalg.norm(solver.read(name).reshape(L, -1), axis=1)
        assert np.allclose(norms, ref[f"{name}/norms"], rtol=1e-8, atol=0), name


@pytest.mark.parametrize("nlevel", [4, 1])
def test_box_update_matches_full_transform(nlevel):
    """The final vort2flux by the box update (the trial's spectral planes plus the transform
    of the change of Gamma, zero outside boxes around the body) and the trial psi on the
    panels' box only, against the full transforms: 512 x 512 cells, h = 0.01, dt = 0.001 (23 Gauss-Seidel half-sweeps on level
    1), cylinder of radius 0.5 off centre. Fields within 1e-12, forces within 1e-12; a box
    that missed part of the change would leave errors of the change's size."""
    prob = cases.cylinder(nx=512, ny=512, nlevel=nlevel, h=0.01, x0=(-1.6, -2.6), Re=100.0,
                          dt=0.001)
    print("case box update, levels", nlevel, flush=True)
    a = ib.create_solver(prob)
    b = ib.create_solver(prob, box_update="off")
    assert b.box_update() is None
    box = a.box_update(1)
    print("level-1 box", box, [a.box_update(l) for l in range(2, nlevel + 1)], flush=True)
    if nlevel == 1:
        # one level: the box update does not pay (it needs at most half the passes' work)
        assert box is None
        return
    assert box is not None and 0 < box[1] < 511 and 0 < box[3] < 511
    pbox = a.box_update(0)  # the trial psi only around the panels
    assert pbox is not None and pbox[1] < box[1] and pbox[3] < box[3]
    fa = a.forces(a.advance(6))
    fb = b.forces(b.advance(6))
    assert cases.relerr(fa, fb) <= 1e-12
    for name in ("Gamma", "psi", "q", "Ftilde"):
        assert cases.relerr(a.read(name), b.read(name)) <= 1e-12, name


def test_force_history_across_pinned_chunks():
    """advance(n) hands each step's total force over through a pinned buffer flushed every
    1024 steps: 1100 steps in one call give the same force history, bit for bit, as 1100
    calls of one step each (on a small cylinder, 60^2 cells, 3 levels)."""
    prob = cases.cylinder()
    print("case cylinder 60^2, 1100 steps", flush=True)
    whole = ib.create_solver(prob)
    f_whole = whole.forces(whole.advance(1100))
    single = ib.create_solver(prob)
    f_single = np.concatenate([single.forces(single.advance(1)) for _ in range(1100)])
    assert f_whole.shape == (1100, 2)
    assert np.array_equal(f_whole, f_single)
    assert np.all(np.isfinite(f_whole)) and np.abs(f_whole).max() > 0
    assert np.array_equal(whole.read("Gamma"), single.read("Gamma"))
