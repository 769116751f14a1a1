"""Batched moving bodies, moving grids and time-varying freestreams (one graph per step).

A moving body's CG (cg!, coupling.jl:194-202) runs inside the captured step as a budget
of iterations whose launches return at once after the stopping test is met, continued on
the host's side up to cg_maxiter when the budget falls short, and each step's
body motion (bodies.jl:43-51), freestream (solver.jl:165) and grid motion (fluid.jl:34-67)
come from device tables the batch fills per call. Every solver advanced in a batch must
end bitwise equal to the same solver advanced alone with the same inputs: fields, forces
and CG iteration counts, in FP64 and mixed precision, with the previous-step,
extrapolated, learned and preconditioned starts.
"""
import math

import numpy as np
import pytest

import cases
import immersedbodies as ib
from immersedbodies import curves
from immersedbodies.warmstart import WarmStartNet

pytestmark = pytest.mark.gpu

FIELDS = ("q", "Gamma", "psi", "nonlin1", "nonlin2", "Ftilde", "fb", "xb", "ub")
CG = dict(maxiter=200, reltol=1e-10)


def gusty_cylinder():
    """Stationary cylinder in a freestream that changes every step."""
    p = cases.cylinder(nx=64, ny=64, nlevel=3, h=0.05, x0=(-1.6, -1.6), Re=100.0)
    p.fluid.freestream_vel = lambda t: (1.0 + 0.5 * math.sin(10 * t), 0.2 * math.cos(7 * t))
    return p


def _moving_problems():
    return [
        (cases.moving_cylinder(), {}),
        (cases.plunging_ellipses(nx=256, ny=128, nlevel=3, h=1.0 / 16), {"ws": "extrapolate"}),
        (cases.plunging_ellipses(nx=192, ny=96, nlevel=2, h=1.0 / 12, nbody=2, seed=3),
         {"precond": True}),
        (cases.pitching_plate(nx=64, ny=64, nlevel=3), {"ws": "extrapolate"}),
        (cases.rotating_frame(nx=64, ny=64, nlevel=3), {}),
        (gusty_cylinder(), {}),
        (cases.no_body(), {}),
    ]


def _make(prob, precision, opts):
    s = ib.create_solver(prob, precision=precision, box_update="off", cg=CG, cg_tol_mode=1,
                         cg_precond=opts.get("precond", False))
    if "ws" in opts:
        s.set_warm_start(opts["ws"], net=opts.get("net"))
    return s


def _run_pair(problems, precision, warm=3, steps=9):
    """Two sets of the same solvers, each advanced `warm` steps alone (history, the
    preconditioner); then one set `steps` steps alone, the other batched (in two calls, so
    both graph parities start a call)."""
    alone = [_make(p, precision, o) for p, o in problems]
    batched = [_make(p, precision, o) for p, o in problems]
    for s in alone + batched:
        s.advance(warm)
    f_alone, it_alone = [], []
    for s in alone:
        st = s.advance(steps)
        f_alone.append(s.forces(st))
        it_alone.append(s.cg_iters(st))
    batch = ib.Batch(batched)
    f1 = batch.advance(steps // 2)
    it1 = batch.cg_iters()
    f2 = batch.advance(steps - steps // 2)
    it2 = batch.cg_iters()
    batch.close()
    f_b = np.concatenate([f1, f2])
    it_b = np.concatenate([it1, it2])
    return alone, batched, np.stack(f_alone, 1), np.stack(it_alone, 1), f_b, it_b


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_moving_bodies_equal_alone(precision):
    probs = _moving_problems()
    print("case batched moving bodies", precision, flush=True)
    alone, batched, f_a, it_a, f_b, it_b = _run_pair(probs, precision)
    print("CG iterations per step (alone):", it_a.mean(0), flush=True)
    assert np.array_equal(it_b, it_a)
    assert (it_a[:, :3] > 0).all()  # the moving-body problems ran their CG
    assert np.array_equal(f_b, f_a)
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert a.state.index == b.state.index
        for name in FIELDS:
            x, y = a.read(name), b.read(name)
            assert np.array_equal(x, y), (k, name, np.abs(x - y).max() if x.size else 0)
    for s in alone + batched:
        s.close()


def test_batched_learned_warm_start_equals_alone():
    """A small network trained on the pitching plate; the learned guess in the batch (kept
    or dropped on the device) gives the alone run's iterates bit for bit."""
    net = WarmStartNet(channels=64, layers=3, kernel=5, max_points=64, capacity=500,
                       max_batch=16, seed=5)
    rec = ib.create_solver(cases.pitching_plate(nx=64, ny=64, nlevel=3, k=0.4), box_update="off",
                           cg=CG, cg_tol_mode=1)
    rec.set_warm_start("extrapolate", net=net, record=True)
    rec.advance(80)
    rec.close()
    net.train(200, batch=16, lr=1e-3)
    probs = [(cases.pitching_plate(nx=64, ny=64, nlevel=3, k=0.6), {"ws": "learned", "net": net}),
             (cases.moving_cylinder(), {"ws": "learned", "net": net})]
    print("case batched learned warm start", flush=True)
    alone, batched, f_a, it_a, f_b, it_b = _run_pair(probs, "fp64", warm=4, steps=8)
    print("CG iterations per step:", it_a.mean(0), flush=True)
    assert np.array_equal(it_b, it_a)
    assert np.array_equal(f_b, f_a)
    for a, b in zip(alone, batched):
        for name in FIELDS:
            assert np.array_equal(a.read(name), b.read(name)), name


def test_batched_moving_grid_and_gusts_match_oracle():
    """The moving grid and the time-varying freestream batched against the FP64 oracle."""
    probs = [cases.rotating_frame(nx=64, ny=64, nlevel=3), gusty_cylinder()]
    sol = [ib.create_solver(p) for p in probs]
    batch = ib.Batch(sol)
    f = batch.advance(6)
    batch.close()
    for k, p in enumerate(probs):
        orc = cases.oracle_for(p)
        fo = cases.oracle_advance(orc, p, 6)
        assert cases.relerr(f[:, k], fo) < 1e-9, k
        assert cases.relerr(sol[k].read("Gamma"), orc.field("Gamma")) < 1e-10, k


def test_batch_refusals():
    mc = cases.moving_cylinder(nx=128, ny=128, h=0.025, x0=(-1.6, -1.6))
    s = ib.create_solver(mc, box_update="force", cg=CG)
    with pytest.raises(ib._lib.IbpmError, match="box_update"):
        ib.Batch([s])
    s = ib.create_solver(cases.moving_cylinder(), box_update="off", cg=CG)
    s.set_warm_start("extrapolate")
    with pytest.raises(ib._lib.IbpmError, match="history"):
        ib.Batch([s])
    s = ib.create_solver(cases.moving_cylinder(), box_update="off", cg=CG, cg_precond=True)
    with pytest.raises(ib._lib.IbpmError, match="preconditioner"):
        ib.Batch([s])


def test_batch_reports_panels_leaving_the_region():
    """A body that drifts out of the grid fails the batched call with the reference's
    "outside allowed region" error (coupling.jl:39-41), naming the problem."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid((-1.0, -1.5), (60, 60), 0.05), 3)
    fluid = ib.Fluid(grid=grid, Re=100.0, freestream_vel=lambda t: (1.0, 0.0))
    pan = curves.panels(grid, curves.Circle(r=0.3, center=(0.0, 0.0)))
    motion = lambda t: ib.RigidBodyTransform(pos=(40.0 * t, 0.0), vel=(40.0, 0.0), angle=0.0,
                                             angular_vel=0.0)
    prob = ib.Problem(fluid=fluid, bodies=ib.Bodies(cyl=ib.MovingRigidBody(motion, pan)),
                      scheme=ib.CNAB(dt=0.004))
    s = ib.create_solver(prob, box_update="off", cg=CG)
    batch = ib.Batch([ib.create_solver(cases.no_body()), s])
    with pytest.raises(ib._lib.IbpmError, match="problem 1.*outside allowed region"):
        batch.advance(20)
    batch.close()


def test_multibody_ensemble_mixed_learned_batched():
    """The combined workload at reduced size: 4 problems of 2-4 plunging ellipses
    (cases.multibody_ensemble on 512 x 256 cells, 4 levels), mixed-precision sine
    contractions and the learned CG warm start, batched. Each problem ends bitwise equal
    to itself advanced alone, and its forces stay within 1e-6 of the FP64 path with the
    reference's plain CG from the previous step's solution."""
    probs = cases.multibody_ensemble(4, nx=512, ny=256, nlevel=4, h=1.0 / 32)
    cg = dict(maxiter=500, reltol=1e-10)
    net = WarmStartNet(channels=64, layers=4, kernel=5, max_points=256, capacity=2000,
                       max_batch=16, seed=3)
    print("case multibody ensemble 4 x 512x256: record", flush=True)
    for p in probs[:2]:
        s = ib.create_solver(p, box_update="off", cg=cg, cg_tol_mode=1)
        s.set_warm_start("extrapolate", net=net, record=True)
        s.advance(100)
        s.close()
    net.train(400, batch=16, lr=1e-3)

    def make(p):
        s = ib.create_solver(p, precision="mixed", box_update="off", cg=cg, cg_tol_mode=1)
        s.set_warm_start("extrapolate")
        return s

    alone = [make(p) for p in probs]
    batched = [make(p) for p in probs]
    warm, steps = 3, 12
    f_warm = [s.forces(s.advance(warm)) for s in alone]
    for s in batched:
        s.advance(warm)
    for s in alone + batched:
        s.set_warm_start("learned", net=net)
    f_alone, it_alone = [], []
    for s in alone:
        st = s.advance(steps)
        f_alone.append(s.forces(st))
        it_alone.append(s.cg_iters(st))
    print("case multibody ensemble: batched", flush=True)
    batch = ib.Batch(batched)
    f_b = batch.advance(steps)
    it_b = batch.cg_iters()
    batch.close()
    print("CG iterations per step (learned):", np.mean(it_alone, 1), flush=True)
    assert np.array_equal(f_b, np.stack(f_alone, 1))
    assert np.array_equal(it_b, np.stack(it_alone, 1))
    for a, b in zip(alone, batched):
        for name in ("Gamma", "psi", "Ftilde", "xb"):
            assert np.array_equal(a.read(name), b.read(name)), name
    for k, p in enumerate(probs):
        ref = ib.create_solver(p, cg=dict(maxiter=5000, reltol=1e-10), cg_tol_mode=1)
        fr = ref.forces(ref.advance(warm + steps))
        ref.close()
        got = np.concatenate([f_warm[k], f_alone[k]])
        e = np.linalg.norm(got - fr, axis=1) / np.linalg.norm(fr, axis=1)
        print(f"problem {k}: force rel err vs FP64 plain CG, max {e.max():.2e}", flush=True)
        assert e.max() <= 1e-6, (k, e)
