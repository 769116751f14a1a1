"""Preconditioned CG of moving bodies (cg_precond): the inverse of B at an earlier body
configuration, applied in each body's frame. CG keeps the reference's stopping test on
the true residual, so forces match the oracle's unpreconditioned cg! to the CG
tolerance, in far fewer iterations."""
import numpy as np
import pytest

import cases
import immersedbodies as ib

pytestmark = pytest.mark.gpu


def test_precond_moving_cylinder_matches_oracle():
    """Translating and turning cylinder: forces and fields against the oracle."""
    prob = cases.moving_cylinder()
    print("case moving_cylinder cg_precond", flush=True)
    s = ib.create_solver(prob, cg_precond=True)
    orc = cases.oracle_for(prob)
    st = s.advance(6)
    f_ref = cases.oracle_advance(orc, prob, 6)
    assert cases.relerr(s.forces(st), f_ref) <= 1e-8
    assert list(s.pc_assembled(st)) == [1, 0, 0, 0, 0, 0]
    assert all(x.cg_converged for x in st)
    it = s.cg_iters(st)
    print("pcg iterations", it)
    assert it[0] <= 2  # the assembly step: P is B's exact inverse
    for name in ("Gamma", "psi", "Ftilde"):
        assert cases.relerr(s.read(name), orc.field(name)) <= 1e-9, name


def test_precond_plunging_ellipses_fewer_iterations():
    """Three plunging ellipses (256 x 128, 3 levels): forces against the oracle, and the
    iterations against the unpreconditioned CG of the same solver settings."""
    prob = cases.plunging_ellipses(nx=256, ny=128, nlevel=3, h=1.0 / 16)
    print("case plunging_ellipses 256x128 cg_precond", flush=True)
    cg = dict(maxiter=2000, reltol=1e-10)
    plain = ib.create_solver(prob, cg=cg, cg_tol_mode=1)
    pre = ib.create_solver(prob, cg=cg, cg_tol_mode=1, cg_precond=True)
    orc = cases.oracle_for(prob)
    n = 30
    sp, sq = plain.advance(n), pre.advance(n)
    f_ref = cases.oracle_advance(orc, prob, 5)
    assert cases.relerr(pre.forces(sq)[:5], f_ref) <= 1e-7
    assert cases.relerr(pre.forces(sq), plain.forces(sp)) <= 1e-7
    ip, iq = plain.cg_iters(sp)[1:], pre.cg_iters(sq)[1:]
    print("mean iterations plain / preconditioned", ip.mean(), iq.mean())
    assert iq.mean() * 4 < ip.mean()


def test_precond_pitching_plate_rotation():
    """A pitching plate turns up to 10 degrees away from its assembly angle: the
    preconditioner is applied in the plate's frame and stays effective."""
    prob = cases.pitching_plate(nx=128, ny=128, nlevel=3, h=1.0 / 32, x0=(-2.0, -2.0), npts=33,
                                k=2.0)
    print("case pitching_plate cg_precond", flush=True)
    cg = dict(maxiter=2000, reltol=1e-10)
    plain = ib.create_solver(prob, cg=cg, cg_tol_mode=1)
    pre = ib.create_solver(prob, cg=cg, cg_tol_mode=1, cg_precond=True)
    n = 120  # a quarter of the pitching period at k = 2, dt = 0.002 covers 10 degrees
    sp, sq = plain.advance(n), pre.advance(n)
    assert all(x.cg_converged for x in sq)
    assert cases.relerr(pre.forces(sq), plain.forces(sp)) <= 1e-7
    ip, iq = plain.cg_iters(sp)[1:], pre.cg_iters(sq)[1:]
    print("mean / max iterations plain", ip.mean(), ip.max(), "preconditioned", iq.mean(), iq.max())
    assert iq.mean() * 3 < ip.mean()


def test_precond_refresh():
    """cg_precond_refresh = 1 reassembles after every step that needed more than one
    iteration; the step after an assembly converges at once."""
    prob = cases.moving_cylinder()
    s = ib.create_solver(prob, cg_precond=True, cg_precond_refresh=1)
    st = s.advance(4)
    pa, it = s.pc_assembled(st), s.cg_iters(st)
    print("assembled", pa, "iterations", it)
    assert pa[0] == 1
    for k in range(1, 4):
        assert pa[k] == (1 if it[k - 1] > 1 else 0)


def test_precond_mixed_precision_moving_cylinder():
    """Mixed precision with a moving body and the preconditioned CG: forces within the
    mixed path's 1e-6 of the oracle."""
    prob = cases.moving_cylinder()
    print("case moving_cylinder mixed cg_precond", flush=True)
    s = ib.create_solver(prob, precision="mixed", cg_precond=True)
    orc = cases.oracle_for(prob)
    st = s.advance(6)
    f_ref = cases.oracle_advance(orc, prob, 6)
    assert all(x.cg_converged for x in st)
    assert cases.relerr(s.forces(st), f_ref) <= 1e-6
