"""No-slip after every GPU step (couple_surface_rigid + project_circ, coupling.jl:217-251).

The projection makes the flux interpolated to the panels equal the panel velocity at the
new time: E(q_1 + q0_1) = ub h, to rounding with the dense B^-1 of static bodies and to
the CG tolerance for moving ones. E is evaluated on the host by the oracle's regT from the
GPU's own q, q0, xb and ub, so the check holds the HIP path to the reference's coupling
independently of the oracle's trajectory.
"""
import numpy as np
import pytest

import cases
import immersedbodies as ib
import oracle as O
from test_oracle_pins import no_slip_residual

pytestmark = pytest.mark.gpu


def _grid(prob):
    fg = prob.fluid.grid
    return O.Grid(fg.base.n[0], fg.base.n[1], fg.nlevel, fg.base.x0[0], fg.base.x0[1],
                  fg.base.h)


def _errors(prob, nsteps, **kw):
    solver = ib.create_solver(prob, **kw)
    g = _grid(prob)
    errs = []
    for _ in range(nsteps):
        solver.advance(1)
        errs.append(no_slip_residual(g, solver.read("q")[:g.nq], solver.read("q0")[:g.nq],
                                     solver.read("xb"), solver.read("ub")))
    solver.close()
    return errs


@pytest.mark.parametrize("box", ["auto", "force", "off"])
def test_no_slip_static_fp64(box):
    prob = cases.cylinder(nx=96, ny=96, nlevel=3, h=0.025, x0=(-1.2, -1.2), Re=100.0,
                          dt=0.002, r=0.4)
    print("case static cylinder 96^2", box, flush=True)
    errs = _errors(prob, 4, box_update=box)
    print("no-slip", errs)
    assert max(errs) < 1e-10, errs


def test_no_slip_static_mixed():
    prob = cases.cylinder(nx=96, ny=96, nlevel=3, h=0.025, x0=(-1.2, -1.2), Re=100.0,
                          dt=0.002, r=0.4)
    print("case static cylinder 96^2 mixed", flush=True)
    errs = _errors(prob, 4, precision="mixed", refine_sweeps=2)
    print("no-slip", errs)
    assert max(errs) < 1e-7, errs


@pytest.mark.parametrize("box", ["auto", "force"])
def test_no_slip_moving_cg(box):
    prob = cases.moving_cylinder()
    print("case moving cylinder", box, flush=True)
    errs = _errors(prob, 4, box_update=box, cg={"reltol": 1e-12})
    print("no-slip", errs)
    assert max(errs) < 1e-9, errs


def test_no_slip_plunging_ellipses_preconditioned():
    prob = cases.plunging_ellipses(nx=256, ny=128, nlevel=3, h=1.0 / 16)
    print("case plunging ellipses 256x128 precond", flush=True)
    errs = _errors(prob, 3, cg={"reltol": 1e-10}, cg_precond=True)
    print("no-slip", errs)
    assert max(errs) < 1e-7, errs
