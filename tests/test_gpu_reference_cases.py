"""The reference's own solver tests (test/solver.jl) through the GPU path: the six
grid-motion x body combinations solved to t = 3 dt on its 150 x 150, 3-level grid (150 is
not divisible by 4: single-fold sine passes and the reference's truncated coarsify
ranges), checked against the oracle; an odd-sized grid; and the "outside allowed region"
error of a 15 x 20 grid."""
import numpy as np
import pytest

import cases
import immersedbodies as ib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("box", ["auto", "force"])
@pytest.mark.parametrize("motion", ["static grid", "moving grid"])
@pytest.mark.parametrize("bodies", ["single cyl", "double line", "moving cyl"])
def test_reference_solver_cases(motion, bodies, box):
    """box "force": the box updates (of the step, and of B for the moving cylinder) on
    these small grids too, the boxes following the moving cylinder."""
    prob = cases.reference_solver_problem(motion, bodies)
    print("case reference solver", motion, bodies, box, flush=True)
    state = ib.State(prob)
    solver = ib.create_solver(state, box_update=box)
    dt = ib.timestep(prob)
    forces = ib.solve_(None, solver, t=3 * dt)
    assert abs(state.t - 3 * dt) < 1e-12  # @test state.t ≈ 3 * timestep(prob)
    orc = cases.oracle_for(prob)
    f_ref = cases.oracle_advance(orc, prob, len(forces))
    moving = bodies == "moving cyl"
    assert cases.relerr(forces, f_ref) <= (1e-8 if moving else 1e-9)
    for name in ("Gamma", "psi", "q", "Ftilde"):
        assert cases.relerr(solver.read(name), orc.field(name)) <= (1e-9 if moving else 1e-10), name


@pytest.mark.parametrize("precision,box", [("fp64", "auto"), ("fp64", "force"), ("mixed", "auto")])
def test_odd_grid_matches_oracle(precision, box):
    """61 x 53 cells (odd lengths), 3 levels, a cylinder; FP64 also with the box updates
    forced (odd coarsify ranges)."""
    prob = cases.cylinder(nx=61, ny=53, nlevel=3, h=0.05, x0=(-1.5, -1.3), npts=40)
    print("case odd grid", precision, box, flush=True)
    solver = ib.create_solver(prob, precision=precision, box_update=box)
    orc = cases.oracle_for(prob)
    stats = solver.advance(4)
    f_ref = cases.oracle_advance(orc, prob, 4)
    assert cases.relerr(solver.forces(stats), f_ref) <= (1e-9 if precision == "fp64" else 1e-6)
    if precision == "fp64":
        for name in ("Gamma", "psi", "q"):
            assert cases.relerr(solver.read(name), orc.field(name)) <= 1e-10, name


def test_reference_outside_region_error():
    """@test_throws "outside allowed region" create_solver(state) (test/solver.jl:78)."""
    prob = cases.reference_outside_region_problem()
    with pytest.raises(RuntimeError, match="outside allowed region"):
        ib.create_solver(prob)
