"""Host-side tests (no GPU): the C-ABI library loads and exports what include/ibpm.h
declares, and the Problem mirror behaves like the reference's problem layer
(test/problem.jl of the reference)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import immersedbodies as ib
from immersedbodies import _lib, curves

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ibpm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ibpm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert syms, "no declarations found"
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.EXPORTS)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libibpm.so")
    with pytest.raises(ImportError):
        _lib.load()


def test_grid_coordinates_and_levels():
    # test/problem.jl:4-21
    grid = ib.fluid_grid(h=0.1, xlims=(-2, 2), ylims=(0, 1), nlevel=3)
    base = grid.base
    assert ib.gridstep(grid) == ib.gridstep(base) == 0.1
    assert ib.subdomain(grid, 1) == base
    xs, ys = ib.coords(ib.subdomain(grid, 2))
    assert np.allclose(xs, np.arange(-4, 4.0001, 0.2)) and np.allclose(ys, np.arange(-0.5, 1.5001, 0.2))
    xs, ys = ib.coords(base, "u")
    assert np.allclose(ys, np.arange(0.05, 0.96, 0.1))
    xs, ys = ib.coords(base, "v")
    assert np.allclose(xs, np.arange(-1.95, 1.96, 0.1))
    xs, ys = ib.coords(base, "gamma")
    assert np.allclose(xs, np.arange(-1.9, 1.91, 0.1)) and np.allclose(ys, np.arange(0.1, 0.91, 0.1))


def test_default_gridstep_and_scheme():
    # test/problem.jl:23-27, :40-47
    fluid = ib.Fluid(grid=ib.fluid_grid(xlims=(0, 4), ylims=(0, 4), nlevel=1), Re=123.0)
    assert ib.gridstep(fluid.grid) == pytest.approx(ib.default_gridstep(123.0))
    grid = ib.fluid_grid(h=0.1, xlims=(0, 1), ylims=(0, 1), nlevel=2)
    sch = ib.default_scheme(grid, Umax=1.0, cfl=0.2)
    assert 1.0 * sch.dt / ib.gridstep(grid) <= 0.2
    with pytest.raises(ValueError):
        ib.CNAB(dt=0.1, n=3)


def test_panel_ranges():
    # test/problem.jl:29-38
    line1 = ib.StaticBody(ib.Panels(np.stack([np.arange(0, 1.01, 0.1), np.zeros(11)], 1), np.full(11, 0.1)))
    line2 = ib.StaticBody(ib.Panels(np.stack([np.zeros(21), np.arange(0, 2.01, 0.1)], 1), np.full(21, 0.1)))
    bodies = ib.Bodies(line1=line1, line2=line2)
    assert bodies.panel_range("line1") == range(0, 11)
    assert bodies.panel_range("line2") == range(11, 32)
    assert bodies["line1"] is line1 and bodies.static


def test_curves_partition():
    grid = ib.fluid_grid(h=0.02, xlims=(-1, 3), ylims=(-2, 2), nlevel=3)
    p = curves.panels(grid, curves.Circle(r=0.5))
    assert p.n == round(math.pi / 0.04)
    seg = curves.partition(curves.LineSegment((-0.5, 0), (0.5, 0)), 200)
    assert seg.n == 200 and np.allclose(seg.xb[[0, -1], 0], [-0.5, 0.5])
    e = curves.partition(curves.ellipse(0.5, 0.2, angle=0.3), 64)
    r = np.hypot(*e.xb.T)
    assert r.max() <= 0.5 + 1e-12 and r.min() >= 0.2 - 1e-3


def test_state_time_index():
    prob = ib.Problem(fluid=ib.Fluid(grid=ib.fluid_grid(h=0.1, xlims=(0, 1.6), ylims=(0, 1.6), nlevel=2), Re=10.0),
                      bodies=ib.Bodies(), scheme=ib.CNAB(dt=0.01))
    st = ib.State(prob)
    assert st.index == 0 and st.t == pytest.approx(-0.01)
    st.inc_time_index()
    assert st.t == 0.0


def test_bodies_accessors():
    """test/problem.jl:30-40: two lines, their panel ranges (0-based here) and no FSI."""
    line1 = ib.StaticBody(ib.Panels(np.stack([np.arange(11) * 0.1, np.zeros(11)], 1), np.full(11, 0.1)))
    line2 = ib.StaticBody(ib.Panels(np.stack([np.zeros(21), np.arange(21) * 0.1], 1), np.full(21, 0.1)))
    bodies = ib.Bodies(line1=line1, line2=line2)
    assert not ib.any_fsi(bodies)
    assert ib.panel_range(bodies, "line1") == range(0, 11)
    assert ib.panel_range(bodies, "line2") == range(11, 32)
    assert bodies["line1"] is line1 and ib.eachbody(bodies)["line2"] is line2


def test_problem_with_moving_grid():
    """test/problem.jl:52-70: a Problem on a MovingGrid with a circle and a segment
    discretised by Panels(grid, curve)."""
    from immersedbodies import curves

    fluid = ib.Fluid(grid=ib.fluid_grid(h=0.2, xlims=(-1.0, 2.0), ylims=(-1.5, 1.5), nlevel=5),
                     Re=10.0, freestream_vel=lambda t: (1.0, 0.0),
                     grid_motion=ib.MovingGrid(lambda t: ib.GridVelocity(
                         center=(0.0, 1.0), vel=(-1.0, 0.0), angle=t, angular_vel=1.0)))
    bodies = ib.Bodies(
        cyl=ib.StaticBody(curves.panels(fluid.grid, curves.Circle(r=0.5, center=(0.0, 0.0)))),
        line=ib.StaticBody(curves.panels(fluid.grid, curves.LineSegment((1.0, -0.3), (1.0, 0.3)))))
    prob = ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=0.1))
    assert ib.gridstep(prob) == 0.2
    assert ib.timestep(prob) == 0.1
    assert prob.fluid.grid_motion(0.5).angle == 0.5
