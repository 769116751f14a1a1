"""Solution savers and state accessors: the assertions of the reference's test/io.jl
"HDF5" testset, run against SaveNpz (the same layout in a .npz file; HDF5 is not
installed here), with a dummy solver that only advances the time index."""
import numpy as np
import pytest

import immersedbodies as ib
from immersedbodies import curves


def dummy_problem():
    """test/io.jl:1-13."""
    fluid = ib.Fluid(grid=ib.fluid_grid(h=0.5, xlims=(-1.0, 1.0), ylims=(-1.0, 1.0), nlevel=2),
                     Re=100.0, freestream_vel=lambda t: (1.0, 0.0))
    bodies = ib.Bodies(cyl=ib.StaticBody(curves.panels(fluid.grid, curves.Circle(r=0.5))))
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=1.0))


class TestSolver(ib.AbstractSolver):
    """A minimal solver that only advances the time index (test/io.jl:22-26)."""

    __test__ = False

    def __init__(self, state):
        self.state = state

    def advance(self):
        self.state.inc_time_index()

    def getstate(self):
        return self.state


PROB = dummy_problem()
DT = ib.timestep(PROB)


def _save(filename):
    def setx0(x, st):
        x[()] = st.index

    def setx1(x, st):
        x[0] = st.index
        x[1] = 1.23

    def setx2(x, st):
        x[...] = st.index * np.arange(1, 7).reshape((2, 3), order="F")

    def sety(y, st):
        y[...] = st.index

    def setz(z, st):
        z[()] = 125

    ycoords = [(np.linspace(0.0, 0.5, 6), np.linspace(0.0, 0.1, 2)),
               (np.linspace(0.0, 1.0, 6), np.linspace(0.0, 0.2, 2)),
               (np.linspace(0.0, 2.0, 6), np.linspace(0.0, 0.4, 2))]
    return ib.SaveNpz(
        filename,
        ("prob", PROB),
        ("foo/bar", ib.SolutionValues(ib.TimestepRange(step=5),
                                      ("x0", ib.ArrayValue(setx0, np.int64, ())),
                                      ("x1", ib.ArrayValue(setx1, np.float64, (2,))))),
        ("a", ib.SolutionValues(ib.TimestepRange(start=4, step=2, length=3),
                                ("x2", ib.ArrayValue(setx2, np.float64, (2, 3))),
                                ("y", ib.MultiDomainValue(sety, ycoords)))),
        ("b", ib.SolutionValues(ib.TimestepRange(start=20, step=1, length=5),
                                ("z0", ib.ArrayValue(setz, np.int64, ())))),
    )


def _check(filename, n):
    f = np.load(filename, allow_pickle=False)
    inds = np.arange(1, n + 1, 5)
    assert np.allclose(f["foo/bar/_timestep_time"], DT * (inds - 1))
    assert np.array_equal(f["foo/bar/_timestep_index"], inds)
    assert np.array_equal(f["foo/bar/x0"], inds)
    assert np.allclose(f["foo/bar/x1"], np.stack([inds, np.full(inds.size, 1.23)]))
    inds = np.arange(4, min(8, n) + 1, 2)
    assert np.allclose(f["a/_timestep_time"], DT * (inds - 1))
    assert np.array_equal(f["a/_timestep_index"], inds)
    base = np.arange(1, 7).reshape((2, 3), order="F")
    assert np.allclose(f["a/x2"], np.stack([i * base for i in inds], axis=-1))
    assert np.allclose(f["a/y"], np.broadcast_to(inds, (6, 2, 3, inds.size)))
    inds = np.arange(20, n + 1)
    assert np.allclose(f["b/_timestep_time"], DT * (inds - 1))
    assert np.array_equal(f["b/_timestep_index"], inds)
    assert np.array_equal(f["b/z0"], np.full(inds.size, 125))
    c = f["a/y@coords"]
    assert c.shape == (2, 3)
    assert np.allclose(c["start"], 0.0)
    assert np.allclose(c["step"], [[0.1, 0.2, 0.4], [0.1, 0.2, 0.4]])
    assert np.array_equal(c["length"], [[6, 6, 6], [2, 2, 2]])


def test_save_extends_and_overwrites(tmp_path):
    """Solving over new timesteps extends the file, solving over old ones overwrites them,
    and entries the saver does not own are kept (test/io.jl:121-166)."""
    filename = str(tmp_path / "soln.npz")
    np.savez(filename, leave_me_alone=np.array(True))
    state = ib.State(PROB)
    for i, n, nmax in [(0, 6, 6), (2, 22, 22), (5, 10, 22)]:
        state.index = i
        state.t = DT * (i - 1)
        ib.solve_(None, TestSolver(state), t=(n - 1) * DT, save=_save(filename))
        _check(filename, nmax)
    f = np.load(filename, allow_pickle=False)
    assert bool(f["leave_me_alone"]) is True
    g = PROB.fluid.grid
    rec = f["prob/fluid"]
    assert rec["grid"]["h"] == g.base.h and rec["grid"]["nlevel"] == g.nlevel
    assert (rec["grid"]["xlims"]["min"], rec["grid"]["xlims"]["max"]) == ib.extents(g.base)[0]
    assert (rec["grid"]["ylims"]["min"], rec["grid"]["ylims"]["max"]) == ib.extents(g.base)[1]
    assert rec["Re"] == PROB.fluid.Re
    body = PROB.bodies["cyl"]
    assert str(f["prob/bodies/cyl@type"]) == "StaticBody"
    assert np.array_equal(f["prob/bodies/cyl/positions"], body.panels.xb)
    assert np.array_equal(f["prob/bodies/cyl/lengths"], body.panels.ds)
    assert str(f["prob/scheme@type"]) == "CNAB"
    assert f["prob/scheme"]["dt"] == PROB.scheme.dt


def test_save_rejects_a_changed_problem(tmp_path):
    filename = str(tmp_path / "soln.npz")
    state = ib.State(PROB)
    ib.solve_(None, TestSolver(state), t=2 * DT, save=ib.SaveNpz(filename, ("prob", PROB)))
    other = ib.Problem(fluid=PROB.fluid, bodies=PROB.bodies, scheme=ib.CNAB(dt=0.5))
    with pytest.raises(ValueError, match="incorrect value"):
        ib.solve_(None, TestSolver(ib.State(other)), t=1.0, save=ib.SaveNpz(filename, ("prob", other)))


def test_save_hdf5_needs_hdf5(tmp_path):
    state = ib.State(PROB)
    with pytest.raises(RuntimeError, match="h5py"):
        ib.solve_(None, TestSolver(state), t=2 * DT, save=ib.SaveHDF5(str(tmp_path / "x.h5"), ("prob", PROB)))


def test_accessors_and_groups(tmp_path):
    """x_velocity, y_velocity, vorticity and the boundary accessors against the state
    arrays they restate (problem.jl:710-866), and fluid_group / body_group saving them."""
    prob = ib.Problem(fluid=ib.Fluid(grid=ib.fluid_grid(h=0.1, xlims=(-1.0, 1.0), ylims=(-0.8, 0.8), nlevel=2),
                                     Re=100.0, freestream_vel=lambda t: (1.0, 0.0)),
                      bodies=PROB.bodies, scheme=ib.CNAB(dt=0.01))
    st = ib.State(prob)
    g = prob.fluid.grid
    nx, ny = g.base.n
    rng = np.random.default_rng(3)
    st.q = rng.standard_normal(st.q.shape)
    st.q0 = rng.standard_normal(st.q0.shape)
    st.Gamma = rng.standard_normal(st.Gamma.shape)
    st.fb = rng.standard_normal(st.fb.shape)
    u, v, w = ib.x_velocity(st), ib.y_velocity(st), ib.vorticity(st)
    assert u.shape == (nx + 1, ny, 2) and v.shape == (nx, ny + 1, 2) and w.shape == (nx - 1, ny - 1, 2)
    nu = (nx + 1) * ny
    for lev in (1, 2):
        hc = ib.gridstep(g, lev)
        qq = st.q[lev - 1] + st.q0[lev - 1]
        assert np.allclose(u[:, :, lev - 1].ravel(order="F"), qq[:nu] / hc)
        assert np.allclose(v[:, :, lev - 1].ravel(order="F"), qq[nu:] / hc)
        assert np.allclose(w[:, :, lev - 1].ravel(order="F"), st.Gamma[lev - 1] / hc ** 2)
    # the i index runs along x: u[i, j] sits at coords(grid, x_velocity)[level] = (x[i], y[j])
    x, y = ib.coords(g, ib.x_velocity)[0]
    assert (len(x), len(y)) == u.shape[:2]
    assert ib.gridsize(g, "u") == (nx, ny - 1, 2)  # the reference's formula
    assert np.allclose(ib.boundary_total_force(st, "cyl"), st.fb.sum(axis=0))
    filename = str(tmp_path / "fields.npz")
    save = ib.SaveNpz(filename,
                      ("fluid", ib.fluid_group(prob, ib.TimestepRange(step=2), [ib.vorticity_, ib.x_velocity_])),
                      ("bodies", ib.body_group(prob, ib.TimestepRange(step=1), [ib.boundary_force_, ib.boundary_total_force_])))
    st.index, st.t = 0, -prob.scheme.dt
    ib.solve_(None, TestSolver(st), t=3 * prob.scheme.dt, save=save)
    f = np.load(filename, allow_pickle=False)
    assert f["fluid/vorticity"].shape == (nx - 1, ny - 1, 2, 2)
    assert np.allclose(f["fluid/vorticity"][..., 0], w)
    assert np.array_equal(f["fluid/_timestep_index"], [1, 3])
    assert f["bodies/cyl/boundary_force"].shape == (st.fb.shape[0], 2, 4)
    assert np.allclose(f["bodies/cyl/boundary_total_force"][:, 2], st.fb.sum(axis=0))
    assert f["fluid/x_velocity@coords"]["length"][0, 0] == nx + 1
