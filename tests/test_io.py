"""Checkpoint format of io.jl:1-71 (save_state / load_state!)."""
import io

import numpy as np
import pytest

import cases
import immersedbodies as ib
from immersedbodies.io import STATE_FILE_HEADER, load_state, load_state_, save_state


def _random_state(prob, seed=0):
    rng = np.random.default_rng(seed)
    st = ib.State(prob, t0=0.25)
    st.set_time_index(7)
    for name in ("q", "q0", "Gamma", "psi"):
        setattr(st, name, rng.standard_normal(getattr(st, name).shape))
    st.nonlin = [rng.standard_normal(a.shape) for a in st.nonlin]
    n = st.xb.shape[0]
    st.xb = rng.standard_normal((n, 2))
    st.ds = rng.standard_normal(n)
    st.ub = rng.standard_normal((n, 2))
    st.Ftilde = rng.standard_normal(2 * n)
    st.freestream_vel = rng.standard_normal(2)
    return st


def test_round_trip_is_exact():
    # test/io.jl:28-49
    prob = cases.cylinder(nx=16, ny=16, nlevel=2, h=0.2, x0=(-1.6, -1.6), npts=8, r=0.5)
    st = _random_state(prob)
    buf = io.BytesIO()
    save_state(buf, st)
    buf.seek(0)
    st2 = load_state(buf, prob)
    assert st2.index == 7 and st2.t0 == 0.25 and st2.t == st.t
    for name in ("q", "q0", "Gamma", "psi", "xb", "ds", "ub", "Ftilde", "freestream_vel"):
        assert np.array_equal(getattr(st, name), getattr(st2, name)), name
    for a, b in zip(st.nonlin, st2.nonlin):
        assert np.array_equal(a, b)
    h, dt = ib.gridstep(prob), ib.timestep(prob)
    assert np.allclose(st2.fb, st.Ftilde.reshape(2, -1).T * h / dt, rtol=0, atol=0)


def test_byte_layout():
    prob = cases.cylinder(nx=16, ny=16, nlevel=2, h=0.2, x0=(-1.6, -1.6), npts=8, r=0.5)
    st = _random_state(prob, 1)
    buf = io.BytesIO()
    save_state(buf, st)
    b = buf.getvalue()
    H = len(STATE_FILE_HEADER)
    assert b[:H] == b"@ImmersedBodies.jl State v1@"
    assert int(np.frombuffer(b[H:H + 8], "<i8")[0]) == 7
    assert float(np.frombuffer(b[H + 8:H + 16], "<f8")[0]) == 0.25
    g = prob.fluid.grid
    L, nq, nG, n = g.nlevel, g.inds.nq, g.inds.nG, 8
    assert len(b) == H + 16 + 8 * (2 * L * nq + 4 * L * nG + 2 * n + n + 2 * n + 2 * n + 2)
    # q is column-major nq x nlevel: level 2 starts after nq values
    q = np.frombuffer(b[H + 16:H + 16 + 8 * L * nq], "<f8")
    assert np.array_equal(q[nq:2 * nq], st.q[1])
    # xb is column-major npanel x 2: all x first
    off = H + 16 + 8 * (2 * L * nq + 4 * L * nG)
    xb = np.frombuffer(b[off:off + 16 * n], "<f8")
    assert np.array_equal(xb[:n], st.xb[:, 0])


def test_missing_header_is_an_error():
    prob = cases.cylinder(nx=16, ny=16, nlevel=2, h=0.2, x0=(-1.6, -1.6), npts=8, r=0.5)
    with pytest.raises(ValueError, match="missing header"):
        load_state_(io.BytesIO(b"garbage" * 20), ib.State(prob))


@pytest.mark.gpu
def test_gpu_resume_matches_uninterrupted_run():
    prob = cases.cylinder()
    a = ib.create_solver(prob)
    fa = a.forces(a.advance(6))
    b = ib.create_solver(prob)
    b.advance(3)
    buf = io.BytesIO()
    save_state(buf, b.getstate())
    b.close()
    buf.seek(0)
    st = load_state(buf, prob)
    c = ib.create_solver(st)
    fc = c.forces(c.advance(3))
    assert np.array_equal(fa[3:], fc)
    assert np.array_equal(a.read("Gamma"), c.read("Gamma"))


@pytest.mark.gpu
def test_gpu_solve_with_saver(tmp_path):
    """solve_ with SaveNpz on the GPU solver: the host state is refreshed from the device
    at the saved steps (also inside a chunk of steps), so the saved vorticity and forces
    are the solver's."""
    prob = cases.cylinder()
    solver = ib.create_solver(prob)
    filename = str(tmp_path / "run.npz")
    save = ib.SaveNpz(filename, ("prob", prob),
                      ("fluid", ib.fluid_group(prob, ib.TimestepRange(start=2, step=2), [ib.vorticity_])),
                      ("totals", ib.body_group(prob, ib.TimestepRange(step=1), [ib.boundary_total_force_])))
    dt = ib.timestep(prob)
    forces = ib.solve_(None, solver, t=5 * dt, save=save, chunk=4)  # 6 steps from t = -dt
    assert solver.state.index == 6
    f = np.load(filename, allow_pickle=False)
    assert np.array_equal(f["fluid/_timestep_index"], [2, 4, 6])
    L = prob.fluid.grid.nlevel
    g = solver.read("Gamma").reshape(L, -1)
    w = f["fluid/vorticity"][..., 2]
    for lev in range(L):
        assert np.allclose(w[:, :, lev].ravel(order="F"), g[lev] / ib.gridstep(prob, lev + 1) ** 2,
                           rtol=1e-14, atol=0)
    # the host sums fb in another order than the GPU's force kernel: round-off of the drag
    scale = np.abs(forces).max()
    assert np.allclose(f["totals/cyl/boundary_total_force"].T, forces, rtol=0, atol=1e-13 * scale)
