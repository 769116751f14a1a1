"""Pins of the FP64 CPU oracle (no GPU).

The reference ships no golden vectors for this path (its solver tests assert only the
time, test/solver.jl:62), so every oracle function is pinned here by something the
mathematics fixes, chosen so that a plausible slip (a wrong boundary offset, a swapped
index, a sign, a missing scale) fails:
  DST-I            scipy.fft.dst(type=1) (pocketfft, independent of FFTW), explicit sums
  Poisson/Helmholtz dense solves of the assembled 5-point operator
  rot/curl         exact identities, differences of an extended vertex field
  get_bc/coarsify  exactness on linear fields in physical coordinates of the levels
  apply_bc         rot(curl(psi_full)) consistency of a field with boundary values
  nonlinear        extended-field form of the boundary branches; second-order
                   convergence to -u.grad(omega)
  delta_h          zeroth/first moments and symmetry of the Yang3 kernel
  E, E^T           exact adjointness; interpolation of linear fields
  step             irrotational uniform flow, y-symmetry (zero lift), CG vs dense B^-1
"""
import math

import numpy as np
import pytest
import scipy.fft as sf

import oracle as O
import cases


def grid(nx=16, ny=12, L=2, x0=-1.0, y0=-0.75, h=0.125):
    return O.Grid(nx, ny, L, x0, y0, h)


def as2d(g, a):
    """plane (Nx*Ny, x fastest) -> array [i, j] with i = x index (0-based I-1)."""
    return np.asarray(a).reshape(g.Ny, g.Nx).T


def flat(A):
    return np.ascontiguousarray(np.asarray(A).T.reshape(-1))


def K_apply(g, x):
    X = np.pad(as2d(g, x), 1)
    K = 4 * X[1:-1, 1:-1] - X[:-2, 1:-1] - X[2:, 1:-1] - X[1:-1, :-2] - X[1:-1, 2:]
    return flat(K)


def K_dense(g):
    n = g.nG
    M = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1
        M[:, k] = K_apply(g, e)
    return M


# ---------------------------------------------------------------- DST-I
@pytest.mark.parametrize("N", [1, 2, 7, 15, 31, 63, 199, 255, 383, 511])
def test_dst1d_matches_scipy(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((3, N))
    ref = sf.dst(x, type=1, axis=1)
    assert np.abs(O.dst1d(x) - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_dst1d_explicit_sum():
    N = 9
    rng = np.random.default_rng(1)
    x = rng.standard_normal(N)
    j = np.arange(1, N + 1)
    S = 2 * np.sin(np.pi * np.outer(j, j) / (N + 1))
    assert np.allclose(O.dst1d(x[None])[0], S @ x, rtol=0, atol=1e-13)


def test_dst2d_is_separable_and_involutive():
    g = grid()
    rng = np.random.default_rng(2)
    x = rng.standard_normal(g.nG)
    y = O.dst2d(g, x)
    X = as2d(g, x)
    ref = sf.dst(sf.dst(X, type=1, axis=0), type=1, axis=1)
    assert np.abs(as2d(g, y) - ref).max() < 1e-13 * np.abs(ref).max()
    z = O.dst2d(g, y) / (4 * g.nx * g.ny)  # S S = 2(N+1) I per dimension
    assert np.abs(z - x).max() < 1e-14 * 10


# ---------------------------------------------------------------- inverses
def test_lap_eigs_closed_form():
    g = grid()
    lam = as2d(g, O.lap_eigs(g))
    i = np.arange(1, g.Nx + 1)[:, None]
    j = np.arange(1, g.Ny + 1)[None, :]
    ref = 4 * np.sin(np.pi * i / (2 * g.nx)) ** 2 + 4 * np.sin(np.pi * j / (2 * g.ny)) ** 2
    assert np.abs(lam - ref).max() < 1e-13


def test_poisson_inverse_matches_dense_solve():
    g = grid()
    rng = np.random.default_rng(3)
    b = rng.standard_normal(g.nG)
    x = O.lap_inv(g, O.diag_poisson(g), b)
    ref = np.linalg.solve(K_dense(g), b)
    assert np.abs(x - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("lev", [1, 2, 3])
def test_A_and_Ainv_match_dense(lev):
    g = grid(L=3)
    Re, dt = 100.0, 0.004
    hc = 2.0 ** (lev - 1) * g.h
    c = dt / (2 * Re * hc * hc)
    rng = np.random.default_rng(4 + lev)
    x = rng.standard_normal(g.nG)
    A = np.eye(g.nG) - c * K_dense(g)
    Ainv = np.linalg.inv(np.eye(g.nG) + c * K_dense(g))
    yA = O.lap_inv(g, O.diag_A(g, Re, dt, lev, False), x)
    yI = O.lap_inv(g, O.diag_A(g, Re, dt, lev, True), x)
    assert np.abs(yA - A @ x).max() < 1e-13 * np.abs(x).max() * 10
    assert np.abs(yI - Ainv @ x).max() < 1e-13 * np.abs(x).max() * 10


# ---------------------------------------------------------------- rot / curl / bc
def vertex_coords(g, lev):
    """Physical coordinates of the vertices (0..nx, 0..ny) of level lev (subdomain,
    problem.jl:424-429)."""
    h = 2.0 ** (lev - 1) * g.h
    x0 = g.x0 + g.nx * (g.h - h) / 2
    y0 = g.y0 + g.ny * (g.h - h) / 2
    return x0 + h * np.arange(g.nx + 1), y0 + h * np.arange(g.ny + 1), h


def bc_from_boundary(g, F):
    """bc vector in the reference layout from a full vertex field F[i', j'] (0..nx, 0..ny):
    left = F[0, :], right = F[nx, :], bottom = F[:, 0], top = F[:, ny]."""
    return np.concatenate([F[0, :], F[g.nx, :], F[:, 0], F[:, g.ny]])


def test_rot_curl_is_K():
    g = grid()
    rng = np.random.default_rng(5)
    psi = rng.standard_normal(g.nG)
    q = O.curl(g, psi, np.zeros(g.nbc))
    assert np.abs(O.rot(g, q) - K_apply(g, psi)).max() < 1e-13


def test_curl_with_bc_is_difference_of_extended_field():
    g = grid()
    rng = np.random.default_rng(6)
    F = rng.standard_normal((g.nx + 1, g.ny + 1))
    q = O.curl(g, flat(F[1:-1, 1:-1]), bc_from_boundary(g, F))
    u = q[:g.nu].reshape(g.ny, g.nx + 1).T  # u[i-1, j-1], i = 1..nx+1, j = 1..ny
    v = q[g.nu:].reshape(g.ny + 1, g.nx).T  # v[i-1, j-1], i = 1..nx, j = 1..ny+1
    # u(i, j) = F(i-1, j) - F(i-1, j-1); v(i, j) = F(i-1, j-1) - F(i, j-1)
    assert np.abs(u - (F[:, 1:] - F[:, :-1])).max() < 1e-14
    assert np.abs(v - (F[:-1, :] - F[1:, :])).max() < 1e-14


def test_apply_bc_closes_the_poisson_problem():
    g = grid()
    rng = np.random.default_rng(7)
    F = rng.standard_normal((g.nx + 1, g.ny + 1))
    bc = bc_from_boundary(g, F)
    q = O.curl(g, flat(F[1:-1, 1:-1]), bc)
    Gam = O.rot(g, q)  # = 5-point K of the full field at interior vertices
    psi = O.lap_inv(g, O.diag_poisson(g), O.apply_bc(g, Gam, bc, 1.0))
    assert np.abs(psi - flat(F[1:-1, 1:-1])).max() < 1e-12


def test_get_bc_is_exact_on_linear_fields():
    g = grid(nx=16, ny=12, L=2)
    a, b, c = 0.3, -1.7, 0.9
    f = lambda x, y: a + b * x + c * y
    xc, yc, _ = vertex_coords(g, 2)
    coarse = flat(f(xc[1:-1, None], yc[None, 1:-1]))
    bc = O.get_bc(g, coarse)
    xf, yf, _ = vertex_coords(g, 1)
    F = f(xf[:, None], yf[None, :])
    assert np.abs(bc - bc_from_boundary(g, F)).max() < 1e-13


def test_coarsify_is_exact_on_linear_fields():
    g = grid(nx=16, ny=12, L=2)
    f = lambda x, y: 0.5 - 0.25 * x + 1.5 * y
    xf, yf, _ = vertex_coords(g, 1)
    fine = flat(f(xf[1:-1, None], yf[None, 1:-1]))
    sentinel = np.full(g.nG, 123.0)
    out = as2d(g, O.coarsify(g, sentinel, fine))
    xc, yc, _ = vertex_coords(g, 2)
    Ic = np.arange(1, g.nx)  # interior coarse vertex indices
    Jc = np.arange(1, g.ny)
    cen_i = (Ic >= g.nx // 4 + 1) & (Ic <= 3 * g.nx // 4 - 1)
    cen_j = (Jc >= g.ny // 4 + 1) & (Jc <= 3 * g.ny // 4 - 1)
    ref = 4 * f(xc[1:-1, None], yc[None, 1:-1])  # circulation on the coarse grid is 4x
    mask = cen_i[:, None] & cen_j[None, :]
    assert np.abs(out[mask] - ref[mask]).max() < 1e-13
    assert (out[~mask] == 123.0).all()


# ---------------------------------------------------------------- nonlinear
def _nonlinear_extended(g, q, q0, Ge, lev):
    """rot(fq)/hc^2 from the general (interior) formulas on the extended field Ge[i', j']."""
    nx, ny = g.nx, g.ny
    qq = q + q0
    u = qq[:g.nu].reshape(ny, nx + 1).T
    v = qq[g.nu:].reshape(ny + 1, nx).T
    Qx = np.zeros((nx + 2, ny + 2))  # 1-based Qx[i, j]
    Qy = np.zeros((nx + 2, ny + 3))
    for j in range(2, ny + 1):
        for i in range(1, nx + 2):
            Qx[i, j] = (u[i - 1, j - 1] + u[i - 1, j - 2]) / 2
    for j in range(1, ny + 2):
        for i in range(2, nx + 1):
            Qy[i, j] = (v[i - 1, j - 1] + v[i - 2, j - 1]) / 2
    fqx = np.zeros((nx + 2, ny + 2))
    fqy = np.zeros((nx + 2, ny + 2))
    for j in range(1, ny + 1):
        for i in range(2, nx + 1):
            fqx[i, j] = (Qy[i, j + 1] * Ge[i - 1, j] + Qy[i, j] * Ge[i - 1, j - 1]) / 2
    for j in range(2, ny + 1):
        for i in range(1, nx + 1):
            fqy[i, j] = -(Qx[i + 1, j] * Ge[i, j - 1] + Qx[i, j] * Ge[i - 1, j - 1]) / 2
    out = np.zeros((nx - 1, ny - 1))
    for J in range(1, ny):
        for I in range(1, nx):
            out[I - 1, J - 1] = (fqx[I + 1, J] - fqx[I + 1, J + 1]) + (fqy[I + 1, J + 1] - fqy[I, J + 1])
    hc = 2.0 ** (lev - 1) * g.h
    return flat(out) / hc ** 2


def test_nonlinear_boundary_branches_use_the_ring():
    g = grid(nx=12, ny=8)
    rng = np.random.default_rng(8)
    q = rng.standard_normal(g.nq)
    q0 = rng.standard_normal(g.nq)
    Ge = rng.standard_normal((g.nx + 1, g.ny + 1))
    got = O.nonlinear(g, q, q0, flat(Ge[1:-1, 1:-1]), bc_from_boundary(g, Ge), 2)
    ref = _nonlinear_extended(g, q, q0, Ge, 2)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_nonlinear_vanishes_without_vorticity():
    g = grid()
    rng = np.random.default_rng(9)
    got = O.nonlinear(g, rng.standard_normal(g.nq), rng.standard_normal(g.nq), np.zeros(g.nG),
                      np.zeros(g.nbc), 1)
    assert np.abs(got).max() == 0.0


def test_nonlinear_converges_to_advection():
    """Uniform flow (U, V) advecting omega = sin(x) sin(2y): nonlin -> -h^2 u.grad(omega)
    (the term is a circulation rate: Gamma = omega h^2)."""
    U, V = 0.7, -0.4
    errs = []
    for n in (32, 64):
        h = 2 * np.pi / n
        g = O.Grid(n, n, 1, 0.0, 0.0, h)
        xs = h * np.arange(n + 1)
        om = lambda x, y: np.sin(x) * np.sin(2 * y)
        Ge = om(xs[:, None], xs[None, :]) * h * h  # circulation
        q0 = np.concatenate([np.full(g.nu, U * h), np.full(g.nq - g.nu, V * h)])
        got = O.nonlinear(g, np.zeros(g.nq), q0, flat(Ge[1:-1, 1:-1]), bc_from_boundary(g, Ge), 1)
        X, Y = xs[1:-1, None], xs[None, 1:-1]
        ref = flat(-(U * np.cos(X) * np.sin(2 * Y) + V * np.sin(X) * 2 * np.cos(2 * Y)))
        errs.append(np.abs(got / (h * h) - ref).max())
    assert errs[1] < 0.3 * errs[0] and errs[1] < 2e-2, errs


# ---------------------------------------------------------------- delta_h, E, E^T
def test_delta_h_moments_and_symmetry():
    for r in np.linspace(-0.5, 0.5, 23):
        ks = np.arange(-5, 6)
        d = np.array([O.delta_h(k, r, 1.0) for k in ks])
        assert abs(d.sum() - 1.0) < 1e-8
        assert abs(((ks - r) * d).sum()) < 1e-8
        assert abs(O.delta_h(r, 0.0, 1.0) - O.delta_h(-r, 0.0, 1.0)) < 1e-15
    assert O.delta_h(3.0001, 0.0, 1.0) == 0.0
    # continuity across the branch points r = 1, 2, 3
    for r in (1.0, 2.0, 3.0):
        assert abs(O.delta_h(r - 1e-9, 0, 1) - O.delta_h(r + 1e-9, 0, 1)) < 1e-7


def _panels(g, n=23, r=0.45, c=(0.1, -0.05)):
    t = 2 * np.pi * np.arange(n) / n
    return np.stack([c[0] + r * np.cos(t), c[1] + r * np.sin(t)], axis=1)


def test_reg_and_regT_are_adjoint():
    g = grid(nx=32, ny=28, L=2, x0=-1.0, y0=-0.9, h=0.0625)
    xb = _panels(g)
    coords, w = O.reg_update(g, xb)
    rng = np.random.default_rng(10)
    q = rng.standard_normal(g.nq)
    f = rng.standard_normal(2 * xb.shape[0])
    lhs = O.regT(g, coords, w, q) @ f
    rhs = q @ O.reg(g, coords, w, f)
    assert abs(lhs - rhs) < 1e-13 * abs(lhs) * 10


def test_regT_interpolates_linear_fields():
    g = grid(nx=32, ny=28, L=2, x0=-1.0, y0=-0.9, h=0.0625)
    xb = _panels(g)
    coords, w = O.reg_update(g, xb)
    fu = lambda x, y: 0.3 + 1.2 * x - 0.7 * y
    fv = lambda x, y: -0.2 + 0.4 * x + 0.9 * y
    xs, ys, h = vertex_coords(g, 1)
    u = fu(xs[:, None], (ys[:-1] + h / 2)[None, :])  # u points: vertex x, cell-centre y
    v = fv((xs[:-1] + h / 2)[:, None], ys[None, :])  # v points: cell-centre x, vertex y
    q = np.concatenate([flat(u), flat(v)])
    fb = O.regT(g, coords, w, q)
    n = xb.shape[0]
    assert np.abs(fb[:n] - fu(xb[:, 0], xb[:, 1])).max() < 1e-7
    assert np.abs(fb[n:] - fv(xb[:, 0], xb[:, 1])).max() < 1e-7


def test_reg_update_rejects_edge_panels():
    g = grid(nx=16, ny=12, L=2, x0=-1.0, y0=-0.75, h=0.125)
    with pytest.raises(ValueError, match="outside allowed region"):
        O.reg_update(g, np.array([[-0.95, 0.0]]))


# ---------------------------------------------------------------- B and the step
def test_B_is_nearly_symmetric_positive_definite_and_Binv_is_its_inverse():
    prob = cases.cylinder(nx=32, ny=32, nlevel=2, h=0.0625, x0=(-1.0, -1.0), npts=16, r=0.4)
    orc = cases.oracle_for(prob)
    m = 2 * orc.n
    B = np.stack([orc.apply_B(np.eye(m)[k]) for k in range(m)], axis=1)
    asym = np.linalg.norm(B - B.T) / np.linalg.norm(B)
    assert asym < 0.2, asym
    assert np.linalg.eigvalsh((B + B.T) / 2).min() > 0
    assert np.abs(orc.Binv() @ B - np.eye(m)).max() < 1e-9


def test_cg_reaches_the_dense_solution():
    prob = cases.moving_cylinder(nx=40, ny=40, nlevel=2, h=0.075, x0=(-1.5, -1.5))
    orc = cases.oracle_for(prob, cg_reltol=1e-13)
    cases.oracle_advance(orc, prob, 1)
    m = 2 * orc.n
    B = np.stack([orc.apply_B(np.eye(m)[k]) for k in range(m)], axis=1)
    F = orc.field("Ftilde")
    Ff = B @ F
    # the right-hand side the CG solved is B F up to its stopping tolerance
    x = np.linalg.solve(B, Ff)
    assert np.abs(x - F).max() < 1e-9 * np.abs(F).max()
    assert orc.cg_iters() > 0


def test_uniform_flow_without_bodies_stays_irrotational():
    prob = cases.no_body()
    orc = cases.oracle_for(prob)
    cases.oracle_advance(orc, prob, 4)
    assert np.abs(orc.field("Gamma")).max() < 1e-13
    assert np.abs(orc.field("q")).max() < 1e-13


def test_symmetric_cylinder_has_no_lift():
    prob = cases.cylinder()
    orc = cases.oracle_for(prob)
    f = cases.oracle_advance(orc, prob, 5)
    assert (np.abs(f[:, 1]) < 1e-9 * np.abs(f[:, 0])).all()
    assert (f[:, 0] > 0).all()


def test_time_index_convention():
    prob = cases.cylinder(nx=32, ny=32, nlevel=2, h=0.1, x0=(-1.6, -1.6), npts=16)
    orc = cases.oracle_for(prob)
    assert orc.index == 0 and abs(orc.t + prob.scheme.dt) < 1e-15
    cases.oracle_advance(orc, prob, 4)  # solve!(; t = 3 dt) runs 4 steps
    assert orc.index == 4 and abs(orc.t - 3 * prob.scheme.dt) < 1e-15


# ---- moving grid (base_flux!, fluid.jl:34-67) ----

def _q0_after_one_step(prob):
    orc = cases.oracle_for(prob)
    cases.oracle_advance(orc, prob, 1)
    g = prob.fluid.grid
    nx, ny = g.base.n
    q0 = orc.field("q0").reshape(g.nlevel, -1)
    return q0, nx, ny


def test_moving_grid_translation_is_galilean():
    """A grid translating at u0 under freestream U is the static grid under U - u0."""
    import immersedbodies as ib

    mov = cases.rotating_frame(nx=48, ny=48, omega=0.0, vel=(0.3, -0.2), U=(1.0, 0.1))
    sta = cases.rotating_frame(nx=48, ny=48, omega=0.0, vel=(0.0, 0.0), U=(0.7, 0.3))
    sta = ib.Problem(fluid=ib.Fluid(grid=sta.fluid.grid, Re=sta.fluid.Re,
                                    freestream_vel=lambda t: (0.7, 0.3)),
                     bodies=sta.bodies, scheme=sta.scheme)
    a, b = cases.oracle_for(mov), cases.oracle_for(sta)
    fa = cases.oracle_advance(a, mov, 5)
    fb = cases.oracle_advance(b, sta, 5)
    assert cases.relerr(fa, fb) < 1e-13
    for name in ("Gamma", "psi", "q", "q0"):
        assert cases.relerr(a.field(name), b.field(name)) < 1e-13, name


def test_moving_grid_rotated_frame_sees_rotated_freestream():
    """Grid turned by theta (no motion): the base flux is the freestream's components in
    the grid's axes, h (U . e1, U . e2) with e1 = (cos, sin), e2 = (-sin, cos)."""
    import immersedbodies as ib

    th, U = 0.4, (1.0, 0.25)
    p = cases.rotating_frame(nx=32, ny=32, nlevel=2, omega=0.0, vel=(0.0, 0.0), U=U, body=False)
    p.fluid.grid_motion = ib.MovingGrid(lambda t: ib.GridVelocity(angle=th))
    q0, nx, ny = _q0_after_one_step(p)
    nu = (nx + 1) * ny
    for lev in range(2):
        hc = p.fluid.grid.base.h * 2 ** lev
        assert np.allclose(q0[lev, :nu], hc * (U[0] * np.cos(th) + U[1] * np.sin(th)), rtol=1e-14)
        assert np.allclose(q0[lev, nu:], hc * (-U[0] * np.sin(th) + U[1] * np.cos(th)), rtol=1e-14)


def test_moving_grid_spinning_frame_sees_counter_rotation():
    """Grid spinning at Omega about the origin, fluid at rest: in the grid's frame the fluid
    turns the other way, velocity (Omega y, -Omega x) at each face."""
    import immersedbodies as ib

    om = 0.7
    p = cases.rotating_frame(nx=32, ny=32, nlevel=2, omega=om, vel=(0.0, 0.0), U=(0.0, 0.0),
                             body=False)
    p.fluid.grid_motion = ib.MovingGrid(lambda t: ib.GridVelocity(angular_vel=om))
    q0, nx, ny = _q0_after_one_step(p)
    nu = (nx + 1) * ny
    for lev in range(2):
        sub = ib.subdomain(p.fluid.grid, lev + 1)
        hc = sub.h
        xu, yu = ib.coords(sub, "u")
        xv, yv = ib.coords(sub, "v")
        qx = q0[lev, :nu].reshape(ny, nx + 1)
        qy = q0[lev, nu:].reshape(ny + 1, nx)
        assert np.allclose(qx, hc * om * yu[:, None] * np.ones((1, nx + 1)), rtol=1e-12, atol=1e-15)
        assert np.allclose(qy, -hc * om * xv[None, :] * np.ones((ny + 1, 1)), rtol=1e-12, atol=1e-15)


def test_moving_grid_base_flux_is_divergence_free():
    """Rigid motion: the discrete divergence of the base flux vanishes in every cell."""
    p = cases.rotating_frame(nx=40, ny=36, nlevel=3, omega=0.5, vel=(0.2, 0.1), U=(1.0, -0.3),
                             body=False)
    q0, nx, ny = _q0_after_one_step(p)
    nu = (nx + 1) * ny
    for lev in range(3):
        qx = q0[lev, :nu].reshape(ny, nx + 1)
        qy = q0[lev, nu:].reshape(ny + 1, nx)
        div = (qx[:, 1:] - qx[:, :-1]) + (qy[1:, :] - qy[:-1, :])
        assert np.abs(div).max() < 1e-13 * np.abs(q0[lev]).max()
