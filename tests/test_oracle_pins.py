"""Math checks that pin the oracle's multi-level and coupling compositions (no GPU).

The operators these compositions are built from (DST-I, the inverses, rot/curl, get_bc,
coarsify, apply_bc, the nonlinear term, delta_h, E/E^T, B) are pinned one by one in
test_oracle.py. What that leaves open is how the reference strings them together, and a
slip there (a scale factor, the level a boundary is read from, a sign) would pass every
operator check. Each check below is fixed by the mathematics of the composition:

  vort2flux (fluid.jl:175-206)
    closure    every level's psi closes its own Poisson problem with the boundary values
               of the next coarser psi: rot(curl(psi_l, get_bc(psi_l+1))) = Gamma_l, with
               Gamma_l the coarsify chain of the finer levels (zero boundary on level L);
               the fluxes are curl(psi_l, get_bc(psi_l+1))
    free space a compact vortex's streamfunction approaches the unbounded-domain solution
               -(G/4 pi)(ln r^2 + E1(r^2/s^2)) as levels are added (each level doubles the
               domain, so the truncation error falls about 4x per level)
  get_trial_state (fluid.jl:457-507)
    harmonic   a harmonic (linear) vorticity sampled as hc^2 omega on every level is a
               steady state of the Crank-Nicolson diffusion when each level's boundary
               comes from the next coarser level: one trial step at rest keeps Gamma_l on
               every level below L (fac 0.25 dt/(Re hc^2) of the boundary source)
    advection  uniform flow over that vorticity gives the exact nonlinear term
               -hc^2 U.grad(omega) on every vertex of the levels below L, the edges
               included (the coarser ring scaled by 0.25 into the level's units)
  couple_surface_rigid + project_circ (coupling.jl:217-251)
    no-slip    after every step the flux interpolated to the panels is the panel
               velocity: E(q_1 + q0_1) = ub h, to rounding with the dense B^-1 and to the
               CG tolerance for moving bodies

Each check is run on the oracle and on variants of it built with one planted mistake
(the IBO_MUTATION list in ibpm_oracle.c); the check must pass on the first and fail on
each variant it guards against.
"""
import numpy as np
import pytest
import scipy.special as ss

import cases
import oracle as O


def as2d(g, a):
    return np.asarray(a).reshape(g.Ny, g.Nx).T


def flat(A):
    return np.ascontiguousarray(np.asarray(A).T.reshape(-1))


def vertex_xy(g, lev):
    """Interior vertex coordinates of level ``lev`` (subdomain, problem.jl:424-429)."""
    h = 2.0 ** (lev - 1) * g.h
    x0 = g.x0 + g.nx * (g.h - h) / 2
    y0 = g.y0 + g.ny * (g.h - h) / 2
    x = x0 + h * np.arange(1, g.nx)
    y = y0 + h * np.arange(1, g.ny)
    return x[:, None], y[None, :], h


def centre_mask(g):
    """Interior vertices of a level that coarsify! overwrites from the finer level."""
    i = np.arange(1, g.nx)
    j = np.arange(1, g.ny)
    ci = (i >= g.nx // 4 + 1) & (i <= 3 * g.nx // 4 - 1)
    cj = (j >= g.ny // 4 + 1) & (j <= 3 * g.ny // 4 - 1)
    return ci[:, None] & cj[None, :]


# ------------------------------------------------------------------ the checks

def v2f_closure_error(g, G0, psi, q, Gafter):
    """Largest violation of the vort2flux telescoping, relative to max |Gamma|."""
    L = g.nlevel
    scale = np.abs(G0).max()
    err = 0.0
    for lev in range(1, L + 1):
        bc = O.get_bc(g, psi[lev]) if lev < L else np.zeros(g.nbc)
        ql = O.curl(g, psi[lev - 1], bc)
        err = max(err, np.abs(O.rot(g, ql) - Gafter[lev - 1]).max() / scale)
        err = max(err, np.abs(q[lev - 1] - ql).max() / max(np.abs(ql).max(), 1e-300))
        if lev >= 2:
            chain = O.coarsify(g, G0[lev - 1], Gafter[lev - 2])
            err = max(err, np.abs(Gafter[lev - 1] - chain).max() / scale)
    return err


def run_v2f(g, seed=11):
    rng = np.random.default_rng(seed)
    G0 = rng.standard_normal((g.nlevel, g.nG))
    psi, q, Gafter = O.vort2flux(g, G0)
    return G0, psi, q, Gafter


def lamb_oseen(x, y, c, s):
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2
    om = np.exp(-r2 / s ** 2) / (np.pi * s ** 2)
    with np.errstate(divide="ignore"):
        psi = -(np.log(r2) + ss.exp1(r2 / s ** 2)) / (4 * np.pi)
    # r -> 0 limit of ln r^2 + E1(r^2/s^2) is ln s^2 - Euler gamma
    psi = np.where(r2 < 1e-300, -(np.log(s ** 2) - np.euler_gamma) / (4 * np.pi), psi)
    return om, psi


def free_space_errors(levels=(1, 2, 3, 4), n=64, c=(0.35, -0.2), s=0.25):
    """Level-1 streamfunction of a unit Lamb-Oseen vortex against the free-space one, up
    to a constant, for each level count."""
    errs = []
    for L in levels:
        g = O.Grid(n, n, L, -2.0, -2.0, 4.0 / n)
        G = np.zeros((L, g.nG))
        for lev in range(1, L + 1):
            X, Y, h = vertex_xy(g, lev)
            G[lev - 1] = flat(lamb_oseen(X, Y, c, s)[0] * h * h)
        psi = O.vort2flux(g, G)[0]
        X, Y, _ = vertex_xy(g, 1)
        d = as2d(g, psi[0]) - lamb_oseen(X, Y, c, s)[1]
        errs.append(float(np.abs(d - d.mean()).max()))
    return errs


def linear_state(orc, g, coef=(0.4, 1.3, -0.8)):
    """Gamma_l = hc^2 (a + b x + c y) on every level, q = 0, nonlinear history 0."""
    a, b, c = coef
    G = np.zeros((g.nlevel, g.nG))
    for lev in range(1, g.nlevel + 1):
        X, Y, h = vertex_xy(g, lev)
        G[lev - 1] = flat(h * h * (a + b * X + c * Y))
    orc.set_field("Gamma", G.reshape(-1))
    orc.set_field("q", np.zeros(g.nq * g.nlevel))
    orc.set_field("nonlin1", np.zeros(g.nG * g.nlevel))
    orc.set_field("nonlin2", np.zeros(g.nG * g.nlevel))
    return G


def trial_problem(nx=32, ny=24, L=3, h=0.1, Re=50.0, dt=0.05):
    g = O.Grid(nx, ny, L, -1.6, -1.2, h)
    return g, O.OracleSolver(g, Re, dt, np.zeros((0, 2)), bodies=[])


def trial_harmonic_error(**kw):
    """One step at rest from a linear vorticity: change of Gamma on the levels below L
    (and in the centre of level L), relative to max |Gamma|."""
    g, orc = trial_problem(**kw)
    G0 = linear_state(orc, g)
    orc.advance((0.0, 0.0))
    G1 = orc.field("Gamma").reshape(g.nlevel, g.nG)
    L = g.nlevel
    err = np.abs(G1[:L - 1] - G0[:L - 1]).max()
    m = centre_mask(g)
    err = max(err, np.abs(as2d(g, G1[L - 1])[m] - as2d(g, G0[L - 1])[m]).max())
    orc.close()
    return err / np.abs(G0).max()


def trial_advection_error(U=(0.6, -0.35), coef=(0.4, 1.3, -0.8), **kw):
    """Nonlinear term of one step in uniform flow over a linear vorticity against
    -hc^2 U.grad(omega), on every vertex of the levels below L."""
    g, orc = trial_problem(**kw)
    linear_state(orc, g, coef)
    orc.advance(U)
    # _cycle! (fluid.jl:510) leaves the term just computed in nonlin[2]
    N = orc.field("nonlin2").reshape(g.nlevel, g.nG)
    err = 0.0
    for lev in range(1, g.nlevel):
        h = 2.0 ** (lev - 1) * g.h
        exact = -h * h * (U[0] * coef[1] + U[1] * coef[2])
        err = max(err, np.abs(N[lev - 1] - exact).max() / abs(exact))
    orc.close()
    return err


def no_slip_errors(prob, nsteps, **kw):
    """max over steps of |E(q_1 + q0_1) - ub h| / |E(q0_1)| after each step."""
    orc = cases.oracle_for(prob, **kw)
    fg = prob.fluid.grid
    g = O.Grid(fg.base.n[0], fg.base.n[1], fg.nlevel, fg.base.x0[0], fg.base.x0[1], fg.base.h)
    errs = []
    for _ in range(nsteps):
        cases.oracle_advance(orc, prob, 1)
        errs.append(no_slip_residual(g, orc.field("q")[:g.nq], orc.field("q0")[:g.nq],
                                     orc.field("xb"), orc.field("ub")))
    orc.close()
    return errs


def no_slip_residual(g, q1, q01, xb, ub):
    n = xb.size // 2
    coords, w = O.reg_update(g, np.asarray(xb).reshape(2, n).T)
    Eq = O.regT(g, coords, w, q1 + q01)
    ref = np.abs(O.regT(g, coords, w, q01)).max()
    return float(np.abs(Eq - np.asarray(ub) * g.h).max() / ref)


# ------------------------------------------------------------------ on the oracle

@pytest.mark.parametrize("shape", [(32, 24, 2), (32, 24, 3), (40, 32, 4), (30, 26, 3)])
def test_vort2flux_closes_every_level(shape):
    nx, ny, L = shape
    g = O.Grid(nx, ny, L, -1.0, -0.8, 0.0625)
    assert v2f_closure_error(g, *run_v2f(g)) < 1e-11


def test_vort2flux_approaches_free_space():
    errs = free_space_errors()
    # each added level doubles the domain: the truncation error falls ~4x per level
    # until the O(h^2) core error remains
    assert errs[1] < 0.35 * errs[0] and errs[2] < 0.35 * errs[1], errs
    assert errs[3] < 0.6 * errs[2] and errs[3] < 2e-3, errs


def test_trial_state_keeps_a_harmonic_vorticity():
    assert trial_harmonic_error() < 1e-13
    assert trial_harmonic_error(nx=40, ny=32, L=4, h=0.05, Re=20.0, dt=0.02) < 1e-13


def test_trial_state_advects_a_linear_vorticity_exactly():
    assert trial_advection_error() < 1e-12
    assert trial_advection_error(U=(-0.2, 0.9), coef=(1.0, -0.5, 0.7), nx=40, ny=32, L=4,
                                 h=0.05) < 1e-12


def test_no_slip_holds_after_every_step_static():
    prob = cases.cylinder(nx=48, ny=48, nlevel=3, h=0.05, x0=(-1.2, -1.2), Re=100.0,
                          dt=0.004, r=0.4)
    errs = no_slip_errors(prob, 4)
    assert max(errs) < 1e-10, errs


def test_no_slip_holds_after_every_step_moving():
    prob = cases.moving_cylinder(nx=48, ny=48, nlevel=2, h=0.06, x0=(-1.44, -1.44))
    errs = no_slip_errors(prob, 3, cg_reltol=1e-12)
    assert max(errs) < 1e-9, errs


# ------------------------------------------------------------------ planted mistakes

@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("oracle_variants"))
    return {k: O.build_variant(k, d) for k in range(1, 9)}


def test_real_oracle_has_no_planted_mistake(tmp_path):
    """The default library is the IBO_MUTATION 0 build."""
    path = O.build_variant(0, str(tmp_path))
    g = O.Grid(32, 24, 3, -1.0, -0.8, 0.0625)
    with O.variant(path):
        a = run_v2f(g)
    b = run_v2f(g)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("mutation", [3, 4, 8])
def test_vort2flux_closure_catches(variants, mutation):
    g = O.Grid(32, 24, 3, -1.0, -0.8, 0.0625)
    with O.variant(variants[mutation]):
        res = run_v2f(g)
    assert v2f_closure_error(g, *res) > 1e-3


@pytest.mark.parametrize("mutation", [3, 4])
def test_vort2flux_free_space_catches(variants, mutation):
    good = free_space_errors(levels=(4,))[0]
    with O.variant(variants[mutation]):
        bad = free_space_errors(levels=(4,))[0]
    assert bad > 10 * good, (good, bad)


@pytest.mark.parametrize("mutation", [1, 7])
def test_trial_harmonic_catches(variants, mutation):
    with O.variant(variants[mutation]):
        assert trial_harmonic_error() > 1e-4


@pytest.mark.parametrize("mutation", [2])
def test_trial_advection_catches(variants, mutation):
    with O.variant(variants[mutation]):
        assert trial_advection_error() > 1e-2


@pytest.mark.parametrize("mutation", [5])
def test_no_slip_static_catches(variants, mutation):
    prob = cases.cylinder(nx=48, ny=48, nlevel=3, h=0.05, x0=(-1.2, -1.2), Re=100.0,
                          dt=0.004, r=0.4)
    with O.variant(variants[mutation]):
        assert max(no_slip_errors(prob, 2)) > 1e-3


@pytest.mark.parametrize("mutation", [5, 6])
def test_no_slip_moving_catches(variants, mutation):
    prob = cases.moving_cylinder(nx=48, ny=48, nlevel=2, h=0.06, x0=(-1.44, -1.44))
    with O.variant(variants[mutation]):
        assert max(no_slip_errors(prob, 2, cg_reltol=1e-12)) > 1e-3
