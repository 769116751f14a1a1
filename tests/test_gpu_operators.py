"""Operator-level GPU parity (through ibpm_apply): each building block of the step against
the oracle function it restates, in FP64 and in mixed precision.

Tolerances (relative 2-norm): FP64 1e-12 for single solves, 1e-11 for the multi-level
vort2flux; mixed precision with 2 refinement sweeps 1e-9; without refinement the
3xBF16 products must still give 1e-4 (the raw split-product accuracy)."""
import numpy as np
import pytest

import cases
import immersedbodies as ib
import oracle as O

pytestmark = pytest.mark.gpu


def _grid_and_solver(precision="fp64", refine=1, nx=64, ny=48, L=3):
    h = 0.05
    prob = cases.no_body(nx=nx, ny=ny, nlevel=L, h=h, x0=(-nx * h / 2, -ny * h / 2), Re=100.0,
                         dt=0.004)
    g = prob.fluid.grid
    og = O.Grid(nx, ny, L, g.base.x0[0], g.base.x0[1], g.base.h)
    return prob, og, ib.create_solver(prob, precision=precision, refine_sweeps=refine)


@pytest.mark.parametrize("precision,refine,tol", [("fp64", 0, 1e-12), ("mixed", 2, 1e-9),
                                                  ("mixed", 0, 1e-4)])
def test_poisson_and_helmholtz(precision, refine, tol):
    prob, og, s = _grid_and_solver(precision, refine)
    print("poisson/helmholtz", precision, refine, flush=True)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(og.nG)
    got = s.apply("poisson", b)
    ref = O.lap_inv(og, O.diag_poisson(og), b)
    e = cases.relerr(got, ref)
    print("poisson rel err", e)
    assert e < tol
    for lev in (1, 3):
        got = s.apply("helmholtz", b, level=lev)
        ref = O.lap_inv(og, O.diag_A(og, prob.fluid.Re, prob.scheme.dt, lev, True), b)
        e = cases.relerr(got, ref)
        print("helmholtz", lev, "rel err", e)
        assert e < tol
    s.close()


@pytest.mark.parametrize("precision,refine,tol", [("fp64", 0, 1e-12), ("mixed", 2, 1e-9)])
@pytest.mark.parametrize("nx,ny", [(42, 38), (41, 37), (45, 52)])
def test_poisson_and_helmholtz_any_length(precision, refine, tol, nx, ny):
    """Lengths not divisible by 4 (single fold) and odd lengths (single fold without a
    middle element)."""
    prob, og, s = _grid_and_solver(precision, refine, nx, ny, 2)
    print("poisson/helmholtz", nx, ny, precision, flush=True)
    b = np.random.default_rng(nx).standard_normal(og.nG)
    assert cases.relerr(s.apply("poisson", b), O.lap_inv(og, O.diag_poisson(og), b)) < tol
    ref = O.lap_inv(og, O.diag_A(og, prob.fluid.Re, prob.scheme.dt, 2, True), b)
    assert cases.relerr(s.apply("helmholtz", b, level=2), ref) < tol
    s.close()


@pytest.mark.parametrize("tiles", [20, 21, 22, 23, -1])
def test_poisson_third_fold_partial_tiles(tiles):
    """The third-fold sine passes where Q/2 is not a multiple of the tile's 32 or 64 rows:
    nx = 320 gives Q = 80 along x (40 row pairs: a whole and a partial 32-row tile, the
    64-row variants taking 32 rows), ny = 256 gives Q = 64 along y (32 row pairs: one
    whole 32-row tile, the 64-row variants taking 32 rows); against the oracle at 1e-12."""
    prob, og, s = _grid_and_solver("fp64", 0, 320, 256, 2)
    print("poisson third fold, tiles", tiles, flush=True)
    s.close()
    s = ib.create_solver(prob, precision="fp64", dst_tiles=tiles)
    assert s.dst_tiles(1)[3] == 8 and s.dst_tiles(1, False)[3] == 8
    b = np.random.default_rng(320).standard_normal(og.nG)
    e = cases.relerr(s.apply("poisson", b), O.lap_inv(og, O.diag_poisson(og), b))
    print("rel err", e)
    assert e < 1e-12
    s.close()


@pytest.mark.parametrize("dt", [0.004, 0.1, 0.5])
def test_helmholtz_gauss_seidel_and_sine_transforms(dt):
    """The two Helmholtz methods against the oracle's sine-transform solve at 1e-12: with
    h = 0.05 and Re = 100, k = dt / (2 Re h^2) on level 1 is 0.008 (13 Gauss-Seidel
    half-sweeps), 0.2 (49) and 1.0 (beyond 64: the sine transforms; level 2's 0.25 still
    Gauss-Seidel)."""
    nx, ny, L, h = 64, 48, 3, 0.05
    prob = cases.no_body(nx=nx, ny=ny, nlevel=L, h=h, x0=(-nx * h / 2, -ny * h / 2), Re=100.0,
                         dt=dt)
    og = O.Grid(nx, ny, L, -nx * h / 2, -ny * h / 2, h)
    b = np.random.default_rng(5).standard_normal(og.nG)
    got = {}
    for mode in ("auto", "sine"):
        s = ib.create_solver(prob, helmholtz=mode)
        for lev in (1, 2, 3):
            got[mode, lev] = s.apply("helmholtz", b, level=lev)
        s.close()
    for lev in (1, 2, 3):
        ref = O.lap_inv(og, O.diag_A(og, prob.fluid.Re, dt, lev, True), b)
        for mode in ("auto", "sine"):
            e = cases.relerr(got[mode, lev], ref)
            print("dt", dt, "level", lev, mode, "rel err", e)
            assert e < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_step_gauss_seidel_matches_sine_transforms(precision):
    """Steps with the Gauss-Seidel Helmholtz solves end within rounding of those with the
    sine transforms (FP64: 1e-12 on fields, 1e-11 on forces; mixed: the forces of both
    within 1e-7 of each other)."""
    prob = cases.cylinder(nx=96, ny=64, nlevel=3, h=0.05, x0=(-1.6, -1.6), Re=100.0)
    out = {}
    for mode in ("auto", "sine"):
        s = ib.create_solver(prob, precision=precision, helmholtz=mode)
        st = s.advance(6)
        out[mode] = (s.forces(st), s.read("Gamma"), s.read("psi"), s.read("q"))
        s.close()
    fa, fs = out["auto"][0], out["sine"][0]
    if precision == "fp64":
        assert cases.relerr(fa, fs) < 1e-11
        for a, b in zip(out["auto"][1:], out["sine"][1:]):
            assert cases.relerr(a, b) < 1e-12
    else:
        assert cases.relerr(fa, fs) < 1e-7


@pytest.mark.parametrize("precision,refine,tol", [("fp64", 0, 1e-11), ("mixed", 2, 1e-9),
                                                  ("mixed", 0, 1e-4)])
@pytest.mark.parametrize("shape", [(64, 48, 3), (32, 32, 1), (40, 56, 4), (42, 38, 3), (41, 37, 3),
                                   (45, 52, 2)])
def test_vort2flux(precision, refine, tol, shape):
    nx, ny, L = shape
    prob, og, s = _grid_and_solver(precision, refine, nx, ny, L)
    print("vort2flux", shape, precision, refine, flush=True)
    rng = np.random.default_rng(2)
    G = rng.standard_normal((L, og.nG))
    got = s.apply("vort2flux", G).reshape(L, -1)
    psi, _, _ = O.vort2flux(og, G)
    for l in range(L):
        e = cases.relerr(got[l], psi[l])
        print("level", l + 1, "rel err", e)
        assert e < tol
    s.close()


@pytest.mark.parametrize("precision,nx", [("fp64", 96), ("fp64", 128), ("fp64", 256), ("mixed", 96),
                                          ("mixed", 128)])
def test_tile_variants_give_identical_results(precision, nx):
    """Every sine-pass tile variant of one fold sums in the same order: identical fields
    after steps. The single, double and (FP64) third folds agree to rounding in FP64 and,
    through the mixed path's FP64 refinement, to 1e-8 in mixed precision (its split
    products round differently in the two folds). With nx = 96 (Q = 24 and 16 along x and
    y) the FP64 direct-copy variants fall back to register staging (Q not a multiple of 16
    along x) and the mixed double fold runs a partial K chunk; with nx = 128 (Q = 32, 16)
    all double-fold variants run and the third fold falls back to the double fold (Q/2
    below its tiles); with nx = 256 (Q = 64 along x) the third fold runs along x on 32-row
    tiles (its 64-row variants take those) and falls back along y."""
    prob = cases.cylinder(nx=nx, ny=64, nlevel=3, h=0.05, x0=(-1.6, -1.6), Re=100.0)
    variants = range(24 if precision == "fp64" else 20)
    ref = {}
    for v in variants:
        s = ib.create_solver(prob, precision=precision, dst_tiles=v)
        s.advance(4)
        got = np.concatenate([s.read("Gamma"), s.read("psi"), s.read("Ftilde")])
        fold = s.dst_tiles(3)[3]
        s.close()
        ref.setdefault(fold, got)
        assert np.array_equal(got, ref[fold]), v
    assert set(ref) == ({2, 4, 8} if precision == "fp64" and nx == 256 else {2, 4})
    assert cases.relerr(ref[4], ref[2]) < (1e-13 if precision == "fp64" else 1e-8)
    if 8 in ref:
        assert cases.relerr(ref[8], ref[4]) < 1e-13
    # autotuned: each precision is tuned over the deepest fold that applies, so the result
    # is bitwise that of any fixed variant of that fold
    deepest = max(ref)
    s = ib.create_solver(prob, precision=precision)
    s.advance(4)
    got = np.concatenate([s.read("Gamma"), s.read("psi"), s.read("Ftilde")])
    tiles = [s.dst_tiles(1), s.dst_tiles(3), s.dst_tiles(3, False)]
    print("autotuned tiles (TB, TM, PF, fold, KC)", tiles)
    s.close()
    assert tiles[0][3] == deepest and tiles[1][3] == deepest and tiles[2][3] == 4
    assert np.array_equal(got, ref[deepest])
    if precision == "fp64":
        # the direct-copy variant with K chunk 16 runs along x only where Q is a multiple
        # of 16
        s = ib.create_solver(prob, precision=precision, dst_tiles=16)
        assert s.dst_tiles(3) == ((32, 64, 0, 4, 16) if nx >= 128 else (32, 64, 1, 4, 16))
        assert s.dst_tiles(3, False) == (32, 64, 0, 4, 16)
        s.close()
