"""GPU tests of the learned CG warm start: the network kernels against the CPU oracle
(oracle/warmstart_net.py), the Adam kernel, and what the trained network gains on a
pitching plate over CG from the previous step's F̃b (the reference) and from the
extrapolated guess alone."""
import math

import numpy as np
import pytest

import cases
import immersedbodies as ib
from immersedbodies.warmstart import WarmStartNet
import warmstart_net as WN

pytestmark = pytest.mark.gpu


def _rows(seed, lens):
    rng = np.random.default_rng(seed)
    R = sum(lens)
    X = rng.standard_normal((R, 10)).astype(np.float32)
    seg = np.repeat(np.arange(len(lens)), lens).astype(np.int32)
    T = rng.standard_normal((R, 2)).astype(np.float32)
    return X, seg, T


def _random_params(net, seed):
    rng = np.random.default_rng(seed)
    layers = []
    for (co, ci, k) in net.layer_shapes():
        layers.append((rng.uniform(-1, 1, (co, ci, k)) * math.sqrt(6.0 / (ci * k)),
                       rng.uniform(-0.1, 0.1, co)))
    p = WN.flatten(layers, 1.0).astype(np.float32)
    net.set_params(p)
    # the oracle sees exactly the float32 values the GPU holds
    return WN.unflatten(p.astype(np.float64), net.layer_shapes())[0]


def _relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


LENS = [(33,), (17, 40, 9, 1), (200, 64, 131)]


@pytest.mark.parametrize("lens", LENS)
def test_forward_matches_oracle(lens):
    net = WarmStartNet(channels=64, layers=4, kernel=5, max_points=256, capacity=4, max_batch=4)
    layers = _random_params(net, 11)
    X, seg, _ = _rows(1, lens)
    y = net.forward(X, seg)
    _, yo = WN.forward(layers, X, seg, quant=True)
    print("forward rel err", _relerr(y, yo))
    assert _relerr(y, yo) < 2e-3
    # without the bf16 rounding the oracle differs by the bf16 quantisation itself
    _, yf = WN.forward(layers, X, seg, quant=False)
    assert _relerr(y, yf) < 5e-2


@pytest.mark.parametrize("lens", LENS)
@pytest.mark.parametrize("kernel", [3, 5, 7])
def test_gradients_match_oracle(lens, kernel):
    net = WarmStartNet(channels=64, layers=4, kernel=kernel, max_points=256, capacity=4, max_batch=4)
    layers = _random_params(net, 12)
    X, seg, T = _rows(2, lens)
    g, loss = net.gradients(X, seg, T)
    lo, go = WN.loss_and_grads(layers, X, seg, T, quant=True)
    gflat = WN.flatten(go, 0.0)
    print("loss", loss, lo, "grad rel err", _relerr(g, gflat))
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert _relerr(g, gflat) < 5e-3
    # layer by layer (weights and biases)
    gl, _ = WN.unflatten(g.astype(np.float64), net.layer_shapes())
    for (dW, db), (oW, ob) in zip(gl, go):
        assert _relerr(dW, oW) < 1e-2
        assert _relerr(db, ob) < 1e-2


def test_adam_step_matches_oracle():
    net = WarmStartNet(channels=64, layers=3, kernel=5, max_points=256, capacity=4, max_batch=4)
    layers = _random_params(net, 13)
    X, seg, T = _rows(3, (50, 30))
    p0 = net.get_params().astype(np.float64)
    g, _ = net.gradients(X, seg, T)
    lr = 1e-3
    net.step(X, seg, T, lr)  # first Adam step after set_params
    p1 = net.get_params().astype(np.float64)
    q, _, _ = WN.adam(p0[:-1], g[:-1].astype(np.float64), 0.0, 0.0, 1, lr)
    assert np.max(np.abs((p1[:-1] - p0[:-1]) - (q - p0[:-1]))) < 1e-3 * lr
    assert p1[-1] == p0[-1]  # output scale untouched


def test_params_round_trip_and_errors():
    net = WarmStartNet(channels=32, layers=2, kernel=3, max_points=16, capacity=2, max_batch=1)
    p = np.random.default_rng(4).standard_normal(net.param_count()).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    with pytest.raises(ib._lib.IbpmError):
        net.set_params(p[:-1])
    with pytest.raises(ib._lib.IbpmError):
        net.train(10, batch=1)  # empty training set
    X, seg, _ = _rows(5, (17,))
    with pytest.raises(ib._lib.IbpmError):
        net.forward(X, seg)  # 17 rows > max_batch * max_points


def _plate(k):
    # 256^2 cells on [-2, 2]^2, h = 1/64, 3 levels, plate of length 1 at the reference's
    # panel spacing 2h (33 points)
    return cases.pitching_plate(nx=256, ny=256, nlevel=3, h=1.0 / 64, x0=(-2.0, -2.0), npts=33,
                                k=k, dt=0.002)


CG = dict(maxiter=5000, reltol=1e-8)


def _solver(k, precision="fp64"):
    return ib.create_solver(_plate(k), precision=precision, cg=CG, cg_tol_mode=1)


def test_recording_and_training_reduce_loss():
    net = WarmStartNet(channels=64, layers=4, kernel=5, max_points=64, capacity=1000, max_batch=32)
    s = _solver(0.5)
    s.set_warm_start("extrapolate", net=net, record=True)
    s.advance(120)
    assert net.dataset_size() == 120
    loss = net.train(400, batch=32, lr=1e-3)
    print("loss first/last", loss[:10].mean(), loss[-10:].mean())
    assert np.all(np.isfinite(loss))
    assert loss[-20:].mean() < 0.5 * loss[:20].mean()


def test_learned_warm_start_gain():
    """Train on three reduced frequencies, test on a fourth: CG iterations per step with
    the reference's previous-step guess, the extrapolated guess, and the learned one."""
    net = WarmStartNet(channels=128, layers=4, kernel=5, max_points=64, capacity=4000,
                       max_batch=64, seed=7)
    for k in (0.3, 0.6, 0.9):
        s = _solver(k)
        s.set_warm_start("extrapolate", net=net, record=True)
        s.advance(400)
        s.close()
    loss = net.train(3000, batch=64, lr=1e-3)
    print("training loss first/last", loss[:50].mean(), loss[-50:].mean())

    res = {}
    forces = {}
    for mode in ("previous", "extrapolate", "learned"):
        s = _solver(0.75)
        s.set_warm_start("extrapolate")
        s.advance(100)  # same warm-up for every mode
        s.set_warm_start(mode, net=net if mode == "learned" else None)
        st = s.advance(200)
        res[mode] = (s.cg_iters(st)[3:].mean(), s.b_applies(st)[3:].mean(),
                     np.mean([x.ws_used for x in st][3:]))
        forces[mode] = s.forces(st)
        s.close()
    print("mean CG iterations / B applications / learned guess kept:", res)
    prev_it, prev_b, _ = res["previous"]
    ext_it, ext_b, _ = res["extrapolate"]
    lrn_it, lrn_b, kept = res["learned"]
    assert lrn_it < ext_it < prev_it
    # the network's extra B application included, less than half the reference's work
    assert lrn_b < 0.5 * prev_b
    assert kept > 0.9
    # CG runs to the same tolerance: forces agree whatever the guess
    for mode in ("extrapolate", "learned"):
        d = np.abs(forces[mode] - forces["previous"]).max() / np.abs(forces["previous"]).max()
        assert d < 1e-6, (mode, d)


def test_learned_warm_start_gain_plate_512():
    """The pitching-plate workload at its full grid: 512 x 512 cells on [-2,6] x [-4,4]
    (h = 1/64), 4 levels, Re 200, dt 0.002, the plate at the reference's panel spacing 2h
    (33 points), CG to 1e-8 of ||b||. Trained on three reduced frequencies, tested on a
    fourth: learned < extrapolated < previous-step CG iterations, and forces within the CG
    tolerance of the reference's guess."""
    net = WarmStartNet(channels=128, layers=4, kernel=5, max_points=64, capacity=4000,
                       max_batch=64, seed=11)

    def solver(k):
        return ib.create_solver(cases.pitching_plate_512(k=k), cg=CG, cg_tol_mode=1)

    print("case pitching plate 512^2 x 4, record", flush=True)
    for k in (0.3, 0.6, 0.9):
        s = solver(k)
        s.set_warm_start("extrapolate", net=net, record=True)
        s.advance(400)
        s.close()
    loss = net.train(3000, batch=64, lr=1e-3)
    print("training loss first/last", loss[:50].mean(), loss[-50:].mean(), flush=True)
    res, forces = {}, {}
    for mode in ("previous", "extrapolate", "learned"):
        s = solver(0.75)
        s.set_warm_start("extrapolate")
        s.advance(100)
        s.set_warm_start(mode, net=net if mode == "learned" else None)
        st = s.advance(200)
        assert all(x.cg_converged for x in st), mode
        res[mode] = (s.cg_iters(st)[3:].mean(), s.b_applies(st)[3:].mean())
        forces[mode] = s.forces(st)
        s.close()
    print("mean CG iterations / B applications:", res, flush=True)
    assert res["learned"][0] < res["extrapolate"][0] < res["previous"][0]
    assert res["learned"][1] < 0.5 * res["previous"][1]
    for mode in ("extrapolate", "learned"):
        d = np.abs(forces[mode] - forces["previous"]).max() / np.abs(forces["previous"]).max()
        assert d < 1e-6, (mode, d)


def test_stated_200_point_plate():
    """The plate as first stated, 200 points on the 512^2 grid (spacing 0.32 h): B is
    numerically singular there, and plain CG from the reference's guess either fails its
    1e-8 ||b||-relative test within 3000 iterations or needs far more iterations than the
    33-point plate at the reference's 2h spacing."""
    cg = dict(maxiter=3000, reltol=1e-8)
    res = {}
    for npts in (33, 200):
        s = ib.create_solver(cases.pitching_plate_512(k=0.6, npts=npts), cg=cg, cg_tol_mode=1)
        st = s.advance(2)
        res[npts] = [(x.cg_iters, bool(x.cg_converged), x.cg_residual) for x in st]
        s.close()
    print("CG per step (iterations, converged, residual):", res, flush=True)
    assert all(c for _, c, _ in res[33])
    it33 = max(i for i, _, _ in res[33])
    assert all((not c) or i > 10 * it33 for i, c, _ in res[200]), res
