"""Pins the CPU oracle of the warm-start network (oracle/warmstart_net.py) before it is
used to check the GPU kernels: float64 forward/backward against torch autograd on
segmented sequences (an independent implementation), finite differences, Adam against
torch.optim.Adam, and the bfloat16 rounding against torch's."""
import numpy as np
import pytest
import torch

import warmstart_net as WN


def _net(rng, shapes):
    return [(rng.standard_normal(s) * (2.0 / (s[1] * s[2])) ** 0.5, rng.standard_normal(s[0]) * 0.1)
            for s in shapes]


def _torch_forward(layers, X, lens):
    """Each sequence separately through torch Conv1d with zero padding."""
    outs, o = [], 0
    mods = []
    for i, (W, b) in enumerate(layers):
        c = torch.nn.Conv1d(W.shape[1], W.shape[0], W.shape[2], padding=W.shape[2] // 2).double()
        c.weight.data = torch.tensor(W)
        c.bias.data = torch.tensor(b)
        mods.append(c)
    for n in lens:
        a = torch.tensor(X[o:o + n].T[None])
        for i, c in enumerate(mods):
            a = c(a)
            if i < len(mods) - 1:
                a = torch.relu(a)
        outs.append(a[0].T)
        o += n
    return torch.cat(outs, 0), mods


SHAPES = [(32, 10, 5), (32, 32, 5), (2, 32, 5)]


def _case(seed=0, lens=(7, 12, 5, 1)):
    rng = np.random.default_rng(seed)
    layers = _net(rng, SHAPES)
    R = sum(lens)
    X = rng.standard_normal((R, 10))
    seg = np.repeat(np.arange(len(lens)), lens)
    T = rng.standard_normal((R, 2))
    return layers, X, seg, T, lens


def test_forward_matches_torch():
    layers, X, seg, T, lens = _case()
    _, y = WN.forward(layers, X, seg, quant=False)
    yt, _ = _torch_forward(layers, X, lens)
    assert np.max(np.abs(y - yt.detach().numpy())) < 1e-12


def test_gradients_match_torch_autograd():
    layers, X, seg, T, lens = _case(1)
    loss, grads = WN.loss_and_grads(layers, X, seg, T, quant=False)
    yt, mods = _torch_forward(layers, X, lens)
    lt = ((yt - torch.tensor(T)) ** 2).mean()
    lt.backward()
    assert abs(loss - lt.item()) < 1e-12 * max(1.0, abs(loss))
    for (dW, db), c in zip(grads, mods):
        assert np.max(np.abs(dW - c.weight.grad.numpy())) < 1e-12
        assert np.max(np.abs(db - c.bias.grad.numpy())) < 1e-12


def test_gradients_finite_differences():
    layers, X, seg, T, lens = _case(2)
    _, grads = WN.loss_and_grads(layers, X, seg, T, quant=False)
    rng = np.random.default_rng(3)
    eps = 1e-6
    for li in range(len(layers)):
        W, b = layers[li]
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in W.shape)
            Wp = W.copy()
            Wp[idx] += eps
            Wm = W.copy()
            Wm[idx] -= eps
            lp, _ = WN.loss_and_grads(layers[:li] + [(Wp, b)] + layers[li + 1:], X, seg, T, quant=False)
            lm, _ = WN.loss_and_grads(layers[:li] + [(Wm, b)] + layers[li + 1:], X, seg, T, quant=False)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grads[li][0][idx]) < 1e-7 * max(1.0, abs(fd))


def test_sequences_do_not_mix():
    """A row's output depends only on rows of its own sequence."""
    layers, X, seg, T, lens = _case(4)
    _, y = WN.forward(layers, X, seg, quant=False)
    X2 = X.copy()
    X2[seg == 0] += 1.0
    _, y2 = WN.forward(layers, X2, seg, quant=False)
    assert np.all(y2[seg != 0] == y[seg != 0])
    assert np.any(y2[seg == 0] != y[seg == 0])


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    p = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-2)
    q, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t in range(1, 6):
        g = rng.standard_normal(50)
        p.grad = torch.tensor(g)
        opt.step()
        q, m, v = WN.adam(q, g, m, v, t, 1e-2)
    assert np.max(np.abs(q - p.detach().numpy())) < 1e-12


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-6, 6, 10000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, -0.0]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(WN.bf16(x), ref)
    # ties go to even
    assert WN.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert WN.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6


@pytest.mark.parametrize("quant", [False, True])
def test_flatten_round_trip(quant):
    layers, *_ = _case(7)
    p = WN.flatten(layers, 2.5)
    back, s = WN.unflatten(p, SHAPES)
    assert s == 2.5
    for (W, b), (W2, b2) in zip(layers, back):
        assert np.array_equal(W, W2) and np.array_equal(b, b2)
