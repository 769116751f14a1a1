"""CPU oracle of the CG warm-start network (TEST INFRASTRUCTURE ONLY).

Imported by ``tests/`` as the checker of the library's network kernels
(``immersedbodies.jl_amd/csrc/net_kernels.hip``); product code never imports it. The
network has no counterpart in ImmersedBodies.jl (its cg! starts from the previous F̃b,
coupling.jl:194-202); this file restates the library's definition:

* rows r = body points, channels contiguous; a convolution of odd width K is
  ``Y[r] = b + sum_k X[r + k - p] @ W[:, :, k].T`` with ``p = K // 2``, where a shifted
  row only counts inside the row's own sequence (``seg``);
* ReLU after every layer but the last;
* loss ``mean((y - t)^2)`` over rows x 2 outputs;
* Adam with bias correction, ``p -= lr * (m / bc1) / (sqrt(v / bc2) + eps)``.

With ``quant=True`` values are rounded to bfloat16 (round to nearest even) exactly where
the GPU stores bfloat16: network inputs, weights, hidden activations, and the gradient
rows of the backward pass. Sums run in float64 (the GPU accumulates in float32), so the
two agree to float32 summation error except for rare one-ulp bfloat16 rounding flips.
``quant=False`` is plain float64 and is what the tests pin against torch autograd and
finite differences.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even) via float32, returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _valid(seg, K):
    """valid[k][r]: row r + k - p exists and lies in r's sequence; src[k][r] its index."""
    R = len(seg)
    p = K // 2
    r = np.arange(R)
    src, ok = [], []
    for k in range(K):
        s = r + k - p
        good = (s >= 0) & (s < R)
        sc = np.clip(s, 0, R - 1)
        good &= seg[sc] == seg
        src.append(sc)
        ok.append(good)
    return src, ok


def conv(X, W, b, seg):
    """X: R x Cin, W: Cout x Cin x K, b: Cout or None -> R x Cout (float64)."""
    K = W.shape[2]
    src, ok = _valid(np.asarray(seg), K)
    Y = np.zeros((X.shape[0], W.shape[0]))
    for k in range(K):
        Xs = np.where(ok[k][:, None], X[src[k]], 0.0)
        Y += Xs @ W[:, :, k].T
    if b is not None:
        Y += b
    return Y


def forward(layers, X, seg, quant=True):
    """layers: [(W, b), ...]. Returns (inputs of every layer, output)."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    acts = [q(X)]
    A = acts[0]
    for i, (W, b) in enumerate(layers):
        Z = conv(A, q(W), np.asarray(b, dtype=np.float64), seg)
        if i < len(layers) - 1:
            A = q(np.maximum(Z, 0.0))
            acts.append(A)
        else:
            return acts, Z


def loss_and_grads(layers, X, seg, T, quant=True):
    """Loss mean((y - T)^2) and [(dW, db), ...] of every layer."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    seg = np.asarray(seg)
    acts, y = forward(layers, X, seg, quant)
    R = y.shape[0]
    d = y - np.asarray(T, dtype=np.float64)
    loss = float(np.sum(d * d) / (2.0 * R))
    dZ = q(d / R)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W = q(layers[i][0])
        A = acts[i]
        K = W.shape[2]
        src, ok = _valid(seg, K)
        dW = np.zeros(W.shape)
        for k in range(K):
            As = np.where(ok[k][:, None], A[src[k]], 0.0)
            dW[:, :, k] = dZ.T @ As
        grads[i] = (dW, dZ.sum(0))
        if i > 0:
            # dA[s] = sum over (r, k) with s = r + k - p of dZ[r] @ W[:, :, k]
            dA = np.zeros(A.shape)
            for k in range(K):
                contrib = np.where(ok[k][:, None], dZ @ W[:, :, k], 0.0)
                np.add.at(dA, src[k][ok[k]], contrib[ok[k]])
            dZ = q(np.where(A > 0.0, dA, 0.0))
    return loss, grads


def flatten(layers, out_scale=0.0):
    parts = []
    for W, b in layers:
        parts += [np.asarray(W, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()]
    return np.concatenate(parts + [np.array([out_scale])])


def unflatten(p, shapes):
    out, o = [], 0
    for (co, ci, k) in shapes:
        W = np.asarray(p[o:o + co * ci * k], dtype=np.float64).reshape(co, ci, k)
        o += co * ci * k
        b = np.asarray(p[o:o + co], dtype=np.float64)
        o += co
        out.append((W, b))
    return out, float(p[o])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t = 1-based step count); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    bc1 = 1 - b1 ** t
    bc2 = 1 - b2 ** t
    return p - lr * (m / bc1) / (np.sqrt(v / bc2) + eps), m, v
