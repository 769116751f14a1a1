"""Learned warm start of the CG surface-force solve of moving bodies.

No counterpart in ImmersedBodies.jl, whose ``cg!`` starts from the previous step's F̃b
(coupling.jl:194-202, :246-247). ``WarmStartNet`` wraps the library's 1-D convolutional
network (``ibpm_net_*`` in include/ibpm.h): it records training samples while a solver
steps, trains on the GPU with the library's own forward/backward kernels, and then
predicts the correction of the extrapolated guess. CG always runs to its stopping test,
so forces do not depend on the guess beyond the CG tolerance.

Typical use::

    net = WarmStartNet(channels=128, layers=4, kernel=5, max_points=256, capacity=20000)
    solver.set_warm_start("extrapolate", net=net, record=True)
    solver.advance(2000)                 # collect samples
    net.train(iters=3000, batch=64, lr=1e-3)
    solver.set_warm_start("learned", net=net)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

MODES = {"previous": _lib.IBPM_WS_PREVIOUS, "extrapolate": _lib.IBPM_WS_EXTRAPOLATE,
         "learned": _lib.IBPM_WS_LEARNED}


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class WarmStartNet:
    """GPU 1-D conv network over body points (BF16 weights/activations, FP32 accumulation,
    FP32 master weights trained with Adam)."""

    def __init__(self, channels=128, layers=4, kernel=5, max_points=1024, capacity=20000,
                 max_batch=64, seed=1234, device=0):
        L = _lib.load()
        c = _lib.NetConfig()
        L.ibpm_net_default_config(C.byref(c))
        c.channels, c.layers, c.kernel = int(channels), int(layers), int(kernel)
        c.max_points, c.capacity, c.max_batch = int(max_points), int(capacity), int(max_batch)
        c.seed, c.device = int(seed), int(device)
        h = C.c_void_p()
        _lib.check(L.ibpm_net_create(C.byref(c), C.byref(h)))
        self._h = h
        self.channels, self.layers, self.kernel = c.channels, c.layers, c.kernel
        self.max_points, self.capacity, self.max_batch = c.max_points, c.capacity, c.max_batch

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().ibpm_net_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- training set --
    def dataset_size(self) -> int:
        return int(_lib.load().ibpm_net_dataset_size(self._h))

    def clear_dataset(self):
        _lib.check(_lib.load().ibpm_net_clear_dataset(self._h))

    def train(self, iters: int, batch: int = 64, lr: float = 1e-3) -> np.ndarray:
        """Adam on random minibatches of the recorded samples; returns the loss per step."""
        loss = np.zeros(max(int(iters), 1))
        _lib.check(_lib.load().ibpm_net_train(self._h, int(iters), int(batch), float(lr),
                                              loss.ctypes.data_as(C.POINTER(C.c_double))))
        return loss[:int(iters)]

    # -- parameters (per layer: weight [cout][cin][k], bias [cout]; then output scale) --
    def param_count(self) -> int:
        return int(_lib.load().ibpm_net_param_count(self._h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count(), dtype=np.float32)
        _lib.check(_lib.load().ibpm_net_get_params(self._h, _fp(p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _lib.check(_lib.load().ibpm_net_set_params(self._h, _fp(p), p.size))

    def layer_shapes(self):
        """[(cout, cin, kernel), ...] of the convolutions."""
        shapes = []
        for l in range(self.layers):
            cin = _lib.NET_FEATURES if l == 0 else self.channels
            cout = _lib.NET_OUTPUTS if l == self.layers - 1 else self.channels
            shapes.append((cout, cin, self.kernel))
        return shapes

    def unflatten(self, p):
        """Split a flat parameter vector into [(W, b), ...] and the output scale."""
        out, o = [], 0
        for (co, ci, k) in self.layer_shapes():
            W = np.asarray(p[o:o + co * ci * k]).reshape(co, ci, k)
            o += co * ci * k
            b = np.asarray(p[o:o + co])
            o += co
            out.append((W, b))
        return out, float(p[o])

    # -- host rows --
    def forward(self, x, seg) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        seg = np.ascontiguousarray(seg, dtype=np.int32)
        y = np.zeros((x.shape[0], _lib.NET_OUTPUTS), dtype=np.float32)
        _lib.check(_lib.load().ibpm_net_forward(self._h, _fp(x), _ip(seg), x.shape[0], _fp(y)))
        return y

    def gradients(self, x, seg, target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        seg = np.ascontiguousarray(seg, dtype=np.int32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.param_count(), dtype=np.float32)
        loss = C.c_double()
        _lib.check(_lib.load().ibpm_net_gradients(self._h, _fp(x), _ip(seg), _fp(t), x.shape[0],
                                                  _fp(g), g.size, C.byref(loss)))
        return g, loss.value

    def step(self, x, seg, target, lr: float) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        seg = np.ascontiguousarray(seg, dtype=np.int32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_double()
        _lib.check(_lib.load().ibpm_net_step(self._h, _fp(x), _ip(seg), _fp(t), x.shape[0],
                                             float(lr), C.byref(loss)))
        return loss.value
