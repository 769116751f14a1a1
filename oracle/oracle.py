"""ctypes view of the FP64 CPU oracle (``libibpm_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke`` and the
CPU-baseline leg of ``bench.py`` as the checker. Product code never imports it.

The oracle restates the reference's FP64 algorithm (``src/fluid.jl``,
``src/coupling.jl``, ``src/solver.jl:78-190``, ``src/bodies.jl``); see the header of
``ibpm_oracle.c`` for the per-function citations.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libibpm_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def build_variant(mutation: int, outdir: str) -> str:
    """Build the oracle with planted mistake ``mutation`` (the IBO_MUTATION list at the top
    of ibpm_oracle.c) into ``outdir``; returns the library path. Used only by the tests
    that show each math check of the oracle fails on the slip it guards against."""
    out = os.path.join(outdir, f"libibpm_oracle_m{int(mutation)}.so")
    subprocess.run(["make", "-s", "-C", _HERE, f"MUTATION={int(mutation)}", f"OUT={out}",
                    "variant"], check=True, stdout=subprocess.DEVNULL)
    return out


@contextlib.contextmanager
def variant(path: str):
    """Within the block, every function of this module (and every OracleSolver created in
    it) uses the oracle library at ``path``."""
    global _lib
    saved = _lib
    _lib = _bind(C.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


class _Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nlevel", C.c_int),
                ("x0", C.c_double), ("y0", C.c_double), ("h", C.c_double)]


class _Config(C.Structure):
    _fields_ = [("grid", _Grid), ("Re", C.c_double), ("dt", C.c_double), ("t0", C.c_double),
                ("npanel", C.c_int), ("xb", _dp), ("ds", _dp), ("nbody", C.c_int),
                ("body_start", _ip), ("body_count", _ip), ("body_moving", _ip),
                ("cg_maxiter", C.c_int), ("cg_reltol", C.c_double), ("cg_abstol", C.c_double),
                ("cg_tol_mode", C.c_int), ("timing_only", C.c_int), ("grid_moving", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(C.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    g = C.POINTER(_Grid)
    L.ibo_dst2d.argtypes = [g, _dp]
    L.ibo_dst1d_batch.argtypes = [C.c_int, C.c_int, _dp]
    L.ibo_lap_eigs.argtypes = [g, _dp]
    L.ibo_lap_inv.argtypes = [g, _dp, _dp, _dp]
    L.ibo_diag_poisson.argtypes = [g, _dp]
    L.ibo_diag_A.argtypes = [g, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
    L.ibo_rot.argtypes = [g, _dp, _dp]
    L.ibo_curl.argtypes = [g, _dp, _dp, _dp]
    L.ibo_coarsify.argtypes = [g, _dp, _dp]
    L.ibo_get_bc.argtypes = [g, _dp, _dp]
    L.ibo_apply_bc.argtypes = [g, _dp, _dp, C.c_double]
    L.ibo_nonlinear.argtypes = [g, _dp, _dp, _dp, _dp, _dp, C.c_int]
    L.ibo_vort2flux.argtypes = [g, _dp, _dp, _dp]
    L.ibo_delta_h.argtypes = [C.c_double, C.c_double, C.c_double]
    L.ibo_delta_h.restype = C.c_double
    L.ibo_reg_update.argtypes = [g, C.c_int, _dp, _ip, _dp]
    L.ibo_reg_update.restype = C.c_int
    L.ibo_regT.argtypes = [g, C.c_int, _ip, _dp, _dp, _dp]
    L.ibo_reg.argtypes = [g, C.c_int, _ip, _dp, _dp, _dp]
    L.ibo_create.argtypes = [C.POINTER(_Config), C.c_char_p, C.c_size_t]
    L.ibo_create.restype = C.c_void_p
    L.ibo_destroy.argtypes = [C.c_void_p]
    L.ibo_advance.argtypes = [C.c_void_p, _dp, _dp, C.c_char_p, C.c_size_t]
    L.ibo_advance.restype = C.c_int
    L.ibo_field.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_size_t)]
    L.ibo_field.restype = _dp
    L.ibo_index.argtypes = [C.c_void_p]
    L.ibo_index.restype = C.c_int
    L.ibo_time.argtypes = [C.c_void_p]
    L.ibo_time.restype = C.c_double
    L.ibo_set_index.argtypes = [C.c_void_p, C.c_int]
    L.ibo_last_cg_iters.argtypes = [C.c_void_p]
    L.ibo_last_cg_iters.restype = C.c_int
    L.ibo_last_cg_residual.argtypes = [C.c_void_p]
    L.ibo_last_cg_residual.restype = C.c_double
    L.ibo_apply_B.argtypes = [C.c_void_p, _dp, _dp]
    L.ibo_Binv.argtypes = [C.c_void_p]
    L.ibo_Binv.restype = _dp
    L.ibo_set_threads.argtypes = [C.c_int]
    L.ibo_set_grid_motion.argtypes = [C.c_void_p, _dp]
    return L


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _pi(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


class Grid:
    """Geometry of a multi-domain grid: nx x ny cells, nlevel levels, base step h."""

    def __init__(self, nx, ny, nlevel, x0, y0, h):
        self.nx, self.ny, self.nlevel = int(nx), int(ny), int(nlevel)
        self.x0, self.y0, self.h = float(x0), float(y0), float(h)
        self.c = _Grid(self.nx, self.ny, self.nlevel, self.x0, self.y0, self.h)

    @property
    def Nx(self):
        return self.nx - 1

    @property
    def Ny(self):
        return self.ny - 1

    @property
    def nG(self):
        return (self.nx - 1) * (self.ny - 1)

    @property
    def nu(self):
        return (self.nx + 1) * self.ny

    @property
    def nq(self):
        return self.nu + self.nx * (self.ny + 1)

    @property
    def nbc(self):
        return 2 * (self.nx + 1) + 2 * (self.ny + 1)

    def ref(self):
        return C.byref(self.c)


# ---- single-operator wrappers (inputs are 1-D float64 arrays in reference layout) ----

def dst2d(g: Grid, x):
    y = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib().ibo_dst2d(g.ref(), _p(y))
    return y


def dst1d(x):
    """RODFT00 of each row of a 2-D array (rows contiguous)."""
    y = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib().ibo_dst1d_batch(y.shape[-1], y.size // y.shape[-1], _p(y))
    return y


def lap_eigs(g):
    y = np.zeros(g.nG)
    lib().ibo_lap_eigs(g.ref(), _p(y))
    return y


def diag_poisson(g):
    y = np.zeros(g.nG)
    lib().ibo_diag_poisson(g.ref(), _p(y))
    return y


def diag_A(g, Re, dt, lev, implicit):
    y = np.zeros(g.nG)
    lib().ibo_diag_A(g.ref(), Re, dt, lev, int(implicit), _p(y))
    return y


def lap_inv(g, diag, b):
    x = np.zeros(g.nG)
    lib().ibo_lap_inv(g.ref(), _p(np.ascontiguousarray(diag)), _p(x),
                      _p(np.ascontiguousarray(b, dtype=np.float64)))
    return x


def rot(g, q):
    G = np.zeros(g.nG)
    lib().ibo_rot(g.ref(), _p(G), _p(np.ascontiguousarray(q, dtype=np.float64)))
    return G


def curl(g, psi, bc):
    q = np.zeros(g.nq)
    lib().ibo_curl(g.ref(), _p(q), _p(np.ascontiguousarray(psi, dtype=np.float64)),
                   _p(np.ascontiguousarray(bc, dtype=np.float64)))
    return q


def coarsify(g, Gc, G):
    out = np.ascontiguousarray(Gc, dtype=np.float64).copy()
    lib().ibo_coarsify(g.ref(), _p(out), _p(np.ascontiguousarray(G, dtype=np.float64)))
    return out


def get_bc(g, r):
    bc = np.zeros(g.nbc)
    lib().ibo_get_bc(g.ref(), _p(bc), _p(np.ascontiguousarray(r, dtype=np.float64)))
    return bc


def apply_bc(g, r, bc, fac):
    out = np.ascontiguousarray(r, dtype=np.float64).copy()
    lib().ibo_apply_bc(g.ref(), _p(out), _p(np.ascontiguousarray(bc, dtype=np.float64)), fac)
    return out


def nonlinear(g, q, q0, G, Gbc, lev):
    out = np.zeros(g.nG)
    f = lambda a: _p(np.ascontiguousarray(a, dtype=np.float64))
    lib().ibo_nonlinear(g.ref(), _p(out), f(q), f(q0), f(G), f(Gbc), lev)
    return out


def vort2flux(g, G):
    """Returns (psi, q, G_after) for G of shape (nlevel, nG) (rows = levels)."""
    Gm = np.ascontiguousarray(G, dtype=np.float64).copy()
    psi = np.zeros((g.nlevel, g.nG))
    q = np.zeros((g.nlevel, g.nq))
    lib().ibo_vort2flux(g.ref(), _p(psi), _p(q), _p(Gm))
    return psi, q, Gm


def delta_h(rf, rb, dr):
    return lib().ibo_delta_h(float(rf), float(rb), float(dr))


def reg_update(g, xb):
    """xb: (n, 2) panel positions -> (coords (n,2) int32, weights (n,2,7,7) [p,c,b,a])."""
    xb = np.asarray(xb, dtype=np.float64)
    n = xb.shape[0]
    flat = np.ascontiguousarray(xb.T.reshape(-1))
    coords = np.zeros(2 * n, dtype=np.int32)
    w = np.zeros(n * 98)
    rc = lib().ibo_reg_update(g.ref(), n, _p(flat), _pi(coords), _p(w))
    if rc != 0:
        raise ValueError(f"Panel index={-rc} outside allowed region")
    return coords.reshape(n, 2), w.reshape(n, 2, 7, 7)


def regT(g, coords, w, q1):
    n = coords.shape[0]
    fb = np.zeros(2 * n)
    lib().ibo_regT(g.ref(), n, _pi(np.ascontiguousarray(coords.reshape(-1), dtype=np.int32)),
                   _p(np.ascontiguousarray(w.reshape(-1))), _p(fb),
                   _p(np.ascontiguousarray(q1, dtype=np.float64)))
    return fb


def reg(g, coords, w, fb):
    n = coords.shape[0]
    q1 = np.zeros(g.nq)
    lib().ibo_reg(g.ref(), n, _pi(np.ascontiguousarray(coords.reshape(-1), dtype=np.int32)),
                  _p(np.ascontiguousarray(w.reshape(-1))), _p(q1),
                  _p(np.ascontiguousarray(fb, dtype=np.float64)))
    return q1


FIELDS = {"q": 0, "q0": 1, "Gamma": 2, "psi": 3, "nonlin1": 4, "nonlin2": 5, "xb": 6,
          "ub": 7, "Ftilde": 8, "fb": 9, "total_force": 10, "ds": 11, "freestream": 12}


class OracleSolver:
    """FP64 CPU restatement of ``CnabSolver`` (solver.jl:78-190).

    Arguments mirror the HIP solver: the grid, Re, dt, the panel list (n, 2), and the
    body partition (start, count, moving) of the panel list.
    """

    def __init__(self, grid: Grid, Re, dt, xb, ds=None, bodies=None, t0=0.0,
                 cg_maxiter=5000, cg_reltol=1e-12, cg_abstol=0.0, cg_tol_mode=0,
                 timing_only=False, grid_moving=False):
        xb = np.asarray(xb, dtype=np.float64).reshape(-1, 2)
        n = xb.shape[0]
        self.grid, self.n = grid, n
        if bodies is None:
            bodies = [(0, n, False)] if n else []
        self._xb = np.ascontiguousarray(xb.T.reshape(-1))
        self._ds = np.ascontiguousarray(np.zeros(n) if ds is None else np.asarray(ds, float))
        self._bs = np.array([b[0] for b in bodies] or [0], dtype=np.int32)
        self._bc = np.array([b[1] for b in bodies] or [0], dtype=np.int32)
        self._bm = np.array([int(b[2]) for b in bodies] or [0], dtype=np.int32)
        self.nbody = len(bodies)
        cfg = _Config(grid.c, Re, dt, t0, n, _p(self._xb), _p(self._ds), self.nbody,
                      _pi(self._bs), _pi(self._bc), _pi(self._bm), cg_maxiter, cg_reltol,
                      cg_abstol, cg_tol_mode, int(timing_only), int(grid_moving))
        err = C.create_string_buffer(512)
        self._L = lib()
        self._h = self._L.ibo_create(C.byref(cfg), err, 512)
        if not self._h:
            raise RuntimeError(err.value.decode())
        self.Re, self.dt, self.t0 = Re, dt, t0

    def close(self):
        if getattr(self, "_h", None):
            self._L.ibo_destroy(self._h)
            self._h = None

    __del__ = close

    @property
    def index(self):
        return self._L.ibo_index(self._h)

    @property
    def t(self):
        return self._L.ibo_time(self._h)

    def next_time(self):
        return self.t0 + self.dt * self.index

    def advance(self, freestream=(1.0, 0.0), motion=None, grid_motion=None):
        """One step. grid_motion: GridVelocity of a moving grid at the new time (centre x,
        y, velocity x, y, angle, angular_vel)."""
        if grid_motion is not None:
            gm = np.ascontiguousarray(np.asarray(grid_motion, dtype=np.float64).reshape(6))
            self._L.ibo_set_grid_motion(self._h, _p(gm))
        fs = np.ascontiguousarray(np.asarray(freestream, dtype=np.float64)[:2])
        m = np.zeros(6 * max(self.nbody, 1)) if motion is None else \
            np.ascontiguousarray(np.asarray(motion, dtype=np.float64).reshape(-1))
        err = C.create_string_buffer(512)
        rc = self._L.ibo_advance(self._h, _p(fs), _p(m), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())

    def field(self, name):
        n = C.c_size_t()
        ptr = self._L.ibo_field(self._h, FIELDS[name], C.byref(n))
        return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()

    def set_field(self, name, values):
        """Overwrite a state array (q, q0, Gamma, psi, nonlin1, nonlin2, ...) in place, in
        the layout ``field`` returns."""
        assert name != "total_force"
        n = C.c_size_t()
        ptr = self._L.ibo_field(self._h, FIELDS[name], C.byref(n))
        v = np.asarray(values, dtype=np.float64).reshape(-1)
        assert v.size == n.value
        np.ctypeslib.as_array(ptr, shape=(n.value,))[:] = v

    def cg_iters(self):
        return self._L.ibo_last_cg_iters(self._h)

    def cg_residual(self):
        return self._L.ibo_last_cg_residual(self._h)

    def apply_B(self, z):
        x = np.zeros(2 * self.n)
        self._L.ibo_apply_B(self._h, _p(x), _p(np.ascontiguousarray(z, dtype=np.float64)))
        return x

    def Binv(self):
        p = self._L.ibo_Binv(self._h)
        if not p:
            return None
        m = 2 * self.n
        return np.ctypeslib.as_array(p, shape=(m * m,)).reshape(m, m, order="F").copy()


def set_threads(n: int):
    lib().ibo_set_threads(int(n))
