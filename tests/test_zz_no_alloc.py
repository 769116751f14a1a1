"""ibpm_apply_noalloc through the C ABI against ibpm_apply.

The no-allocation variant runs the same kernels as ibpm_apply on device buffers, with its
intermediate planes in the library's static scratch buffer instead of the solver's. Its
results must equal ibpm_apply's bit for bit (FP64 and mixed precision), on consecutive
calls whose second input is smaller (the buffer keeps the first call's contents), and it
must reject an operator whose scratch exceeds IBPM_NOALLOC_SCRATCH_BYTES.
"""
import ctypes as C

import numpy as np
import pytest

import cases
import immersedbodies as ib
from immersedbodies import _lib

pytestmark = pytest.mark.gpu

OPS = {"poisson": _lib.IBPM_OP_POISSON, "helmholtz": _lib.IBPM_OP_HELMHOLTZ,
       "vort2flux": _lib.IBPM_OP_VORT2FLUX}


def _hip():
    """The HIP runtime libibpm links (already loaded with it): the caller's allocations."""
    _lib.load()
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


class DeviceArray:
    """float64 device buffer owned by the test (hipMalloc), with host copies."""
    H2D, D2H = 1, 2

    def __init__(self, n):
        self.H = _hip()
        self.n = int(n)
        self.ptr = C.c_void_p()
        assert self.H.hipMalloc(C.byref(self.ptr), self.n * 8) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.H.hipMemcpy(self.ptr, a.ctypes.data, self.n * 8, self.H2D) == 0

    def download(self):
        out = np.empty(self.n)
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, self.n * 8, self.D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.H.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _noalloc(solver, op, x, level=1):
    L = _lib.load()
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    din, dout = DeviceArray(a.size), DeviceArray(a.size)
    din.upload(a)
    dout.upload(np.full(a.size, np.nan))
    rc = L.ibpm_apply_noalloc(solver._h, OPS[op], int(level), din.ptr, dout.ptr, a.size)
    _lib.check(rc)
    _lib.check(L.ibpm_synchronize(solver._h))
    out = dout.download()
    assert np.array_equal(din.download(), a), "input modified"
    din.free()
    dout.free()
    return out


def _check_ops(solver, nlevel, seed, ops=("vort2flux", "poisson", "helmholtz")):
    rng = np.random.default_rng(seed)
    n = solver.read("Gamma").size // nlevel
    for op in ops:
        x = rng.standard_normal(n * nlevel if op == "vort2flux" else n)
        level = 2 if op == "helmholtz" else 1
        ref = solver.apply(op, x, level=level)
        got = _noalloc(solver, op, x, level=level)
        print(op, "max |diff|", np.abs(got - ref).max(), flush=True)
        assert np.array_equal(got, ref), op


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_apply_twice_second_smaller(precision):
    big = ib.create_solver(cases.no_body(nx=256, ny=256, nlevel=3, h=0.0125, x0=(-1.6, -1.6)),
                           precision=precision)
    small = ib.create_solver(cases.no_body(nx=96, ny=80, nlevel=2, h=0.03, x0=(-1.44, -1.2)),
                             precision=precision)
    print("case 256^2 x 3 then 96x80 x 2", precision, flush=True)
    _check_ops(big, 3, seed=1)
    _check_ops(small, 2, seed=2)
    # and back: the larger call after the smaller one
    _check_ops(big, 3, seed=3, ops=("vort2flux",))
    big.close()
    small.close()


def test_noalloc_rejects_operators_beyond_its_scratch():
    s = ib.create_solver(cases.no_body(nx=1024, ny=1024, nlevel=5, h=1.0 / 128, x0=(-4.0, -4.0)))
    print("case 1024^2 x 5 vort2flux over the scratch limit", flush=True)
    n = s.read("Gamma").size
    din, dout = DeviceArray(n), DeviceArray(n)
    L = _lib.load()
    rc = L.ibpm_apply_noalloc(s._h, _lib.IBPM_OP_VORT2FLUX, 1, din.ptr, dout.ptr, n)
    assert rc == _lib.IBPM_E_ARG
    assert b"bytes of scratch" in L.ibpm_last_error()
    # overlapping buffers are refused too
    rc = L.ibpm_apply_noalloc(s._h, _lib.IBPM_OP_POISSON, 1, din.ptr, din.ptr, n // 5)
    assert rc == _lib.IBPM_E_ARG
    din.free()
    dout.free()
    # the single-plane solves of this grid fit (two 8.4 MB planes)
    x = np.random.default_rng(4).standard_normal(n // 5)
    assert np.array_equal(_noalloc(s, "poisson", x), s.apply("poisson", x))
    s.close()
