"""Batched ensembles: many independent solvers advanced together on one GPU.

No counterpart in ImmersedBodies.jl, where each CnabSolver advances alone
(solver.jl:155-190). ``Batch`` wraps ``ibpm_batch_*`` (include/ibpm.h): one step of every
solver is captured into a hipGraph and each batch step is one replay. Problems may
differ in grid size, levels, Reynolds number and bodies (static bodies or none); every
solver ends in the state it would reach advanced alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .problem import timestep


class Batch:
    def __init__(self, solvers, freestream=None):
        """solvers: HipCnabSolver list (static bodies). freestream: nprob x 2 velocities,
        fixed for the batch; default: each problem's freestream_vel at its next time."""
        self.solvers = list(solvers)
        n = len(self.solvers)
        if freestream is None:
            fs = []
            for s in self.solvers:
                st = s.state
                fs.append(st.prob.fluid.freestream_vel(st.t0 + timestep(st.prob) * st.index))
            freestream = fs
        self.freestream = np.ascontiguousarray(np.asarray(freestream, dtype=np.float64).reshape(n, 2))
        handles = (C.c_void_p * n)(*[s._h.value for s in self.solvers])
        h = C.c_void_p()
        _lib.check(_lib.load().ibpm_batch_create(
            n, handles, self.freestream.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h)))
        self._h = h

    def advance(self, nsteps: int) -> np.ndarray:
        """Advance every problem nsteps steps; returns the total forces, nsteps x nprob x 2."""
        n = len(self.solvers)
        forces = np.zeros((max(int(nsteps), 0), n, 2))
        _lib.check(_lib.load().ibpm_batch_advance(
            self._h, int(nsteps), forces.ctypes.data_as(C.POINTER(C.c_double))))
        for s, fs in zip(self.solvers, self.freestream):
            s.state.set_time_index(s.state.index + int(nsteps))
            s.state.freestream_vel = fs.copy()
        return forces

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().ibpm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
