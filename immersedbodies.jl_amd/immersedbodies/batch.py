"""Batched ensembles: many independent solvers advanced together on one GPU.

No counterpart in ImmersedBodies.jl, where each CnabSolver advances alone
(solver.jl:155-190). ``Batch`` wraps ``ibpm_batch_*`` (include/ibpm.h): one step of every
solver is captured into hipGraphs; a batch step is one replay, or two around a host check
of the CG state when bodies move (a CG the captured budget of iterations leaves short of
its stopping test continues up to its own maxiter in between). Problems may
differ in grid size, levels, Reynolds number, time step and bodies (static or moving)
and may have static or moving grids; every solver ends in the state it would reach
advanced alone. Each step's freestream, grid motion and body motion are evaluated here
on the host at every problem's own new time, as ``HipCnabSolver.advance`` does, and the
graph reads them from device tables.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .problem import MovingGrid, timestep


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Batch:
    def __init__(self, solvers):
        """solvers: HipCnabSolver list. Moving-body solvers need box_update="off", no
        warm-start recording, three steps of history for the extrapolated or learned
        guess and, when preconditioned, one step taken and no preconditioner refresh (the
        library names what is missing)."""
        self.solvers = list(solvers)
        n = len(self.solvers)
        handles = (C.c_void_p * n)(*[s._h.value for s in self.solvers])
        h = C.c_void_p()
        _lib.check(_lib.load().ibpm_batch_create(n, handles, None, C.byref(h)))
        self._h = h
        self._nbody = [len(s.state.prob.bodies.sections) for s in self.solvers]
        self._moving = any(s._moving for s in self.solvers)
        self.last_stats = None

    def _inputs(self, nsteps):
        n = len(self.solvers)
        fs = np.zeros((nsteps, n, 8))
        mrow = 6 * sum(self._nbody)
        mot = np.zeros((nsteps, mrow)) if self._moving else None
        for p, s in enumerate(self.solvers):
            st = s.state
            prob = st.prob
            dt = timestep(prob)
            gm = prob.fluid.grid_motion
            off = 6 * sum(self._nbody[:p])
            nb = self._nbody[p]
            for k in range(nsteps):
                t = st.t0 + dt * (st.index + k)  # time after inc_time_index!
                fs[k, p, :2] = prob.fluid.freestream_vel(t)
                if isinstance(gm, MovingGrid):
                    fs[k, p, 2:] = gm(t).as_array()
                if mot is not None and s._moving:
                    mot[k, off:off + 6 * nb] = prob.bodies.motion_array(t)[:6 * nb]
        return fs, mot

    def advance(self, nsteps: int) -> np.ndarray:
        """Advance every problem nsteps steps; returns the total forces, nsteps x nprob x 2
        (per-step statistics in ``last_stats``, nsteps x nprob)."""
        n = len(self.solvers)
        nsteps = max(int(nsteps), 0)
        fs, mot = self._inputs(nsteps)
        forces = np.zeros((nsteps, n, 2))
        stats = (_lib.StepStats * max(nsteps * n, 1))()
        _lib.check(_lib.load().ibpm_batch_advance(
            self._h, nsteps, _dp(fs), _dp(mot) if mot is not None else None, _dp(forces),
            stats))
        for p, s in enumerate(self.solvers):
            s.state.set_time_index(s.state.index + nsteps)
            if nsteps:
                s.state.freestream_vel = fs[-1, p, :2].copy()
        self.last_stats = [[stats[k * n + p] for p in range(n)] for k in range(nsteps)]
        return forces

    def cg_iters(self) -> np.ndarray:
        """CG iterations of the last advance, nsteps x nprob."""
        return np.array([[x.cg_iters for x in row] for row in self.last_stats])

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().ibpm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
