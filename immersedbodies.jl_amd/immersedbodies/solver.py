"""Solver protocol of ImmersedBodies.jl (solver.jl) over the HIP library.

``create_solver``, ``advance``, ``getstate``, ``solve`` and ``solve_`` (``solve!``) keep
the reference's names, argument meanings and time-index convention (solver.jl:1-76,
problem.jl:703-708). ``HipCnabSolver`` stands where ``CnabSolver`` (solver.jl:78-190)
stands: its state lives in HBM and is copied to the host ``State`` only when
``getstate`` (or ``sync_state``) is called; forces are returned every step.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Optional

import numpy as np

from . import _lib
from .problem import MovingGrid, Problem, State, timestep


class AbstractSolver:
    """Solves a Problem (solver.jl:6): implement ``advance`` and ``getstate``."""

    def advance(self):
        raise NotImplementedError

    def getstate(self) -> State:
        raise NotImplementedError


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class HipCnabSolver(AbstractSolver):
    """CNAB IBPM solver running on the GPU through libibpm.

    Keyword arguments:
      precision: "fp64" (sine contractions on FP64 MFMA) or "mixed" (3xBF16 split products
                 with FP32 accumulation and FP64 residual refinement)
      refine_sweeps: refinement sweeps per solve in mixed precision
      cg: dict of maxiter / reltol / abstol for moving bodies (solver.jl:137 defaults
          5000 / 1e-12 / 0); cg_tol_mode 0 is IterativeSolvers' reltol*||r0||,
          1 is reltol*||b||
      cg_precond: moving bodies: precondition cg! with the inverse of B assembled at the
                  first step (2n applications of B), applied in each body's frame; CG
                  still stops on the true residual. cg_precond_refresh > 0 reassembles
                  after a step that needed more iterations than that.
      device: HIP device ordinal
      dst_tiles: sine-pass tile variant; -1 autotunes at creation (over the deepest fold
                 the lengths allow), 0..23 fix it (mixed 0..19): for v < 16, v % 4 picks
                 (TB, TM) = (64,64), (64,32), (32,64), (32,32), FP64 loads run 1
                 (v % 8 < 4) or 2 K chunks ahead, v >= 8 folds the sine matrix twice;
                 16..19 fold twice and copy the sine tables straight into LDS, (TB, TM,
                 KC) = (32,64,16), (32,64,8), (32,32,16), (32,32,8); 20..23 (FP64) fold
                 three times with the same tiles
      helmholtz: "auto" solves the implicit Helmholtz systems by red-black Gauss-Seidel
                 in FP64 where 64 half-sweeps bring the error below 2^-56 max|x| (Jacobi
                 contraction 4k/(1+4k) <= 0.54, k = dt / (2 Re h^2)), by the sine
                 transforms elsewhere; "sine" always by the sine transforms
      box_update: "auto" (FP64, Gauss-Seidel on level 1): the final vort2flux of a
                  step transforms only the change of Gamma since the trial one, which is
                  zero outside a box around the bodies, the trial forms psi_1 only where
                  the panels read it, and B runs on the same boxes (which follow moving
                  bodies); "off"; "force": wherever valid, cheaper or not (tests)
    """

    def __init__(self, state: State, *, precision="fp64", refine_sweeps=2, cg=None,
                 cg_tol_mode=0, device=0, dst_tiles=-1, cg_precond=False,
                 cg_precond_refresh=0, helmholtz="auto", box_update="auto"):
        L = _lib.load()
        self.state = state
        prob = state.prob
        g = prob.fluid.grid
        bodies = prob.bodies
        n = bodies.npanel
        self._n = n
        self._nbody = len(bodies.sections)
        self._moving = not bodies.static
        self._grid_moving = isinstance(prob.fluid.grid_motion, MovingGrid)
        xb = np.ascontiguousarray(bodies.xb().T.reshape(-1))  # column-major n x 2
        ds = np.ascontiguousarray(bodies.ds())
        bs = np.array([s[2] for s in bodies.sections] or [0], dtype=np.int32)
        bc = np.array([s[3] for s in bodies.sections] or [0], dtype=np.int32)
        bm = np.array([int(s[1].moving) for s in bodies.sections] or [0], dtype=np.int32)
        self._keep = (xb, ds, bs, bc, bm)
        p = _lib.Problem(g.base.n[0], g.base.n[1], g.nlevel, g.base.x0[0], g.base.x0[1],
                         g.base.h, prob.fluid.Re, timestep(prob), state.t0, n,
                         _dp(xb) if n else None, _dp(ds) if n else None, self._nbody,
                         bs.ctypes.data_as(C.POINTER(C.c_int)),
                         bc.ctypes.data_as(C.POINTER(C.c_int)),
                         bm.ctypes.data_as(C.POINTER(C.c_int)),
                         int(self._grid_moving))
        o = _lib.Options()
        L.ibpm_default_options(C.byref(o))
        o.precision = {"fp64": _lib.IBPM_FP64, "mixed": _lib.IBPM_MIXED}[precision]
        o.refine_sweeps = int(refine_sweeps)
        cg = dict(cg or {})
        o.cg_maxiter = int(cg.get("maxiter", 5000))
        o.cg_reltol = float(cg.get("reltol", 1e-12))
        o.cg_abstol = float(cg.get("abstol", 0.0))
        o.cg_tol_mode = int(cg_tol_mode)
        o.device = int(device)
        o.dst_tiles = int(dst_tiles)
        o.cg_precond = int(bool(cg_precond))
        o.cg_precond_refresh = int(cg_precond_refresh)
        o.helmholtz = {"auto": 0, "sine": 1}[helmholtz]
        o.box_update = {"auto": 0, "off": 1, "force": 2}[box_update]
        h = C.c_void_p()
        _lib.check(L.ibpm_create(C.byref(p), C.byref(o), C.byref(h)))
        self._h = h
        self.precision = precision

        self.last_stats = None
        # a State created before the solver may carry a non-zero index (resume)
        if state.index != 0:
            _lib.check(L.ibpm_set_index(self._h, state.index))
            self.write_state(state)

    # -- lifetime --
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().ibpm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping --
    def advance(self, nsteps: int = 1):
        """advance! (solver.jl:155-190), ``nsteps`` times. Freestream and body motion are
        evaluated on the host at each new time and passed to the library."""
        st = self.state
        prob = st.prob
        dt = timestep(prob)
        w = 8 if self._grid_moving else 2  # freestream, then the grid's GridVelocity
        fs = np.zeros(w * nsteps)
        mot = np.zeros(6 * max(self._nbody, 1) * nsteps) if self._moving else None
        for k in range(nsteps):
            t = st.t0 + dt * (st.index + k)  # time after inc_time_index!
            fs[w * k:w * k + 2] = prob.fluid.freestream_vel(t)
            if self._grid_moving:
                fs[w * k + 2:w * k + 8] = prob.fluid.grid_motion(t).as_array()
            if self._moving:
                mot[6 * self._nbody * k:6 * self._nbody * (k + 1)] = \
                    prob.bodies.motion_array(t)[:6 * self._nbody]
        stats = (_lib.StepStats * nsteps)()
        rc = _lib.load().ibpm_advance(self._h, nsteps, _dp(fs),
                                      _dp(mot) if mot is not None else None, stats)
        _lib.check(rc)
        st.set_time_index(st.index + nsteps)
        st.freestream_vel = fs[w * (nsteps - 1):w * (nsteps - 1) + 2].copy() if nsteps \
            else st.freestream_vel
        self.last_stats = stats
        return stats

    def set_body_motion(self, body: int, t: float, transform):
        """Stage body ``body``'s RigidBodyTransform at time ``t`` (the next step's time) for
        the next ``advance_staged``; the library checks the time."""
        pos = (C.c_double * 2)(*transform.pos)
        vel = (C.c_double * 2)(*transform.vel)
        _lib.check(_lib.load().ibpm_set_body_motion(self._h, int(body), float(t), pos, vel,
                                                    float(transform.angle),
                                                    float(transform.angular_vel)))

    def advance_staged(self):
        """One advance! with the motion staged by set_body_motion (the freestream and grid
        motion still come from the problem)."""
        st = self.state
        prob = st.prob
        t = st.t0 + timestep(prob) * st.index
        w = 8 if self._grid_moving else 2
        fs = np.zeros(w)
        fs[:2] = prob.fluid.freestream_vel(t)
        if self._grid_moving:
            fs[2:] = prob.fluid.grid_motion(t).as_array()
        stats = (_lib.StepStats * 1)()
        _lib.check(_lib.load().ibpm_advance(self._h, 1, _dp(fs), None, stats))
        st.set_time_index(st.index + 1)
        st.freestream_vel = fs[:2].copy()
        self.last_stats = stats
        return stats

    def forces(self, stats=None):
        stats = self.last_stats if stats is None else stats
        return np.array([[s.force[0], s.force[1]] for s in stats])

    def cg_iters(self, stats=None):
        stats = self.last_stats if stats is None else stats
        return np.array([s.cg_iters for s in stats])

    def b_applies(self, stats=None):
        stats = self.last_stats if stats is None else stats
        return np.array([s.b_applies for s in stats])

    def pc_assembled(self, stats=None):
        stats = self.last_stats if stats is None else stats
        return np.array([s.pc_assembled for s in stats])

    def dst_tiles(self, nplanes: int, xdir: bool = True):
        """(TB, TM, PF, fold, KC): tile shape, prefetch depth (0: tables copied straight
        into LDS), fold (2 or 4) and K chunk of the sine passes over ``nplanes`` planes
        (autotuned)."""
        tb, tm, pf, fold, kc = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _lib.check(_lib.load().ibpm_dst_tiles(self._h, int(nplanes), int(bool(xdir)),
                                              C.byref(tb), C.byref(tm), C.byref(pf),
                                              C.byref(fold), C.byref(kc)))
        return tb.value, tm.value, pf.value, fold.value, kc.value

    def box_update(self, level: int = 1):
        """None when the steps transform Gamma again in the final vort2flux; else the box
        (first column, columns, first row, rows) of the change of Gamma on ``level``.
        level 0: the box of level 1 on which the trial vort2flux computes psi (None when it
        runs the inverse transforms)."""
        box = (C.c_int * 4)()
        on = _lib.load().ibpm_box_update(self._h, int(level), box)
        return tuple(box) if on else None

    def set_warm_start(self, mode: str = "previous", net=None, record: bool = False):
        """Initial guess of the moving-body CG: "previous" (the reference's previous F̃b),
        "extrapolate" (quadratic extrapolation of the last three F̃b) or "learned"
        (extrapolation corrected by ``net``, a ``WarmStartNet``). ``record`` appends one
        training sample per step to ``net``."""
        from .warmstart import MODES

        self._net = net  # keep the network alive while the solver uses it
        _lib.check(_lib.load().ibpm_set_warm_start(self._h, MODES[mode],
                                                   net.handle if net is not None else None,
                                                   int(bool(record))))

    # -- state exchange --
    def read(self, name: str) -> np.ndarray:
        L = _lib.load()
        f = _lib.FIELDS[name]
        n = L.ibpm_field_size(self._h, f)
        out = np.zeros(n)
        _lib.check(L.ibpm_read(self._h, f, _dp(out), n))
        return out

    def write(self, name: str, values):
        L = _lib.load()
        f = _lib.FIELDS[name]
        a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        _lib.check(L.ibpm_write(self._h, f, _dp(a), a.size))

    def apply(self, op: str, x, level: int = 1) -> np.ndarray:
        """Apply one building block of the step on the GPU: "poisson" (Delta_inv),
        "helmholtz" (Ainv of a 1-based level) or "vort2flux" (psi of all levels)."""
        code = {"poisson": _lib.IBPM_OP_POISSON, "helmholtz": _lib.IBPM_OP_HELMHOLTZ,
                "vort2flux": _lib.IBPM_OP_VORT2FLUX}[op]
        a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
        out = np.zeros_like(a)
        _lib.check(_lib.load().ibpm_apply(self._h, code, int(level), _dp(a), _dp(out), a.size))
        return out

    def sync_state(self) -> State:
        st = self.state
        g = st.prob.fluid.grid
        Lv = g.nlevel
        n = self._n
        st.q = self.read("q").reshape(Lv, -1)
        st.q0 = self.read("q0").reshape(Lv, -1)
        st.Gamma = self.read("Gamma").reshape(Lv, -1)
        st.psi = self.read("psi").reshape(Lv, -1)
        st.nonlin = [self.read("nonlin1").reshape(Lv, -1), self.read("nonlin2").reshape(Lv, -1)]
        st.xb = self.read("xb").reshape(2, n).T.copy() if n else np.zeros((0, 2))
        st.ub = self.read("ub").reshape(2, n).T.copy() if n else np.zeros((0, 2))
        st.ds = self.read("ds") if n else np.zeros(0)
        st.Ftilde = self.read("Ftilde")
        st.fb = self.read("fb").reshape(2, n).T.copy() if n else np.zeros((0, 2))
        st.freestream_vel = self.read("freestream")
        return st

    def write_state(self, st: State):
        """Push a host State (e.g. from load_state) to the device."""
        n = self._n
        self.write("q", st.q)
        self.write("q0", st.q0)
        self.write("Gamma", st.Gamma)
        self.write("psi", st.psi)
        self.write("nonlin1", st.nonlin[0])
        self.write("nonlin2", st.nonlin[1])
        if n:
            self.write("xb", st.xb.T)
            self.write("ub", st.ub.T)
            self.write("ds", st.ds)
            self.write("Ftilde", st.Ftilde)
            self.write("fb", st.fb.T)
        self.write("freestream", st.freestream_vel)
        _lib.check(_lib.load().ibpm_set_index(self._h, st.index))

    def getstate(self) -> State:
        return self.sync_state()


CnabSolver = HipCnabSolver  # the reference's name for the CNAB solver (solver.jl:78)


def create_solver(x, **kw) -> HipCnabSolver:
    """create_solver(prob_or_state; kw...) (solver.jl:28,152)."""
    state = State(x) if isinstance(x, Problem) else x
    return HipCnabSolver(state, **kw)


def advance(solver: AbstractSolver, nsteps: int = 1):
    """advance!(solver)."""
    return solver.advance(nsteps)


def getstate(solver: AbstractSolver) -> State:
    return solver.getstate()


def solve_(f: Optional[Callable], solver: AbstractSolver, *, t: float, save=None,
           log: bool = False, chunk: int = 1):
    """solve!(f, solver; t, save, log) (solver.jl:31-63): n = round((t - state.t)/dt) steps,
    calling f(state, i, n) after each step and handing the state to the ``save`` saver
    (a SolutionSaver, saving.py) at the timesteps it wants. ``log`` prints every tenth step
    as the reference does (its default is on; here off).

    For the GPU solver the host State is refreshed from the device only where a saver
    wants it (call ``solver.sync_state()`` inside ``f`` to read fields there), and
    ``chunk`` > 1 advances groups of steps between callbacks (stopping early at a saved
    timestep). Returns the force history of the GPU solver (one row per step)."""
    from .saving import SaveNothing

    hip = isinstance(solver, HipCnabSolver)
    st = solver.state if hip else solver.getstate()
    dt = timestep(st.prob)
    n = int(round((t - st.t) / dt))
    saver = (save if save is not None else SaveNothing()).init(st.index + n)
    forces = []
    try:
        i = 0
        while i < n:
            if hip:
                k = min(max(1, int(chunk)), n - i)
                for idx in range(st.index + 1, st.index + k):
                    if saver.wants(idx):
                        k = idx - st.index
                        break
                forces.append(solver.forces(solver.advance(k)))
            else:
                k = 1
                solver.advance()
            i += k
            if saver.wants(st.index):
                if hip:
                    solver.sync_state()
                saver.update(st)
            if f is not None:
                f(st, i, n)
            if log and i % 10 == 1:
                print("i = ", st.index, "  t = ", st.t)
    finally:
        saver.finalize()
    return np.concatenate(forces) if forces else np.zeros((0, 2))


def solve(f: Optional[Callable], prob: Problem, *, t: float, solver_kw=None, **kw):
    """solve(f, prob; t, solver_kw) (solver.jl:71-76)."""
    solver = create_solver(State(prob), **(solver_kw or {}))
    return solve_(f, solver, t=t, **kw)
