"""Problem description: the host-side mirror of ImmersedBodies.jl's ``problem.jl``.

The names, argument meanings and errors follow the reference so a user of the Julia
package finds the same objects (paths relative to the reference repository):

* ``CartesianGrid``, ``MultiDomainGrid``, ``fluid_grid``, ``gridstep``, ``subdomain``,
  ``coords``, ``GridIndices``                               problem.jl:269-472
* ``Fluid``, ``default_gridstep``                            problem.jl:534-556
* ``StaticGrid``, ``MovingGrid``, ``GridVelocity``           problem.jl:476-523
* ``CNAB``, ``default_scheme``, ``timestep``                 problem.jl:566-596
* ``Panels``, ``StaticBody``, ``MovingRigidBody``,
  ``RigidBodyTransform``, ``Bodies``, ``any_fsi``,
  ``panel_range``, ``eachbody``                              problem.jl:1-261
* ``Problem``, ``State``                                     problem.jl:598-708
* ``boundary_total_force``                                   problem.jl:853-856
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np


def _floor_sig1(x: float) -> float:
    """floor(x; sigdigits=1) of Julia."""
    if x == 0:
        return 0.0
    e = math.floor(math.log10(abs(x)))
    p = 10.0 ** e
    return math.floor(x / p) * p


@dataclass(frozen=True)
class CartesianGrid:
    """Cartesian grid with lower-left corner ``x0``, cell counts ``n`` and step ``h``
    (problem.jl:269-281)."""

    x0: Tuple[float, float]
    n: Tuple[int, int]
    h: float

    @staticmethod
    def from_limits(h, xlims, ylims) -> "CartesianGrid":
        x0 = (float(xlims[0]), float(ylims[0]))
        n = (int(round((xlims[1] - xlims[0]) / h)), int(round((ylims[1] - ylims[0]) / h)))
        return CartesianGrid(x0, n, float(h))


@dataclass(frozen=True)
class GridIndices:
    """Sizes and boundary offsets of the flattened fields (problem.jl:363-389)."""

    nx: int
    ny: int

    @property
    def nu(self):
        return (self.nx + 1) * self.ny

    @property
    def nv(self):
        return self.nx * (self.ny + 1)

    @property
    def nq(self):
        return self.nu + self.nv

    @property
    def nG(self):
        return (self.nx - 1) * (self.ny - 1)

    @property
    def L(self):
        return 0

    @property
    def R(self):
        return self.ny + 1

    @property
    def B(self):
        return 2 * (self.ny + 1)

    @property
    def T(self):
        return 2 * (self.ny + 1) + self.nx + 1


@dataclass(frozen=True)
class MultiDomainGrid:
    """``nlevel`` grids with a common centre, level ``l`` having step ``2^(l-1) h``
    (problem.jl:398-429)."""

    base: CartesianGrid
    nlevel: int

    @property
    def inds(self) -> GridIndices:
        return GridIndices(*self.base.n)

    @property
    def h(self):
        return self.base.h


@dataclass(frozen=True)
class MultiDomainExtents:
    """The region of a MultiDomainGrid without its step (problem.jl:450-457); Fluid picks
    the step from the Reynolds number."""

    xlims: Tuple[float, float]
    ylims: Tuple[float, float]
    nlevel: int


def fluid_grid(*, h=None, xlims, ylims, nlevel):
    """``fluid_grid(; h, xlims, ylims, nlevel)`` (problem.jl:459-475). Without ``h`` the
    step is chosen by ``Fluid`` from the Reynolds number."""
    if h is None:
        return MultiDomainExtents(tuple(xlims), tuple(ylims), int(nlevel))
    return MultiDomainGrid(CartesianGrid.from_limits(h, xlims, ylims), int(nlevel))


# grid point kinds (GridPoints, problem.jl:297-301) as the names coords / gridsize take
GridVertices = "vertices"
GridU = "u"
GridV = "v"
GridΓ = GridGamma = "gamma"


def gridstep(grid, level: int = 1) -> float:
    """Grid step on level ``level`` (problem.jl:416-417)."""
    if isinstance(grid, CartesianGrid):
        return grid.h
    if isinstance(grid, Problem):
        return gridstep(grid.fluid.grid, level)
    return 2.0 ** (level - 1) * grid.base.h


def subdomain(grid: MultiDomainGrid, level: int) -> CartesianGrid:
    """Level ``level`` as a CartesianGrid (problem.jl:424-429)."""
    b = grid.base
    h = 2.0 ** (level - 1) * b.h
    x0 = tuple(b.x0[i] + b.n[i] * (b.h - h) / 2 for i in range(2))
    return CartesianGrid(x0, b.n, h)


def coords(grid, pts="vertices", levels=None):
    """Coordinates of grid points: ``vertices``, ``u``, ``v`` or ``gamma``, or those of an
    accessor (x_velocity, y_velocity, vorticity) (problem.jl:311-338, 432-441, 776-787).
    For a MultiDomainGrid, one (x, y) pair per level in ``levels`` (1-based, all by
    default)."""
    if callable(pts):
        pts = GRID_POINTS[pts]
    if isinstance(grid, MultiDomainGrid):
        return [coords(subdomain(grid, l), pts) for l in _level_list(grid, levels)]
    xs = grid.x0[0] + grid.h * np.arange(grid.n[0] + 1)
    ys = grid.x0[1] + grid.h * np.arange(grid.n[1] + 1)
    mid = lambda r: r[:-1] + grid.h / 2
    if pts == "vertices":
        return xs, ys
    if pts == "u":
        return xs, mid(ys)
    if pts == "v":
        return mid(xs), ys
    if pts == "gamma":
        return xs[1:-1], ys[1:-1]
    raise ValueError(pts)


def extents(grid: CartesianGrid):
    """((xmin, xmax), (ymin, ymax)) of a grid (problem.jl:290-295)."""
    return tuple((grid.x0[i], grid.x0[i] + grid.n[i] * grid.h) for i in range(2))


def gridsize(grid, pts: str = "vertices", levels=None):
    """gridsize(grid, pts, levels) as the reference defines it (problem.jl:304-340,
    443-448): n for vertices, (n1, n2 - 1) for u, (n1 - 1, n2) for v and n - 2 for gamma,
    plus the number of levels for a MultiDomainGrid. These are one less in each direction
    than the arrays ``coords`` describes (n + 1 vertices, the flux and circulation arrays),
    exactly as in the reference; the accessors below size their arrays from ``coords``."""
    if isinstance(grid, MultiDomainGrid):
        lev = _level_list(grid, levels)
        return gridsize(grid.base, pts) + (len(lev),)
    n1, n2 = grid.n
    return {"vertices": (n1, n2), "u": (n1, n2 - 1), "v": (n1 - 1, n2),
            "gamma": (n1 - 2, n2 - 2)}[pts]


def _level_list(grid, levels):
    """1-based levels: None (all), an int, or an iterable."""
    if levels is None:
        return list(range(1, grid.nlevel + 1))
    if isinstance(levels, (int, np.integer)):
        return [int(levels)]
    return [int(l) for l in levels]


def default_gridstep(Re: float) -> float:
    """floor(2/Re; sigdigits=1) (problem.jl:556)."""
    return _floor_sig1(2.0 / Re)


class GridMotion:
    """Motion of the grid in the global frame: StaticGrid or MovingGrid (problem.jl:492-502)."""


@dataclass
class StaticGrid(GridMotion):
    """A grid that is static in the global frame (problem.jl:508-513)."""


@dataclass(frozen=True)
class GridVelocity:
    """Motion and orientation of the grid relative to the global frame (problem.jl:476-490):
    centre of rotation, velocity of that centre, angle and angular velocity."""

    center: Tuple[float, float] = (0.0, 0.0)
    vel: Tuple[float, float] = (0.0, 0.0)
    angle: float = 0.0
    angular_vel: float = 0.0

    def as_array(self):
        return np.array([self.center[0], self.center[1], self.vel[0], self.vel[1], self.angle,
                         self.angular_vel], dtype=np.float64)


@dataclass
class MovingGrid(GridMotion):
    """A grid that moves in the global frame: ``motion(t)`` returns its GridVelocity
    (problem.jl:515-523)."""

    motion: Callable[[float], GridVelocity]

    def __call__(self, t):
        return self.motion(t)


@dataclass
class Fluid:
    """``Fluid(; grid, Re, freestream_vel, grid_motion)`` (problem.jl:534-552)."""

    grid: object
    Re: float
    freestream_vel: Callable[[float], Tuple[float, float]] = lambda t: (0.0, 0.0)
    grid_motion: object = field(default_factory=StaticGrid)

    def __post_init__(self):
        if isinstance(self.grid, MultiDomainExtents):
            e = self.grid
            self.grid = MultiDomainGrid(
                CartesianGrid.from_limits(default_gridstep(self.Re), e.xlims, e.ylims), e.nlevel)
        if not isinstance(self.grid_motion, (StaticGrid, MovingGrid)):
            raise TypeError("grid_motion must be a StaticGrid or a MovingGrid")


class AbstractScheme:
    """A time-stepping scheme (problem.jl:559-564)."""


@dataclass(frozen=True)
class CNAB(AbstractScheme):
    """2-step Crank-Nicolson / Adams-Bashforth scheme (problem.jl:566-577)."""

    dt: float
    n: int = 2

    def __post_init__(self):
        if self.n != 2:
            raise ValueError("only 2-step CNAB is currently supported")

    @property
    def beta(self):
        return (1.5, -0.5)


def default_scheme(grid, *, Umax, cfl=0.1) -> CNAB:
    """dt = floor(cfl h / Umax; sigdigits=1) (problem.jl:590-596)."""
    return CNAB(dt=_floor_sig1(cfl * gridstep(grid) / Umax))


def timestep(x) -> float:
    if isinstance(x, CNAB):
        return x.dt
    return x.scheme.dt


@dataclass
class Panels:
    """Panel positions ``xb`` (n x 2) and lengths ``ds`` (n) (problem.jl:9-12)."""

    xb: np.ndarray
    ds: np.ndarray

    def __post_init__(self):
        self.xb = np.asarray(self.xb, dtype=np.float64).reshape(-1, 2)
        self.ds = np.asarray(self.ds, dtype=np.float64).reshape(-1)

    @property
    def n(self):
        return self.xb.shape[0]


@dataclass(frozen=True)
class RigidBodyTransform:
    """Position, velocity, angle and angular velocity of a rigid body (problem.jl:233-239)."""

    pos: Tuple[float, float]
    vel: Tuple[float, float]
    angle: float
    angular_vel: float

    def as_array(self):
        return np.array([self.pos[0], self.pos[1], self.vel[0], self.vel[1], self.angle,
                         self.angular_vel], dtype=np.float64)


class AbstractBody:
    """A body immersed in the fluid (problem.jl:47-53)."""


class PresetBody(AbstractBody):
    """A body whose motion is prescribed (problem.jl:70-75)."""


class FsiBody(AbstractBody):
    """A body moved by the fluid (problem.jl:77-92). The reference defines no concrete
    one; neither does this package."""


@dataclass
class PanelState:
    """Positions xb (n x 2), lengths ds and velocities ub (n x 2) of panels
    (problem.jl:21-45)."""

    xb: np.ndarray
    ds: np.ndarray
    ub: np.ndarray

    @property
    def n(self):
        return self.xb.shape[0]


@dataclass
class StaticBody(PresetBody):
    """A body fixed at ``panels`` (problem.jl:220-226)."""

    panels: Panels

    @property
    def moving(self):
        return False


@dataclass
class MovingRigidBody(PresetBody):
    """A rigid body at ``motion(t)`` applied to ``panels`` (problem.jl:241-261)."""

    motion: Callable[[float], RigidBodyTransform]
    panels: Panels

    @property
    def moving(self):
        return True


class Bodies:
    """``Bodies(; name=body, ...)`` (problem.jl:136-178): panels of all bodies are
    concatenated in the order given."""

    def __init__(self, **bodies):
        self.byname: Dict[str, object] = dict(bodies)
        self.sections: List[Tuple[str, object, int, int]] = []
        off = 0
        for name, body in bodies.items():
            n = body.panels.n
            self.sections.append((name, body, off, n))
            off += n
        self.npanel = off
        self.static = all(not b.moving for b in bodies.values())

    def panel_range(self, name):
        for nm, _, off, n in self.sections:
            if nm == name:
                return range(off, off + n)
        raise KeyError(name)

    def __getitem__(self, name):
        return self.byname[name]

    def xb(self):
        if not self.sections:
            return np.zeros((0, 2))
        return np.concatenate([b.panels.xb for _, b, _, _ in self.sections], axis=0)

    def ds(self):
        if not self.sections:
            return np.zeros(0)
        return np.concatenate([b.panels.ds for _, b, _, _ in self.sections])

    def motion_array(self, t: float) -> np.ndarray:
        """6 values per body at time t (zeros for static bodies)."""
        out = np.zeros(6 * max(len(self.sections), 1))
        for k, (_, b, _, _) in enumerate(self.sections):
            if b.moving:
                out[6 * k:6 * k + 6] = b.motion(t).as_array()
        return out


@dataclass(frozen=True)
class PanelSection:
    """A body's place among the concatenated panels (problem.jl:94-130): its 1-based
    index among the bodies, the body, and its 0-based panel range."""

    index: int
    body: object
    panel_range: range


def panel_section(bodies: Bodies, name) -> PanelSection:
    """The PanelSection of body ``name`` (problem.jl:198-203)."""
    for k, (nm, body, off, n) in enumerate(bodies.sections):
        if nm == name:
            return PanelSection(k + 1, body, range(off, off + n))
    raise KeyError(name)


def any_fsi(bodies: Bodies) -> bool:
    """Whether any body is an FSI body (problem.jl:186-191). The reference declares the
    abstract FsiBody (problem.jl:75-92) but defines no concrete one, so this is False."""
    return False


def panel_range(bodies: Bodies, name) -> range:
    """0-based index range of body ``name`` in the concatenated panels (problem.jl:103-111;
    the reference's 1-based range shifted by one)."""
    return bodies.panel_range(name)


def eachbody(bodies: Bodies):
    """Mapping of body names to bodies (problem.jl:180-184)."""
    return dict(bodies.byname)


def n_panels(x) -> int:
    if isinstance(x, Bodies):
        return x.npanel
    return x.panels.n


@dataclass
class Problem:
    """``Problem(; fluid, bodies, scheme)`` (problem.jl:598-603)."""

    fluid: Fluid
    bodies: Bodies
    scheme: CNAB


class State:
    """Host copy of the solver state (problem.jl:626-665). Arrays use the reference's
    column-major layout, stored here as (nlevel, n) row blocks: ``q[l]`` is the flux
    column of level l+1."""

    def __init__(self, prob: Problem, t0: float = 0.0):
        g = prob.fluid.grid
        ind = g.inds
        L = g.nlevel
        n = prob.bodies.npanel
        self.prob = prob
        self.t0 = float(t0)
        self.q = np.zeros((L, ind.nq))
        self.q0 = np.zeros((L, ind.nq))
        self.Gamma = np.zeros((L, ind.nG))
        self.psi = np.zeros((L, ind.nG))
        self.nonlin = [np.zeros((L, ind.nG)), np.zeros((L, ind.nG))]
        self.xb = np.zeros((n, 2))
        self.ds = np.zeros(n)
        self.ub = np.zeros((n, 2))
        self.Ftilde = np.zeros(2 * n)
        self.fb = np.zeros((n, 2))
        self.freestream_vel = np.zeros(2)
        self.set_time_index(0)

    @property
    def panels(self) -> PanelState:
        """state.panels of the reference: the panel arrays of this State."""
        return PanelState(self.xb, self.ds, self.ub)

    @property
    def Γ(self):
        return self.Gamma

    @property
    def ψ(self):
        return self.psi

    def set_time_index(self, index: int):
        """t = t0 + dt (index - 1) (problem.jl:703-708)."""
        self.index = int(index)
        self.t = self.t0 + timestep(self.prob) * (self.index - 1)

    def inc_time_index(self):
        self.set_time_index(self.index + 1)


# ---- state accessors (problem.jl:710-866) ----
# Arrays keep the reference's index order: u[i, j, level] (i along x), level 0-based in
# the last axis; ``levels`` (1-based) selects which level slices are filled.

def _flux_level(state, lev, axis):
    g = state.prob.fluid.grid
    nx, ny = g.base.n
    nu = (nx + 1) * ny
    q = state.q[lev - 1] + state.q0[lev - 1]
    if axis == 0:
        return q[:nu].reshape((nx + 1, ny), order="F")
    return q[nu:].reshape((nx, ny + 1), order="F")


def _velocity_(out, state, axis, levels=None):
    g = state.prob.fluid.grid
    for lev in _level_list(g, levels):
        out[:, :, lev - 1] = _flux_level(state, lev, axis) / gridstep(g, lev)
    return out


def x_velocity_(out, state, levels=None):
    """x_velocity!(u, state, [levels]): (q + q0) / h_level on the u-faces (problem.jl:716-748)."""
    return _velocity_(out, state, 0, levels)


def y_velocity_(out, state, levels=None):
    """y_velocity!(v, state, [levels]) on the v-faces (problem.jl:726-748)."""
    return _velocity_(out, state, 1, levels)


def x_velocity(state, levels=None):
    """x velocity, (nx + 1) x ny x nlevel (problem.jl:710-748)."""
    g = state.prob.fluid.grid
    nx, ny = g.base.n
    return x_velocity_(np.zeros((nx + 1, ny, g.nlevel), order="F"), state, levels)


def y_velocity(state, levels=None):
    """y velocity, nx x (ny + 1) x nlevel (problem.jl:720-748)."""
    g = state.prob.fluid.grid
    nx, ny = g.base.n
    return y_velocity_(np.zeros((nx, ny + 1, g.nlevel), order="F"), state, levels)


def vorticity_(out, state, levels=None):
    """vorticity!(w, state, [levels]): Gamma / h_level^2 (problem.jl:761-771)."""
    g = state.prob.fluid.grid
    nx, ny = g.base.n
    for lev in _level_list(g, levels):
        out[:, :, lev - 1] = state.Gamma[lev - 1].reshape((nx - 1, ny - 1), order="F") / gridstep(g, lev) ** 2
    return out


def vorticity(state, levels=None):
    """Vorticity, (nx - 1) x (ny - 1) x nlevel (problem.jl:750-759)."""
    g = state.prob.fluid.grid
    nx, ny = g.base.n
    return vorticity_(np.zeros((nx - 1, ny - 1, g.nlevel), order="F"), state, levels)


GRID_POINTS = {x_velocity: "u", x_velocity_: "u", y_velocity: "v", y_velocity_: "v",
               vorticity: "gamma", vorticity_: "gamma"}


def _select(state, sel):
    """Panel selection: all (None), 0-based indices or range, or a body name."""
    if sel is None:
        return slice(None)
    if isinstance(sel, str):
        r = state.prob.bodies.panel_range(sel)
        return slice(r.start, r.stop)
    return sel


def boundary_pos(state, sel=None):
    """Panel positions, npanel x 2 (problem.jl:789-803)."""
    return state.xb[_select(state, sel)]


def boundary_len(state, sel=None):
    """Panel lengths (problem.jl:805-815)."""
    return state.ds[_select(state, sel)]


def boundary_vel(state, sel=None):
    """Panel velocities, npanel x 2 (problem.jl:817-827)."""
    return state.ub[_select(state, sel)]


def boundary_force(state, sel=None):
    """Force on each panel by the fluid, npanel x 2 (problem.jl:829-839)."""
    return state.fb[_select(state, sel)]


def boundary_total_force(state: State, sel=None):
    """Total [x, y] force on the panels, all or a selection (problem.jl:845-856)."""
    return boundary_force(state, sel).sum(axis=0)


boundary_ds = boundary_len  # the name the reference exports for boundary_len


def _copy_into(out, val):
    out[...] = val
    return out


def boundary_pos_(out, state, sel=None):
    return _copy_into(out, boundary_pos(state, sel))


def boundary_len_(out, state, sel=None):
    return _copy_into(out, boundary_len(state, sel))


boundary_ds_ = boundary_len_


def boundary_vel_(out, state, sel=None):
    return _copy_into(out, boundary_vel(state, sel))


def boundary_force_(out, state, sel=None):
    return _copy_into(out, boundary_force(state, sel))


def boundary_total_force_(out, state, sel=None):
    return _copy_into(out, boundary_total_force(state, sel))


def arraysize(func, npanel: int):
    """Array size of a boundary accessor for ``npanel`` panels (problem.jl:858-862)."""
    if func in (boundary_len, boundary_len_, boundary_ds, boundary_ds_):
        return (npanel,)
    if func in (boundary_total_force, boundary_total_force_):
        return (2,)
    return (npanel, 2)
