"""Curve discretisation into panels: the host-side mirror of ImmersedBodies.jl's ``Curves``
module (curves.jl). Names, argument meanings and errors follow the reference:

* ``Curve``, ``isclosed``, ``arclength``, ``partition``           curves.jl:17-82
* ``Panels(grid, curve)`` at spacing 2h (here ``panels``)          curves.jl:53-55
* ``Transformation``, ``translate``, ``rotate``, ``scale``,
  ``TransformedCurve``                                             curves.jl:86-174
* ``LineSegment``, ``Circle``                                      curves.jl:176-231
* ``ParameterizedCurve``                                           curves.jl:233-332
* ``NacaParams``, ``NACA4``, ``NACA``, ``naca`` (``naca"…"``),
  ``leading_edge``, ``trailing_edge``                              curves.jl:334-413

Julia's ``curve |> t`` pipe is written ``curve | t`` (or ``t(curve)``); a transformation
applied to a transformation composes them, the argument's first.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .problem import CartesianGrid, Fluid, MultiDomainGrid, Panels


class Curve:
    """A curve ``c(t)``, 0 <= t <= 1 (curves.jl:17-36). Subclasses define ``__call__``,
    ``arclength`` and ``isclosed``, and partition either by a count or by a spacing."""

    def isclosed(self) -> bool:
        raise NotImplementedError

    def arclength(self) -> float:
        raise NotImplementedError

    def __call__(self, t):
        raise NotImplementedError

    # A curve that partitions by count inherits the spacing form and vice versa
    # (curves.jl:74-82).
    def partition_n(self, n: int) -> Panels:
        return self.partition_ds(self.arclength() / (n - (not self.isclosed())))

    def partition_ds(self, ds: float) -> Panels:
        return self.partition_n(int(round(self.arclength() / ds)) + (not self.isclosed()))

    def __or__(self, t):
        return t(self)

    @property
    def closed(self) -> bool:
        return self.isclosed()


def isclosed(curve: Curve) -> bool:
    """True when the curve's start and end points coincide (curves.jl:38-44)."""
    return curve.isclosed()


def arclength(curve: Curve) -> float:
    """Total arclength of a curve (curves.jl:57-62)."""
    return curve.arclength()


def partition(curve: Curve, n_or_ds) -> Panels:
    """``partition(curve, n::Integer)`` or ``partition(curve, ds::AbstractFloat)``
    (curves.jl:64-82): about equal segments, the curve's endpoints preserved. The count
    from a spacing is ``round(arclength/ds)`` (ties to even), plus one for an open curve."""
    if isinstance(n_or_ds, (bool, np.bool_)):
        raise TypeError("partition: give a panel count (int) or a spacing (float)")
    if isinstance(n_or_ds, (int, np.integer)):
        return curve.partition_n(int(n_or_ds))
    return curve.partition_ds(float(n_or_ds))


def panels(grid, curve: Curve) -> Panels:
    """``Panels(grid, curve)`` / ``Panels(fluid, curve)``: panels spaced 2h of the base
    grid (curves.jl:53-55)."""
    if isinstance(grid, Fluid):
        grid = grid.grid
    if isinstance(grid, MultiDomainGrid):
        grid = grid.base
    if not isinstance(grid, CartesianGrid):
        raise TypeError("panels: expected a Fluid, MultiDomainGrid or CartesianGrid")
    return partition(curve, 2.0 * grid.h)


# ---- transformations (curves.jl:86-174) ----

_I2 = np.eye(2)


@dataclass(frozen=True)
class Transformation:
    """``v -> offset + rotation · scale · v``: translation, rotation and uniform scaling
    (curves.jl:86-106)."""

    offset: np.ndarray
    rotation: np.ndarray
    scale: float

    def __call__(self, x):
        if isinstance(x, Transformation):
            return Transformation(self(x.offset), self.rotation @ x.rotation,
                                  self.scale * x.scale)
        if isinstance(x, TransformedCurve):
            return TransformedCurve(x.base, self(x.transform))
        if isinstance(x, Curve):
            return TransformedCurve(x, self)
        v = np.asarray(x, dtype=np.float64).reshape(2)
        return self.offset + self.rotation @ (self.scale * v)

    def __ror__(self, x):
        return self(x)

    def __or__(self, other):
        # t1 | t2: t1 first, then t2 (Python never tries the reflected method between
        # two operands of one type)
        if isinstance(other, Transformation):
            return other(self)
        return NotImplemented


def _rot(theta: float) -> np.ndarray:
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[c, -s], [s, c]])


def translate(v) -> Transformation:
    """Translation by ``v`` (curves.jl:114-120)."""
    return Transformation(np.asarray(v, dtype=np.float64).reshape(2), _I2, 1.0)


def rotate(theta: float) -> Transformation:
    """Counter-clockwise rotation by ``theta`` radians (curves.jl:122-127)."""
    return Transformation(np.zeros(2), _rot(float(theta)), 1.0)


def scale(k: float) -> Transformation:
    """Uniform scaling by ``k`` (curves.jl:129-134)."""
    return Transformation(np.zeros(2), _I2, float(k))


class TransformedCurve(Curve):
    """A curve moved by a Transformation (curves.jl:136-174). Its panels are the base
    curve's, transformed, with lengths times the scale."""

    def __init__(self, base: Curve, transform: Transformation):
        self.base = base
        self.transform = transform

    def isclosed(self):
        return self.base.isclosed()

    def arclength(self):
        return abs(self.transform.scale) * self.base.arclength()

    def __call__(self, t):
        return self.transform(self.base(t))

    def partition_n(self, n: int) -> Panels:
        p = self.base.partition_n(n)
        t = self.transform
        xb = t.offset[None, :] + (t.scale * p.xb) @ t.rotation.T
        return Panels(xb, p.ds * t.scale)


# ---- basic curves (curves.jl:176-231) ----

class LineSegment(Curve):
    """The segment from ``p1`` to ``p2`` (curves.jl:176-196)."""

    def __init__(self, p1, p2):
        self.p1 = np.asarray(p1, dtype=np.float64).reshape(2)
        self.p2 = np.asarray(p2, dtype=np.float64).reshape(2)

    def isclosed(self):
        return False

    def arclength(self):
        return float(np.hypot(*(self.p2 - self.p1)))

    def __call__(self, t):
        return self.p1 + float(t) * (self.p2 - self.p1)

    def partition_n(self, n: int) -> Panels:
        xb = np.stack([np.linspace(self.p1[0], self.p2[0], n),
                       np.linspace(self.p1[1], self.p2[1], n)], axis=1)
        return Panels(xb, np.full(n, self.arclength() / (n - 1)))


class Circle(Curve):
    """Circle of radius ``r`` about ``center`` (curves.jl:198-231): every panel gets the
    chord length between neighbouring points."""

    def __init__(self, r=1.0, center=(0.0, 0.0)):
        self.r = float(r)
        self.center = np.asarray(center, dtype=np.float64).reshape(2)

    def isclosed(self):
        return True

    def arclength(self):
        return 2 * math.pi * self.r

    def __call__(self, t):
        s = 2 * math.pi * float(t)
        return self.center + self.r * np.array([math.cos(s), math.sin(s)])

    def partition_n(self, n: int) -> Panels:
        t = 2 * math.pi / n * np.arange(n)
        xs = self.center[0] + self.r * np.cos(t)
        ys = self.center[1] + self.r * np.sin(t)
        ds = math.hypot(xs[1] - xs[0], ys[1] - ys[0])
        return Panels(np.stack([xs, ys], axis=1), np.full(n, ds))


# ---- parameterised curves (curves.jl:233-332) ----

def _point_array_lengths(points: np.ndarray, closed: bool) -> np.ndarray:
    """Each point's length: the mean of its two neighbouring segments, the end points of
    an open curve taking their one segment (curves.jl:303-332)."""
    n = points.shape[0]
    if n < 2:
        raise ValueError("cannot determine length with single point")
    lens = np.hypot(*(points[1:] - points[:-1]).T)
    out = np.empty(n)
    if closed:
        endlen = float(np.hypot(*(points[0] - points[-1])))
        out[0] = (endlen + lens[0]) / 2
        out[-1] = (lens[-1] + endlen) / 2
    else:
        out[0] = lens[0]
        out[-1] = lens[-1]
    out[1:-1] = (lens[:-1] + lens[1:]) / 2
    return out


class ParameterizedCurve(Curve):
    """Curve ``f(t)``, 0 <= t <= 1, re-parameterised by normalised arclength measured on
    ``n_sample`` samples and interpolated linearly both ways (curves.jl:233-301). It is
    closed when its end points agree to 1e-10 of its length."""

    def __init__(self, f, n_sample=100):
        self.f = f
        t = np.linspace(0.0, 1.0, n_sample)
        pts = np.array([np.asarray(f(ti), dtype=np.float64) for ti in t])
        s = np.concatenate([[0.0], np.cumsum(np.hypot(*(pts[1:] - pts[:-1]).T))])
        self.arclen = float(s[-1])
        self.s = s / self.arclen
        self.t = t
        self._closed = float(np.hypot(*(pts[-1] - pts[0]))) / self.arclen < 1e-10

    def isclosed(self):
        return self._closed

    def arclength(self):
        return self.arclen

    def param(self, u: float) -> float:
        """Parameter at normalised arclength ``u``."""
        return float(np.interp(u, self.s, self.t))

    def inv_param(self, t: float) -> float:
        """Normalised arclength at parameter ``t``."""
        return float(np.interp(t, self.t, self.s))

    def __call__(self, u):
        return np.asarray(self.f(self.param(float(u))), dtype=np.float64)

    def partition_n(self, n: int) -> Panels:
        us = np.linspace(0.0, 1.0, n + 1)[:n] if self.isclosed() else np.linspace(0.0, 1.0, n)
        pts = np.array([self(u) for u in us])
        return Panels(pts, _point_array_lengths(pts, self.isclosed()))


def ellipse(a, b, center=(0.0, 0.0), angle=0.0, n_sample=400) -> ParameterizedCurve:
    """Ellipse with semi-axes a, b rotated counter-clockwise by ``angle``, as a
    ParameterizedCurve of the angle."""
    c, s = math.cos(angle), math.sin(angle)
    cx, cy = center

    def f(t):
        th = 2 * math.pi * t
        x, y = a * math.cos(th), b * math.sin(th)
        return (cx + c * x - s * y, cy + s * x + c * y)

    return ParameterizedCurve(f, n_sample=n_sample)


# ---- NACA airfoils (curves.jl:334-413) ----

class _NacaParamsType(type):
    def __call__(cls, *args, **kw):
        if cls is NacaParams:  # NacaParams(spec): dispatch on the number of digits
            (spec,) = args
            spec = str(spec)
            if not spec or not all(ch in "0123456789" for ch in spec):
                raise ValueError("NACA specification must only contain digits")
            if len(spec) == 4:
                return NACA4.from_spec(spec)
            raise ValueError(f"{len(spec)}-digit NACA airfoil not defined")
        return super().__call__(*args, **kw)


class NacaParams(metaclass=_NacaParamsType):
    """Parameters of a NACA section; ``NacaParams(spec)`` parses a digit string
    (curves.jl:334-350)."""


class NACA4(NacaParams):
    """Four-digit section: camber ``m``, its chordwise place ``p``, thickness ``t``
    (curves.jl:370-383)."""

    def __init__(self, m: float, p: float, t: float):
        self.m, self.p, self.t = float(m), float(p), float(t)

    @staticmethod
    def from_spec(spec: str) -> "NACA4":
        return NACA4(int(spec[0]) / 100, int(spec[1]) / 10, int(spec[2:4]) / 100)

    def point(self, s: float) -> np.ndarray:
        """Airfoil point at parameter ``s``: the chord runs x = (1 + cos 2πs)/2, upper
        surface for s < 1/2, thickness laid off normal to the camber line, with the
        closed-trailing-edge thickness polynomial (curves.jl:385-413; the NACA 4-digit
        equations as tabulated by airfoiltools.com)."""
        m, p, t = self.m, self.p, self.t
        x = 0.5 * (1 + math.cos(2 * math.pi * s))
        if x < p:
            yc = m / p ** 2 * (2 * p * x - x ** 2)
            dyc = 2 * m / p ** 2 * (p - x)
        else:
            yc = m / (1 - p) ** 2 * (1 - 2 * p + 2 * p * x - x ** 2)
            dyc = 2 * m / (1 - p) ** 2 * (p - x)
        yt = 5 * t * (0.2969 * math.sqrt(x) - 0.126 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3
                      - 0.1036 * x ** 4)
        th = math.atan(dyc)
        dz = (-yt * math.sin(th), yt * math.cos(th))
        sg = 1.0 if s < 0.5 else -1.0
        return np.array([x + sg * dz[0], yc + sg * dz[1]])


class NACA(Curve):
    """Closed airfoil curve of unit chord from the trailing edge (1, 0) over the upper
    surface to the leading edge (0, 0) and back (curves.jl:352-368)."""

    def __init__(self, params: NacaParams):
        self.params = params
        self.curve = ParameterizedCurve(params.point)
        self.leading_edge = self.curve.inv_param(0.5)

    def isclosed(self):
        return True

    def arclength(self):
        return self.curve.arclength()

    def __call__(self, s):
        return self.curve(s)

    def partition_n(self, n: int) -> Panels:
        return self.curve.partition_n(n)


def naca(spec: str) -> NACA:
    """``naca"0012"``: the airfoil of a digit string (curves.jl:345-350)."""
    return NACA(NacaParams(spec))


def leading_edge(airfoil: NACA) -> float:
    """Curve parameter of the leading edge."""
    return airfoil.leading_edge


def trailing_edge(airfoil: NACA) -> float:
    """Curve parameter of the trailing edge."""
    return 0.0
