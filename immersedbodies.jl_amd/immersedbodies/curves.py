"""Curve discretisation into panels, following ImmersedBodies.jl's ``Curves`` module
(curves.jl). Only the curves the accelerated path's workloads need are provided:
``LineSegment`` (curves.jl:176-196), ``Circle`` (:203-231) and ``ParameterizedCurve``
(:239-301) used for ellipses, plus ``partition`` (:64-82) and ``panels`` =
``Panels(grid, curve)`` at spacing 2h (:53-55).
"""
from __future__ import annotations

import math

import numpy as np

from .problem import MultiDomainGrid, Panels


class LineSegment:
    def __init__(self, p1, p2):
        self.p1 = np.asarray(p1, dtype=np.float64)
        self.p2 = np.asarray(p2, dtype=np.float64)
        self.closed = False

    def arclength(self):
        return float(np.hypot(*(self.p2 - self.p1)))

    def partition_n(self, n: int) -> Panels:
        xb = np.stack([np.linspace(self.p1[0], self.p2[0], n),
                       np.linspace(self.p1[1], self.p2[1], n)], axis=1)
        ds = np.full(n, self.arclength() / (n - 1))
        return Panels(xb, ds)


class Circle:
    def __init__(self, r=1.0, center=(0.0, 0.0)):
        self.r = float(r)
        self.center = np.asarray(center, dtype=np.float64)
        self.closed = True

    def arclength(self):
        return 2 * math.pi * self.r

    def partition_n(self, n: int) -> Panels:
        t = 2 * math.pi / n * np.arange(n)
        xs = self.center[0] + self.r * np.cos(t)
        ys = self.center[1] + self.r * np.sin(t)
        ds = math.hypot(xs[1] - xs[0], ys[1] - ys[0])
        return Panels(np.stack([xs, ys], axis=1), np.full(n, ds))


def _point_array_lengths(points: np.ndarray, closed: bool) -> np.ndarray:
    n = points.shape[0]
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


class ParameterizedCurve:
    """Curve f(t), 0 <= t <= 1, re-parameterised by normalised arclength from
    ``n_sample`` samples with linear interpolation (curves.jl:239-301)."""

    def __init__(self, f, n_sample=100):
        self.f = f
        t = np.linspace(0.0, 1.0, n_sample)
        pts = np.array([f(ti) for ti in t], dtype=np.float64)
        s = np.concatenate([[0.0], np.cumsum(np.hypot(*(pts[1:] - pts[:-1]).T))])
        self.arclen = float(s[-1])
        self.s = s / self.arclen
        self.t = t
        self.closed = float(np.hypot(*(pts[-1] - pts[0]))) / self.arclen < 1e-10

    def arclength(self):
        return self.arclen

    def __call__(self, u):
        return np.asarray(self.f(float(np.interp(u, self.s, self.t))), dtype=np.float64)

    def partition_n(self, n: int) -> Panels:
        us = np.linspace(0.0, 1.0, n + 1)[:n] if self.closed else np.linspace(0.0, 1.0, n)
        pts = np.array([self(u) for u in us])
        return Panels(pts, _point_array_lengths(pts, self.closed))


def ellipse(a, b, center=(0.0, 0.0), angle=0.0, n_sample=400) -> ParameterizedCurve:
    """Ellipse with semi-axes a, b rotated counter-clockwise by ``angle``."""
    c, s = math.cos(angle), math.sin(angle)
    cx, cy = center

    def f(t):
        th = 2 * math.pi * t
        x, y = a * math.cos(th), b * math.sin(th)
        return (cx + c * x - s * y, cy + s * x + c * y)

    return ParameterizedCurve(f, n_sample=n_sample)


def partition(curve, n_or_ds) -> Panels:
    """``partition(curve, n::Integer)`` or ``partition(curve, ds::AbstractFloat)``
    (curves.jl:74-82): endpoints are preserved for open curves."""
    if isinstance(n_or_ds, (int, np.integer)):
        return curve.partition_n(int(n_or_ds))
    n = int(round(curve.arclength() / float(n_or_ds))) + (0 if curve.closed else 1)
    return curve.partition_n(n)


def panels(grid, curve) -> Panels:
    """``Panels(grid, curve)``: panels spaced 2h (curves.jl:53-55)."""
    h = grid.base.h if isinstance(grid, MultiDomainGrid) else grid.h
    return partition(curve, 2.0 * h)
