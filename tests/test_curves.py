"""Curves module: the assertions of the reference's test/curves.jl, plus closed forms of
the NACA section and of transformed panels."""
import math

import numpy as np
import pytest

from immersedbodies import curves
from immersedbodies.curves import (NACA, NACA4, Circle, LineSegment, NacaParams,
                                   ParameterizedCurve, arclength, isclosed, leading_edge,
                                   partition, rotate, scale, trailing_edge, translate)


def test_parameterized_curve_reparameterises_by_arclength():
    """An unevenly parameterised circle partitions into evenly spaced points
    (test/curves.jl:5-33)."""
    f = lambda t: (math.cos(2 * math.pi * t), math.sin(2 * math.pi * t))
    curve = ParameterizedCurve(lambda s: f(s ** 2))
    assert isclosed(curve)
    assert np.allclose(curve(0), f(0))
    n = 20
    p = partition(curve, n)
    even = np.array([f(k / n) for k in range(n)])
    assert np.allclose(p.xb, even, rtol=1e-3, atol=1e-3)
    assert np.allclose(p.ds, 2 * math.pi / n, rtol=0.01)


def test_open_parameterized_curve():
    f = lambda t: (t * t, 2 * t)
    curve = ParameterizedCurve(f)
    assert not isclosed(curve)
    assert np.allclose(curve(1), f(1))


def test_transformation_composition():
    """translate, then rotate 90 degrees, then scale 2 (test/curves.jl:44-52)."""
    t = translate((1, 2)) | rotate(math.radians(90)) | scale(2)
    assert np.allclose(t((-1, 2)), [-8, 0])
    assert np.allclose(t((-1, -2)), [0, 0], atol=1e-15)
    assert np.allclose(t([0, 0]), [-4, 2])
    assert np.allclose(translate((2, 3))((1, 5)), [3, 8])
    assert np.allclose(rotate(1.0)((1, 0)), [math.cos(1.0), math.sin(1.0)])
    assert np.allclose(scale(3)((5, 10)), [15, 30])


def test_transformed_line_keeps_its_endpoints():
    line = LineSegment((0, 1), (1, 1)) | translate((1, 0)) | rotate(math.pi) | scale(2)
    assert np.allclose(line(0), [-2, -2])
    assert np.allclose(line(1), [-4, -2])
    pts = partition(line, 3).xb
    assert np.allclose(pts[0], [-2, -2]) and np.allclose(pts[-1], [-4, -2])


@pytest.mark.parametrize("curve", [LineSegment((1, 2), (-5, 3)), Circle()])
def test_transformed_curve(curve):
    tr = translate((1, 0)) | rotate(math.radians(90)) | scale(3)
    c2 = tr(curve)
    assert isclosed(c2) == isclosed(curve)
    assert math.isclose(arclength(c2), 3 * arclength(curve))
    for t in (0.3, 0.6):
        assert np.allclose(c2(t), tr(curve(t)))
    # panels: the base curve's points moved, lengths scaled
    p0, p1 = partition(curve, 12), partition(c2, 12)
    assert np.allclose(p1.xb, np.array([tr(x) for x in p0.xb]))
    assert np.allclose(p1.ds, 3 * p0.ds)


def test_naca_params():
    with pytest.raises(ValueError, match="only contain digits"):
        NacaParams("hi")
    with pytest.raises(ValueError, match="3-digit NACA airfoil"):
        NacaParams("001")
    p = NacaParams("3412")
    assert isinstance(p, NACA4)
    assert math.isclose(p.m, 0.03) and math.isclose(p.p, 0.40) and math.isclose(p.t, 0.12)
    assert isclosed(NACA(p))


def test_naca_2412_edges_and_panels():
    """Leading edge at (0, 0), trailing edge at (1, 0), near-even panels and the sign
    change of y at the leading edge (test/curves.jl:88-108)."""
    airfoil = curves.naca("2412")
    s_le, s_te = leading_edge(airfoil), trailing_edge(airfoil)
    assert np.allclose(airfoil(s_le), [0, 0], atol=1e-8)
    assert np.allclose(airfoil(s_te), [1, 0], atol=1e-8)
    n = 50
    p = partition(airfoil, n)
    assert np.allclose(p.ds, arclength(airfoil) / n, rtol=0.1)
    i1 = math.floor(n * s_le)  # 0-based index of the reference's 1 + floor(n s_le)
    assert np.sign(p.xb[i1, 1]) != np.sign(p.xb[i1 + 1, 1])


def test_naca_symmetric_thickness_closed_form():
    """A symmetric section (m = 0) has y = ±5t(0.2969√x − 0.1260x − 0.3516x² + 0.2843x³
    − 0.1036x⁴) with the upper surface first; its maximum thickness is t near x = 0.3."""
    a = curves.naca("0012")
    for s in (0.05, 0.2, 0.37):
        x, y = a.params.point(s)
        assert math.isclose(x, 0.5 * (1 + math.cos(2 * math.pi * s)))
        yt = 0.6 * (0.2969 * math.sqrt(x) - 0.126 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3
                    - 0.1036 * x ** 4)
        assert math.isclose(y, yt, rel_tol=1e-14)
        assert math.isclose(a.params.point(1 - s)[1], -yt, rel_tol=1e-12)
    xs = np.linspace(0, 1, 2001)
    thick = 2 * np.array([NACA4(0, 0, 0.12).point(0.5 * math.acos(2 * x - 1) / math.pi)[1]
                          for x in xs])
    assert abs(thick.max() - 0.12) < 2e-4 and abs(xs[thick.argmax()] - 0.3) < 0.01


def test_spacing_partition_counts():
    """partition(curve, ds): round(arclength/ds) panels, plus one for an open curve."""
    assert partition(LineSegment((0, 0), (1, 0)), 0.1).xb.shape[0] == 11
    assert partition(Circle(r=1.0), 2 * math.pi / 40).xb.shape[0] == 40
    a = curves.naca("0012") | translate((-0.5, 0.0)) | rotate(-math.radians(20))
    from immersedbodies import fluid_grid

    g = fluid_grid(h=0.01, xlims=(-1.0, 2.0), ylims=(-0.5, 0.5), nlevel=3)
    p = curves.panels(g, a)
    assert p.xb.shape[0] == round(arclength(a) / 0.02)
    assert np.allclose(p.ds.sum(), arclength(a), rtol=1e-3)
