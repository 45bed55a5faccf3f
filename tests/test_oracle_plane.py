"""Pins of the oracle's single-plane-symmetric cross-section
(ShapeGenerator.cpp:35-318, ShapeAlignment.cu:11-152, Utils.cpp:59-122,
TubeDepth.cu:493-640).  The reference's Boost.Geometry calls have no pinned
outputs, so each helper is checked against closed forms:

* the shoelace area against regular polygons (n/2 r^2 sin(2 pi / n));
* the segment crossing against hand-solved pairs;
* the vertical chord of an ellipse polyline against b sqrt(1 - y^2/a^2);
* the grown contour: area reaches the section area within one 0.01 mm growth
  step; the max chord is the end-to-end span a1 + c2 + a2 with the contour
  left unrotated; the height is twice the common semi-minor axis (the 3:1
  ellipse and the circle share it by construction); the junction lies where
  the ellipse and the circle cross, y = c a / (a + b); the contour is mirror
  symmetric about its mid-height;
* the tube depth of a one-section tube equals half the contour's vertical
  chord at the cell faces, i.e. the closed-form ellipse/circle height.
"""
import math

import numpy as np
import pytest

import oracle as orc
from vocalsim import tubes

GROW = np.float32(1e-3) * np.float32(0.01)


def test_shoelace_area_of_regular_polygons():
    for n, r in ((3, 1.0), (6, 0.01), (100, 0.02), (997, 0.005)):
        th = 2 * np.pi * np.arange(n) / n
        y = (r * np.cos(th)).astype(np.float32)
        z = (r * np.sin(th)).astype(np.float32)
        exact = 0.5 * n * r * r * math.sin(2 * math.pi / n)
        assert orc.polygon_area(y, z) == pytest.approx(exact, rel=2e-6)
        # orientation does not matter, nor an explicit closing point
        assert orc.polygon_area(y[::-1], z[::-1]) == pytest.approx(exact, rel=2e-6)
        assert orc.polygon_area(np.append(y, y[0]), np.append(z, z[0])) == pytest.approx(exact, rel=2e-6)


def test_segment_crossing_closed_forms():
    assert orc.segment_cross([0, 0, 2, 2], [0, 2, 2, 0]) == pytest.approx((1.0, 1.0))
    assert orc.segment_cross([0, 0, 4, 0], [1, -1, 1, 3]) == pytest.approx((1.0, 0.0))
    assert orc.segment_cross([0, 0, 1, 0], [2, -1, 2, 1]) is None          # beyond the end
    assert orc.segment_cross([0, 0, 1, 1], [0, 1, 1, 2]) is None           # parallel
    y, z = orc.segment_cross([-1.0, 0.5, 3.0, 2.5], [0.0, 3.0, 2.0, -1.0])
    # lines z = 0.5 y + 1 and z = -2 y + 3 cross at y = 0.8
    assert (y, z) == pytest.approx((0.8, 1.4), rel=1e-6)


def test_vertical_chord_of_an_ellipse_polyline():
    a, b, c = 0.02, 0.007, 0.05
    th = np.linspace(0, 2 * np.pi, 4001)
    y = (c + a * np.cos(th)).astype(np.float32)
    z = (c + b * np.sin(th)).astype(np.float32)
    for u in (-0.019, -0.01, 0.0, 0.004, 0.0185):
        lo, hi = orc.vertical_chord(y, z, np.float32(c + u))
        exact = 2 * b * math.sqrt(1 - (u / a) ** 2)
        # polyline sag: at most (arc step)^2 / 8 of the local radius of curvature
        assert hi - lo == pytest.approx(exact, abs=2e-6)
    assert orc.vertical_chord(y, z, np.float32(c + 1.5 * a)) is None


@pytest.mark.parametrize("area", [0.25, 1.0, 2.84, 5.0, 8.0])
def test_grown_contour_structure(area):
    y, z, info = orc.plane_contour(area)
    A = np.float32(area) * np.float32(1e-2) * np.float32(1e-2)
    n = y.size
    assert n == 3997
    D = info["diameter"]
    # area reached within one growth step (perimeter <= pi * diameter for the
    # convex-ish union, and the step grows every semi-axis by GROW)
    assert info["area"] >= A
    assert info["area"] - A <= math.pi * D * GROW * 1.05
    # semi-axes after `steps` growth steps (float accumulation as in the loop)
    x = A / np.float32(4.0)
    l1 = np.float32(math.sqrt(np.float32(3.0) * x / (math.pi * 3.0)))
    l2 = np.float32(math.sqrt(x / math.pi))
    a1, b1, a2, b2 = l1 * np.float32(3.0), l1, l2, l2
    c2 = np.float32(0.0) + a1 + np.float32(0.003) * np.float32(1e-3)
    for _ in range(int(info["steps"])):
        a1, b1, a2, b2 = a1 + GROW, b1 + GROW, a2 + GROW, b2 + GROW
    # the maximum chord is the axis span, and the contour is only shifted
    assert D == pytest.approx(float(c2 + a2 + a1), rel=2e-6)
    assert info["minY"] == pytest.approx(0.05, abs=1e-7)
    assert y.max() - y.min() == pytest.approx(D, rel=2e-6)
    # height: the ellipse and the circle share the semi-minor axis
    assert b1 == pytest.approx(float(b2), rel=1e-6)
    assert info["maxZ"] - info["minZ"] == pytest.approx(2 * float(b1), rel=1e-5)
    # mirror symmetry about the mid-height
    mid = 0.5 * (info["maxZ"] + info["minZ"])
    k = np.arange(1999)
    assert np.allclose(y[k], y[n - 1 - k], rtol=0, atol=0)
    assert np.allclose(z[k] - mid, mid - z[n - 1 - k], rtol=0, atol=2e-8)
    # junction (point 999 of the upper half) where ellipse and circle cross
    a, b, c = float(a1), float(b1), float(c2)
    yj = c * a / (a + b)
    zj = math.sqrt(b * b - (yj - c) ** 2)
    shift_y = 0.05 + a
    assert y[999] - shift_y == pytest.approx(yj, abs=3e-7)
    assert z[999] - (0.05 + b) == pytest.approx(zj, abs=3e-7)


def test_one_section_tube_depth_is_half_the_vertical_chord():
    area = 2.0
    spec = tubes.TubeSpec(lengths_cm=np.array([8.0]), areas_cm2=np.array([area]),
                          shape=tubes.SHAPE_PLANE_SYMMETRIC, num_steps=10, pml=0)
    t = orc.OracleTube(spec, 64)
    y, z, info = orc.plane_contour(area)
    d = t.dims
    dX = t.field("depth_x")
    up = t.field("upper_wall").astype(int)
    lo = t.field("lower_wall").astype(int)
    ds = np.float32(t.scalars["ds_f32"])
    # the sagittal width in cells is the max chord, rounded and made odd downwards
    dc = int(round(np.float32(info["diameter"]) / ds))
    dc = dc - 1 if dc % 2 == 0 else dc
    j = 5
    x = d["tube_start_x"] + j
    air = lo[j] - up[j] - 1
    assert air == dc
    # closed-form height of the grown contour (ellipse left of the junction,
    # circle right of it), in the unshifted frame
    A = np.float32(area) * np.float32(1e-2) * np.float32(1e-2)
    x4 = A / np.float32(4.0)
    l1 = np.float32(math.sqrt(np.float32(3.0) * x4 / (math.pi * 3.0)))
    a1, b1 = l1 * np.float32(3.0), l1
    c2 = np.float32(0.0) + a1 + np.float32(0.003) * np.float32(1e-3)
    a2 = np.float32(math.sqrt(x4 / math.pi))
    for _ in range(int(info["steps"])):
        a1, b1, a2 = a1 + GROW, b1 + GROW, a2 + GROW
    a, b, c = float(a1), float(b1), float(c2)
    yj = c * a / (a + b)
    for cc in range(1, air + 1):
        y0 = np.float32(info["minY"]) + np.float32((cc - 0.5) * ds)
        lo_z, hi_z = orc.vertical_chord(y, z, y0)
        want = (hi_z - lo_z) / 2.0
        got = dX[lo[j] - cc, x]
        # interior columns of a one-section tube are unchanged by the x
        # smoothing; the 1 mm floor clamps the thinnest cells
        assert got == pytest.approx(max(want, 1e-3), rel=1e-12, abs=0)
        u = float(y0) - 0.05 - a
        h = b * math.sqrt(max(0.0, 1 - (u / a) ** 2)) if u <= yj else math.sqrt(max(0.0, b * b - (u - c) ** 2))
        assert want == pytest.approx(h, abs=2e-6)
