"""Boundary entities: the oracle's projections (oracle/entity_oracle.c)
pinned by what the geometry fixes, and the compute core's host geometry
(cpp_core.geometry_project_batch, the FP64 instantiation of the projections
the constrained sweep runs) against the oracle on every kind (reference
tests/cpp/test_curves.cpp:45-503, tests/cpp/test_3d.cpp, src/geometry.hpp).

Pins: a NURBS circle is the circle, a degree-1 B-spline its polyline and a
clamped degree-2 B-spline on three points the quadratic Bezier on them; an
ellipse arc with equal axes is the circle; plane, sphere and cylinder feet
are the orthogonal / radial projections; a bilinear patch is its plane; a
NURBS (quadratic x linear) patch is its cylinder; a composite keeps its
nearest segment; every parametrization's frame is the finite difference of
its points; feet satisfy (C - q).C' = 0 inside the trim; the trims clamp and
wrap; the wire checks refuse bad records."""

from __future__ import annotations

import numpy as np
import pytest

from egg_amd.smoothing.entities import encode_entity

S2 = np.sqrt(0.5)
NURBS_CIRCLE = {"degree": 2, "knots": [0, 0, 0, .25, .25, .5, .5, .75, .75, 1, 1, 1],
                "ctrl": [[1, 0], [1, 1], [0, 1], [-1, 1], [-1, 0], [-1, -1], [0, -1], [1, -1], [1, 0]],
                "weights": [1, S2, 1, S2, 1, S2, 1, S2, 1], "closed": True}
# a quadratic x linear NURBS patch: the quarter cylinder x^2 + y^2 = 1, x, y >= 0, z in [0, 2]
NURBS_CYL = {"degree": (2, 1), "knots_u": [0, 0, 0, 1, 1, 1], "knots_v": [0, 0, 1, 1],
             "ctrl": [[[1, 0, 0], [1, 0, 2]], [[1, 1, 0], [1, 1, 2]], [[0, 1, 0], [0, 1, 2]]],
             "weights": [[1, 1], [S2, S2], [1, 1]]}


def _blob(oracle, kind, spec):
    tag, _, rec, seg = encode_entity(kind, spec)
    return oracle.entity_record(tag, rec, seg or [])


def _core(kind, spec, Q):
    from egg_amd._cpp import cpp_core

    tag, _, rec, seg = encode_entity(kind, spec)
    segd = [{"data": a, "off": np.array([0, a.size], np.int32)} for a in (seg or [])]
    return cpp_core.geometry_project_batch(np.atleast_2d(np.asarray(Q, float)), tag, rec, segd)


CASES = [
    ("circlearc", (0.2, -0.1, 1.3, -0.5, 2.0, 0)),
    ("circlearc", (0.0, 0.0, 1.0, -np.pi, np.pi, 1)),
    ("ellipsearc", (0.1, 0.2, 1.5, 0.7, 0.4, 0.3, 4.0, 0)),
    ("quadbezier", (0, 0, 1, 2, 3, 1, 0.1, 0.9)),
    ("cubicbezier", (0, 0, 1, 2, 2, -1, 3, 1, 0, 1)),
    ("bspline", NURBS_CIRCLE),
    ("bspline", {"degree": 3, "knots": [0, 0, 0, 0, 0.3, 0.6, 1, 1, 1, 1],
                 "ctrl": [[0, 0], [1, 1], [2, -1], [3, 0.5], [4, 0], [5, 1]]}),
    ("composite", [("lineseg", (0, 0, 1, 0)), ("circlearc", (1, 1, 1, -np.pi / 2, 0, 0)),
                   ("bspline", {"degree": 1, "knots": [0, 0, 1, 1], "ctrl": [[2, 1], [2, 3]]})]),
    ("lineseg", (0, 0, 2, 1)), ("circle", (0.5, 0.5, 1.0)), ("ellipse", (0, 0, 2, 1)),
    ("linerail", (0, 0, 2, 1)),
    ("sphere", (0.1, 0, 0.2, 1.5, 1, 0, 0, 0, 1, 0)),
    ("plane", (0, 0, 0.5, 1, 0, 0, 0, S2, S2)),
    ("cylinder", (0, 0, 0, 1, 0, 0, 0, 1, 0, 0.8)),
    ("line3", (0, 0, 0, 1, 2, 3, 0.2, 0.8)),
    ("linerail3", (0, 0, 0, 1, 2, 3)),
    ("bsplinesurf", NURBS_CYL),
    ("bsplinesurf", {"degree": (1, 1), "knots_u": [0, 0, 1, 1], "knots_v": [0, 0, 1, 1],
                     "ctrl": [[[0, 0, 0.5], [0, 1, 0.5]], [[1, 0, 0.5], [1, 1, 0.5]]],
                     "trim": [[[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.1, 0.9]]]}),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_core_matches_oracle(oracle, case):
    """The compute core's FP64 projection and tangent basis equal the
    oracle's on points around the entity."""
    kind, spec = CASES[case]
    blob = _blob(oracle, kind, spec)
    dim = 3 if kind in ("sphere", "plane", "cylinder", "line3", "linerail3", "bsplinesurf") else 2
    Q = np.random.default_rng(case).uniform(-2.5, 3.5, (24, dim))
    feet, basis = _core(kind, spec, Q)
    for i, q in enumerate(Q):
        f, B = oracle.entity_frame(blob, q)
        np.testing.assert_allclose(feet[i], f, atol=1e-9)
        np.testing.assert_allclose(basis[i], B, atol=1e-7)
        # the basis is orthonormal
        np.testing.assert_allclose(B.T @ B, np.eye(B.shape[1]), atol=1e-12)


def test_nurbs_circle_is_the_circle(oracle):
    blob = _blob(oracle, "bspline", NURBS_CIRCLE)
    for u in np.linspace(0, 1, 41):
        assert np.linalg.norm(oracle.entity_param(blob, 0, [u])) == pytest.approx(1.0, abs=1e-14)
    for q in np.random.default_rng(1).uniform(-2, 2, (20, 2)):
        f, B = oracle.entity_frame(blob, q)
        np.testing.assert_allclose(f, q / np.linalg.norm(q), atol=1e-12)
        assert abs(B[:, 0] @ f) < 1e-10  # tangent perpendicular to the radius


def test_degree1_bspline_is_its_polyline_and_bezier_its_bspline(oracle):
    pts = np.array([[0, 0], [1, 1], [2, 0], [3, 1]], float)
    poly = _blob(oracle, "bspline", {"degree": 1, "knots": [0, 0, 1 / 3, 2 / 3, 1, 1], "ctrl": pts})
    segs = [_blob(oracle, "lineseg", (*pts[i], *pts[i + 1])) for i in range(3)]
    P = np.array([[0, 0], [1, 2], [3, 1]], float)
    bez = _blob(oracle, "quadbezier", (*P.ravel(), 0, 1))
    bsp = _blob(oracle, "bspline", {"degree": 2, "knots": [0, 0, 0, 1, 1, 1], "ctrl": P})
    # the degree-1 curve interpolates its control points linearly between
    # the knots
    for u in np.linspace(0, 1, 31):
        i = min(int(u * 3), 2)
        w = u * 3 - i
        np.testing.assert_allclose(oracle.entity_param(poly, 0, [u]), (1 - w) * pts[i] + w * pts[i + 1], atol=1e-14)
    # a point just off a segment's interior has that segment's foot (the
    # sampled seed lands on its span; far points may stop at a local foot,
    # as the reference's seeded Newton does)
    rng = np.random.default_rng(2)
    for _ in range(20):
        i = rng.integers(3)
        w = rng.uniform(0.15, 0.85)
        e = pts[i + 1] - pts[i]
        q = pts[i] + w * e + 0.05 * np.array([-e[1], e[0]]) / np.linalg.norm(e)
        np.testing.assert_allclose(oracle.entity_frame(poly, q)[0], oracle.entity_frame(segs[i], q)[0], atol=1e-12)
    for t in np.linspace(-0.2, 1.2, 15):  # past the ends the end pieces extend
        np.testing.assert_allclose(oracle.entity_param(bez, 0, [t]), oracle.entity_param(bsp, 0, [t]), atol=1e-14)
        np.testing.assert_allclose(oracle.entity_param(bez, 1, [t]), oracle.entity_param(bsp, 1, [t]), atol=1e-13)


def test_ellipse_arc_with_equal_axes_is_the_circle(oracle):
    blob = _blob(oracle, "ellipsearc", (0.3, -0.2, 1.2, 1.2, 0.7, 0, 2 * np.pi, 1))
    for q in np.random.default_rng(3).uniform(-2, 2, (20, 2)):
        d = q - (0.3, -0.2)
        np.testing.assert_allclose(oracle.entity_frame(blob, q)[0], (0.3, -0.2) + 1.2 * d / np.linalg.norm(d),
                                   atol=1e-9)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_frames_are_derivatives_and_feet_stationary(oracle, case):
    kind, spec = CASES[case]
    if kind in ("composite", "lineseg", "circle", "ellipse", "linerail", "linerail3"):
        return  # no parametrization
    blob = _blob(oracle, kind, spec)
    k = 2 if kind in ("sphere", "plane", "cylinder", "bsplinesurf") else 1
    rng = np.random.default_rng(10 + case)
    lo, hi = (0.05, 0.95)
    for _ in range(6):
        q = rng.uniform(lo, hi, k)
        J = oracle.entity_param(blob, 1, q)
        for c in range(k):
            h = 1e-6
            e = np.zeros(k)
            e[c] = h
            fd = (oracle.entity_param(blob, 0, q + e) - oracle.entity_param(blob, 0, q - e)) / (2 * h)
            np.testing.assert_allclose(J[:, c], fd, atol=2e-7)
    if kind in ("sphere", "plane", "cylinder", "line3"):
        return
    # a foot strictly inside the trim makes the residual perpendicular to
    # the entity (the Newton foot's stationarity)
    tr = {6: (4, 5), 7: (6, 7), 8: (7, 8), 9: (9, 10), 10: (3, 4)}.get(int(blob[0]))
    dim = 3 if kind == "bsplinesurf" else 2
    for q in rng.uniform(-1, 3, (10, dim)):
        f, B = oracle.entity_frame(blob, q)
        t = oracle.entity_param(blob, 2, q)
        if tr is not None and not (blob[tr[0]] < t[0] < blob[tr[1]]):
            continue
        if kind == "bsplinesurf" and np.any((t <= 1e-9) | (t >= 1 - 1e-9)):
            continue  # the foot iterate stopped on the patch's edge (a constrained minimum)
        if not np.allclose(f, oracle.entity_param(blob, 0, t), atol=1e-12):
            continue  # clamped onto the trim
        assert np.all(np.abs((f - q) @ B) < 1e-7 * (1 + np.linalg.norm(f - q)))


def test_analytic_surfaces_and_3d_curves(oracle):
    rng = np.random.default_rng(4)
    sph = _blob(oracle, "sphere", (0.1, 0, 0.2, 1.5, 1, 0, 0, 0, 1, 0))
    pl = _blob(oracle, "plane", (0, 0, 0.5, 1, 0, 0, 0, S2, S2))
    cyl = _blob(oracle, "cylinder", (0, 0, 0, 1, 0, 0, 0, 1, 0, 0.8))
    ln = _blob(oracle, "line3", (0, 0, 0, 1, 2, 3, 0.2, 0.8))
    rail = _blob(oracle, "linerail3", (0, 0, 0, 1, 2, 3))
    n = np.array([0, -S2, S2])
    for q in rng.uniform(-2, 2, (20, 3)):
        c = np.array([0.1, 0, 0.2])
        np.testing.assert_allclose(oracle.entity_frame(sph, q)[0], c + 1.5 * (q - c) / np.linalg.norm(q - c),
                                   atol=1e-12)
        np.testing.assert_allclose(oracle.entity_frame(pl, q)[0], q - ((q - (0, 0, 0.5)) @ n) * n, atol=1e-12)
        r = np.hypot(q[0], q[1])
        np.testing.assert_allclose(oracle.entity_frame(cyl, q)[0], (0.8 * q[0] / r, 0.8 * q[1] / r, q[2]),
                                   atol=1e-12)
        d = np.array([1, 2, 3.0])
        t = np.clip(q @ d / (d @ d), 0.2, 0.8)
        np.testing.assert_allclose(oracle.entity_frame(ln, q)[0], t * d, atol=1e-12)
        tr = q @ d / (d @ d)
        np.testing.assert_allclose(oracle.entity_frame(rail, q)[0], tr * d if 0 <= tr <= 1 else q, atol=1e-12)


def test_bspline_surfaces(oracle):
    cyl = _blob(oracle, "bsplinesurf", NURBS_CYL)
    rng = np.random.default_rng(5)
    for q in rng.uniform(0.2, 1.8, (20, 3)):  # points whose radial foot lies on the quarter patch
        r = np.hypot(q[0], q[1])
        np.testing.assert_allclose(oracle.entity_frame(cyl, q)[0], (q[0] / r, q[1] / r, q[2]), atol=1e-8)
    flat = {"degree": (1, 1), "knots_u": [0, 0, 1, 1], "knots_v": [0, 0, 1, 1],
            "ctrl": [[[0, 0, 0.5], [0, 1, 0.5]], [[1, 0, 0.5], [1, 1, 0.5]]]}
    plain = _blob(oracle, "bsplinesurf", flat)
    trimmed = _blob(oracle, "bsplinesurf", {**flat, "trim": [[[0.1, 0.1], [0.9, 0.1], [0.9, 0.9], [0.1, 0.9]]]})
    for q in rng.uniform(-0.5, 1.5, (20, 3)):
        np.testing.assert_allclose(oracle.entity_frame(plain, q)[0], (*np.clip(q[:2], 0, 1), 0.5), atol=1e-9)
        inside = np.all((q[:2] >= 0.1) & (q[:2] <= 0.9))
        f = oracle.entity_frame(trimmed, q)[0]
        if inside:
            np.testing.assert_allclose(f, (*q[:2], 0.5), atol=1e-9)
        else:  # clamped onto the trim loop: the nearest point of the square's boundary
            u = np.clip(q[:2], 0.1, 0.9)
            np.testing.assert_allclose(f, (*u, 0.5), atol=1e-9)


def test_trims_clamp_and_wrap(oracle):
    arc = _blob(oracle, "circlearc", (0, 0, 1, 0.25, 1.25, 0))
    for ang in np.linspace(-np.pi, np.pi, 37):
        q = 2 * np.array([np.cos(ang), np.sin(ang)])
        f = oracle.entity_frame(arc, q)[0]
        if 0.25 <= ang <= 1.25:
            np.testing.assert_allclose(f, q / 2, atol=1e-12)
        else:  # the nearer end in angle (the re-representation nearest the arc's midpoint)
            dm = np.remainder(ang - 0.75 + np.pi, 2 * np.pi) - np.pi
            end = 1.25 if dm > 0 else 0.25
            np.testing.assert_allclose(f, (np.cos(end), np.sin(end)), atol=1e-12)
    # an arc across the angle seam keeps a query near the seam on its own side
    seam = _blob(oracle, "circlearc", (0, 0, 1, 1.5 * np.pi, 2 * np.pi, 0))
    np.testing.assert_allclose(oracle.entity_frame(seam, (1.0, -0.05))[0], np.array([1.0, -0.05]) / np.hypot(1, 0.05),
                               atol=1e-12)


def test_composite_keeps_its_nearest_segment(oracle):
    segs = [("lineseg", (0, 0, 1, 0)), ("circlearc", (1, 1, 1, -np.pi / 2, 0, 0)),
            ("cubicbezier", (2, 1, 2.2, 1.7, 1.8, 2.3, 2, 3, 0, 1))]
    comp = _blob(oracle, "composite", segs)
    parts = [_blob(oracle, k, s) for k, s in segs]
    for q in np.random.default_rng(6).uniform(-0.5, 3.5, (30, 2)):
        cands = [oracle.entity_frame(p, q) for p in parts]
        best = min(cands, key=lambda fb: np.sum((fb[0] - q) ** 2))
        f, B = oracle.entity_frame(comp, q)
        np.testing.assert_allclose(f, best[0], atol=1e-14)
        np.testing.assert_allclose(B, best[1], atol=1e-14)


def test_wire_checks():
    from egg_amd._cpp import cpp_core

    bad = [
        ("bspline", {"degree": 8, "knots": np.r_[np.zeros(9), np.ones(9)], "ctrl": np.zeros((9, 2))}, "degree"),
        ("bspline", {"degree": 2, "knots": [0, 0, 0, 0.5, 0.2, 1, 1, 1], "ctrl": np.zeros((5, 2))}, "non-decreasing"),
        ("bspline", {"degree": 1, "knots": [0, 0, 1, 1], "ctrl": [[0, 0], [1, 0]], "weights": [1, -1]}, "positive"),
        ("bsplinesurf", {"degree": (4, 1), "knots_u": np.r_[np.zeros(5), np.ones(5)], "knots_v": [0, 0, 1, 1],
                         "ctrl": np.zeros((5, 2, 3))}, "degree"),
    ]
    for kind, spec, msg in bad:
        tag, _, rec, seg = encode_entity(kind, spec)
        segd = [{"data": a, "off": np.array([0, a.size], np.int32)} for a in seg]
        dim = 3 if kind == "bsplinesurf" else 2
        with pytest.raises(ValueError, match=msg):
            cpp_core.geometry_project_batch(np.zeros((1, dim)), tag, rec, segd)
    with pytest.raises(ValueError, match="2D curve"):
        encode_entity("composite", [("sphere", (0,) * 10)])
    # a 3D kind in a 2D query
    tag, _, rec, _ = encode_entity("plane", (0, 0, 0, 1, 0, 0, 0, 1, 0))
    with pytest.raises(ValueError, match="3D entity in a 2D"):
        cpp_core.geometry_project_batch(np.zeros((1, 2)), tag, rec, [])
