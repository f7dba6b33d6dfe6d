"""Pins the closest-point oracle (oracle_tri_closest / oracle_closest), CPU only.

The oracle restates Eberly's region method of point_triangle_squared_distance
(modules/core/include/lagrange/utils/point_triangle_squared_distance.h).  It is checked against
  * closed forms on the unit right triangle, one point in each of the seven regions of the (s, t)
    plane (a wrong branch or swapped edge gives a different distance or point);
  * an independent characterisation on random triangles: the closest point is the projection onto
    the plane when that lies inside the triangle, otherwise the nearest of the three edge points;
  * consistency: |q - closest|^2 equals the returned squared distance, and the brute force over a
    mesh equals the minimum of per-triangle distances computed in numpy.
"""
import numpy as np
import pytest

T0 = np.array([0.0, 0.0, 0.0])
T1 = np.array([1.0, 0.0, 0.0])
T2 = np.array([0.0, 1.0, 0.0])


@pytest.mark.parametrize("q,d2,cp", [
    ((0.25, 0.25, 2.0), 4.0, (0.25, 0.25, 0.0)),    # region 0: above the interior
    ((1.0, 1.0, 0.0), 0.5, (0.5, 0.5, 0.0)),        # region 1: beyond the hypotenuse
    ((-1.0, 2.0, 0.0), 2.0, (0.0, 1.0, 0.0)),       # region 2: vertex v2
    ((-1.0, 0.5, 0.0), 1.0, (0.0, 0.5, 0.0)),       # region 3: edge s = 0
    ((-1.0, -1.0, 0.0), 2.0, (0.0, 0.0, 0.0)),      # region 4: vertex v0
    ((0.5, -1.0, 0.0), 1.0, (0.5, 0.0, 0.0)),       # region 5: edge t = 0
    ((2.0, -1.0, 0.0), 2.0, (1.0, 0.0, 0.0)),       # region 6: vertex v1
    ((2.0, 0.5, 1.0), 2.25, (1.0, 0.0, 0.0)),       # region 1 clamped to vertex v1, off the plane
])
def test_regions_closed_form(oracle, q, d2, cp):
    d, st = oracle.tri_closest(T0, T1, T2, q)
    assert d == pytest.approx(d2, abs=1e-15)
    p = T0 + st[0] * (T1 - T0) + st[1] * (T2 - T0)
    np.testing.assert_allclose(p, cp, atol=1e-15)


def test_degenerate_triangle_uses_edges(oracle):
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([3.0, 0, 0])  # collinear
    d, st = oracle.tri_closest(a, b, c, [2.0, 1.0, 0.0])
    assert d == pytest.approx(1.0)
    p = a + st[0] * (b - a) + st[1] * (c - a)
    np.testing.assert_allclose(p, [2.0, 0.0, 0.0], atol=1e-15)


def _independent(v0, v1, v2, q):
    """Plane projection if inside, else nearest point on the three edges."""
    n = np.cross(v1 - v0, v2 - v0)
    n2 = n @ n
    best = np.inf
    if n2 > 0:
        p = q - ((q - v0) @ n) / n2 * n
        # barycentric coordinates of p
        w = np.array([np.cross(v2 - v1, p - v1) @ n, np.cross(v0 - v2, p - v2) @ n,
                      np.cross(v1 - v0, p - v0) @ n]) / n2
        if (w >= 0).all():
            best = (q - p) @ (q - p)
    for a, b in ((v0, v1), (v0, v2), (v1, v2)):
        ab = b - a
        u = np.clip(((q - a) @ ab) / (ab @ ab), 0.0, 1.0)
        x = a + u * ab
        best = min(best, (q - x) @ (q - x))
    return best


def test_random_triangles_match_independent_method(oracle):
    rng = np.random.default_rng(17)
    for _ in range(3000):
        v = rng.normal(size=(3, 3))
        q = rng.normal(size=3) * rng.choice([0.1, 1.0, 5.0])
        d, st = oracle.tri_closest(v[0], v[1], v[2], q)
        want = _independent(v[0], v[1], v[2], q)
        assert d == pytest.approx(want, rel=1e-9, abs=1e-12)
        p = v[0] + st[0] * (v[1] - v[0]) + st[1] * (v[2] - v[0])
        assert (q - p) @ (q - p) == pytest.approx(d, rel=1e-9, abs=1e-12)
        assert st[0] >= -1e-12 and st[1] >= -1e-12 and st[0] + st[1] <= 1 + 1e-12


def test_mesh_brute_force(oracle):
    from lagrange_amd import workloads as W
    V, F = W.displaced_sphere(6, 3, delete_fraction=0.05).f32()
    rng = np.random.default_rng(2)
    Q = rng.uniform(-1.5, 1.5, size=(200, 3)).astype(np.float32)
    d2, fi, cp = oracle.closest(V, F, Q)
    Vd = V.astype(np.float64)
    for i in range(0, 200, 7):
        per = [_independent(Vd[f[0]], Vd[f[1]], Vd[f[2]], Q[i].astype(np.float64)) for f in F]
        assert d2[i] == pytest.approx(min(per), rel=1e-9)
        assert per[fi[i]] == pytest.approx(d2[i], rel=1e-9)
        diff = Q[i].astype(np.float64) - cp[i]
        assert diff @ diff == pytest.approx(d2[i], rel=1e-9, abs=1e-14)
    # empty mesh
    d2e, fie, _ = oracle.closest(V, F[:0], Q[:3])
    assert np.isinf(d2e).all() and (fie == -1).all()
