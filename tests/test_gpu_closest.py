"""GPU closest point / unsigned distance (fwn_closest_point) through the C ABI.

Checked against the brute-force oracle (every facet, Eberly's method in double; pinned in
tests/test_closest_oracle.py): the squared distance agrees to 1e-12 relative (the same per-
triangle formula in the same precision; only the traversal order differs), the facet agrees
wherever the second-nearest facet is farther by more than that tolerance, and the closest point
lies at the returned distance.  Also: points on the surface have distance ~0; every tree shape
gives the same distances; device pointers; empty mesh.
"""
import numpy as np
import pytest

from lagrange_amd import _lib as L
from lagrange_amd import workloads as W
from lagrange_amd.winding import FastWindingNumber

pytestmark = pytest.mark.gpu


def _queries(V, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = V.min(0), V.max(0)
    ext = hi - lo
    return rng.uniform(lo - 0.3 * ext, hi + 0.3 * ext, size=(n, 3)).astype(np.float32)


def _check(e, V, F, Q, oracle):
    d2, fi, cp = e.closest_point(Q)
    od2, ofi, ocp = oracle.closest(V, F, Q)
    np.testing.assert_allclose(d2, od2, rtol=1e-12, atol=1e-300)
    # facet: equal unless another facet is (numerically) as close
    diff = fi != ofi
    if diff.any():
        Vd = V.astype(np.float64)
        for i in np.where(diff)[0]:
            d_g, _ = oracle.tri_closest(Vd[F[fi[i], 0]], Vd[F[fi[i], 1]], Vd[F[fi[i], 2]], Q[i])
            assert d_g == pytest.approx(od2[i], rel=1e-12), i
    r = Q.astype(np.float64) - cp
    np.testing.assert_allclose((r * r).sum(1), d2, rtol=1e-9, atol=1e-15)
    return d2, fi, cp


def test_matches_brute_force_open_mesh(oracle, open_mesh):
    V, F = open_mesh
    e = FastWindingNumber((V, F))
    _check(e, V, F, _queries(V, 20000, 3), oracle)


@pytest.mark.parametrize("leaf_max,width", [(1, 2), (4, 4), (8, 8)])
def test_tree_shapes_give_same_distances(oracle, leaf_max, width):
    V, F = W.displaced_sphere(16, 4, delete_fraction=0.05).f32()
    Q = _queries(V, 5000, 5)
    e = FastWindingNumber((V, F), leaf_max=leaf_max, tree_width=width)
    d2, _, _ = _check(e, V, F, Q, oracle)
    ref = FastWindingNumber((V, F)).closest_point(Q)[0]
    np.testing.assert_array_equal(d2, ref)


def test_points_on_surface(ico5):
    V, F = ico5
    e = FastWindingNumber((V, F))
    rng = np.random.default_rng(8)
    f = rng.integers(0, F.shape[0], 4000)
    u, v = rng.random(4000), rng.random(4000)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    Vd = V.astype(np.float64)
    P = (Vd[F[f, 0]] + u[:, None] * (Vd[F[f, 1]] - Vd[F[f, 0]]) + v[:, None] * (Vd[F[f, 2]] - Vd[F[f, 0]]))
    d2, fi, _ = e.closest_point(P.astype(np.float32))
    assert d2.max() < 1e-12
    # the facet found contains the point up to float rounding
    assert np.mean(fi == f) > 0.98


def test_device_pointers_empty_and_errors(ico5):
    import torch
    V, F = ico5
    e = FastWindingNumber((V, F))
    Q = _queries(V, 3000, 9)
    d2, fi, cp = e.closest_point(Q)
    dd, dfi, dcp = e.closest_point(torch.from_numpy(Q).cuda())
    np.testing.assert_array_equal(dd.cpu().numpy(), d2)
    np.testing.assert_array_equal(dfi.cpu().numpy(), fi)
    np.testing.assert_array_equal(dcp.cpu().numpy(), cp)
    d2b, fib = e.closest_point(Q, with_point=False)
    np.testing.assert_array_equal(d2b, d2)
    assert e.closest_point(Q[:0])[0].shape == (0,)
    empty = FastWindingNumber((V, F[:0]))
    d2e, fie, _ = empty.closest_point(Q[:10])
    assert np.isinf(d2e).all() and (fie == -1).all()
