"""Pins the CPU oracle (test infrastructure) before it is trusted as the checker.

Each check is chosen so a plausible slip fails it:
  * triangle solid angle: closed forms (octant = pi/2, cube face from the centre = 2pi/3,
    orientation flip negates) and a second, independent numpy FP64 implementation;
  * brute force: closed mesh gives 4pi inside / 0 outside; additive over facet partitions;
    invariant under translation + uniform scaling of mesh and point together;
  * moments: divergence-theorem identities for closed meshes (N = 0, M1 = Vol I, and the
    symmetrised second moment equal to the volume-centroid expression), which pin the
    per-triangle covariance term and the parallel-axis shifts;
  * far field: the error of the order-0/1/2 expansion of one node decays with log-log slope
    -3/-4/-5 in distance (a dropped or wrong-signed term changes the slope);
  * restated engine: accuracy vs the brute force, and every facet appears once in its tree.
"""
import itertools

import numpy as np
import pytest

from lagrange_amd import workloads as W
from oracle import oracle as oracle_mod

FOUR_PI = 4 * np.pi

CUBE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
CUBE_F = np.array([[0, 3, 2], [2, 1, 0], [4, 5, 6], [6, 7, 4], [1, 2, 6], [6, 5, 1], [4, 7, 3], [3, 0, 4],
                   [2, 3, 7], [7, 6, 2], [0, 1, 5], [5, 4, 0]], np.int32)


def test_triangle_known_answers(oracle):
    assert oracle.tri_f64([1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]) == pytest.approx(np.pi / 2, abs=1e-15)
    assert oracle.tri_f64([1, 0, 0], [0, 0, 1], [0, 1, 0], [0, 0, 0]) == pytest.approx(-np.pi / 2, abs=1e-15)
    # each face of the unit cube subtends 4pi/6 from the centre
    c = np.array([0.5, 0.5, 0.5])
    for f in range(6):
        a, b = CUBE_F[2 * f], CUBE_F[2 * f + 1]
        s = sum(oracle.tri_f64(*(CUBE_V[t].astype(np.float64) for t in tri), c) for tri in (a, b))
        assert s == pytest.approx(2 * np.pi / 3, abs=1e-14)
    # query on a vertex: no contribution
    assert oracle.tri_f64([1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 0]) == 0.0
    # tiny far triangle: solid angle ~ A cos(theta) / d^2
    h = 1e-3
    v = np.array([[10, 0, 0], [10, h, 0], [10, 0, h]], np.float64)
    want = 0.5 * h * h / 100.0 * (1.0 if np.cross(v[1] - v[0], v[2] - v[0])[0] > 0 else -1.0)
    assert oracle.tri_f64(v[0], v[1], v[2], [0, 0, 0]) == pytest.approx(want, rel=1e-3)


def test_two_bruteforces_agree(oracle, ico5):
    V, F = W.displaced_sphere(8, 3, delete_fraction=0.05).f32()
    q = np.concatenate([W.uniform_queries(V.min(0), V.max(0), 300, 5, 1.3),
                        W.near_surface_queries(W.Mesh(V.astype(np.float64), F.astype(np.int64)), 100, 6)])
    a = oracle.bruteforce(V, F, q)
    b = oracle.bruteforce_numpy(V, F, q)
    assert np.abs(a - b).max() <= 1e-11


def test_closed_mesh_and_invariances(oracle, ico5):
    V, F = ico5
    q = W.uniform_queries(np.full(3, -1.5), np.full(3, 1.5), 2000, 7)
    r = np.linalg.norm(q.astype(np.float64), axis=1)
    om = oracle.bruteforce(V, F, q)
    inside = r < 0.98
    outside = r > 1.02
    assert np.abs(om[inside] - FOUR_PI).max() <= 1e-9
    assert np.abs(om[outside]).max() <= 1e-9
    # cube (modules/conftest.py cube_triangular): 4pi inside, 0 outside
    cq = np.array([[0.5, 0.5, 0.5], [0.1, 0.9, 0.2], [1.5, 0.5, 0.5], [-3, 2, 1]], np.float32)
    np.testing.assert_allclose(oracle.bruteforce(CUBE_V, CUBE_F, cq), [FOUR_PI, FOUR_PI, 0, 0], atol=1e-12)
    # additive over a partition of the facets
    Vo, Fo = W.displaced_sphere(10, 4, delete_fraction=0.05).f32()
    q2 = W.uniform_queries(Vo.min(0), Vo.max(0), 300, 8, 1.2)
    whole = oracle.bruteforce(Vo, Fo, q2)
    parts = oracle.bruteforce(Vo, Fo[::2], q2) + oracle.bruteforce(Vo, Fo[1::2], q2)
    assert np.abs(whole - parts).max() <= 1e-11
    # orientation flip negates
    assert np.abs(oracle.bruteforce(Vo, Fo[:, ::-1].copy(), q2) + whole).max() <= 1e-11
    # uniform scaling of mesh and points together by a power of two is exact in fp32
    assert np.abs(oracle.bruteforce(Vo * 4, Fo, q2 * 4) - whole).max() <= 1e-10
    # translation: equal up to the fp32 rounding of the moved coordinates
    t = np.array([4.0, -2.0, 1.0], np.float32)
    assert np.abs(oracle.bruteforce(Vo + t, Fo, q2 + t) - whole).max() <= 1e-4


def _sym3(X):
    """Fully symmetrised components (xxx yyy zzz xxy xxz yyx yyz zzx zzy xyz) of X[j,k,i]."""
    idx = [(0, 0, 0), (1, 1, 1), (2, 2, 2), (0, 0, 1), (0, 0, 2), (1, 1, 0), (1, 1, 2), (2, 2, 0), (2, 2, 1), (0, 1, 2)]
    out = []
    for a, b, c in idx:
        perms = set(itertools.permutations((a, b, c)))
        out.append(np.mean([X[p] for p in perms]))
    return np.array(out)


@pytest.mark.parametrize("mesh", ["cube", "displaced"])
def test_moment_identities_closed_mesh(oracle, mesh):
    if mesh == "cube":
        V, F = CUBE_V, CUBE_F
    else:
        V, F = W.displaced_sphere(12, 9, amplitude=0.25).f32()
    rec = oracle.moments(V, F, np.arange(F.shape[0]))
    p = rec[1:4]
    N = rec[4:7]
    M1 = rec[7:16].reshape(3, 3)
    S = rec[16:26]
    Vd = V.astype(np.float64)
    a, b, c = Vd[F[:, 0]], Vd[F[:, 1]], Vd[F[:, 2]]
    tet = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    vol = tet.sum()
    vcen = (tet[:, None] * (a + b + c) / 4.0).sum(0) / vol
    scale = np.abs(rec[7:26]).max()
    assert np.abs(N).max() <= 1e-12 * max(1.0, scale)
    np.testing.assert_allclose(M1, vol * np.eye(3), atol=1e-10 * scale)
    g = vol * (vcen - p)  # int (x - p) dV
    X = np.zeros((3, 3, 3))
    for j in range(3):
        for k in range(3):
            for i in range(3):
                X[j, k, i] = (i == j) * g[k] + (i == k) * g[j]
    np.testing.assert_allclose(S, _sym3(X), atol=1e-10 * max(scale, 1e-3))


def _pack_slot(rec, order=2, thr=0.0):
    """One-slot tree around a moment record (same packing as the engine's slot format), keeping
    the expansion up to `order`; thr is the far-field threshold word (0: always far)."""
    M = rec[7:16]
    S = rec[16:26]
    g = np.zeros(23)
    g[0:3] = rec[1:4]
    g[3] = thr
    g[4:7] = rec[4:7]
    if order >= 1:
        g[7:13] = -3 * np.array([M[0], M[4], M[8], M[1] + M[3], M[5] + M[7], M[6] + M[2]])
        g[7:10] += M[0] + M[4] + M[8]  # tr joins the quadratic's diagonal
    if order >= 2:
        C = -4.5 * np.array([S[0] + S[5] + S[7], S[3] + S[1] + S[8], S[4] + S[6] + S[2]])
        # C joins the cubic monomials xxx yyy zzz xxy xxz yyx yyz zzx zzy (xyz has none)
        g[13:23] = np.array([7.5] * 3 + [22.5] * 6 + [45.0]) * S + np.append(C[[0, 1, 2, 1, 2, 0, 2, 0, 1]], 0.0)
    w = np.zeros(24, np.uint32)
    w[:23] = g.astype(np.float32).view(np.uint32)
    w[23] = oracle_mod.LINK_LEAF  # a leaf of one triangle (never reached: thr 0 keeps it far)
    return w


def test_far_field_convergence_order(oracle):
    m = W.icosphere(16)
    V, F = m.f32()
    c = V[F].mean(axis=1)
    ids = np.where(np.linalg.norm(c - np.array([0, 0, 1.0], np.float32), axis=1) < 0.3)[0]
    rec = oracle.moments(V, F, ids)
    slot = _pack_slot(rec)
    tris = np.zeros(12, np.uint32)
    r = np.sqrt(rec[26])
    d = np.array([0.3, -0.2, 1.0])
    d /= np.linalg.norm(d)
    ks = np.array([8.0, 16.0, 32.0])
    q = (rec[1:4] + d[None, :] * (r * ks)[:, None]).astype(np.float32)
    exact = oracle.bruteforce(V, F[ids], q)
    # the full expansion, evaluated by the oracle's far-field code (threshold 0 -> always far)
    far2 = oracle.tree_eval(slot, 1, 0, 1, tris, q).astype(np.float64)
    e2 = np.abs(far2 - exact)
    slope2 = np.polyfit(np.log(ks), np.log(e2), 1)[0]
    assert slope2 <= -4.5, slope2
    # order 1 only (order-2 words zeroed): slope -4; order 0 only: slope -3
    e1 = np.abs(oracle.tree_eval(_pack_slot(rec, 1), 1, 0, 1, tris, q) - exact)
    e0 = np.abs(oracle.tree_eval(_pack_slot(rec, 0), 1, 0, 1, tris, q) - exact)
    assert -4.5 < np.polyfit(np.log(ks), np.log(e1), 1)[0] <= -3.5
    assert -3.5 < np.polyfit(np.log(ks), np.log(e0), 1)[0] <= -2.5
    assert np.all(e2 < e1) and np.all(e1 < e0)


def test_restated_engine_accuracy_and_structure(oracle, ico5, open_mesh):
    for V, F in (ico5, open_mesh):
        q = W.uniform_queries(V.min(0), V.max(0), 3000, 9, 1.3)
        bf = oracle.bruteforce(V, F, q)
        R = oracle.Restated(V, F)
        om = R.solid_angle(q)
        err = np.abs(om - bf) / FOUR_PI
        assert err.max() <= 4e-3 and err.mean() <= 1e-3
        ids = R.tris.reshape(-1, 12)[:, 9].astype(np.int64)
        assert np.array_equal(np.sort(ids), np.arange(F.shape[0]))
        ins = oracle.inside_from_omega(om)
        w = bf / FOUR_PI
        ok = np.abs(w - 0.5) > 5e-3
        assert np.array_equal(ins[ok], (w > 0.5)[ok])


def test_near_field_fp64_option(oracle, open_mesh):
    V, F = open_mesh
    M = W.Mesh(V.astype(np.float64), F.astype(np.int64))
    q = W.near_surface_queries(M, 500, 13)
    R = oracle.Restated(V, F)
    bf = oracle.bruteforce(V, F, q)
    diag = float(np.linalg.norm(V.max(0) - V.min(0)))
    a = R.solid_angle(q)
    b, st = R.solid_angle(q, 2.0, stats=True, tau=1e-4 * diag)
    assert st[:, 3].sum() > 0
    assert np.abs(b - bf).max() / FOUR_PI <= 4e-3 and np.abs(a - bf).max() / FOUR_PI <= 4e-3


def test_lbvh_order_is_sorted_permutation(oracle):
    ms = [W.icosphere(4), W.torus(20, 8)]
    order, keys = oracle.lbvh_sorted_order([m.f32()[0] for m in ms], [m.f32()[1] for m in ms])
    assert np.array_equal(np.sort(order), np.arange(order.size))
    assert np.all(np.diff(keys.astype(np.float64)) >= 0)
    assert np.all((keys[:ms[0].num_facets] >> 32) == 0) and np.all((keys[ms[0].num_facets:] >> 32) == 1)


def _interleave_bits(x, y, z):
    """30-bit Morton code by the definition, one bit at a time: bit b of x, y, z goes to bits
    3b + 2, 3b + 1, 3b."""
    k = 0
    for b in range(10):
        k |= ((x >> b) & 1) << (3 * b + 2) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b)
    return k


def test_lbvh_morton_keys_known_answers(oracle):
    """Known answers of the Morton key of lbvh_sorted_order: degenerate triangles whose centroids
    are exact multiples of 1/1024 in a unit centroid box, so each quantises to its own integer
    coordinate (1.0 clamps to 1023).  (1,0,0) -> 4, (0,1,0) -> 2, (0,0,1) -> 1, (3,5,6) -> 238 by
    hand, (1023,1023,1023) -> 2^30 - 1; random coordinates against the bit-by-bit definition; the
    order is the stable sort of the keys; a second mesh's keys carry mesh id 1 in bits 32+."""
    rng = np.random.default_rng(19)
    coords = [(0, 0, 0), (1024, 1024, 1024), (1, 0, 0), (0, 1, 0), (0, 0, 1), (3, 5, 6), (3, 5, 6)]
    coords += [tuple(int(v) for v in rng.integers(0, 1024, 3)) for _ in range(200)]
    C = np.array(coords, np.float32) / np.float32(1024)
    V = np.repeat(C, 3, axis=0)
    F = np.arange(V.shape[0], dtype=np.int32).reshape(-1, 3)
    order, keys = oracle.lbvh_sorted_order([V, V], [F, F])
    n = F.shape[0]
    q = np.minimum(np.array(coords, np.int64), 1023)
    want = np.array([_interleave_bits(*c) for c in q], np.uint64)
    assert want[2] == 4 and want[3] == 2 and want[4] == 1 and want[5] == 238 and want[1] == 2 ** 30 - 1
    inv = np.empty(2 * n, np.int64)
    inv[order] = np.arange(2 * n)
    got = keys[inv]  # key of each input facet
    np.testing.assert_array_equal(got[:n], want)
    np.testing.assert_array_equal(got[n:], want | np.uint64(1 << 32))
    np.testing.assert_array_equal(order[:n], np.argsort(want, kind="stable"))
