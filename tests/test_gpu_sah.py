"""GPU parity of the SAH tree builder (tree_builder = FWN_TREE_SAH) against the CPU restatement
of the reference's 4-ary box-area tree (oracle.Restated), through the C ABI.

With leaf_max 1 the GPU builds the restated tree itself:
  * integer work bit-exact: the sorted triangle order (original facet ids in tree order), and
    for every node, slot by slot, leaf / internal, a leaf's triangle range, and the child's slot
    count;
    node count and depth;
  * every slot's float data within 2e-5 of its largest word (the GPU merges moments pairwise
    in FP32 below fp64_moment_min triangles, the restatement merges up to 4 children in FP64);
  * queries: the GPU engine against the CPU traversal of the restated tree (same rule, FP64 near
    field with the same tau): identical per-point work counters except where a point sits on a
    slot's beta boundary within the slot data's rounding (at most 0.5 % of the points here, each
    off by at most one slot's truncation error), and |d omega| <= 2e-4 sr elsewhere.
"""
import numpy as np
import pytest

from lagrange_amd import workloads as W
from lagrange_amd import _lib as L
from lagrange_amd.winding import FastWindingNumber, MultiFastWindingNumber

pytestmark = pytest.mark.gpu

SAH = dict(tree_builder=L.FWN_TREE_SAH, leaf_max=1)


def _compare_trees(gs, groot, gns, rs, rns, width):
    from oracle import oracle as O
    return O.compare_slot_trees(gs, groot, gns, rs, rns, width)


def _meshes():
    rng = np.random.default_rng(7)
    ico = W.icosphere(16).f32()
    opn = W.displaced_sphere(40, 3, amplitude=0.15, delete_fraction=0.05).f32()
    # a soup with repeated triangles (coincident centroids: middle splits) and slivers
    V = rng.standard_normal((300, 3)).astype(np.float32)
    F = rng.integers(0, 300, (700, 3)).astype(np.int32)
    F = np.concatenate([F, np.repeat(F[:5], 40, axis=0)])
    tiny = [(ico[0], ico[1][:k]) for k in (2, 3, 5, 17)]
    return [("ico16", ico), ("open40", opn), ("soup", (V, F))] + [(f"first{k}", t) for k, t in zip((2, 3, 5, 17), tiny)]


@pytest.mark.parametrize("name,mesh", _meshes(), ids=[m[0] for m in _meshes()])
def test_sah_tree_is_the_restated_tree(oracle, name, mesh):
    V, F = mesh
    e = FastWindingNumber((V, F), **SAH)
    info = e.tree_info()
    slots, tris, roots, _ = e.export_tree()
    R = oracle.Restated(V, F)
    got = tris.reshape(-1, 12)[:, 9].astype(np.int64)
    want = R.tris.reshape(-1, 12)[:, 9].astype(np.int64)
    np.testing.assert_array_equal(got, want)
    assert int(roots[0, 1]) == R.root_ns
    nodes = _compare_trees(slots, int(roots[0, 0]), int(roots[0, 1]), R.slots, R.root_ns, 4)
    assert nodes == info["num_nodes"] == R.num_nodes
    assert info["max_depth"] == R.depth


@pytest.mark.parametrize("width", [2, 8])
def test_sah_tree_other_widths(oracle, width):
    V, F = W.displaced_sphere(24, 4, delete_fraction=0.05).f32()
    e = FastWindingNumber((V, F), tree_width=width, **SAH)
    slots, tris, roots, _ = e.export_tree()
    R = oracle.Restated(V, F, width=width)
    np.testing.assert_array_equal(tris.reshape(-1, 12)[:, 9], R.tris.reshape(-1, 12)[:, 9])
    _compare_trees(slots, int(roots[0, 0]), int(roots[0, 1]), R.slots, R.root_ns, width)


def test_sah_queries_match_restated_tree(oracle, open_mesh):
    V, F = open_mesh
    lo, hi = V.min(0), V.max(0)
    q = np.concatenate([W.uniform_queries(lo, hi, 8000, 21, 1.3),
                        W.near_surface_queries(W.Mesh(V.astype(np.float64), F.astype(np.int64)), 3000, 22)])
    e = FastWindingNumber((V, F), **SAH)
    om, st = e.query_stats(q)
    _, _, _, tau = e.export_tree()
    R = oracle.Restated(V, F)
    ref, rst = R.solid_angle(q, 2.0, tau=float(tau[0]), stats=True)
    same = np.all(st == rst, axis=1)
    assert same.mean() >= 0.995, same.mean()
    d = np.abs(om.astype(np.float64) - ref)
    assert d[same].max() <= 2e-4
    assert d.max() <= 4 * np.pi * 1e-3
    # and the accuracy of the restated tree against the FP64 brute force
    wb = oracle.bruteforce(V, F, q[::7]) / (4 * np.pi)
    assert np.abs(om[::7] / (4 * np.pi) - wb).max() <= 4e-3


def test_sah_buckets_and_multi_mesh(oracle):
    """Leaf buckets (leaf_max 4) on the SAH tree and a multi-mesh SAH build: each mesh of the
    batch gets the same tree and results as its own engine."""
    meshes = [W.icosphere(8).f32(), W.displaced_sphere(20, 5, delete_fraction=0.05).f32(),
              (W.icosphere(4).f32()[0], W.icosphere(4).f32()[1][:1])]
    rng = np.random.default_rng(3)
    pts = [rng.uniform(-1.3, 1.3, (2000 + 500 * k, 3)).astype(np.float32) for k in range(len(meshes))]
    mm = MultiFastWindingNumber(meshes, tree_builder=L.FWN_TREE_SAH, leaf_max=4)
    q_off = np.zeros(len(pts) + 1, np.int64)
    q_off[1:] = np.cumsum([p.shape[0] for p in pts])
    om = mm.solid_angle_batch(points=np.concatenate(pts), q_off=q_off)
    for k, (m, p) in enumerate(zip(meshes, pts)):
        e = FastWindingNumber(m, tree_builder=L.FWN_TREE_SAH, leaf_max=4)
        np.testing.assert_array_equal(om[q_off[k]:q_off[k + 1]], e.solid_angle_batch(p))
    V, F = meshes[1]
    e = FastWindingNumber((V, F), tree_builder=L.FWN_TREE_SAH, leaf_max=4)
    slots, tris, roots, tau = e.export_tree()
    leaf, _, _, _, count, _ = oracle.decode_links(slots)
    assert np.all(count[leaf] <= 4) and count[leaf].max() > 1
    q = pts[1]
    om1, st1 = e.query_stats(q)
    ref, rst = oracle.tree_eval(slots, 4, int(roots[0, 0]), int(roots[0, 1]), tris, q, float(tau[0]), stats=True)
    np.testing.assert_array_equal(st1, rst)
    assert np.abs(om1.astype(np.float64) - ref).max() <= 2e-4


def test_sah_large_mesh_matches_restated_order(oracle):
    """The level-synchronous phase (segments above 128 triangles, several chunks per segment)
    on a 200k-triangle mesh: same triangle order and tree as the restatement."""
    V, F = W.displaced_sphere(100, 4, amplitude=0.15, delete_fraction=0.05).f32()
    e = FastWindingNumber((V, F), **SAH)
    slots, tris, roots, _ = e.export_tree()
    R = oracle.Restated(V, F)
    np.testing.assert_array_equal(tris.reshape(-1, 12)[:, 9], R.tris.reshape(-1, 12)[:, 9])
    _compare_trees(slots, int(roots[0, 0]), int(roots[0, 1]), R.slots, R.root_ns, 4)


def _random_mesh(rng, kind, n):
    if kind == "soup":
        c = rng.uniform(-1, 1, (n, 3))
        V = (c[:, None, :] + rng.normal(0, 0.05, (n, 3, 3))).reshape(-1, 3)
        F = np.arange(3 * n).reshape(n, 3)
    elif kind == "sheet":  # a height field: centroids on a 2D manifold, many equal-extent ties
        k = max(2, int(np.sqrt(n / 2)))
        x, y = np.meshgrid(np.linspace(0, 1, k + 1), np.linspace(0, 1, k + 1))
        z = 0.1 * np.sin(6 * x) * np.cos(4 * y)
        V = np.stack([x, y, z], -1).reshape(-1, 3)
        i = np.arange(k)[:, None] * (k + 1) + np.arange(k)[None, :]
        a, b, c2, d = i.ravel(), i.ravel() + 1, i.ravel() + k + 1, i.ravel() + k + 2
        F = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c2], 1)])
    elif kind == "line":  # centroids along one axis with duplicates
        xs = np.repeat(rng.uniform(0, 10, n // 3 + 1), 3)[:n]
        V = np.stack([np.stack([xs, 0 * xs, 0 * xs], 1), np.stack([xs + 0.01, 0 * xs, 0 * xs], 1),
                      np.stack([xs, 0.01 + 0 * xs, 0 * xs], 1)], 1).reshape(-1, 3)
        F = np.arange(3 * n).reshape(n, 3)
    else:  # a displaced sphere with holes
        m = W.displaced_sphere(max(2, int(np.sqrt(n / 20))), int(rng.integers(100)), delete_fraction=0.05)
        return m.f32()
    return V.astype(np.float32), F.astype(np.int32)


@pytest.mark.parametrize("seed", range(8))
def test_sah_random_meshes_match_restated(oracle, seed):
    """Seeded random meshes of four kinds and sizes from 2 to ~6,000 triangles: the default
    engine's tree equals the restated tree bit for bit, and at a random accuracy scale its
    results follow the restated tree's CPU traversal (same per-point work for >= 99 % of points)."""
    rng = np.random.default_rng(1000 + seed)
    kind = ["soup", "sheet", "line", "sphere"][seed % 4]
    n = int(rng.choice([2, 7, 130, 600, 6000]))
    V, F = _random_mesh(rng, kind, n)
    beta = float(rng.choice([1.5, 2.0, 3.0]))
    e = FastWindingNumber((V, F), accuracy_scale=beta, **SAH)
    slots, tris, roots, tau = e.export_tree()
    R = oracle.Restated(V, F, beta=beta)  # the threshold words carry beta^2 r^2
    np.testing.assert_array_equal(tris.reshape(-1, 12)[:, 9], R.tris.reshape(-1, 12)[:, 9])
    _compare_trees(slots, int(roots[0, 0]), int(roots[0, 1]), R.slots, R.root_ns, 4)
    lo, hi = V.min(0), V.max(0)
    q = (lo + (hi - lo) * rng.uniform(-0.2, 1.2, (3000, 3))).astype(np.float32)
    om, st = e.query_stats(q)
    ref, rst = R.solid_angle(q, beta, tau=float(tau[0]), stats=True)
    same = np.all(st == rst, axis=1)
    assert same.mean() >= 0.99, (kind, n, beta, same.mean())
    assert np.abs(om.astype(np.float64) - ref)[same].max() <= 2e-4
