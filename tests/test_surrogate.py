"""Learned surrogate: CPU checks of the FP64 reference network (finite differences) and GPU
checks of the engine's kernels against it, plus the trained network's gain.

Gradient tolerance (GPU vs emulating reference): relative L2 error <= 2e-2 per parameter block.
Both sides round the same values to BF16/FP16; what remains is FP32 accumulation order and rare
BF16 rounding-boundary flips.  A transposed operand, a missing ReLU mask or a wrong trilinear
weight gives O(1) errors."""
import numpy as np
import pytest

from oracle import surrogate_ref as R

LEVELS = 16


def _grid(T, base=16, top=2048):
    g = np.exp(np.log(top / base) / (LEVELS - 1))
    res = [int(np.floor(base * g ** l + 0.5)) for l in range(LEVELS)]
    dense = [(r + 1) ** 3 <= T for r in res]
    return res, dense


def _rand_params(T, seed):
    rng = np.random.default_rng(seed)
    table = rng.uniform(-0.5, 0.5, LEVELS * T * 2).astype(np.float32)
    sizes = [32 * 128, 128, 128 * 128, 128, 128 * 128, 128, 128 * 128, 128, 128, 1]
    mlp = np.concatenate([rng.uniform(-1, 1, s) * (np.sqrt(6 / (32 if i == 0 else 128)) if i % 2 == 0 else 0.1)
                          for i, s in enumerate(sizes)]).astype(np.float32)
    return table, mlp


def test_reference_gradients_finite_difference():
    T = 1 << 10
    res, dense = _grid(T)
    table, mlp = _rand_params(T, 1)
    rng = np.random.default_rng(2)
    P = rng.uniform(0, 1, (64, 3)).astype(np.float32)
    y = rng.uniform(0, 1, 64)
    lo, inv = np.zeros(3), np.ones(3)
    loss, pred, gt, gm = R.forward_backward(P, y, lo, inv, res, dense, T, table.astype(np.float64),
                                            mlp.astype(np.float64), emulate=False)

    def L(t, m):
        return R.forward_backward(P, y, lo, inv, res, dense, T, t, m, emulate=False)[0]

    h = 1e-6
    t64, m64 = table.astype(np.float64), mlp.astype(np.float64)
    for i in rng.choice(np.flatnonzero(np.abs(gm) > 1e-8), 12, replace=False):
        mp, mm = m64.copy(), m64.copy()
        mp[i] += h
        mm[i] -= h
        fd = (L(t64, mp) - L(t64, mm)) / (2 * h)
        assert fd == pytest.approx(gm[i], rel=1e-4, abs=1e-9)
    for i in rng.choice(np.flatnonzero(np.abs(gt) > 1e-8), 12, replace=False):
        tp, tm = t64.copy(), t64.copy()
        tp[i] += h
        tm[i] -= h
        fd = (L(tp, m64) - L(tm, m64)) / (2 * h)
        assert fd == pytest.approx(gt[i], rel=1e-4, abs=1e-9)


def _blocks(g):
    sizes = [32 * 128, 128, 128 * 128, 128, 128 * 128, 128, 128 * 128, 128, 128, 1]
    out, off = [], 0
    for s in sizes:
        out.append(g[off:off + s])
        off += s
    return out


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.gpu
def test_gpu_gradients_match_reference():
    from lagrange_amd import workloads as W
    from lagrange_amd.surrogate import Surrogate
    from lagrange_amd.winding import FastWindingNumber
    V, F = W.icosphere(8).f32()
    e = FastWindingNumber((V, F))
    s = Surrogate(e, log2_table=12, grid_res=64, n_samples=4096, batch=4096, steps=0)
    info = s.info()
    T = 1 << 12
    table, mlp = _rand_params(T, 3)
    s.set_params(table, mlp)
    rng = np.random.default_rng(4)
    lo, hi = V.min(0), V.max(0)
    P = (lo + (hi - lo) * rng.uniform(-0.05, 1.05, (1000, 3))).astype(np.float32)
    y = rng.uniform(0, 1, 1000).astype(np.float32)
    loss, pred, gt, gm = s.gradients(P, y)
    inv = np.where(hi > lo, 1 / (hi - lo), 0).astype(np.float32)
    rl, rp, rgt, rgm = R.forward_backward(P, y, lo, inv, info["res"], info["dense"], T, table.astype(np.float64),
                                          mlp.astype(np.float64), emulate=True)
    assert _rel(pred, rp) <= 1e-3
    assert loss == pytest.approx(rl, rel=1e-3)
    # the fused inference kernels (register-resident MFMA layers) give the training forward's values
    # (FP32 sums in another order; a BF16 rounding-boundary flip of one activation moves a point by ~1e-3)
    fused, _ = s.predict(P)
    d = np.abs(fused - pred)
    assert np.median(d) <= 1e-6 and d.max() <= 1e-2
    for g, r in zip(_blocks(gm), _blocks(rgm)):
        assert _rel(g, r) <= 2e-2, (_rel(g, r), g.size)
    assert _rel(gt, rgt) <= 2e-2


@pytest.mark.gpu
def test_trained_surrogate_gain_and_exact_fallback():
    """Trained vs untrained network on the same fallback rule: the trained one answers most
    uniform points itself with at most 1e-4 misclassification; fallback points carry exact values."""
    from lagrange_amd import workloads as W
    from lagrange_amd.surrogate import Surrogate
    from lagrange_amd.winding import FastWindingNumber
    m = W.displaced_sphere(48, 2, amplitude=0.15, delete_fraction=0.05)
    V, F = m.f32()
    e = FastWindingNumber((V, F))
    q = W.uniform_queries(V.min(0), V.max(0), 300_000, 91)
    om, ins_exact = e.solid_angle_batch(q, with_inside=True)
    untrained = Surrogate(e, n_samples=1_000_000, steps=0, grid_res=256)
    _, ins_u, nf_u = untrained.query(q)
    s = Surrogate(e, n_samples=2_000_000, steps=800, grid_res=256)
    info = s.train()
    w, ins, nf = s.query(q)
    pred, fb = s.predict(q)
    # fallback points are exactly the exact engine's values
    np.testing.assert_array_equal(w[fb], (om[fb].astype(np.float64) / (4 * np.pi)).astype(np.float32))
    assert np.array_equal(ins[fb], ins_exact[fb])
    mis = np.mean(ins != ins_exact)
    mis_u = np.mean(ins_u != ins_exact)
    print(f"loss {info['last_loss']:.3e} fallback {nf / q.shape[0]:.3f} (untrained {nf_u / q.shape[0]:.3f}) "
          f"misclassified {mis:.2e} (untrained {mis_u:.2e})")
    assert nf / q.shape[0] <= 0.35 and nf < nf_u
    assert mis <= 1e-4
    assert info["last_loss"] < 0.05


@pytest.mark.gpu
def test_sorted_inference_equals_unsorted_chunks():
    """Large batches are encoded in Morton order and scattered back: per point the result is the
    same as for small batches, which run in input order."""
    from lagrange_amd import workloads as W
    from lagrange_amd.surrogate import Surrogate
    from lagrange_amd.winding import FastWindingNumber
    V, F = W.displaced_sphere(16, 2, delete_fraction=0.05).f32()
    e = FastWindingNumber((V, F))
    s = Surrogate(e, n_samples=200_000, steps=50, grid_res=128, batch=65536)
    s.train()
    q = W.uniform_queries(V.min(0), V.max(0), 150_000, 33)
    big, fb_big = s.predict(q)
    parts = [s.predict(q[i:i + 40_000]) for i in range(0, q.shape[0], 40_000)]
    np.testing.assert_array_equal(big, np.concatenate([p[0] for p in parts]))
    np.testing.assert_array_equal(fb_big, np.concatenate([p[1] for p in parts]))
    w, ins, nf = s.query(q)
    assert nf == int(fb_big.sum())


@pytest.mark.gpu
def test_c5_surrogate_at_config_size(oracle):
    """The surrogate's own workload at full size: the 953,344-triangle open displaced sphere, the
    default options (16 levels x 2^19 x 2 FP16, MLP 4 x 128, 20,000,000 training points labelled
    by the exact engine, band 0.2 around w = 1/2, distance-bound grid), 67,108,864 points uniform
    in the box, answered by surrogate + exact fallback:
      * inside/outside misclassification against the exact engine <= 1e-4;
      * every fallback point carries the exact engine's value bit for bit, and the fallback count
        is the number of flagged points;
      * on a 512-point sample the flags equal the CPU FP64 brute force wherever it is more than
        1e-3 from w = 1/2."""
    import torch
    from lagrange_amd import workloads as W
    from lagrange_amd.surrogate import Surrogate
    from lagrange_amd.winding import FastWindingNumber
    V, F = W.large_displaced_mesh().f32()
    assert F.shape[0] == 953_344
    e = FastWindingNumber((V, F))
    s = Surrogate(e)
    o = s.options
    assert (o.levels, o.log2_table, o.hidden_layers, o.n_samples) == (16, 19, 4, 20_000_000)
    info = s.train()
    n = 1 << 26
    qd = torch.from_numpy(W.uniform_queries(V.min(0), V.max(0), n, 5, 1.0)).cuda()
    w, ins, nf = s.query(qd)
    om, ins_e = e.solid_angle_batch(qd, with_inside=True)
    pred, fb = s.predict(qd)
    assert nf == int(fb.sum().item())
    mis = int((ins != ins_e).sum().item())
    print(f"C5: train {info['train_ms'] / 1e3:.1f} s, loss {info['last_loss']:.2e}, fallback {nf / n:.4f}, "
          f"misclassified {mis} of {n}", flush=True)
    assert mis <= 1e-4 * n
    exact_w = (om.double() / (4 * np.pi)).float()
    assert torch.equal(w[fb], exact_w[fb])
    assert torch.equal(ins[fb], ins_e[fb])
    # sample indices in int64 from a float64 linspace (a float32 one rounds n - 1 = 2^26 - 1 up to n)
    idx = torch.from_numpy(np.linspace(0, n - 1, 512).astype(np.int64)).to(qd.device)
    assert int(idx.max()) < n
    wb = oracle.bruteforce(V, F, qd[idx].cpu().numpy()) / (4 * np.pi)
    band = np.abs(wb - 0.5) > 1e-3
    np.testing.assert_array_equal(ins[idx].cpu().numpy()[band], (wb > 0.5)[band])
