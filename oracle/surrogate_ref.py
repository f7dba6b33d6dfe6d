"""CPU restatement of the surrogate's forward and backward pass (TEST INFRASTRUCTURE ONLY).

Same network as lagrange_amd/csrc/fwn_surrogate.hip: multi-resolution hash grid (dense levels
indexed x + (r+1)(y + (r+1) z); hashed levels in 4 x 4 x 4 vertex blocks: block b =
(x/4) ^ 2654435761 (y/4) ^ 805459861 (z/4) mod T/64, entry 64 b + 16 (z%4) + 4 (y%4) + x%4;
trilinear weights), 32 -> 128 x 4 -> 1 ReLU MLP, mean squared error.  Arithmetic is FP64; with
``emulate=True`` the values the engine rounds are rounded the same way (FP16 table, BF16 encoded
features, BF16 weights and hidden activations, BF16 output-layer data gradient and hidden data
gradients), so the engine's gradients can be compared tightly.  ``emulate=False`` is the plain
FP64 network used for the finite-difference check of these formulas.
"""
from __future__ import annotations

import numpy as np

H = 128


def bf16(x):
    """Round float64/float32 values to bfloat16 (nearest even) and return float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float64)


def layout(mlp):
    """Views of the packed MLP vector: W0 (32x128) b0 W1 b1 W2 b2 W3 b3 w4 (128) b4 (1)."""
    out, off = [], 0
    for k in (32, H, H, H):
        W = mlp[off:off + k * H].reshape(k, H)
        off += k * H
        b = mlp[off:off + H]
        off += H
        out.append((W, b))
    w4 = mlp[off:off + H]
    b4 = mlp[off + H:off + H + 1]
    return out, w4, b4


def level_tables(res, dense, T):
    return [(int(r), bool(d)) for r, d in zip(res, dense)]


def corners(p, res, dense, T):
    """Per point: 8 corner indices and trilinear weights on one level (float32 cell math)."""
    r = np.float32(res)
    s = np.clip(p.astype(np.float32), 0, 1) * r
    i = np.minimum(np.floor(s), r - 1)
    f = (s - i).astype(np.float64)
    i = i.astype(np.uint64)
    idx, w = [], []
    for k in range(8):
        dx, dy, dz = k & 1, (k >> 1) & 1, k >> 2
        x, y, z = i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz
        if dense:
            n = np.uint64(res + 1)
            ix = x + n * (y + n * z)
        else:
            # vertices in 4 x 4 x 4 blocks: the block coordinates hashed, the vertex's place
            # in its block in the low 6 bits
            bx, by, bz = x >> np.uint64(2), y >> np.uint64(2), z >> np.uint64(2)
            b = (bx ^ ((by * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF))
                 ^ ((bz * np.uint64(805459861)) & np.uint64(0xFFFFFFFF))) & np.uint64((T - 1) >> 6)
            ix = (b << np.uint64(6)) | ((z & np.uint64(3)) << np.uint64(4)) | ((y & np.uint64(3)) << np.uint64(2)) \
                | (x & np.uint64(3))
        wx = f[:, 0] if dx else 1 - f[:, 0]
        wy = f[:, 1] if dy else 1 - f[:, 1]
        wz = f[:, 2] if dz else 1 - f[:, 2]
        idx.append(ix.astype(np.int64))
        w.append(wx * wy * wz)
    return np.stack(idx, 1), np.stack(w, 1)


def forward_backward(P, y, lo, inv, res, dense, T, table, mlp, emulate=True):
    """Returns loss, pred, g_table (same shape as table), g_mlp (packed)."""
    P = np.asarray(P, np.float32)
    n = P.shape[0]
    L = len(res)
    p = ((P - np.asarray(lo, np.float32)) * np.asarray(inv, np.float32)).astype(np.float32)
    tab = table.reshape(L, T, 2).astype(np.float64)
    if emulate:
        tab = f16(tab)
    feats = np.zeros((n, 2 * L))
    cache = []
    for l in range(L):
        idx, w = corners(p, res[l], dense[l], T)
        cache.append((idx, w))
        v = tab[l][idx]  # n x 8 x 2
        feats[:, 2 * l] = (w * v[..., 0]).sum(1)
        feats[:, 2 * l + 1] = (w * v[..., 1]).sum(1)
    layers, w4, b4 = layout(mlp.astype(np.float64))
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    X = [rnd(feats)]
    for W, b in layers:
        X.append(rnd(np.maximum(X[-1] @ rnd(W) + b, 0.0)))
    pred = X[-1] @ w4 + b4[0]
    e = pred - np.asarray(y, np.float64)
    loss = np.mean(e * e)
    dpred = 2.0 * e / n
    g_mlp = np.zeros_like(mlp, dtype=np.float64)
    gl, gw4, gb4 = layout(g_mlp)
    gw4[:] = X[-1].T @ dpred
    gb4[:] = dpred.sum()
    dP = rnd(np.outer(dpred, w4) * (X[-1] > 0))
    for l in range(3, -1, -1):
        W, _ = layers[l]
        gl[l][0][:] = X[l].T @ dP
        gl[l][1][:] = dP.sum(0)
        if l > 0:
            dP = rnd((dP @ rnd(W).T) * (X[l] > 0))
        else:
            dX0 = dP @ rnd(W).T
    g_tab = np.zeros((L, T, 2))
    for l in range(L):
        idx, w = cache[l]
        for f in range(2):
            np.add.at(g_tab[l][:, f], idx.reshape(-1), (w * dX0[:, 2 * l + f][:, None]).reshape(-1))
    return loss, pred, g_tab.reshape(-1), g_mlp
