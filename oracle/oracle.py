"""Python access to the CPU oracle.  TEST INFRASTRUCTURE ONLY (tests/, smoke(), and the
CPU-baseline leg of bench.py); the engine never imports this module.

Two independent FP64 brute forces exist: ``bruteforce`` (C, OpenMP) and
``bruteforce_numpy`` (vectorised numpy).  Tests require them to agree to ~1e-12 so a slip in
either is caught.  ``Restated`` is the CPU restatement of the reference engine
(jdumas/WindingNumber's UT_SolidAngle<float,float>, called at
modules/winding/src/FastWindingNumber.cpp:57,66,75): 4-ary box-area BVH, order-2 far field,
accuracy scale 2, FP32.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
U8P = C.POINTER(C.c_uint8)

WIDTH_RESTATED = 4
SLOT_WORDS = 24
LINK = 23        # slot word: (child << 4) | child slots, or LINK_LEAF | first << 3 | count - 1
LINK_LEAF = 0x80000000
LINK_EMPTY = 0xFFFFFFFF
TRI_WORDS = 12


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.oracle_tri_f64.restype = C.c_double
        L.oracle_tri_f64.argtypes = [DP, DP, DP, DP]
        L.oracle_bruteforce.argtypes = [FP, C.c_int64, I32P, C.c_int64, FP, C.c_int64, DP, C.c_int]
        L.oracle_moments.argtypes = [FP, I32P, I64P, C.c_int64, DP]
        L.restated_build.restype = C.c_void_p
        L.restated_build.argtypes = [FP, C.c_int64, I32P, C.c_int64]
        L.restated_build_w.restype = C.c_void_p
        L.restated_build_w.argtypes = [FP, C.c_int64, I32P, C.c_int64, C.c_int32]
        L.restated_free.argtypes = [C.c_void_p]
        L.restated_num_nodes.restype = C.c_int64
        L.restated_num_nodes.argtypes = [C.c_void_p]
        L.restated_root_ns.restype = C.c_int32
        L.restated_root_ns.argtypes = [C.c_void_p]
        L.restated_depth.restype = C.c_int32
        L.restated_depth.argtypes = [C.c_void_p]
        L.restated_slots.restype = U32P
        L.restated_slots.argtypes = [C.c_void_p]
        L.restated_tris.restype = U32P
        L.restated_tris.argtypes = [C.c_void_p]
        L.restated_r2.restype = FP
        L.restated_r2.argtypes = [C.c_void_p]
        L.set_thresholds.argtypes = [U32P, C.c_int64, U32P, FP, C.c_float, C.c_float]
        L.tree_eval.argtypes = [U32P, C.c_int32, C.c_int32, C.c_int32, U32P, FP, C.c_int64, C.c_float, FP,
                                I32P, C.c_int]
        L.inside_from_omega.argtypes = [FP, C.c_int64, U8P]
        L.lbvh_sorted_order.argtypes = [C.c_int32, FP, I64P, I32P, I64P, I64P, U64P]
        L.oracle_tri_closest.restype = C.c_double
        L.oracle_tri_closest.argtypes = [DP, DP, DP, DP, DP]
        L.oracle_closest.argtypes = [FP, I32P, C.c_int64, FP, C.c_int64, DP, I32P, DP, C.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def bruteforce(V, F, Q, nthreads: int = 0) -> np.ndarray:
    """FP64 sum of exact triangle solid angles (steradians) at float32 points Q."""
    V, F, Q = _f32(V), _i32(F), _f32(Q)
    out = np.empty(Q.shape[0], np.float64)
    lib().oracle_bruteforce(_p(V, FP), V.shape[0], _p(F, I32P), F.shape[0], _p(Q, FP), Q.shape[0],
                            _p(out, DP), nthreads)
    return out


def tri_closest(v0, v1, v2, q):
    """(squared distance, (s, t)) from q to triangle (v0, v1, v2); closest = v0 + s e0 + t e1."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (v0, v1, v2, q)]
    st = np.zeros(2)
    d = lib().oracle_tri_closest(*[_p(x, DP) for x in a], _p(st, DP))
    return d, st


def closest(V, F, Q, nthreads: int = 0):
    """Brute-force closest point: (squared distance, facet, closest point) per point."""
    V, F, Q = _f32(V), _i32(F), _f32(Q)
    d2 = np.empty(Q.shape[0], np.float64)
    fi = np.empty(Q.shape[0], np.int32)
    cp = np.empty((Q.shape[0], 3), np.float64)
    lib().oracle_closest(_p(V, FP), _p(F, I32P), F.shape[0], _p(Q, FP), Q.shape[0], _p(d2, DP), _p(fi, I32P),
                         _p(cp, DP), nthreads)
    return d2, fi, cp


def tri_f64(v0, v1, v2, q) -> float:
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (v0, v1, v2, q)]
    return lib().oracle_tri_f64(*[_p(x, DP) for x in a])


def bruteforce_numpy(V, F, Q, chunk: int = 256) -> np.ndarray:
    """Second FP64 brute force in numpy (van Oosterom-Strackee), for small cases."""
    V = np.asarray(V, np.float32).astype(np.float64)
    F = np.asarray(F, np.int64)
    Q = np.asarray(Q, np.float32).astype(np.float64)
    A, B, Cc = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    out = np.empty(Q.shape[0])
    for s in range(0, Q.shape[0], chunk):
        q = Q[s:s + chunk, None, :]
        a, b, c = A[None] - q, B[None] - q, Cc[None] - q
        la, lb, lc = (np.linalg.norm(x, axis=2) for x in (a, b, c))
        num = np.einsum("qtd,qtd->qt", a, np.cross(b, c))
        den = (la * lb * lc + np.einsum("qtd,qtd->qt", a, b) * lc + np.einsum("qtd,qtd->qt", b, c) * la
               + np.einsum("qtd,qtd->qt", c, a) * lb)
        om = 2.0 * np.arctan2(num, den)
        om[(la == 0) | (lb == 0) | (lc == 0) | (num == 0)] = 0.0
        out[s:s + chunk] = om.sum(axis=1)
    return out


def moments(V, F, tri_ids) -> np.ndarray:
    """FP64 order-2 moment record (33 doubles) of a triangle set, computed directly."""
    V, F = _f32(V), _i32(F)
    ids = np.ascontiguousarray(tri_ids, dtype=np.int64)
    out = np.empty(33, np.float64)
    lib().oracle_moments(_p(V, FP), _p(F, I32P), _p(ids, I64P), ids.size, _p(out, DP))
    return out


def inside_from_omega(om) -> np.ndarray:
    om = _f32(om)
    out = np.empty(om.shape[0], np.uint8)
    lib().inside_from_omega(_p(om, FP), om.shape[0], _p(out, U8P))
    return out.astype(bool)


def tree_eval(slots, width, root, root_ns, tris, Q, tau=0.0, stats=False, nthreads=0):
    """FP32 traversal of a slot tree (the engine's exported format, or a Restated tree's): each
    slot's far-field threshold and each leaf's FP64 near-field thresholds are read from the
    arrays; tau > 0 enables the FP64 near field."""
    slots = np.ascontiguousarray(slots, dtype=np.uint32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    Q = _f32(Q)
    out = np.empty(Q.shape[0], np.float32)
    st = np.empty((Q.shape[0], 4), np.int32) if stats else None
    lib().tree_eval(_p(slots, U32P), width, root, root_ns, _p(tris, U32P), _p(Q, FP), Q.shape[0], tau,
                    _p(out, FP), _p(st, I32P) if stats else None, nthreads)
    return (out, st) if stats else out


def decode_links(slots):
    """Per slot of a slot array: (is_leaf, child node, child slot count, first triangle, count,
    empty), decoded from the link word."""
    link = np.ascontiguousarray(slots, dtype=np.uint32).reshape(-1, SLOT_WORDS)[:, LINK].astype(np.int64)
    empty = link == LINK_EMPTY
    leaf = ((link & LINK_LEAF) != 0) & ~empty
    inner = ~leaf & ~empty
    child = np.where(inner, link >> 4, -1)
    child_ns = np.where(inner, link & 15, 0)
    first = np.where(leaf, (link & 0x7FFFFFFF) >> 3, -1)
    count = np.where(leaf, (link & 7) + 1, 0)
    return leaf, child, child_ns, first, count, empty


def slot_ranges(slots, width, roots):
    """Triangle range (first, count) covered by every slot of a slot tree, from its leaves
    (internal slots store only their child link).  roots: (root node, root slot count) pairs."""
    leaf, child, child_ns, first, count, empty = decode_links(slots)
    n = first.shape[0]
    first = first.copy()
    count = count.copy()
    # level order from the roots, then internal slots take the union of their child's slots
    levels = []
    cur = np.array([r for r, ns in roots if r >= 0], np.int64)
    while cur.size:
        levels.append(cur)
        sl = (cur[:, None] * width + np.arange(width)[None, :]).ravel()
        sl = sl[sl < n]
        nxt = child[sl]
        cur = np.unique(nxt[nxt >= 0])
    for nodes in reversed(levels):
        sl = (nodes[:, None] * width + np.arange(width)[None, :]).ravel()
        inner = sl[child[sl] >= 0]
        cs = child[inner][:, None] * width + np.arange(width)[None, :]
        occ = ~empty[cs]
        cf = np.where(occ, first[cs], np.iinfo(np.int64).max)
        first[inner] = cf.min(axis=1)
        count[inner] = np.where(occ, count[cs], 0).sum(axis=1)
    return first, count


def compare_slot_trees(gs, groot, gns, rs, rns, width, rtol=2e-5):
    """Walks two slot trees together, one level at a time: every slot's leaf / internal kind,
    leaf triangle range and child slot count equal, float words within rtol of the slot's
    largest word (node numbering may differ).  Returns the number of nodes visited."""
    gs = np.ascontiguousarray(gs, np.uint32).reshape(-1, SLOT_WORDS)
    rs = np.ascontiguousarray(rs, np.uint32).reshape(-1, SLOT_WORDS)
    gf, rf = gs.view(np.float32), rs.view(np.float32)
    gl, gc, gcn, gfi, gco, _ = decode_links(gs)
    rl, rc, rcn, rfi, rco, _ = decode_links(rs)
    assert gns == rns, (gns, rns)
    g = np.array([groot], np.int64)
    r = np.array([0], np.int64)
    ns = np.array([gns], np.int64)
    nodes = 0
    while g.size:
        nodes += g.size
        nxt_g, nxt_r, nxt_ns = [], [], []
        for k in range(width):
            sel = ns > k
            a, b = g[sel] * width + k, r[sel] * width + k
            np.testing.assert_array_equal(gl[a], rl[b])
            np.testing.assert_array_equal(gfi[a], rfi[b])
            np.testing.assert_array_equal(gco[a], rco[b])
            np.testing.assert_array_equal(gcn[a], rcn[b])
            gv, rv = gf[a, :LINK].astype(np.float64), rf[b, :LINK].astype(np.float64)
            scale = np.maximum(np.abs(rv).max(axis=1), 1e-30)
            bad = np.abs(gv - rv).max(axis=1) > rtol * scale
            assert not bad.any(), (a[bad][:5], (np.abs(gv - rv).max(axis=1) / scale)[bad][:5])
            inner = ~gl[a]
            nxt_g.append(gc[a][inner])
            nxt_r.append(rc[b][inner])
            nxt_ns.append(gcn[a][inner])
        g, r, ns = (np.concatenate(x).astype(np.int64) for x in (nxt_g, nxt_r, nxt_ns))
    return nodes


class Restated:
    """CPU restatement of the reference engine (FP32, order 2, beta = 2 by default).  `slots` and
    `tris` carry the thresholds of accuracy scale `beta` and FP64 distance `tau` (0: FP32 only,
    the reference engine); `thresholds()` gives them for other values."""

    def __init__(self, V, F, width: int = WIDTH_RESTATED, beta: float = 2.0, tau: float = 0.0):
        """width 4 restates the reference's 4-ary tree; 2 and 8 exist for comparison."""
        self.V, self.F = _f32(V), _i32(F)
        self.width = int(width)
        L = lib()
        self._h = L.restated_build_w(_p(self.V, FP), self.V.shape[0], _p(self.F, I32P), self.F.shape[0], self.width)
        n = L.restated_num_nodes(self._h)
        self.num_nodes = int(n)
        self.root_ns = int(L.restated_root_ns(self._h))
        self.depth = int(L.restated_depth(self._h))
        nslots = max(n, 1) * self.width
        self._slots0 = np.ctypeslib.as_array(L.restated_slots(self._h), shape=(nslots * SLOT_WORDS,)).copy()
        self._tris0 = np.ctypeslib.as_array(L.restated_tris(self._h),
                                            shape=(max(self.F.shape[0], 1) * TRI_WORDS,)).copy()
        self.r2 = np.ctypeslib.as_array(L.restated_r2(self._h), shape=(nslots,)).copy() if n > 0 else \
            np.zeros(nslots, np.float32)
        L.restated_free(self._h)
        self._h = None
        self.beta, self.tau = float(beta), float(tau)
        self.slots, self.tris = self.thresholds(self.beta, self.tau)
        self._cache = {(self.beta, self.tau): (self.slots, self.tris)}

    def thresholds(self, beta: float, tau: float = 0.0):
        """(slots, tris) copies with the far-field thresholds of accuracy scale beta and the
        leaves' FP64 near-field thresholds of distance tau."""
        slots, tris = self._slots0.copy(), self._tris0.copy()
        r2 = np.ascontiguousarray(self.r2, np.float32)
        lib().set_thresholds(_p(slots, U32P), r2.shape[0], _p(tris, U32P), _p(r2, FP), beta, tau)
        return slots, tris

    def solid_angle(self, Q, beta: float = 2.0, stats: bool = False, nthreads: int = 0, tau: float = 0.0):
        key = (float(beta), float(tau))
        if key not in self._cache:
            self._cache[key] = self.thresholds(*key)
        slots, tris = self._cache[key]
        return tree_eval(slots, self.width, 0, self.root_ns, tris, Q, tau, stats, nthreads)

    def is_inside(self, Q, beta: float = 2.0):
        return inside_from_omega(self.solid_angle(Q, beta))


def lbvh_sorted_order(meshes_V, meshes_F):
    """Facet order (global ids) and 64-bit keys a Morton LBVH over these meshes must produce."""
    V = _f32(np.concatenate(meshes_V))
    F = _i32(np.concatenate(meshes_F))
    v_off = np.zeros(len(meshes_V) + 1, np.int64)
    f_off = np.zeros(len(meshes_F) + 1, np.int64)
    v_off[1:] = np.cumsum([x.shape[0] for x in meshes_V])
    f_off[1:] = np.cumsum([x.shape[0] for x in meshes_F])
    order = np.empty(F.shape[0], np.int64)
    keys = np.empty(F.shape[0], np.uint64)
    lib().lbvh_sorted_order(len(meshes_V), _p(V, FP), _p(v_off, I64P), _p(F, I32P), _p(f_off, I64P),
                            _p(order, I64P), _p(keys, U64P))
    return order, keys
