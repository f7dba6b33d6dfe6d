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
SLOT_WORDS = 32
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
        L.tree_eval.argtypes = [U32P, C.c_int32, C.c_int32, C.c_int32, U32P, FP, C.c_int64, C.c_float,
                                C.c_float, FP, I32P, C.c_int]
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


def tree_eval(slots, width, root, root_ns, tris, Q, beta=2.0, tau=0.0, stats=False, nthreads=0):
    slots = np.ascontiguousarray(slots, dtype=np.uint32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    Q = _f32(Q)
    out = np.empty(Q.shape[0], np.float32)
    st = np.empty((Q.shape[0], 4), np.int32) if stats else None
    lib().tree_eval(_p(slots, U32P), width, root, root_ns, _p(tris, U32P), _p(Q, FP), Q.shape[0], beta, tau,
                    _p(out, FP), _p(st, I32P) if stats else None, nthreads)
    return (out, st) if stats else out


class Restated:
    """CPU restatement of the reference engine (FP32, order 2, beta = 2 by default)."""

    def __init__(self, V, F, width: int = WIDTH_RESTATED):
        """width 4 restates the reference's 4-ary tree; 2 and 8 exist for comparison."""
        self.V, self.F = _f32(V), _i32(F)
        self.width = int(width)
        L = lib()
        self._h = L.restated_build_w(_p(self.V, FP), self.V.shape[0], _p(self.F, I32P), self.F.shape[0], self.width)
        n = L.restated_num_nodes(self._h)
        self.num_nodes = int(n)
        self.root_ns = int(L.restated_root_ns(self._h))
        self.depth = int(L.restated_depth(self._h))
        self.slots = np.ctypeslib.as_array(L.restated_slots(self._h),
                                           shape=(max(n, 1) * self.width * SLOT_WORDS,)).copy()
        self.tris = np.ctypeslib.as_array(L.restated_tris(self._h),
                                          shape=(max(self.F.shape[0], 1) * TRI_WORDS,)).copy()
        L.restated_free(self._h)
        self._h = None

    def solid_angle(self, Q, beta: float = 2.0, stats: bool = False, nthreads: int = 0):
        return tree_eval(self.slots, self.width, 0, self.root_ns, self.tris, Q, beta, 0.0, stats, nthreads)

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
