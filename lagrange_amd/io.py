"""OBJ / PLY mesh input and output (ctypes over ``libfwn_io.so``, include/fwn_io.h).

Counterpart of ``lagrange.io.load_mesh`` / ``save_mesh`` for OBJ and PLY
(modules/io/include/lagrange/io/load_mesh_obj.h, load_mesh_ply.h, save_mesh_obj.h,
save_mesh_ply.h), keeping what the winding-number engine consumes: vertex positions and
polygonal facets.  ``triangulate=True`` splits polygons into fans (LoadOptions::triangulate,
modules/io/include/lagrange/io/types.h:99).  Host code; no GPU involved.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfwn_io.so")

FORMAT_AUTO, FORMAT_OBJ, FORMAT_PLY, FORMAT_PLY_BINARY = 0, 1, 2, 3
_STATUS = {1: "INVALID_ARGUMENT", 2: "FILE_ERROR", 3: "PARSE_ERROR"}

_lib = None


class MeshIOError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{_STATUS.get(status, status)}] {message}")
        self.status = status


def load_lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lagrange_amd.build`")
    L = C.CDLL(LIB_PATH)
    VP, I64, I32 = C.c_void_p, C.c_int64, C.c_int32
    sig = {
        "fwn_mesh_load": ([C.c_char_p, I32, I32, C.POINTER(VP)], I32),
        "fwn_mesh_load_buffer": ([C.c_char_p, I64, I32, I32, C.POINTER(VP)], I32),
        "fwn_mesh_sizes": ([VP, VP, VP, VP], I32),
        "fwn_mesh_copy": ([VP, VP, VP, VP], I32),
        "fwn_mesh_free": ([VP], I32),
        "fwn_mesh_save": ([C.c_char_p, I32, VP, I64, VP, VP, I64], I32),
        "fwn_io_last_error": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(st: int):
    if st != 0:
        raise MeshIOError(st, load_lib().fwn_io_last_error().decode())


class PolygonMesh:
    """Positions (n, 3) float64 and facets in CSR form (offsets, corners)."""

    def __init__(self, vertices, offsets, corners):
        self.vertices = np.asarray(vertices, np.float64).reshape(-1, 3)
        self.offsets = np.asarray(offsets, np.int64)
        self.corners = np.asarray(corners, np.int32)

    @property
    def num_facets(self) -> int:
        return self.offsets.size - 1

    def is_triangle_mesh(self) -> bool:
        return bool((np.diff(self.offsets) == 3).all())

    @property
    def facets(self) -> np.ndarray:
        """(F, 3) int32 facets of a triangle mesh."""
        if not self.is_triangle_mesh():
            raise ValueError("not a triangle mesh: load with triangulate=True")
        return self.corners.reshape(-1, 3)


def _take(h) -> PolygonMesh:
    L = load_lib()
    try:
        nv, nf, nc = C.c_int64(), C.c_int64(), C.c_int64()
        _check(L.fwn_mesh_sizes(h, C.byref(nv), C.byref(nf), C.byref(nc)))
        V = np.empty((nv.value, 3), np.float64)
        off = np.empty(nf.value + 1, np.int64)
        cor = np.empty(nc.value, np.int32)
        _check(L.fwn_mesh_copy(h, V.ctypes.data, off.ctypes.data, cor.ctypes.data))
        return PolygonMesh(V, off, cor)
    finally:
        L.fwn_mesh_free(h)


def load_mesh(path, triangulate: bool = False, fmt: int = FORMAT_AUTO) -> PolygonMesh:
    """Read an OBJ or PLY file (format from the extension unless ``fmt`` is given)."""
    h = C.c_void_p()
    _check(load_lib().fwn_mesh_load(os.fsencode(path), fmt, int(triangulate), C.byref(h)))
    return _take(h)


def load_mesh_buffer(data: bytes, fmt: int, triangulate: bool = False) -> PolygonMesh:
    h = C.c_void_p()
    _check(load_lib().fwn_mesh_load_buffer(data, len(data), fmt, int(triangulate), C.byref(h)))
    return _take(h)


def save_mesh(path, vertices, facets=None, *, offsets=None, corners=None, fmt: int = FORMAT_AUTO):
    """Write positions and facets ((F, k) array, or CSR offsets + corners).  FORMAT_AUTO picks OBJ
    or binary PLY from the extension."""
    V = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1, 3))
    if facets is not None:
        Fa = np.asarray(facets)
        corners = Fa.reshape(-1)
        offsets = np.arange(Fa.shape[0] + 1, dtype=np.int64) * Fa.shape[1]
    off = np.ascontiguousarray(np.asarray(offsets, np.int64))
    cor = np.ascontiguousarray(np.asarray(corners, np.int64).astype(np.int32))
    if fmt == FORMAT_AUTO:
        ext = os.path.splitext(os.fspath(path))[1].lower()
        fmt = {".obj": FORMAT_OBJ, ".ply": FORMAT_PLY_BINARY}.get(ext)
        if fmt is None:
            raise MeshIOError(1, f"unknown mesh file extension: {path}")
    _check(load_lib().fwn_mesh_save(os.fsencode(path), fmt, V.ctypes.data, V.shape[0], off.ctypes.data,
                                    cor.ctypes.data, off.size - 1))
