"""Winding-number voxel signing, the caller side of ``lagrange::volume::mesh_to_volume``.

With ``MeshToVolumeOptions::Sign::WindingNumber`` the reference builds a FastWindingNumber engine
and lets OpenVDB call ``engine.is_inside(float(transform.indexToWorld(ijk)))`` once per voxel
from its worker threads (modules/volume/src/mesh_to_volume.cpp:156-183).  The transform is a
uniform scale by ``voxel_size`` post-translated by ``voxel_size / 2`` (:147-149), and a negative
``voxel_size`` is relative to the mesh's bounding-box diagonal (:132-145).  OpenVDB is not part of
this package: these helpers resolve the voxel size and the index box the same way and hand whole
boxes or lists of voxels to the GPU in one call (``FastWindingNumber.sign_voxel_box`` /
``sign_voxels``), where the voxel centres are generated.
"""
from __future__ import annotations

import numpy as np

DEFAULT_VOXEL_SIZE = -0.01  # MeshToVolumeOptions::voxel_size (mesh_to_volume.h:40)
NARROW_BAND = 3             # exterior / interior bandwidth in voxels (mesh_to_volume.cpp:164-165)


def bbox_diagonal(vertices) -> float:
    """Length of the bounding-box diagonal in the mesh's own scalar type (the reference computes
    ``mesh_bbox<3>(mesh).diagonal().norm()`` in ``Scalar``)."""
    V = np.asarray(vertices)
    if V.ndim != 2 or V.shape[1] != 3 or V.shape[0] == 0:
        raise ValueError("vertices must be a non-empty (N, 3) array")
    t = V.dtype.type if V.dtype in (np.float32, np.float64) else np.float64
    d = (V.max(axis=0) - V.min(axis=0)).astype(t)
    return float(np.sqrt(t(d[0] * d[0] + d[1] * d[1]) + t(d[2] * d[2])))


def resolve_voxel_size(vertices, voxel_size: float = DEFAULT_VOXEL_SIZE) -> float:
    """Absolute voxel size: a negative value is |voxel_size| x bbox diagonal (double)."""
    vs = float(voxel_size)
    if vs < 0:
        vs = abs(vs) * bbox_diagonal(vertices)
    if not (vs > 0 and np.isfinite(vs)):
        raise ValueError("voxel size must be non-zero and finite")
    return vs


def voxel_centres(ijk, voxel_size: float) -> np.ndarray:
    """World-space centres float(voxel_size * ijk + voxel_size / 2), multiply and add rounded
    separately in double, then rounded to float: the points the GPU generates."""
    ijk = np.asarray(ijk, dtype=np.int64).reshape(-1, 3)
    vs = float(voxel_size)
    return (ijk.astype(np.float64) * vs + vs * 0.5).astype(np.float32)


def index_box(vertices, voxel_size: float, band: int = NARROW_BAND):
    """(lo, dims) int32 of the voxel-index box whose centres cover the mesh's bounding box grown
    by ``band`` voxels on every side (a dense superset of OpenVDB's narrow band)."""
    V = np.asarray(vertices, dtype=np.float64)
    vs = float(voxel_size)
    lo = np.floor((V.min(axis=0) - vs * 0.5) / vs).astype(np.int64) - band
    hi = np.ceil((V.max(axis=0) - vs * 0.5) / vs).astype(np.int64) + band
    if (lo < np.iinfo(np.int32).min).any() or (hi > np.iinfo(np.int32).max).any():
        raise ValueError("voxel indices beyond int32: voxel size too small")
    return lo.astype(np.int32), (hi - lo + 1).astype(np.int32)


def winding_sign_volume(engine, vertices, voxel_size: float = DEFAULT_VOXEL_SIZE, band: int = NARROW_BAND,
                        device=None):
    """Inside flags of every voxel of the mesh's index box, as mesh_to_volume's winding-number
    interior test would give them.  Returns (lo, voxel_size, inside[z, y, x])."""
    vs = resolve_voxel_size(vertices, voxel_size)
    lo, dims = index_box(vertices, vs, band)
    return lo, vs, engine.sign_voxel_box(lo, dims, vs, device=device)
