"""Host-side mirror of ``lagrange::winding::FastWindingNumber`` over the GPU engine.

Reference API (modules/winding/include/lagrange/winding/FastWindingNumber.h:24-108):
  * constructed from a triangle mesh; a non-3D or non-triangle mesh raises
    (FastWindingNumber.cpp:91-96, la_runtime_assert -> lagrange::Error);
  * positions are converted to float and indices to int (FastWindingNumber.cpp:39-52);
  * ``solid_angle(pos)`` returns the solid angle in steradians (FastWindingNumber.cpp:69-76);
  * ``is_inside(pos)`` is ``solid_angle / (4 pi) > 0.5`` evaluated in double
    (FastWindingNumber.cpp:66);
  * the object is move-only (copy constructor and assignment deleted, FastWindingNumber.h:73,82).
Additive here: batched calls over arrays (host numpy or device torch tensors) and the
multi-mesh engine ``MultiFastWindingNumber``.  Every call runs on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib as L


class LagrangeError(RuntimeError):
    """Counterpart of lagrange::Error (modules/core/include/lagrange/utils/Error.h:26)."""


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _torch_ready(device) -> None:
    """Wait for torch's current stream on `device`: the engine runs on its own stream, and the
    input tensors (and the memory of freshly allocated outputs) belong to work queued there."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def _mesh_arrays(mesh):
    if isinstance(mesh, tuple) and len(mesh) == 2:
        V, F = mesh
    else:
        V, F = mesh.vertices, mesh.facets
    V = np.asarray(V)
    F = np.asarray(F)
    if V.ndim != 2 or V.shape[1] != 3:
        raise LagrangeError("Fast winding number engine only supports 3D meshes")
    if F.ndim != 2 or F.shape[1] != 3:
        raise LagrangeError("Fast winding number engine only supports triangle meshes")
    # the reference's internal copy: float positions, int indices (FastWindingNumber.cpp:39-52)
    return (np.ascontiguousarray(V, dtype=np.float32),
            np.ascontiguousarray(F.astype(np.int64, copy=False).astype(np.int32), dtype=np.int32))


def _ptr(a) -> int:
    return a.ctypes.data if isinstance(a, np.ndarray) else int(a.data_ptr())


def _points_rows(points) -> int:
    if _is_torch(points):
        return int(points.reshape(-1, 3).shape[0])
    return int(np.asarray(points).reshape(-1, 3).shape[0])


def _check_q_off(q_off, n_meshes: int, n_points: int) -> np.ndarray:
    """Per-mesh point offsets as the C ABI reads them: n_meshes + 1 int64 entries, starting at 0,
    non-decreasing, ending at the number of points.  The library takes the point count from the
    last entry, so a wrong array would make it read past the caller's buffers: reject it here."""
    qo = np.ascontiguousarray(np.asarray(q_off).reshape(-1), dtype=np.int64)
    if qo.shape[0] != n_meshes + 1:
        raise LagrangeError(f"q_off has {qo.shape[0]} entries; {n_meshes + 1} (meshes + 1) are required")
    if qo[0] != 0:
        raise LagrangeError("q_off[0] must be 0")
    if np.any(np.diff(qo) < 0):
        raise LagrangeError("q_off must be non-decreasing")
    if qo[-1] != n_points:
        raise LagrangeError(f"q_off[-1] = {int(qo[-1])} differs from the number of points ({n_points})")
    return qo


def _create_tuned(n_meshes, V, v_off, F, f_off, points, q_off, leaf_sizes, options, opts):
    """Shared body of the tuned constructors: (handle, chosen leaf size, query ms)."""
    if q_off is None and n_meshes != 1:
        raise LagrangeError("per-mesh point offsets are required for a multi-mesh engine")
    qo = None if q_off is None else _check_q_off(q_off, n_meshes, _points_rows(points))
    o = options if options is not None else L.default_options(**opts)
    leaves = np.ascontiguousarray(np.asarray(leaf_sizes, dtype=np.int32).reshape(-1))
    if _is_torch(points) and points.is_cuda:
        import torch
        q = points.contiguous().to(torch.float32)
        _torch_ready(q.device)
        n, ptr, fl = q.shape[0], _ptr(q), L.FWN_DEVICE_POINTERS
    else:
        q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        n, ptr, fl = q.shape[0], q.ctypes.data, 0
    h, leaf, ms = C.c_void_p(), C.c_int32(), C.c_double()
    L.check(L.load().fwn_create_tuned(C.byref(h), n_meshes, V.ctypes.data, v_off.ctypes.data, F.ctypes.data,
                                      f_off.ctypes.data, C.byref(o), ptr, qo.ctypes.data if qo is not None else None,
                                      n, fl, leaves.ctypes.data, leaves.size, C.byref(leaf), C.byref(ms)))
    return h, leaf.value, ms.value


class _EngineBase:
    _handle = None
    _dependents = None

    def _add_dependent(self, obj):
        """Objects built on this engine (surrogates) are closed before the engine itself."""
        if self._dependents is None:
            self._dependents = weakref.WeakSet()
        self._dependents.add(obj)

    def __copy__(self):
        raise TypeError("FastWindingNumber is move-only (copy is deleted in the reference API)")

    __deepcopy__ = lambda self, memo: self.__copy__()  # noqa: E731

    def close(self):
        if self._dependents is not None:
            for d in list(self._dependents):
                d.close()
        if self._handle:
            L.load().fwn_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _h(self):
        if not self._handle:
            raise LagrangeError("engine is not initialised (default-constructed or closed)")
        return self._handle

    # ---- introspection ----------------------------------------------------------------
    def tree_info(self) -> dict:
        info = L.TreeInfo()
        L.check(L.load().fwn_tree_info_get(self._h(), C.byref(info)))
        return {k: getattr(info, k) for k, _ in L.TreeInfo._fields_}

    def export_tree(self):
        """(slots u32 [nodes*width*32], tris u32 [F*12], roots int32 [M,2], tau float32 [M])."""
        info = self.tree_info()
        slots = np.zeros(max(info["num_nodes"], 1) * info["width"] * L.SLOT_WORDS, np.uint32)
        tris = np.zeros(max(info["num_triangles"], 1) * L.TRI_WORDS, np.uint32)
        roots = np.zeros((info["num_meshes"], 2), np.int32)
        tau = np.zeros(info["num_meshes"], np.float32)
        L.check(L.load().fwn_export_tree(self._h(), slots.ctypes.data, tris.ctypes.data, roots.ctypes.data,
                                         tau.ctypes.data))
        return slots, tris, roots, tau

    def set_launch(self, waves_per_block: int, mode: int = L.FWN_LAUNCH_GRID):
        L.check(L.load().fwn_set_launch(self._h(), waves_per_block, mode))

    def autotune(self, points, q_off=None, cache_path: str | None = None):
        """Time every launch shape on these points and keep the fastest: (waves, mode, ms)."""
        w, m, ms = C.c_int32(), C.c_int32(), C.c_double()
        n_meshes = getattr(self, "num_meshes", 1)
        if q_off is None and n_meshes != 1:
            raise LagrangeError("per-mesh point offsets are required for a multi-mesh engine")
        qo = None if q_off is None else _check_q_off(q_off, n_meshes, _points_rows(points))
        if _is_torch(points) and points.is_cuda:
            import torch
            q = points.contiguous().to(torch.float32)
            _torch_ready(q.device)
            n, ptr, fl = q.shape[0], _ptr(q), L.FWN_DEVICE_POINTERS
        else:
            q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
            n, ptr, fl = q.shape[0], q.ctypes.data, 0
        L.check(L.load().fwn_autotune(self._h(), ptr, qo.ctypes.data if qo is not None else None, n, fl,
                                      cache_path.encode() if cache_path else None, C.byref(w), C.byref(m),
                                      C.byref(ms)))
        return w.value, m.value, ms.value

    def last_timing(self):
        tr, tot = C.c_double(), C.c_double()
        L.check(L.load().fwn_last_timing(self._h(), C.byref(tr), C.byref(tot)))
        return tr.value, tot.value


class FastWindingNumber(_EngineBase):
    """Fast winding number engine for one triangle mesh (reference-compatible API)."""

    def __init__(self, mesh=None, *, options: L.Options | None = None, **opts):
        self._handle = None
        if mesh is None:  # default constructor: an empty engine
            return
        V, F = _mesh_arrays(mesh)
        o = options if options is not None else L.default_options(**opts)
        h = C.c_void_p()
        L.check(L.load().fwn_create(C.byref(h), V.ctypes.data, V.shape[0], F.ctypes.data, F.shape[0], C.byref(o)))
        self._handle = h
        self.num_vertices = V.shape[0]
        self.num_facets = F.shape[0]
        # float bounding box of the mesh (mesh_bbox<3> of a float mesh), for sampling
        self.bbox = ((V.min(axis=0), V.max(axis=0)) if V.shape[0] else
                     (np.zeros(3, np.float32), np.zeros(3, np.float32)))

    @classmethod
    def tuned(cls, mesh, points, leaf_sizes=(1, 2, 4, 8), *, options: L.Options | None = None, **opts):
        """Engine built with the leaf size (and launch shape) that answers ``points`` fastest
        (fwn_create_tuned); ``engine.tuning`` holds the choice and its query time."""
        V, F = _mesh_arrays(mesh)
        v_off = np.array([0, V.shape[0]], np.int64)
        f_off = np.array([0, F.shape[0]], np.int64)
        h, leaf, ms = _create_tuned(1, V, v_off, F, f_off, points, None, leaf_sizes, options, opts)
        self = cls.__new__(cls)
        self._handle = h
        self.num_vertices = V.shape[0]
        self.num_facets = F.shape[0]
        self.bbox = ((V.min(axis=0), V.max(axis=0)) if V.shape[0] else
                     (np.zeros(3, np.float32), np.zeros(3, np.float32)))
        self.tuning = {"leaf_max": leaf, "query_ms": ms}
        return self

    # ---- per-point API (as the reference) ------------------------------------------------
    def solid_angle(self, pos) -> float:
        q = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).reshape(3))
        out = C.c_float()
        L.check(L.load().fwn_solid_angle_point(self._h(), q.ctypes.data, C.byref(out)))
        return float(out.value)

    def is_inside(self, pos) -> bool:
        q = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).reshape(3))
        out = C.c_uint8()
        L.check(L.load().fwn_is_inside_point(self._h(), q.ctypes.data, C.byref(out)))
        return bool(out.value)

    # ---- batched API -----------------------------------------------------------------------
    def solid_angle_batch(self, points, with_inside: bool = False):
        """Solid angles of an (N, 3) array; numpy in -> numpy out, torch device tensor in ->
        torch device tensors out (no host round trip)."""
        if _is_torch(points) and points.is_cuda:
            import torch
            q = points.contiguous().to(torch.float32)
            n = q.shape[0]
            om = torch.empty(n, dtype=torch.float32, device=q.device)
            # a torch.bool element is one byte holding 0 or 1: the engine's flag bytes, written in place
            ins = torch.empty(n, dtype=torch.bool, device=q.device) if with_inside else None
            _torch_ready(q.device)
            L.check(L.load().fwn_solid_angle(self._h(), _ptr(q), n, _ptr(om), _ptr(ins) if ins is not None else None,
                                             L.FWN_DEVICE_POINTERS))
            return (om, ins) if with_inside else om
        q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        n = q.shape[0]
        om = np.empty(n, np.float32)
        ins = np.empty(n, np.uint8) if with_inside else None
        L.check(L.load().fwn_solid_angle(self._h(), q.ctypes.data, n, om.ctypes.data,
                                         ins.ctypes.data if ins is not None else None, 0))
        return (om, ins.astype(bool)) if with_inside else om

    def is_inside_batch(self, points):
        return self.solid_angle_batch(points, with_inside=True)[1]

    # ---- voxel signs (mesh_to_volume's interior test, batched) -----------------------------
    def sign_voxel_box(self, lo, dims, voxel_size: float, with_omega: bool = False, device=None):
        """is_inside of every voxel centre float(voxel_size * ijk + voxel_size / 2) in the index box
        lo + [0, dims) (x fastest), as volume::mesh_to_volume's winding-number interior test
        (modules/volume/src/mesh_to_volume.cpp:175-183).  Centres are generated on the GPU.
        Returns an array of shape (dims[2], dims[1], dims[0]) (and omega of the same shape);
        ``device`` = a torch device keeps the results in HBM."""
        lo_a = np.ascontiguousarray(np.asarray(lo, dtype=np.int32).reshape(3))
        dims_a = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(3))
        if (dims_a < 0).any():
            raise LagrangeError("negative box dimension")
        n = int(np.prod(dims_a.astype(np.int64)))
        shape = (int(dims_a[2]), int(dims_a[1]), int(dims_a[0]))
        if device is not None:
            import torch
            ins = torch.empty(n, dtype=torch.bool, device=device)
            om = torch.empty(n, dtype=torch.float32, device=device) if with_omega else None
            _torch_ready(device)
            L.check(L.load().fwn_sign_voxel_box(self._h(), lo_a.ctypes.data, dims_a.ctypes.data, float(voxel_size),
                                                _ptr(ins), _ptr(om) if om is not None else None,
                                                L.FWN_DEVICE_POINTERS))
            ins = ins.reshape(shape)
            return (ins, om.reshape(shape)) if with_omega else ins
        ins = np.empty(n, np.uint8)
        om = np.empty(n, np.float32) if with_omega else None
        L.check(L.load().fwn_sign_voxel_box(self._h(), lo_a.ctypes.data, dims_a.ctypes.data, float(voxel_size),
                                            ins.ctypes.data, om.ctypes.data if om is not None else None, 0))
        ins = ins.astype(bool).reshape(shape)
        return (ins, om.reshape(shape)) if with_omega else ins

    def sign_voxels(self, ijk, voxel_size: float, with_omega: bool = False):
        """is_inside of the centres of an (N, 3) int32 list of voxel indices (see sign_voxel_box);
        numpy in -> numpy out, torch device tensor in -> device tensors out."""
        if _is_torch(ijk) and ijk.is_cuda:
            import torch
            v = ijk.contiguous().to(torch.int32)
            n = v.shape[0]
            ins = torch.empty(n, dtype=torch.bool, device=v.device)
            om = torch.empty(n, dtype=torch.float32, device=v.device) if with_omega else None
            _torch_ready(v.device)
            L.check(L.load().fwn_sign_voxels(self._h(), _ptr(v), n, float(voxel_size), _ptr(ins),
                                             _ptr(om) if om is not None else None, L.FWN_DEVICE_POINTERS))
            return (ins, om) if with_omega else ins
        v = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
        n = v.shape[0]
        ins = np.empty(n, np.uint8)
        om = np.empty(n, np.float32) if with_omega else None
        L.check(L.load().fwn_sign_voxels(self._h(), v.ctypes.data, n, float(voxel_size), ins.ctypes.data,
                                         om.ctypes.data if om is not None else None, 0))
        return (ins.astype(bool), om) if with_omega else ins.astype(bool)

    # ---- the reference examples, batched ---------------------------------------------------
    def orientation_criterion(self, vertices, facets, epsilon: float = 1e-2):
        """Per-facet criterion of modules/winding/examples/fix_orientation.cpp:86-115 for an input
        mesh against this engine's mesh: (omega(g + eps n) - omega(g - eps n)) / (4 pi), float64,
        probes made on the GPU.  Device tensors in -> device tensor out."""
        if _is_torch(vertices) and vertices.is_cuda:
            import torch
            V = vertices.contiguous().to(torch.float32)
            F = facets.contiguous().to(torch.int32)
            out = torch.empty(F.shape[0], dtype=torch.float64, device=V.device)
            _torch_ready(V.device)
            L.check(L.load().fwn_orientation_criterion(self._h(), _ptr(V), V.shape[0], _ptr(F), F.shape[0],
                                                       float(epsilon), _ptr(out), L.FWN_DEVICE_POINTERS))
            return out
        V, F = _mesh_arrays((vertices, facets))
        out = np.empty(F.shape[0], np.float64)
        L.check(L.load().fwn_orientation_criterion(self._h(), V.ctypes.data, V.shape[0], F.ctypes.data, F.shape[0],
                                                   float(epsilon), out.ctypes.data, 0))
        return out

    def sample_inside(self, num_samples: int, seed: int = 0, lo=None, hi=None, device=None):
        """Samples uniform in the box [lo, hi] (default: the mesh's bounding box) kept when inside,
        in sample order (modules/winding/examples/sample_points_in_mesh.cpp:59-76); sample i uses
        the counter-based generator of (seed, i).  Returns an (M, 3) float32 array (or a device
        tensor with ``device``)."""
        lo = np.ascontiguousarray(np.asarray(self.bbox[0] if lo is None else lo, np.float32).reshape(3))
        hi = np.ascontiguousarray(np.asarray(self.bbox[1] if hi is None else hi, np.float32).reshape(3))
        n = int(num_samples)
        cnt = C.c_int64()
        if device is not None:
            import torch
            out = torch.empty((max(n, 1), 3), dtype=torch.float32, device=device)
            _torch_ready(device)
            L.check(L.load().fwn_sample_inside(self._h(), lo.ctypes.data, hi.ctypes.data, n, int(seed), _ptr(out),
                                               C.byref(cnt), L.FWN_DEVICE_POINTERS))
            return out[:cnt.value]
        out = np.empty((max(n, 1), 3), np.float32)
        L.check(L.load().fwn_sample_inside(self._h(), lo.ctypes.data, hi.ctypes.data, n, int(seed), out.ctypes.data,
                                           C.byref(cnt), 0))
        return out[:cnt.value].copy()

    # ---- closest point (bvh::TriangleAABBTree::get_closest_point, batched) ----------------
    def closest_point(self, points, with_point: bool = True):
        """(squared distance float64, facet int32, closest point (N, 3) float64) per point, over
        the engine's tree (modules/bvh/src/TriangleAABBTree.cpp:86-127); an empty mesh gives
        facet -1 and +inf.  numpy in -> numpy out, torch device tensor in -> device tensors out."""
        if _is_torch(points) and points.is_cuda:
            import torch
            q = points.contiguous().to(torch.float32)
            n = q.shape[0]
            d2 = torch.empty(n, dtype=torch.float64, device=q.device)
            fi = torch.empty(n, dtype=torch.int32, device=q.device)
            cp = torch.empty((n, 3), dtype=torch.float64, device=q.device) if with_point else None
            _torch_ready(q.device)
            L.check(L.load().fwn_closest_point(self._h(), _ptr(q), n, _ptr(d2), _ptr(fi),
                                               _ptr(cp) if cp is not None else None, L.FWN_DEVICE_POINTERS))
            return (d2, fi, cp) if with_point else (d2, fi)
        q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        n = q.shape[0]
        d2 = np.empty(n, np.float64)
        fi = np.empty(n, np.int32)
        cp = np.empty((n, 3), np.float64) if with_point else None
        L.check(L.load().fwn_closest_point(self._h(), q.ctypes.data, n, d2.ctypes.data, fi.ctypes.data,
                                           cp.ctypes.data if cp is not None else None, 0))
        return (d2, fi, cp) if with_point else (d2, fi)

    def query_stats(self, points):
        """(omega, stats[N,4]) with per-point work counters: tests, far terms, FP32 and FP64 triangles."""
        q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        om = np.empty(q.shape[0], np.float32)
        st = np.empty((q.shape[0], 4), np.int32)
        L.check(L.load().fwn_query_stats(self._h(), q.ctypes.data, None, q.shape[0], om.ctypes.data,
                                         st.ctypes.data, 0))
        return om, st

    def bruteforce_f64(self, points):
        """FP64 sum of exact triangle solid angles, computed on the GPU (verification hook)."""
        if _is_torch(points) and points.is_cuda:
            import torch
            q = points.contiguous().to(torch.float32)
            out = torch.empty(q.shape[0], dtype=torch.float64, device=q.device)
            _torch_ready(q.device)
            L.check(L.load().fwn_bruteforce_f64(self._h(), _ptr(q), q.shape[0], _ptr(out), L.FWN_DEVICE_POINTERS))
            return out
        q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        out = np.empty(q.shape[0], np.float64)
        L.check(L.load().fwn_bruteforce_f64(self._h(), q.ctypes.data, q.shape[0], out.ctypes.data, 0))
        return out


class MultiFastWindingNumber(_EngineBase):
    """Many meshes in one engine: one build and one query launch for all of them."""

    def __init__(self, meshes, *, options: L.Options | None = None, **opts):
        self._handle = None
        arrs = [_mesh_arrays(m) for m in meshes]
        if not arrs:
            raise LagrangeError("at least one mesh is required")
        V = np.ascontiguousarray(np.concatenate([a[0] for a in arrs]), dtype=np.float32)
        F = np.ascontiguousarray(np.concatenate([a[1] for a in arrs]), dtype=np.int32)
        self.v_off = np.zeros(len(arrs) + 1, np.int64)
        self.f_off = np.zeros(len(arrs) + 1, np.int64)
        self.v_off[1:] = np.cumsum([a[0].shape[0] for a in arrs])
        self.f_off[1:] = np.cumsum([a[1].shape[0] for a in arrs])
        o = options if options is not None else L.default_options(**opts)
        h = C.c_void_p()
        L.check(L.load().fwn_create_multi(C.byref(h), len(arrs), V.ctypes.data, self.v_off.ctypes.data,
                                          F.ctypes.data, self.f_off.ctypes.data, C.byref(o)))
        self._handle = h
        self.num_meshes = len(arrs)

    @classmethod
    def tuned(cls, meshes, points, q_off, leaf_sizes=(1, 2, 4, 8), *, options: L.Options | None = None, **opts):
        """Engine built with the leaf size (and launch shape) that answers the concatenated
        ``points`` (offsets ``q_off``) fastest; ``engine.tuning`` holds the choice."""
        arrs = [_mesh_arrays(m) for m in meshes]
        if not arrs:
            raise LagrangeError("at least one mesh is required")
        V = np.ascontiguousarray(np.concatenate([a[0] for a in arrs]), dtype=np.float32)
        F = np.ascontiguousarray(np.concatenate([a[1] for a in arrs]), dtype=np.int32)
        v_off = np.zeros(len(arrs) + 1, np.int64)
        f_off = np.zeros(len(arrs) + 1, np.int64)
        v_off[1:] = np.cumsum([a[0].shape[0] for a in arrs])
        f_off[1:] = np.cumsum([a[1].shape[0] for a in arrs])
        h, leaf, ms = _create_tuned(len(arrs), V, v_off, F, f_off, points, q_off, leaf_sizes, options, opts)
        self = cls.__new__(cls)
        self._handle = h
        self.v_off, self.f_off = v_off, f_off
        self.num_meshes = len(arrs)
        self.tuning = {"leaf_max": leaf, "query_ms": ms}
        return self

    def solid_angle_batch(self, points_per_mesh=None, *, points=None, q_off=None, with_inside=False):
        """Either a list of (Ni, 3) arrays (one per mesh) or one concatenated array + offsets."""
        if points is None:
            q_off = np.zeros(self.num_meshes + 1, np.int64)
            q_off[1:] = np.cumsum([np.asarray(p).reshape(-1, 3).shape[0] for p in points_per_mesh])
            if _is_torch(points_per_mesh[0]):
                import torch
                points = torch.cat([p.reshape(-1, 3) for p in points_per_mesh])
            else:
                points = np.concatenate([np.asarray(p, np.float32).reshape(-1, 3) for p in points_per_mesh])
        elif q_off is None:
            raise LagrangeError("q_off is required with concatenated points")
        q_off = _check_q_off(q_off, self.num_meshes, _points_rows(points))
        if _is_torch(points) and points.is_cuda:
            import torch
            q = points.contiguous().to(torch.float32)
            n = q.shape[0]
            om = torch.empty(n, dtype=torch.float32, device=q.device)
            # a torch.bool element is one byte holding 0 or 1: the engine's flag bytes, written in place
            ins = torch.empty(n, dtype=torch.bool, device=q.device) if with_inside else None
            _torch_ready(q.device)
            L.check(L.load().fwn_solid_angle_multi(self._h(), _ptr(q), q_off.ctypes.data, _ptr(om),
                                                   _ptr(ins) if ins is not None else None, L.FWN_DEVICE_POINTERS))
            return (om, ins) if with_inside else om
        q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        om = np.empty(q.shape[0], np.float32)
        ins = np.empty(q.shape[0], np.uint8) if with_inside else None
        L.check(L.load().fwn_solid_angle_multi(self._h(), q.ctypes.data, q_off.ctypes.data, om.ctypes.data,
                                               ins.ctypes.data if ins is not None else None, 0))
        return (om, ins.astype(bool)) if with_inside else om

    def query_stats(self, points, q_off):
        q = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        q_off = _check_q_off(q_off, self.num_meshes, q.shape[0])
        om = np.empty(q.shape[0], np.float32)
        st = np.empty((q.shape[0], 4), np.int32)
        L.check(L.load().fwn_query_stats(self._h(), q.ctypes.data, q_off.ctypes.data, q.shape[0], om.ctypes.data,
                                         st.ctypes.data, 0))
        return om, st
