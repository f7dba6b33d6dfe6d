"""Learned surrogate over an exact engine (C ABI fwn_surrogate_*; no reference counterpart).

A 16-level hash-grid encoding (2^19 entries per level, 2 FP16 features) feeds a 4 x 128 ReLU MLP
(BF16 MFMA, FP32 accumulation), trained on the GPU on winding numbers the exact engine generates.
``query`` answers w = omega / (4 pi) and is_inside, falling back to the exact engine near the
surface and where the prediction is within ``band`` of 0.5.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .winding import FastWindingNumber, _is_torch, _ptr, _torch_ready


class SurrogateOptions(C.Structure):
    _fields_ = [
        ("levels", C.c_int32), ("log2_table", C.c_int32), ("base_res", C.c_int32), ("max_res", C.c_int32),
        ("hidden_layers", C.c_int32), ("batch", C.c_int32), ("steps", C.c_int32), ("grid_res", C.c_int32),
        ("lr", C.c_float), ("band", C.c_float), ("dist_rel", C.c_float), ("near_frac", C.c_float),
        ("n_samples", C.c_int64), ("seed", C.c_uint64),
    ]


class SurrogateInfo(C.Structure):
    _fields_ = [
        ("levels", C.c_int32), ("res", C.c_int32 * 16), ("dense", C.c_int32 * 16),
        ("table_floats", C.c_int64), ("mlp_floats", C.c_int64), ("delta", C.c_float),
        ("last_loss", C.c_float), ("train_ms", C.c_double), ("encode_ms", C.c_double), ("mlp_ms", C.c_double),
        ("fallback_ms", C.c_double), ("query_ms", C.c_double),
    ]


def _bind():
    lib = L.load()
    if getattr(lib, "_sg_bound", False):
        return lib
    VP, I64, I32 = C.c_void_p, C.c_int64, C.c_int32
    sig = {
        "fwn_surrogate_options_default": ([C.POINTER(SurrogateOptions)], I32),
        "fwn_surrogate_create": ([C.POINTER(VP), VP, C.POINTER(SurrogateOptions)], I32),
        "fwn_surrogate_destroy": ([VP], I32),
        "fwn_surrogate_train": ([VP], I32),
        "fwn_surrogate_info_get": ([VP, C.POINTER(SurrogateInfo)], I32),
        "fwn_surrogate_query": ([VP, VP, I64, VP, VP, I32, C.POINTER(I64)], I32),
        "fwn_surrogate_predict": ([VP, VP, I64, VP, VP, I32], I32),
        "fwn_surrogate_params": ([VP, VP, VP, VP, VP], I32),
        "fwn_surrogate_gradients": ([VP, VP, VP, I64, C.POINTER(C.c_float), VP, VP, VP], I32),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    lib._sg_bound = True
    return lib


def default_options(**kw) -> SurrogateOptions:
    o = SurrogateOptions()
    L.check(_bind().fwn_surrogate_options_default(C.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Surrogate:
    def __init__(self, engine: FastWindingNumber, options: SurrogateOptions | None = None, **kw):
        lib = _bind()
        self.engine = engine  # keeps the exact engine alive
        o = options if options is not None else default_options(**kw)
        h = C.c_void_p()
        L.check(lib.fwn_surrogate_create(C.byref(h), engine._h(), C.byref(o)))
        self._h = h
        self.options = o
        engine._add_dependent(self)

    _h = None

    def close(self):
        if self._h:
            _bind().fwn_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self):
        L.check(_bind().fwn_surrogate_train(self._h))
        return self.info()

    def info(self) -> dict:
        i = SurrogateInfo()
        L.check(_bind().fwn_surrogate_info_get(self._h, C.byref(i)))
        return {"levels": i.levels, "res": list(i.res)[: i.levels], "dense": list(i.dense)[: i.levels],
                "table_floats": i.table_floats, "mlp_floats": i.mlp_floats, "delta": i.delta,
                "last_loss": i.last_loss, "train_ms": i.train_ms, "encode_ms": i.encode_ms, "mlp_ms": i.mlp_ms,
                "fallback_ms": i.fallback_ms, "query_ms": i.query_ms}

    def query(self, points):
        """(w, inside, n_fallback): numpy in -> numpy out, torch device tensor -> device tensors."""
        lib = _bind()
        nf = C.c_int64()
        if _is_torch(points) and points.is_cuda:
            import torch
            q = points.contiguous().to(torch.float32)
            w = torch.empty(q.shape[0], dtype=torch.float32, device=q.device)
            ins = torch.empty(q.shape[0], dtype=torch.bool, device=q.device)
            _torch_ready(q.device)
            L.check(lib.fwn_surrogate_query(self._h, _ptr(q), q.shape[0], _ptr(w), _ptr(ins), L.FWN_DEVICE_POINTERS,
                                            C.byref(nf)))
            return w, ins, nf.value
        q = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        w = np.empty(q.shape[0], np.float32)
        ins = np.empty(q.shape[0], np.uint8)
        L.check(lib.fwn_surrogate_query(self._h, q.ctypes.data, q.shape[0], w.ctypes.data, ins.ctypes.data, 0,
                                        C.byref(nf)))
        return w, ins.astype(bool), nf.value

    def predict(self, points):
        """(network prediction, fallback flag) per point: numpy in -> numpy out, torch device
        tensor in -> device tensors out."""
        if _is_torch(points) and points.is_cuda:
            import torch
            q = points.contiguous().to(torch.float32)
            p = torch.empty(q.shape[0], dtype=torch.float32, device=q.device)
            fb = torch.empty(q.shape[0], dtype=torch.bool, device=q.device)
            _torch_ready(q.device)
            L.check(_bind().fwn_surrogate_predict(self._h, _ptr(q), q.shape[0], _ptr(p), _ptr(fb),
                                                  L.FWN_DEVICE_POINTERS))
            return p, fb
        q = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        p = np.empty(q.shape[0], np.float32)
        fb = np.empty(q.shape[0], np.uint8)
        L.check(_bind().fwn_surrogate_predict(self._h, q.ctypes.data, q.shape[0], p.ctypes.data, fb.ctypes.data, 0))
        return p, fb.astype(bool)

    def params(self):
        i = self.info()
        t = np.empty(i["table_floats"], np.float32)
        m = np.empty(i["mlp_floats"], np.float32)
        L.check(_bind().fwn_surrogate_params(self._h, t.ctypes.data, m.ctypes.data, None, None))
        return t, m

    def set_params(self, table=None, mlp=None):
        t = None if table is None else np.ascontiguousarray(table, np.float32)
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        L.check(_bind().fwn_surrogate_params(self._h, None, None, t.ctypes.data if t is not None else None,
                                             m.ctypes.data if m is not None else None))

    def gradients(self, points, targets):
        """One forward + backward (no update): loss, predictions, table and MLP gradients."""
        i = self.info()
        q = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        y = np.ascontiguousarray(np.asarray(targets, np.float32))
        loss = C.c_float()
        pred = np.empty(q.shape[0], np.float32)
        gt = np.empty(i["table_floats"], np.float32)
        gm = np.empty(i["mlp_floats"], np.float32)
        L.check(_bind().fwn_surrogate_gradients(self._h, q.ctypes.data, y.ctypes.data, q.shape[0], C.byref(loss),
                                                pred.ctypes.data, gt.ctypes.data, gm.ctypes.data))
        return loss.value, pred, gt, gm
