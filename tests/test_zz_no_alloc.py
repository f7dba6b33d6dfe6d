"""fwn_solid_angle_noalloc, the query entry point that allocates nothing (no counterpart in the
reference; for callers that cannot allocate during a call), against fwn_solid_angle with device
pointers, through the C ABI: omega and flags bit for bit on two calls in a row, the second with
fewer points (the shared scratch array keeps the first call's contents), FP64 near leaves and a
partial last batch included; more than FWN_NOALLOC_MAX_POINTS points is refused before any work."""
import numpy as np
import pytest

from lagrange_amd import _lib as L
from lagrange_amd import workloads as W
from lagrange_amd.winding import FastWindingNumber

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 20


def _call(lib, e, fn, qd, n):
    import torch
    om = torch.empty(n, dtype=torch.float32, device=qd.device)
    ins = torch.empty(n, dtype=torch.uint8, device=qd.device)
    torch.cuda.synchronize()
    if fn == "noalloc":
        st = lib.fwn_solid_angle_noalloc(e._h(), qd.data_ptr(), n, om.data_ptr(), ins.data_ptr())
    else:
        st = lib.fwn_solid_angle(e._h(), qd.data_ptr(), n, om.data_ptr(), ins.data_ptr(), L.FWN_DEVICE_POINTERS)
    L.check(st)
    return om.cpu().numpy(), ins.cpu().numpy()


def test_noalloc_equals_device_call(open_mesh):
    import torch
    V, F = open_mesh
    e = FastWindingNumber((V, F))
    lib = L.load()
    lo, hi = V.min(0), V.max(0)
    M = W.Mesh(V.astype(np.float64), F.astype(np.int64))
    big = np.concatenate([W.uniform_queries(lo, hi, 250_000, 41, 1.3), W.near_surface_queries(M, 50_000, 42)])
    small = np.concatenate([W.near_surface_queries(M, 7_001, 43), W.uniform_queries(lo, hi, 3_000, 44, 1.3)])
    _, st = e.query_stats(big[::7])
    assert st[:, 3].sum() > 0  # FP64 near leaves on the path
    for q in (big, small):
        print(f"noalloc: {q.shape[0]} points", flush=True)
        qd = torch.from_numpy(np.ascontiguousarray(q, np.float32)).cuda()
        om_a, in_a = _call(lib, e, "device", qd, q.shape[0])
        om_b, in_b = _call(lib, e, "noalloc", qd, q.shape[0])
        np.testing.assert_array_equal(om_b.view(np.int32), om_a.view(np.int32))
        np.testing.assert_array_equal(in_b, in_a)
    # the limit is checked before any pointer is read or any kernel is launched
    qd = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    om = torch.empty(4, dtype=torch.float32, device="cuda")
    assert lib.fwn_solid_angle_noalloc(e._h(), qd.data_ptr(), MAX_POINTS + 1, om.data_ptr(), None) == L.FWN_LIMIT_EXCEEDED
    assert "FWN_NOALLOC_MAX_POINTS" in L.load().fwn_last_error().decode()
    assert lib.fwn_solid_angle_noalloc(e._h(), qd.data_ptr(), 0, om.data_ptr(), None) == L.FWN_OK
