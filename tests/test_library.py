"""CPU-side checks of the shared library, the host API and the workload generators (no GPU)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from lagrange_amd import _lib as L
from lagrange_amd import workloads as W


def test_library_exports_every_declared_function():
    lib = L.load()
    names = L.declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_io_library_exports_every_declared_function():
    """include/fwn_io.h is served by libfwn_io.so (host code)."""
    import re
    from lagrange_amd import io as mio
    txt = open(os.path.join(os.path.dirname(L.HEADER), "fwn_io.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(fwn_[a-z0-9_]+)\s*\(", txt)))
    assert len(names) == 7
    mio.load_lib()
    out = subprocess.run(["nm", "-D", "--defined-only", mio.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_default_options_match_reference_defaults():
    o = L.default_options()
    assert o.accuracy_scale == 2.0  # the reference passes no accuracy argument (default 2)
    assert o.near_fp64_rel == pytest.approx(1e-4)
    assert o.tree_width == 4  # the reference's 4-ary tree
    assert o.waves_per_block == 1  # measured fastest; same criterion and accuracy
    # the reference's box-area 4-ary tree layout with single-triangle items
    assert o.tree_builder == L.FWN_TREE_SAH and o.leaf_max == 1


def test_errors_without_device_or_with_bad_arguments():
    lib = L.load()
    h = C.c_void_p()
    o = L.default_options()
    st = lib.fwn_create(C.byref(h), None, 0, None, 0, None)
    # without a GPU the call reports NO_DEVICE; with one an empty mesh is valid
    assert st in (L.FWN_OK, L.FWN_NO_DEVICE)
    if st == L.FWN_OK:
        lib.fwn_destroy(h)
    else:
        assert b"no HIP device" in lib.fwn_last_error()
    assert lib.fwn_create(None, None, 0, None, 0, C.byref(o)) != L.FWN_OK
    assert lib.fwn_solid_angle(None, None, 0, None, None, 0) == L.FWN_INVALID_ARGUMENT
    assert b"engine is NULL" in lib.fwn_last_error()
    assert lib.fwn_options_default(None) == L.FWN_INVALID_ARGUMENT


def test_python_api_validates_meshes_before_any_device_call():
    from lagrange_amd.winding import FastWindingNumber, LagrangeError
    with pytest.raises(LagrangeError, match="3D"):
        FastWindingNumber((np.zeros((4, 2)), np.array([[0, 1, 2]])))
    with pytest.raises(LagrangeError, match="triangle"):
        FastWindingNumber((np.zeros((4, 3)), np.array([[0, 1, 2, 3]])))
    e = FastWindingNumber()
    with pytest.raises(LagrangeError):
        e.is_inside([0, 0, 0])
    import copy
    with pytest.raises(TypeError):
        copy.copy(e)


def test_workload_sizes_and_determinism():
    m = W.icosphere(32)
    assert (m.num_vertices, m.num_facets) == (10242, 20480)
    a = W.large_displaced_mesh(nu=32)
    b = W.large_displaced_mesh(nu=32)
    assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.facets, b.facets)
    assert a.num_facets == 20480 - round(0.05 * 20480)
    q1 = W.uniform_queries(np.zeros(3), np.ones(3), 1000, 3)
    assert np.array_equal(q1, W.uniform_queries(np.zeros(3), np.ones(3), 1000, 3))
    assert q1.min() >= 0 and q1.max() <= 1
    g = W.grid_queries(np.zeros(3), np.ones(3), 4, 1.0)
    assert g.shape == (64, 3) and np.array_equal(g[1] - g[0], np.array([1 / 3, 0, 0], np.float32))


def test_icosphere_closed_and_outward():
    m = W.icosphere(8)
    F = m.facets
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    s = set(map(tuple, e))
    assert len(s) == len(e) and all((b, a) in s for a, b in e)
    V = m.vertices
    vol = np.einsum("ij,ij->i", V[F[:, 0]], np.cross(V[F[:, 1]], V[F[:, 2]])).sum() / 6
    assert vol > 0


def test_near_surface_offsets_in_range():
    m = W.displaced_sphere(16, 3)
    q = W.near_surface_queries(m, 2000, 4, -7, -2).astype(np.float64)
    lo, hi = m.bbox()
    diag = np.linalg.norm(hi - lo)
    r = np.linalg.norm(q, axis=1)
    assert np.all(np.abs(r - 1) < 0.2 + 1e-2 * diag)


def test_bench_refuses_multi_gpu():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_q_off_is_checked_before_any_library_call(monkeypatch):
    """Every wrapper that takes per-mesh point offsets validates them on the host: the library
    reads the point count from q_off[n_meshes], so a short or inconsistent array would make it
    read past the caller's buffers."""
    from lagrange_amd import winding as Wd

    def no_library():
        raise AssertionError("the library was called before the offsets were checked")

    monkeypatch.setattr(L, "load", no_library)
    e = Wd.MultiFastWindingNumber.__new__(Wd.MultiFastWindingNumber)
    e._handle = C.c_void_p(1)  # never dereferenced: every call below must fail first
    e.num_meshes = 3
    pts = np.zeros((10, 3), np.float32)
    bad = {
        "short": np.array([0, 4, 10]),
        "long": np.array([0, 4, 6, 8, 10]),
        "nonzero start": np.array([1, 4, 6, 10]),
        "decreasing": np.array([0, 6, 4, 10]),
        "too large": np.array([0, 4, 6, 11]),
        "too small": np.array([0, 4, 6, 9]),
    }
    for name, qo in bad.items():
        with pytest.raises(Wd.LagrangeError):
            e.solid_angle_batch(points=pts, q_off=qo)
        with pytest.raises(Wd.LagrangeError):
            e.query_stats(pts, qo)
        with pytest.raises(Wd.LagrangeError):
            e.autotune(pts, q_off=qo)
        with pytest.raises(Wd.LagrangeError):
            Wd._create_tuned(3, None, None, None, None, pts, qo, (1,), None, {})
    with pytest.raises(Wd.LagrangeError):
        e.solid_angle_batch(points=pts)  # concatenated points without offsets
    with pytest.raises(Wd.LagrangeError):
        e.autotune(pts)  # a multi-mesh engine needs offsets
    assert np.array_equal(Wd._check_q_off([0, 4, 6, 10], 3, 10), [0, 4, 6, 10])
