#!/usr/bin/env python3
"""Benchmark of the exact winding-number engine on one MI355X.

Workload: the ~1M-triangle open displaced sphere (nu = 224 geodesic sphere, 32 seeded spherical
harmonics up to degree 16, amplitude 0.15, 5% of facets deleted -> 953,344 triangles) and a
256^3 grid of points over its bounding box x1.2 (16,777,216 points).  One step = one batched
solid_angle + is_inside call over all points (Morton sort of the points, traversal, outputs),
with the points already resident in HBM.  The tree is built once before the timed region, as
the reference benchmark constructs its engine before timing queries
(modules/winding/tests/test_fast_winding_number.cpp:99-110).

Also reported (outside the timed region):
  * roofline of the traversal kernel: algorithmic FP32 FLOPs (per-point work counters of the
    same tree and criterion x FLOPs per operation) / its average duration from HIP events on
    the engine's stream;
  * accuracy of the timed results against the engine's FP64 brute force on a strided sample;
  * cpu_baseline: the CPU restatement of the reference engine (oracle/, OpenMP over points)
    on the host cores, over a strided sample of the same points.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FP32 FLOPs per operation of the traversal kernel, counted from fwn_query.hip
# (FMA = 2, each rsqrt / atan2 counted as 1):
#   slot test   : d = p - q (3), |d|^2 (5), compare with the stored beta^2 r2 (1)            = 9
#   far term    : rsq + 2 powers (3), d.N (5), dz^2 (1), order 1 in Horner form (13),
#                 order 2 in Horner form (25), combine + accumulate (6); the constant terms
#                 tr M1 and C.d are folded into the order-1 and order-2 coefficients by the
#                 build                                                                     = 53
#   FP32 tri    : offsets (9), lengths (15), 3 rsq (3), normalise (9), edges (6), cross (9),
#                 triple product (5), dots (15), denominator (3), atan2 (1), x2 + acc (2)    = 77
#   FP64 tri    : the same work in FP64 (reported separately, not in the FP32 figure)       = 77
FLOP_TEST, FLOP_FAR, FLOP_TRI32, FLOP_TRI64 = 9, 53, 77, 77
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
METRIC = "winding queries/s @1M tris: exact GPU vs surrogate at ≤1e-4 inside-error"


def _read_traffic():
    """HBM bytes per traversal launch from the committed rocprofv3 --pmc summary, if any."""
    p = os.path.join(ROOT, "profiles", "k_traverse_large_exact_pmc.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def _host_cpu():
    model, sockets = "unknown", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            if line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return {"model": model, "sockets": max(len(sockets), 1), "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(V, F, q, sample_stride: int, threads: int, acc=None):
    """CPU restatement timed on the host cores. With ``acc`` = (sample indices, FP64 w of the
    sample, beta at 1e-3) it also reports the restatement's own error at its default beta 2 on
    that sample, and its rate at the beta the GPU engine needs for max |dw| <= 1e-3."""
    from oracle import oracle as O
    import numpy as np
    sub = q[::sample_stride].copy()
    t0 = time.perf_counter()
    R = O.Restated(V, F)
    build_s = time.perf_counter() - t0
    R.solid_angle(sub[:4096], nthreads=threads)  # warm-up
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        R.solid_angle(sub, nthreads=threads)
        times.append(time.perf_counter() - t0)
    dt = sorted(times)[2]
    one = sub[::16].copy()
    t1 = []
    for _ in range(3):
        t0 = time.perf_counter()
        R.solid_angle(one, nthreads=1)
        t1.append(time.perf_counter() - t0)
    dt1 = sorted(t1)[1]
    extra = {}
    if acc is not None:
        idx, wb, beta = acc
        wr = R.solid_angle(q[idx], nthreads=threads).astype(np.float64) / (4 * np.pi)
        err = np.abs(wr - wb)
        extra["restated_error_at_beta_2"] = {"sample": int(idx.size), "max_abs_dw": float(err.max()),
                                             "p99_abs_dw": float(np.quantile(err, 0.99)),
                                             "mean_abs_dw": float(err.mean())}
        if beta is not None:
            sb = q[::4 * sample_stride].copy()
            tb = []
            for _ in range(3):
                t0 = time.perf_counter()
                R.solid_angle(sb, beta=beta, nthreads=threads)
                tb.append(time.perf_counter() - t0)
            extra["value_at_matched_beta"] = {"accuracy_scale": beta, "value": sb.shape[0] / sorted(tb)[1],
                                              "sample": f"every {4 * sample_stride}th grid point "
                                                        f"({sb.shape[0]} points), median of 3"}
    return {
        "value": sub.shape[0] / dt,
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"every {sample_stride}th grid point" if sample_stride > 1 else "all grid points") +
                  f" ({sub.shape[0]} points), median of 5 runs; "
                  f"CPU restatement of the reference engine (4-ary box-area BVH, order 2, beta 2, FP32), "
                  f"OpenMP over points; build {build_s:.2f} s single-threaded, not included",
        "single_thread_value": one.shape[0] / dt1,
        "single_thread_sample": f"every {16 * sample_stride}th grid point ({one.shape[0]} points), median of 3",
        "host": _host_cpu(),
        **extra,
    }


def bench_batched(steps: int = 5):
    """64 meshes (displaced spheres / tori, 1k..500k triangles, 10k..2M points each, uniform in
    bbox x1.3): one build and one query launch for all meshes; leaf size and launch shape picked
    by the tuner (fwn_create_tuned over leaf sizes 1, 2, 4, 8, each with its launch autotune)."""
    import numpy as np
    import torch

    from lagrange_amd import workloads as W
    from lagrange_amd.winding import MultiFastWindingNumber

    meshes, pts = W.batched_workload()
    arrs = [(m.vertices, m.facets) for m in meshes]
    q_off = np.zeros(len(pts) + 1, np.int64)
    q_off[1:] = np.cumsum([p.shape[0] for p in pts])
    qd = torch.from_numpy(np.concatenate(pts)).cuda()
    t0 = time.perf_counter()
    tuned = MultiFastWindingNumber.tuned(arrs, qd, q_off, (1, 2, 4, 8))
    tune_wall = time.perf_counter() - t0
    leaf = tuned.tuning["leaf_max"]
    tuned.close()
    t0 = time.perf_counter()
    mm = MultiFastWindingNumber(arrs, leaf_max=leaf)
    build_wall = time.perf_counter() - t0
    waves, mode, _ = mm.autotune(qd, q_off)
    mm.solid_angle_batch(points=qd, q_off=q_off)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        om = mm.solid_angle_batch(points=qd, q_off=q_off)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    n = int(q_off[-1])
    # accuracy on a strided sample of each of 8 meshes vs a per-mesh FP64 brute force engine
    from lagrange_amd.winding import FastWindingNumber
    errs = []
    omh = om.cpu().numpy()
    for i in range(0, len(meshes), 8):
        sub = pts[i][:: max(1, pts[i].shape[0] // 512)]
        wb = FastWindingNumber((meshes[i].vertices, meshes[i].facets)).bruteforce_f64(sub) / (4 * np.pi)
        w = omh[q_off[i]:q_off[i + 1]][:: max(1, pts[i].shape[0] // 512)].astype(np.float64) / (4 * np.pi)
        errs.append(np.abs(w - wb))
    err = np.concatenate(errs)
    info = mm.tree_info()
    return {
        "workload": f"{len(meshes)} meshes, {info['num_triangles']} triangles, {n} points (uniform in each bbox x1.3)",
        "queries_per_s": n / dt,
        "ms_per_step": dt * 1e3,
        "launch": {"waves_per_block": waves, "mode": "persistent" if mode == 1 else "grid", "chosen_by": "autotune"},
        "leaf_max": leaf,
        "leaf_max_chosen_by": "fwn_create_tuned over 1, 2, 4, 8",
        "tuning_wall_s": tune_wall,
        "build_ms": info["build_ms"],
        "build_wall_s": build_wall,
        "max_abs_dw": float(err.max()),
        "mean_abs_dw": float(err.mean()),
    }


def bench_accuracy(V, F, q, qd, steps: int = 10, betas=(2.0, 2.5, 2.75, 3.0, 3.5, 4.0), sample: int = 16384):
    """Accuracy / speed trade of the accuracy scale on the headline workload: for beta = 2, 2.5,
    2.75, 3, ... the engine is rebuilt (the build stores beta^2 r^2 per slot), one step is timed as in
    the headline, and max / p99 |dw| are measured against the FP64 brute force on a strided
    sample of ~16,384 grid points (odd stride). Stops at the first beta whose maximum error is <= 1e-3."""
    import numpy as np
    import torch

    from lagrange_amd.winding import FastWindingNumber

    n = q.shape[0]
    idx = np.arange(0, n, max(n // sample, 1) | 1)  # odd stride: spreads over x of the 2^k grid
    wb = None
    rows = []
    for beta in betas:
        e = FastWindingNumber((V, F), accuracy_scale=beta)
        if wb is None:
            wb = e.bruteforce_f64(q[idx]) / (4 * np.pi)
        for _ in range(2):
            om, ins = e.solid_angle_batch(qd, with_inside=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            om, ins = e.solid_angle_batch(qd, with_inside=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        err = np.abs(om.cpu().numpy()[idx].astype(np.float64) / (4 * np.pi) - wb)
        amb = np.abs(wb - 0.5) <= 1e-3
        ins_bad = (ins.cpu().numpy()[idx] != (wb > 0.5)) & ~amb
        rows.append({"accuracy_scale": beta, "queries_per_s": n / dt, "ms_per_step": dt * 1e3,
                     "max_abs_dw": float(err.max()), "p99_abs_dw": float(np.quantile(err, 0.99)),
                     "mean_abs_dw": float(err.mean()), "inside_mismatch_excl_band": int(ins_bad.sum())})
        e.close()
        del om, ins
        if err.max() <= 1e-3:
            break
    ok = [r for r in rows if r["max_abs_dw"] <= 1e-3]
    return {
        "workload": f"headline mesh and grid ({n} points); error sample every {idx[1] - idx[0]}th point "
                    f"({idx.size} points) vs FP64 brute force",
        "sweep": rows,
        "at_1e-3": ok[0] if ok else None,
    }, idx, wb


def bench_near_surface(steps: int = 5):
    """200,000-triangle displaced sphere, 4M points offset 1e-7..1e-2 x diag along normals."""
    import numpy as np
    import torch

    from lagrange_amd import workloads as W
    from lagrange_amd.winding import FastWindingNumber

    m, q = W.near_surface_workload()
    V, F = m.f32()
    qd = torch.from_numpy(q).cuda()
    # the tuner's engine (leaf size and launch shape for these points), then the default one
    tuned = FastWindingNumber.tuned((V, F), qd, (1, 2, 4, 8))
    tuned.solid_angle_batch(qd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tuned.solid_angle_batch(qd)
    torch.cuda.synchronize()
    dt_tuned = (time.perf_counter() - t0) / steps
    leaf_tuned = tuned.tuning["leaf_max"]
    tuned.close()
    e = FastWindingNumber((V, F))
    e.solid_angle_batch(qd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        om, ins = e.solid_angle_batch(qd, with_inside=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    idx = np.arange(0, q.shape[0], 977)
    wb = e.bruteforce_f64(q[idx]) / (4 * np.pi)
    w = om.cpu().numpy()[idx].astype(np.float64) / (4 * np.pi)
    err = np.abs(w - wb)
    insb = ins.cpu().numpy()[idx]
    amb = np.abs(wb - 0.5) <= 1e-3
    _, st = e.query_stats(q[idx])
    return {
        "workload": f"{F.shape[0]} triangles, {q.shape[0]} points at 1e-7..1e-2 x diag from the surface",
        "queries_per_s": q.shape[0] / dt,
        "ms_per_step": dt * 1e3,
        "sample": int(idx.size),
        "max_abs_dw": float(err.max()),
        "p99_abs_dw": float(np.quantile(err, 0.99)),
        "mean_abs_dw": float(err.mean()),
        "inside_mismatch_excl_band": int(((insb != (wb > 0.5)) & ~amb).sum()),
        "fp64_triangles_per_point": float(st[:, 3].mean()),
        "tuned": {"leaf_max": leaf_tuned, "queries_per_s": q.shape[0] / dt_tuned, "ms_per_step": dt_tuned * 1e3,
                  "chosen_by": "fwn_create_tuned over leaf sizes 1, 2, 4, 8"},
    }


def bench_voxel(mesh, steps: int = 5):
    """Winding-number voxel signs (mesh_to_volume's interior test, batched) on the ~1M-triangle
    mesh: every voxel of the index box around the mesh (bbox + 3-voxel band) at voxel size
    0.002 x bbox diagonal, centres generated on the GPU, flags left in HBM.  Beside it, the same
    voxels sent as host points (centres made on the host, uploaded, flags copied back)."""
    import numpy as np
    import torch

    from lagrange_amd import volume as VOL
    from lagrange_amd.winding import FastWindingNumber

    V, F = mesh.f32()
    e = FastWindingNumber((V, F))
    vs = VOL.resolve_voxel_size(V, -0.002)
    lo, dims = VOL.index_box(V, vs)
    n = int(np.prod(dims.astype(np.int64)))
    dev = torch.device("cuda:0")
    ins = e.sign_voxel_box(lo, dims, vs, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ins = e.sign_voxel_box(lo, dims, vs, device=dev)
    torch.cuda.synchronize()
    t_dev = (time.perf_counter() - t0) / steps
    # host-points path, timed once (centre generation on the host included)
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    t0 = time.perf_counter()
    ijk = np.stack([x.ravel() + lo[0], y.ravel() + lo[1], z.ravel() + lo[2]], axis=1)
    c = VOL.voxel_centres(ijk, vs)
    ins_h = e.is_inside_batch(c)
    t_host = time.perf_counter() - t0
    same = bool(np.array_equal(ins.cpu().numpy().ravel(), ins_h))
    return {
        "workload": f"{F.shape[0]} triangles, voxel box {int(dims[0])}x{int(dims[1])}x{int(dims[2])} "
                    f"({n} voxels, voxel size {vs:.5g} = 0.002 x bbox diagonal)",
        "voxels_per_s": n / t_dev,
        "ms_per_box": t_dev * 1e3,
        "host_points_path_voxels_per_s": n / t_host,
        "inside_fraction": float(ins.float().mean().item()),
        "identical_to_host_points_path": same,
    }


def bench_examples(mesh, steps: int = 5):
    """The reference's two winding examples, batched, on the ~1M-triangle mesh: the
    fix_orientation criterion of every facet (2 probes per facet at epsilon 1e-2, made on the GPU)
    and sample_points_in_mesh with 16,777,216 samples in the bounding box (drawn, tested and
    compacted on the GPU)."""
    import numpy as np
    import torch

    from lagrange_amd.winding import FastWindingNumber

    V, F = mesh.f32()
    e = FastWindingNumber((V, F))
    dV, dF = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()
    crit = e.orientation_criterion(dV, dF, 1e-2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        crit = e.orientation_criterion(dV, dF, 1e-2)
    torch.cuda.synchronize()
    t_or = (time.perf_counter() - t0) / steps
    flipped = int((crit > 0.8).sum().item())
    n = 1 << 24
    dev = torch.device("cuda:0")
    pts = e.sample_inside(n, 1, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pts = e.sample_inside(n, 1, device=dev)
    torch.cuda.synchronize()
    t_s = (time.perf_counter() - t0) / steps
    return {
        "fix_orientation": {"workload": f"{F.shape[0]} facets, epsilon 1e-2 (2 probes per facet)",
                            "facets_per_s": F.shape[0] / t_or, "ms": t_or * 1e3,
                            "facets_over_threshold": flipped},
        "sample_points_in_mesh": {"workload": f"{n} samples uniform in the bbox", "samples_per_s": n / t_s,
                                  "ms": t_s * 1e3, "kept_fraction": pts.shape[0] / n},
    }


def bench_closest(mesh, q, steps: int = 5):
    """Closest point / unsigned distance (bvh::TriangleAABBTree::get_closest_point, batched) for the
    headline grid points on the ~1M-triangle mesh, FP64 point-triangle distances, results in HBM;
    the outputs are checked for consistency on a strided sample."""
    import numpy as np
    import torch

    from lagrange_amd.winding import FastWindingNumber

    V, F = mesh.f32()
    e = FastWindingNumber((V, F), leaf_max=8)  # the box search visits fewer nodes with buckets
    qd = torch.from_numpy(q).cuda()
    d2, fi, cp = e.closest_point(qd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        d2, fi, cp = e.closest_point(qd)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    out = {"workload": f"{F.shape[0]} triangles, {q.shape[0]} grid points (bbox x1.2)",
           "points_per_s": q.shape[0] / dt, "ms": dt * 1e3, "leaf_max": 8}
    # consistency of the three outputs (exactness against a brute force is a test's job):
    # |closest - q|^2 equals the squared distance, and the closest point lies on its facet's plane
    idx = np.arange(0, q.shape[0], q.shape[0] // 4096)
    d2h, fih, cph = d2.cpu().numpy()[idx], fi.cpu().numpy()[idx], cp.cpu().numpy()[idx]
    qq = q[idx].astype(np.float64)
    tri = V[F[fih]].astype(np.float64)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-300)
    out["sample"] = int(idx.size)
    out["max_rel_dev_d2_vs_point"] = float(np.max(np.abs(np.sum((cph - qq) ** 2, 1) - d2h) / np.maximum(d2h, 1e-30)))
    out["max_off_plane"] = float(np.max(np.abs(np.sum((cph - tri[:, 0]) * nrm, 1))))
    out["exactness"] = "tests/test_gpu_closest.py (FP64 brute force)"
    return out


def bench_io(mesh):
    """Host-side mesh input: the ~1M-triangle mesh written as OBJ (text) and binary PLY to a
    temporary directory, then read back by the native reader (best of 3 loads each)."""
    import tempfile

    import numpy as np

    from lagrange_amd import io as mio

    V = mesh.vertices.astype(np.float64)
    F = mesh.facets.astype(np.int32)
    out = {"workload": f"{F.shape[0]} triangles, {V.shape[0]} vertices"}
    with tempfile.TemporaryDirectory() as d:
        for ext, key in ((".obj", "obj"), (".ply", "ply_binary")):
            p = os.path.join(d, "mesh" + ext)
            mio.save_mesh(p, V, F)
            best = float("inf")
            for _ in range(3):
                t0 = time.perf_counter()
                m = mio.load_mesh(p)
                best = min(best, time.perf_counter() - t0)
            same = bool(np.array_equal(m.vertices, V) and np.array_equal(m.facets, F))
            size = os.path.getsize(p)
            out[key] = {"load_s": best, "MB_per_s": size / best / 1e6, "triangles_per_s": F.shape[0] / best,
                        "file_MB": size / 1e6, "round_trip_exact": same}
    return out


def bench_point_calls(mesh, n: int = 10000, threads: int = 16, runs: int = 3):
    """The reference's per-point benchmark (modules/winding/tests/test_fast_winding_number.cpp:85-110)
    through the C++ drop-in class (lagrange_amd/cpp/bench/bench_point_calls.cpp): n serial
    is_inside calls on std::mt19937 points over the mesh's float bounding box, the same from
    `threads` callers at once, the n points in one batched call, and batched calls of k points.
    Beside it the CPU restatement answering the same n points one at a time on one thread (the
    reference's own calling pattern on the host)."""
    import subprocess
    import tempfile

    import numpy as np

    from lagrange_amd import build as B
    from lagrange_amd import io as mio
    from lagrange_amd.examples import mt19937_box_samples
    from oracle import oracle as O

    V = mesh.vertices.astype(np.float64)
    F = mesh.facets.astype(np.int32)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "mesh.ply")
        mio.save_mesh(path, V, F)
        r = subprocess.run([B.CXX_BENCH, path, "-n", str(n), "-t", str(threads), "-r", str(runs)],
                           capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        return {"error": (r.stdout + r.stderr)[-800:]}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    Vf, Ff = mesh.f32()
    pts = mt19937_box_samples(Vf.min(0), Vf.max(0), n)
    R = O.Restated(Vf, Ff)
    R.solid_angle(pts[:256], nthreads=1)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        om = R.solid_angle(pts, nthreads=1)
        ts.append(time.perf_counter() - t0)
    dt = sorted(ts)[1]
    out["cpu_restated_one_thread"] = {"us_per_point": 1e6 * dt / n, "points_per_s": n / dt,
                                      "num_insides": int(O.inside_from_omega(om).sum())}
    bk = out.get("batched_call_us", {})
    # a batched call of one point is a per-point call: the crossover is the smallest k >= 2 at
    # which one batched call of k points beats k per-point calls
    cross = [int(k) for k, us in bk.items() if int(k) >= 2 and us < int(k) * out["serial_us_per_call"]]
    out["crossover_points"] = min(cross) if cross else None
    out["workload"] = (f"{F.shape[0]} triangles; {n} points from std::mt19937 (default seed), uniform in the "
                       f"float bbox; {threads} threads x {n} calls for the threaded rate")
    return out


def bench_surrogate(mesh, steps: int = 3, train_steps: int = 0):
    """Surrogate on the ~1M-triangle mesh: 16 x 2^19 hash grid (2 FP16 features) + 4 x 128 ReLU MLP
    (BF16 MFMA), trained on 20M points labelled by the exact engine (half uniform, half near the
    surface; points inside the distance bound carry no loss); 67,108,864 uniform points in the box answered by surrogate + exact fallback
    (prediction within 0.2 of 0.5, or inside the distance bound), against the exact engine."""
    import numpy as np
    import torch

    from lagrange_amd import workloads as W
    from lagrange_amd.surrogate import Surrogate
    from lagrange_amd.winding import FastWindingNumber

    V, F = mesh.f32()
    e = FastWindingNumber((V, F))
    # train_steps > 0 shortens training (profiling runs of the inference kernels, whose work per
    # point does not depend on the weights); 0 keeps the default 3,000 Adam steps
    s = Surrogate(e, steps=train_steps) if train_steps > 0 else Surrogate(e)
    info = s.train()
    lo, hi = V.min(0), V.max(0)
    n = 1 << 26
    q = W.uniform_queries(lo, hi, n, 5, 1.0)
    qd = torch.from_numpy(q).cuda()
    del q
    om, ins_e = e.solid_angle_batch(qd, with_inside=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        om, ins_e = e.solid_angle_batch(qd, with_inside=True)
    torch.cuda.synchronize()
    t_exact = (time.perf_counter() - t0) / steps
    w, ins_s, nf = s.query(qd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        w, ins_s, nf = s.query(qd)
    torch.cuda.synchronize()
    t_sur = (time.perf_counter() - t0) / steps
    si = s.info()
    mis = int((ins_s != ins_e).sum().item())
    # FP64 check of both on a strided sample
    idx = torch.arange(0, n, n // 2048, device=qd.device)
    wb = e.bruteforce_f64(qd[idx]).cpu().numpy() / (4 * np.pi)
    amb = np.abs(wb - 0.5) <= 1e-3
    mis64 = int(((ins_s[idx].cpu().numpy() != (wb > 0.5)) & ~amb).sum())
    mis64_e = int(((ins_e[idx].cpu().numpy() != (wb > 0.5)) & ~amb).sum())
    mlp_flops = n * 2 * (32 * 128 + 3 * 128 * 128)
    return {
        "workload": f"{F.shape[0]} triangles, {n} points uniform in the bbox; 16 levels x 2^19 x 2 FP16 hash grid, "
                    "MLP 32-128-128-128-128-1 ReLU, 20M training points labelled by the exact engine",
        "exact_queries_per_s": n / t_exact,
        "surrogate_queries_per_s": n / t_sur,
        "surrogate_vs_exact": t_exact / t_sur,
        "fallback_fraction": nf / n,
        "misclassified_vs_exact": mis / n,
        "misclassified_vs_fp64_sample": mis64,
        "exact_misclassified_vs_fp64_sample": mis64_e,
        "fp64_sample": int(idx.numel()),
        "train_s": si["train_ms"] / 1e3,
        "train_loss": si["last_loss"],
        "sort_ms": si["encode_ms"],
        "encode_mlp_ms": si["mlp_ms"],
        "fallback_ms": si["fallback_ms"],
        # k_sg_infer fuses the hash-grid encoding with the MLP: the MLP's FLOPs over the fused
        # kernel's time (encoding included)
        "mfma": {"kernel": "k_sg_infer", "achieved": mlp_flops / (si["mlp_ms"] * 1e-3) / 1e12, "peak": 2500.0,
                 "unit": "TFLOP/s", "frac": mlp_flops / (si["mlp_ms"] * 1e-3) / 1e12 / 2500.0},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--leaf-max", type=int, default=1,
                    help="leaf bucket size (default 1: the engine's default, the reference's tree layout)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tuned", action="store_true",
                    help="skip the tuner's engine beside the headline (keeps a profile to the default engine)")
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--extras", default="batched,near,voxel,examples,closest,io,points,accuracy,surrogate",
                    help="comma list of extra workloads reported beside the headline: batched, near, voxel, "
                         "examples, closest, io, accuracy, surrogate")
    ap.add_argument("--surrogate-train-steps", type=int, default=0,
                    help="Adam steps of the surrogate leg (0: the default 3,000; fewer for profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (scale for rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this engine's workloads run on a single MI355X "
              "(query sharding across GPUs is a design note, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from lagrange_amd import workloads as W
    from lagrange_amd.winding import FastWindingNumber

    mesh, q = W.large_exact_workload(grid_n=args.grid)
    V, F = mesh.f32()
    n = q.shape[0]
    dev = torch.device("cuda:0")
    e = FastWindingNumber((V, F), leaf_max=args.leaf_max)
    info = e.tree_info()
    qd = torch.from_numpy(q).to(dev)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        om, ins = e.solid_angle_batch(qd, with_inside=True)
    torch.cuda.synchronize()
    trav = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        om, ins = e.solid_angle_batch(qd, with_inside=True)
        trav.append(e.last_timing()[0])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = n * args.steps / dt

    # PCIe-inclusive rate of the same call through host pointers (points uploaded, results copied
    # back to pageable host arrays): what a caller holding host buffers sees; never `value`
    e.solid_angle_batch(q, with_inside=True)
    t0 = time.perf_counter()
    for _ in range(3):
        e.solid_angle_batch(q, with_inside=True)
    dt_h = (time.perf_counter() - t0) / 3
    host_rec = {"queries_per_s": n / dt_h, "ms_per_call": dt_h * 1e3,
                "bytes_moved": int(n * (12 + 4 + 1)),
                "note": "host (pageable) arrays in and out through the C ABI, copies included"}

    # the tuner's engine for the same points (leaf size x launch shape), timed the same way
    tuned = om_t = tuned_rec = None
    if not args.no_tuned:
        tuned = FastWindingNumber.tuned((V, F), qd, (1, 2, 4, 8))
        for _ in range(args.warmup):
            tuned.solid_angle_batch(qd, with_inside=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            om_t, _ = tuned.solid_angle_batch(qd, with_inside=True)
        torch.cuda.synchronize()
        dt_t = time.perf_counter() - t0
        tuned_rec = {"leaf_max": tuned.tuning["leaf_max"], "queries_per_s": n * args.steps / dt_t,
                     "ms_per_step": dt_t * 1e3 / args.steps,
                     "chosen_by": "fwn_create_tuned over leaf sizes 1, 2, 4, 8"}

    # algorithmic work of the traversal on this tree and criterion (all points, one extra launch)
    _, st = e.query_stats(q)
    tot = st.astype(np.int64).sum(axis=0)
    flops32 = tot[0] * FLOP_TEST + tot[1] * FLOP_FAR + tot[2] * FLOP_TRI32
    flops64 = tot[3] * FLOP_TRI64
    trav_ms = float(np.mean(trav))
    achieved = flops32 / (trav_ms * 1e-3) / 1e12
    traffic = _read_traffic()

    # accuracy of the timed results vs the engine's FP64 brute force on a strided sample
    idx = np.arange(0, n, 4099)
    w = om.cpu().numpy()[idx].astype(np.float64) / (4 * np.pi)
    wb = e.bruteforce_f64(q[idx]) / (4 * np.pi)
    err = np.abs(w - wb)
    ins_ok = ins.cpu().numpy()[idx] == (wb > 0.5)
    amb = np.abs(wb - 0.5) <= 1e-3
    if tuned is not None:
        tuned_rec["max_abs_dw"] = float(np.abs(om_t.cpu().numpy()[idx].astype(np.float64) / (4 * np.pi) - wb).max())
        tuned.close()
        del om_t

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "exact engine, displaced sphere nu=224 with 5% facets deleted "
                        f"({F.shape[0]} triangles), {args.grid}^3 grid over bbox x1.2 ({n} points)",
            "engine": "exact",
            "accuracy_scale": 2.0,
            "order": 2,
            "leaf_max": args.leaf_max,
            "tree_builder": "sah (binned box-area, 4-ary: the reference's tree layout)",
            "tree_width": info["width"],
            "tree_nodes": info["num_nodes"],
            "tree_depth": info["max_depth"],
            "build_ms": round(info["build_ms"], 3),
            "fp64_near_field_rel": 1e-4,
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_traverse",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "kernel_ms": trav_ms,
            "flop_per_point": flops32 / n,
            "fp64_flop_per_point": flops64 / n,
            "ops_per_point": {"tests": tot[0] / n, "far": tot[1] / n, "tri32": tot[2] / n, "tri64": tot[3] / n},
            "hbm_algorithmic_gbs": n * 17 / (trav_ms * 1e-3) / 1e9,
            "hbm_peak_gbs": PEAK_HBM_GBS,
        },
        "accuracy": {
            "sample": int(idx.size),
            "vs": "FP64 brute force (GPU kernel)",
            "max_abs_dw": float(err.max()),
            "p99_abs_dw": float(np.quantile(err, 0.99)),
            "mean_abs_dw": float(err.mean()),
            "inside_mismatch_excl_band": int((~ins_ok & ~amb).sum()),
        },
        "tuned": tuned_rec,
        "host_arrays": host_rec,
    }
    extras = [x for x in args.extras.split(",") if x and x != "none"]
    if "batched" in extras:
        line["batched"] = bench_batched()
    if "near" in extras:
        line["near_surface"] = bench_near_surface()
    if "voxel" in extras:
        line["voxel"] = bench_voxel(mesh)
    if "examples" in extras:
        line["examples"] = bench_examples(mesh)
    if "closest" in extras:
        line["closest"] = bench_closest(mesh, q)
    if "io" in extras:
        line["io"] = bench_io(mesh)
    if "points" in extras:
        line["point_calls"] = bench_point_calls(mesh)
    acc = None
    if "accuracy" in extras:
        line["accuracy_scale_sweep"], idx_a, wb_a = bench_accuracy(V, F, q, qd)
        hit = line["accuracy_scale_sweep"]["at_1e-3"]
        acc = (idx_a, wb_a, hit["accuracy_scale"] if hit else None)
    if "surrogate" in extras:
        del qd, om, ins
        torch.cuda.empty_cache()
        line["surrogate"] = bench_surrogate(mesh, train_steps=args.surrogate_train_steps)
    if not args.no_cpu:
        # every core this process may run on; a GPU box gives one GPU's process its CPU share
        # through OMP_NUM_THREADS (16 there), which bounds the OpenMP pool the same way
        threads = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
        line["cpu_baseline"] = cpu_baseline(V, F, q, args.cpu_stride, threads, acc)
        line["speedup_vs_cpu"] = value / line["cpu_baseline"]["value"]
        hit = line.get("accuracy_scale_sweep", {}).get("at_1e-3")
        if hit:
            # GPU at the accuracy scale that keeps max |dw| <= 1e-3, against the CPU restatement at
            # the reference default (beta 2, itself less accurate) and at the same beta
            cb = line["cpu_baseline"]
            hit["speedup_vs_cpu_beta_2"] = hit["queries_per_s"] / cb["value"]
            if "value_at_matched_beta" in cb:
                hit["speedup_vs_cpu_same_beta"] = hit["queries_per_s"] / cb["value_at_matched_beta"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
