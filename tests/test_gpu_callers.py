"""GPU tests of the batched example callers through the C ABI (and the C++ example programs).

fix_orientation (modules/winding/examples/fix_orientation.cpp:86-115):
  * bit-exact: the criterion equals (omega(p) - omega(q)) / (4 pi) of the engine's own batched
    query at probes computed on the host with the same float operations (so the device-made
    probes are the same floats);
  * mathematics: for a closed outward mesh the criterion is -1 for correctly oriented facets and
    +1 for flipped ones, so fixing a mesh with flipped facets restores it exactly;
  * the CPU restatement of the reference engine (oracle) gives the same criterion within twice
    the order-2 error bound (2 x 4e-3 in w).
sample_points_in_mesh (modules/winding/examples/sample_points_in_mesh.cpp:59-83):
  * reference sequence: the C++ example (std::mt19937 + std::uniform_real_distribution<float>)
    and the Python restatement write the same .xyz text, and their kept points equal the
    reference's serial loop of per-point is_inside calls bit for bit;
  * the batched counter-based sampler (fwn_sample_inside): kept points equal host-regenerated
    samples filtered by the engine's own is_inside, in sample order, and the kept fraction
    matches the volume ratio mesh / box (divergence theorem) within 4 sigma.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from lagrange_amd import _lib as L
from lagrange_amd import examples as EX
from lagrange_amd import io as mio
from lagrange_amd import workloads as W
from lagrange_amd.winding import FastWindingNumber

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP_BUILD = os.path.join(ROOT, "lagrange_amd", "cpp", "build")


def host_probes(V, F, eps):
    V = np.asarray(V, np.float32)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    u, v = b - a, c - a
    n = np.stack([u[:, 1] * v[:, 2] - u[:, 2] * v[:, 1], u[:, 2] * v[:, 0] - u[:, 0] * v[:, 2],
                  u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]], axis=1)
    z = (n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2]
    pos = z > 0
    n[pos] = n[pos] / np.sqrt(z[pos])[:, None]
    g = ((a + b) + c) / np.float32(3.0)
    t = np.float32(eps) * n
    return g + t, g - t


def host_criterion(om_p, om_q):
    return (np.asarray(om_p, np.float32) - np.asarray(om_q, np.float32)).astype(np.float64) / (4.0 * np.pi)


def test_criterion_bit_exact_vs_host_probes(open_mesh):
    V, F = open_mesh
    e = FastWindingNumber((V, F))
    eps = 1e-2
    crit = e.orientation_criterion(V, F, eps)
    p, q = host_probes(V, F, eps)
    assert p.dtype == np.float32
    want = host_criterion(e.solid_angle_batch(p), e.solid_angle_batch(q))
    np.testing.assert_array_equal(crit, want)


def test_fix_orientation_restores_flipped_facets(ico5):
    V, F = ico5
    rng = np.random.default_rng(5)
    flip = rng.random(F.shape[0]) < 0.25
    Fin = F.copy()
    Fin[flip, 0], Fin[flip, 1] = F[flip, 1], F[flip, 0]
    F2, crit, counts = EX.fix_orientation(V, Fin, (V, F))
    assert np.all(crit[flip] > 0.99) and np.all(crit[~flip] < -0.99)
    np.testing.assert_array_equal(F2, F)
    assert counts == {"pos": int(flip.sum()), "neg": int((~flip).sum()), "tot": F.shape[0]}


def test_criterion_matches_restated_reference(oracle, open_mesh):
    V, F = open_mesh
    e = FastWindingNumber((V, F))
    crit = e.orientation_criterion(V, F, 1e-2)
    p, q = host_probes(V, F, 1e-2)
    R = oracle.Restated(V, F)
    want = host_criterion(R.solid_angle(p), R.solid_angle(q))
    assert np.abs(crit - want).max() <= 8e-3
    band = np.abs(want - 0.8) <= 8e-3
    assert np.array_equal((crit > 0.8)[~band], (want > 0.8)[~band])


def test_criterion_device_pointers_and_errors(ico5):
    import torch
    V, F = ico5
    e = FastWindingNumber((V, F))
    host = e.orientation_criterion(V, F, 0.02)
    dev = e.orientation_criterion(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), 0.02)
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
    assert e.orientation_criterion(V, F[:0], 0.02).shape == (0,)
    bad = F.copy()
    bad[7, 2] = V.shape[0]
    with pytest.raises(L.FWNError) as ex:
        e.orientation_criterion(V, bad, 0.02)
    assert ex.value.status == L.FWN_INVALID_ARGUMENT
    with pytest.raises(L.FWNError):
        e.orientation_criterion(V, F, -1.0)


def _mesh_volume(V, F):
    V = V.astype(np.float64)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0


def test_sample_inside_matches_host_regeneration(ico5):
    V, F = ico5
    e = FastWindingNumber((V, F))
    n, seed = 300_000, 7
    pts = e.sample_inside(n, seed)
    lo, hi = V.min(0), V.max(0)
    u = np.stack([W.rng_uniform(seed, 40 + d, n) for d in range(3)], axis=1).astype(np.float32)
    cand = (hi - lo).astype(np.float32) * u + lo
    keep = e.is_inside_batch(cand)
    np.testing.assert_array_equal(pts, cand[keep])
    ratio = _mesh_volume(V, F) / np.prod((hi - lo).astype(np.float64))
    sigma = np.sqrt(ratio * (1 - ratio) / n)
    assert abs(pts.shape[0] / n - ratio) < 4 * sigma
    # device output and an explicit box
    import torch
    dpts = e.sample_inside(n, seed, device=torch.device("cuda:0"))
    np.testing.assert_array_equal(dpts.cpu().numpy(), pts)
    assert e.sample_inside(0).shape == (0, 3)
    far = e.sample_inside(1000, 1, lo=[5, 5, 5], hi=[6, 6, 6])
    assert far.shape == (0, 3)
    with pytest.raises(L.FWNError):
        e.sample_inside(10, 1, lo=[1, 0, 0], hi=[0, 1, 1])


def test_cpp_example_programs(tmp_path, ico5):
    V, F = ico5
    rng = np.random.default_rng(9)
    flip = rng.random(F.shape[0]) < 0.3
    Fin = F.copy()
    Fin[flip, 0], Fin[flip, 1] = F[flip, 1], F[flip, 0]
    mio.save_mesh(tmp_path / "in.ply", V, Fin)
    mio.save_mesh(tmp_path / "ref.obj", V, F)
    exe = os.path.join(CPP_BUILD, "fix_orientation")
    r = subprocess.run([exe, str(tmp_path / "in.ply"), str(tmp_path / "ref.obj"), str(tmp_path / "out.obj"),
                        "--criterion-json", str(tmp_path / "c.json")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.strip() == f"pos {int(flip.sum())} neg {int((~flip).sum())} tot {F.shape[0]}"
    np.testing.assert_array_equal(mio.load_mesh(tmp_path / "out.obj").facets, F)
    import json
    crit = np.array(json.load(open(tmp_path / "c.json"))["criterions"])
    np.testing.assert_array_equal(crit, FastWindingNumber((V, F)).orientation_criterion(V, Fin, 1e-2))
    # sample_points_in_mesh: the reference's mt19937 sequence, batched is_inside, fmt "{}" text
    exe = os.path.join(CPP_BUILD, "sample_points_in_mesh")
    r = subprocess.run([exe, str(tmp_path / "ref.obj"), str(tmp_path / "s.xyz"), "-n", "50000"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    want = EX.sample_points_in_mesh(FastWindingNumber((V, F)), 50000)
    assert r.stdout.strip() == f"{want.shape[0]} of 50000 samples inside"
    assert open(tmp_path / "s.xyz").read() == EX.xyz_text(want)
    r = subprocess.run([sys.executable, "-m", "lagrange_amd.examples", "sample_points_in_mesh",
                        str(tmp_path / "ref.obj"), str(tmp_path / "p.xyz"), "-n", "50000"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    assert open(tmp_path / "p.xyz").read() == open(tmp_path / "s.xyz").read()


def test_reference_sampling_equals_serial_per_point_loop(ico5):
    """The kept points of the batched sampler equal the reference's loop (one is_inside call per
    sample, in order, modules/winding/examples/sample_points_in_mesh.cpp:69-76) bit for bit."""
    V, F = ico5
    e = FastWindingNumber((V, F))
    n = 10000
    pts = EX.mt19937_box_samples(*e.bbox, n)
    serial = np.array([p for p in pts if e.is_inside(p)], np.float32).reshape(-1, 3)
    kept = EX.sample_points_in_mesh(e, n)
    np.testing.assert_array_equal(kept.view(np.uint32), serial.view(np.uint32))
    assert 0.3 < kept.shape[0] / n < 0.7
