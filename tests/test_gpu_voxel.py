"""GPU tests of the batched voxel signs (fwn_sign_voxel_box / fwn_sign_voxels) through the C ABI.

mesh_to_volume's winding-number interior test (modules/volume/src/mesh_to_volume.cpp:175-183)
is is_inside(float(voxel_size * ijk + voxel_size / 2)) per voxel.  Checks:
  * bit-exact: inside flags and solid angles equal the engine's own batched query at the same
    centres computed on the host (so the device-generated centres are the same floats);
  * the CPU restatement of the reference engine (oracle) gives the same inside flags outside the
    error band |w - 0.5| <= 4e-3 of the order-2 expansion at beta = 2;
  * a closed sphere: voxels well inside are inside, well outside are outside;
  * argument errors (non-positive voxel size, negative dims, multi-mesh engine) raise.
"""
import numpy as np
import pytest

from lagrange_amd import _lib as L
from lagrange_amd import volume as VOL
from lagrange_amd import workloads as W
from lagrange_amd.winding import FastWindingNumber, MultiFastWindingNumber

pytestmark = pytest.mark.gpu


def _box_ijk(lo, dims):
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    return np.stack([x.ravel() + lo[0], y.ravel() + lo[1], z.ravel() + lo[2]], axis=1).astype(np.int32)


@pytest.mark.parametrize("rel", [-0.02, -0.013])
def test_box_matches_batched_query_bit_exact(open_mesh, rel):
    V, F = open_mesh
    e = FastWindingNumber((V, F))
    vs = VOL.resolve_voxel_size(V, rel)
    lo, dims = VOL.index_box(V, vs)
    print(f"box lo {lo} dims {dims} voxel {vs}", flush=True)
    ins, om = e.sign_voxel_box(lo, dims, vs, with_omega=True)
    assert ins.shape == (dims[2], dims[1], dims[0])
    c = VOL.voxel_centres(_box_ijk(lo, dims), vs)
    om_ref, ins_ref = e.solid_angle_batch(c, with_inside=True)
    np.testing.assert_array_equal(om.ravel(), om_ref)
    np.testing.assert_array_equal(ins.ravel(), ins_ref)
    assert 0 < ins.sum() < ins.size


def test_list_host_and_device_match_box(open_mesh):
    import torch
    V, F = open_mesh
    e = FastWindingNumber((V, F))
    vs = VOL.resolve_voxel_size(V, -0.015)
    lo, dims = VOL.index_box(V, vs)
    box = e.sign_voxel_box(lo, dims, vs).ravel()
    ijk = _box_ijk(lo, dims)
    rng = np.random.default_rng(11)
    pick = rng.choice(ijk.shape[0], size=5000, replace=False)
    got = e.sign_voxels(ijk[pick], vs)
    np.testing.assert_array_equal(got, box[pick])
    got_d, om_d = e.sign_voxels(torch.from_numpy(ijk[pick]).cuda(), vs, with_omega=True)
    np.testing.assert_array_equal(got_d.cpu().numpy(), box[pick])
    om_ref = e.solid_angle_batch(VOL.voxel_centres(ijk[pick], vs))
    np.testing.assert_array_equal(om_d.cpu().numpy(), om_ref)
    # device outputs for the box too
    box_d = e.sign_voxel_box(lo, dims, vs, device=torch.device("cuda:0"))
    np.testing.assert_array_equal(box_d.cpu().numpy().ravel(), box)


def test_inside_flags_match_restated_reference(oracle, open_mesh):
    V, F = open_mesh
    e = FastWindingNumber((V, F))
    vs = VOL.resolve_voxel_size(V, -0.03)
    lo, dims = VOL.index_box(V, vs, band=1)
    ins, om = e.sign_voxel_box(lo, dims, vs, with_omega=True)
    c = VOL.voxel_centres(_box_ijk(lo, dims), vs)
    R = oracle.Restated(V, F)
    om_r = R.solid_angle(c)
    w_r = om_r.astype(np.float64) / (4 * np.pi)
    amb = np.abs(w_r - 0.5) <= 4e-3
    mism = (ins.ravel() != (w_r > 0.5)) & ~amb
    assert mism.sum() == 0
    assert np.abs(om.ravel().astype(np.float64) - om_r).max() / (4 * np.pi) <= 4e-3


def test_closed_sphere_interior_and_exterior(ico5):
    V, F = ico5
    e = FastWindingNumber((V, F))
    vs = 0.05
    lo, dims = VOL.index_box(V, vs)
    ins = e.sign_voxel_box(lo, dims, vs).ravel()
    c = VOL.voxel_centres(_box_ijk(lo, dims), vs).astype(np.float64)
    r = np.linalg.norm(c, axis=1)
    assert ins[r < 0.95].all()
    assert not ins[r > 1.05].any()
    # winding_sign_volume: the same box and flags through the helper
    lo2, vs2, ins2 = VOL.winding_sign_volume(e, V, voxel_size=vs)
    np.testing.assert_array_equal(lo2, lo)
    assert vs2 == vs
    np.testing.assert_array_equal(ins2.ravel(), ins)


def test_empty_and_errors(ico5):
    V, F = ico5
    e = FastWindingNumber((V, F))
    assert e.sign_voxel_box([0, 0, 0], [0, 5, 5], 0.1).shape == (5, 5, 0)
    assert e.sign_voxels(np.zeros((0, 3), np.int32), 0.1).shape == (0,)
    for bad in (0.0, -0.1, float("nan"), float("inf")):
        with pytest.raises(L.FWNError) as ex:
            e.sign_voxel_box([0, 0, 0], [2, 2, 2], bad)
        assert ex.value.status == L.FWN_INVALID_ARGUMENT
    with pytest.raises(L.FWNError):
        e.sign_voxel_box([2147483640, 0, 0], [100, 1, 1], 0.1)
    multi = MultiFastWindingNumber([(V, F), (V, F)])
    lo3 = np.zeros(3, np.int32)
    dims3 = np.full(3, 2, np.int32)
    out = np.empty(8, np.uint8)
    st = L.load().fwn_sign_voxel_box(multi._h(), lo3.ctypes.data, dims3.ctypes.data, 0.1, out.ctypes.data, None, 0)
    assert st == L.FWN_INVALID_ARGUMENT
