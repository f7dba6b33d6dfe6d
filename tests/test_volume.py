"""Host logic of the winding-number voxel signing (lagrange_amd.volume), no GPU needed.

Pins against the reference's mesh_to_volume (modules/volume/src/mesh_to_volume.cpp):
  * a negative voxel size is |voxel_size| x bbox diagonal (:132-145);
  * voxel (i, j, k) has centre voxel_size * (i, j, k) + voxel_size / 2, in double, then rounded
    to float (:147-149, :176-181); with a power-of-two voxel size the centres are exact;
  * the index box covers the bounding box grown by the narrow band on every side.
"""
import numpy as np
import pytest

from lagrange_amd import volume as VOL


def test_relative_voxel_size_uses_bbox_diagonal():
    V = np.array([[0, 0, 0], [3, 4, 12]], np.float64)  # diagonal 13
    assert VOL.bbox_diagonal(V) == 13.0
    assert VOL.resolve_voxel_size(V, -0.01) == pytest.approx(0.13, rel=0, abs=1e-15)
    assert VOL.resolve_voxel_size(V, 0.25) == 0.25
    with pytest.raises(ValueError):
        VOL.resolve_voxel_size(V, 0.0)


def test_diagonal_in_float_for_float_meshes():
    V = np.array([[0, 0, 0], [1, 1, 1]], np.float32)
    # sqrt(3) rounded to float, not the double value
    assert VOL.bbox_diagonal(V) == float(np.float32(np.sqrt(np.float32(3.0))))
    assert VOL.bbox_diagonal(V) != float(np.sqrt(3.0))


def test_voxel_centres_exact_for_power_of_two_size():
    ijk = np.array([[0, 0, 0], [1, -1, 7], [-1000, 5, 123456]], np.int32)
    c = VOL.voxel_centres(ijk, 0.25)
    np.testing.assert_array_equal(c, (ijk + 0.5) * 0.25)
    # voxel 0 is centred at half a voxel, voxel -1 at minus half a voxel
    assert c[0, 0] == 0.125 and c[1, 1] == -0.125


def test_voxel_centres_round_double_to_float():
    vs = 0.1
    c = VOL.voxel_centres(np.array([[3, 10, -7]]), vs)
    assert c.dtype == np.float32
    want = [np.float32(3 * vs + vs / 2), np.float32(10 * vs + vs / 2), np.float32(-7 * vs + vs / 2)]
    np.testing.assert_array_equal(c[0], want)
    # the float centre is the nearest float to the double one
    assert abs(float(c[0, 0]) - 0.35) <= np.spacing(np.float32(0.35)) / 2


@pytest.mark.parametrize("vs,band", [(0.1, 3), (0.037, 0), (1.0, 2)])
def test_index_box_covers_bbox_plus_band(vs, band):
    rng = np.random.default_rng(3)
    V = rng.uniform(-1.3, 2.1, size=(50, 3))
    lo, dims = VOL.index_box(V, vs, band)
    assert lo.dtype == np.int32 and (dims > 0).all()
    first = (lo + band).astype(np.float64) * vs + vs / 2
    last = (lo + dims - 1 - band).astype(np.float64) * vs + vs / 2
    assert (first <= V.min(0) + 1e-12).all() and (last >= V.max(0) - 1e-12).all()
    # and it is tight: one voxel less on either side no longer covers
    assert ((first + vs) > V.min(0)).all() and ((last - vs) < V.max(0)).all()
