"""OBJ / PLY reader and writer (libfwn_io.so through ctypes); CPU only.

Modelled on the reference's own I/O tests, which save a mesh and load it back
(modules/io/tests/test_obj.cpp "io/obj", "io/obj empty"; modules/io/tests/test_ply.cpp):
round trips must be exact (17 significant digits in text, doubles in binary).  Hand-written files
pin the parsing rules that a round trip cannot: 1-based and negative relative OBJ indices,
v/t/n corner forms, comments and skipped records, polygon fans, big-endian PLY with float
positions and extra properties / elements, and errors on bad input.
"""
import struct

import numpy as np
import pytest

from lagrange_amd import io as mio
from lagrange_amd import workloads as W


def _sphere():
    return W.icosphere(4)


@pytest.mark.parametrize("ext,fmt", [(".obj", mio.FORMAT_OBJ), (".ply", mio.FORMAT_PLY),
                                      (".ply", mio.FORMAT_PLY_BINARY)])
def test_round_trip_exact(tmp_path, ext, fmt):
    m = _sphere()
    V = m.vertices.astype(np.float64) * (1 + 1e-9 * np.pi)  # digits beyond float precision
    F = m.facets.astype(np.int32)
    p = tmp_path / f"m{fmt}{ext}"
    mio.save_mesh(p, V, F, fmt=fmt)
    got = mio.load_mesh(p)
    np.testing.assert_array_equal(got.vertices, V)
    np.testing.assert_array_equal(got.facets, F)
    assert got.is_triangle_mesh()


@pytest.mark.parametrize("ext", [".obj", ".ply"])
def test_empty_mesh_round_trip(tmp_path, ext):
    p = tmp_path / f"empty{ext}"
    mio.save_mesh(p, np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    got = mio.load_mesh(p)
    assert got.vertices.shape == (0, 3) and got.num_facets == 0


def test_polygons_round_trip_and_fan_triangulation(tmp_path):
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 1.5, 0], [2, 0, 0]], np.float64)
    off = np.array([0, 5, 8])
    cor = np.array([0, 1, 2, 4, 3, 1, 5, 2])
    for ext in (".obj", ".ply"):
        p = tmp_path / f"poly{ext}"
        mio.save_mesh(p, V, offsets=off, corners=cor)
        got = mio.load_mesh(p)
        np.testing.assert_array_equal(got.offsets, off)
        np.testing.assert_array_equal(got.corners, cor)
        assert not got.is_triangle_mesh()
        tri = mio.load_mesh(p, triangulate=True)
        np.testing.assert_array_equal(tri.facets, [[0, 1, 2], [0, 2, 4], [0, 4, 3], [1, 5, 2]])


def test_obj_parsing_rules():
    txt = b"""# comment
mtllib x.mtl
o thing
v 0 0 0
v 1.5 0 0 1.0
vt 0.5 0.5
vn 0 0 1
v +0 1 -0
g group
usemtl m
s off
f 1/1/1 2/1/1 3/1/1
v 1e0 1 0
f -4//1 -2//1 -1//1   # trailing comment
l 1 2
f 1/1 2/1 4/1
"""
    m = mio.load_mesh_buffer(txt, mio.FORMAT_OBJ)
    np.testing.assert_array_equal(m.vertices, [[0, 0, 0], [1.5, 0, 0], [0, 1, 0], [1, 1, 0]])
    np.testing.assert_array_equal(m.facets, [[0, 1, 2], [0, 2, 3], [0, 1, 3]])


@pytest.mark.parametrize("bad", [b"v 0 0 0\nv 1 0 0\nf 1 2 3\n", b"v 0 0 0\nf 0 1 1\n", b"v 0 0 x\n",
                                 b"v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1 2\n", b"v 0 0 0\nf -2 1 1\n"])
def test_obj_errors(bad):
    with pytest.raises(mio.MeshIOError) as e:
        mio.load_mesh_buffer(bad, mio.FORMAT_OBJ)
    assert e.value.status == 3


def test_ply_big_endian_float_with_extra_properties():
    hdr = (b"ply\nformat binary_big_endian 1.0\ncomment made by hand\n"
           b"element vertex 4\nproperty float x\nproperty uchar red\nproperty float y\nproperty float z\n"
           b"element face 2\nproperty uchar flags\nproperty list uchar uint vertex_indices\n"
           b"element edge 1\nproperty int vertex1\nproperty int vertex2\nend_header\n")
    body = b""
    pts = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.25, 0.5, 2.0)]
    for x, y, z in pts:
        body += struct.pack(">fBff", x, 7, y, z)
    body += struct.pack(">BB3I", 1, 3, 0, 1, 2)
    body += struct.pack(">BB4I", 0, 4, 0, 1, 3, 2)
    body += struct.pack(">ii", 0, 1)
    m = mio.load_mesh_buffer(hdr + body, mio.FORMAT_PLY)
    np.testing.assert_array_equal(m.vertices, np.array(pts))
    np.testing.assert_array_equal(m.offsets, [0, 3, 7])
    np.testing.assert_array_equal(m.corners, [0, 1, 2, 0, 1, 3, 2])


def test_ply_ascii_vertex_index_name():
    txt = (b"ply\nformat ascii 1.0\nelement vertex 3\nproperty double x\nproperty double y\n"
           b"property double z\nelement face 1\nproperty list uchar int vertex_index\nend_header\n"
           b"0 0 0\n1 0 0\n0 1 0\n3 2 1 0\n")
    m = mio.load_mesh_buffer(txt, mio.FORMAT_PLY)
    np.testing.assert_array_equal(m.facets, [[2, 1, 0]])


@pytest.mark.parametrize("bad", [
    b"plx\nformat ascii 1.0\nend_header\n",
    b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nend_header\n0 0\n",
    b"ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\nproperty float z\n"
    b"element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0 0\n1 1 1\n3 0 1 2\n",
    b"ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
    b"end_header\n0 0 0\n1 1 1\n",
])
def test_ply_errors(bad):
    with pytest.raises(mio.MeshIOError):
        mio.load_mesh_buffer(bad, mio.FORMAT_PLY)


def test_file_errors(tmp_path):
    with pytest.raises(mio.MeshIOError) as e:
        mio.load_mesh(tmp_path / "missing.obj")
    assert e.value.status == 2
    with pytest.raises(mio.MeshIOError) as e:
        mio.load_mesh(tmp_path / "mesh.stl")
    assert e.value.status == 1
    with pytest.raises(mio.MeshIOError):
        mio.save_mesh(tmp_path / "bad.obj", np.zeros((2, 3)), np.array([[0, 1, 2]]))


def test_loaded_mesh_feeds_engine_arrays(tmp_path):
    """The loaded arrays are what the engine takes: float positions, int32 triangles."""
    m = W.displaced_sphere(8, 1, delete_fraction=0.05)
    p = tmp_path / "d.ply"
    mio.save_mesh(p, m.vertices, m.facets)
    got = mio.load_mesh(p)
    np.testing.assert_array_equal(got.vertices, m.vertices.astype(np.float64))
    np.testing.assert_array_equal(got.facets, m.facets)
    assert got.facets.dtype == np.int32
