"""CPU checks of the C++ drop-in's header layout (no GPU).

The wrapper lagrange_amd/cpp/src/FastWindingNumber.cpp replaces
modules/winding/src/FastWindingNumber.cpp inside a real Lagrange build, so it must include the
headers that declare what it uses, as the reference source does (:13-18): the real
<lagrange/SurfaceMesh.h> declares neither lagrange::Error nor the matrix views
(modules/core/include/lagrange/SurfaceMesh.h:14-18).  The compat headers used here mirror the
real paths and declare only what their real counterparts do, so a missing include fails here as
it would in a Lagrange build.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "lagrange_amd", "cpp")
COMPAT = os.path.join(CPP, "compat")
WRAPPER = os.path.join(CPP, "src", "FastWindingNumber.cpp")
REF_CORE = "/root/reference/modules/core/include"

# compat header -> names it declares (each also declared by the real header of the same path)
LAYOUT = {
    "lagrange/SurfaceMesh.h": ["class SurfaceMesh"],
    "lagrange/SurfaceMeshTypes.h": ["LA_SURFACE_MESH_X"],
    "lagrange/views.h": ["vertex_view", "facet_view"],
    "lagrange/utils/Error.h": ["Error : public std::runtime_error"],
    "lagrange/utils/assert.h": ["la_runtime_assert"],
    "lagrange/utils/value_ptr.h": ["class value_ptr", "make_value_ptr"],
}


def _compile(src_text, tmp_path, name="t.cpp"):
    f = tmp_path / name
    f.write_text(src_text)
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(CPP, "include"), "-I", COMPAT,
           "-I", os.path.join(ROOT, "include"), str(f)]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_compat_headers_mirror_the_real_paths():
    for rel, names in LAYOUT.items():
        txt = open(os.path.join(COMPAT, rel)).read()
        for n in names:
            assert n in txt, (rel, n)
    if not os.path.isdir(REF_CORE):
        pytest.skip("reference checkout absent: path correspondence checked on the build host only")
    for rel, names in LAYOUT.items():
        ref = os.path.join(REF_CORE, rel)
        assert os.path.exists(ref), rel
        txt = open(ref).read()
        for n in names:
            key = n.split(":")[0].strip().split()[-1]
            assert key in txt, (rel, key)


def test_compat_surface_mesh_declares_no_error_and_no_views(tmp_path):
    for use in ("lagrange::Error e(\"x\");",
                "lagrange::SurfaceMesh<float, uint32_t> m; auto v = lagrange::vertex_view(m);"):
        r = _compile("#include <lagrange/SurfaceMesh.h>\nvoid f() { " + use + " }\n", tmp_path)
        assert r.returncode != 0, use
    # the assertion header does not declare Error either (its handler throws it out of line)
    r = _compile("#include <lagrange/utils/assert.h>\nvoid f() { lagrange::Error e(\"x\"); }\n", tmp_path)
    assert r.returncode != 0
    r = _compile("#include <lagrange/utils/assert.h>\nvoid f(int x) { la_runtime_assert(x == 3); "
                 "la_runtime_assert(x > 0, \"positive\"); }\n", tmp_path)
    assert r.returncode == 0, r.stderr


def test_wrapper_compiles_and_needs_each_of_its_includes(tmp_path):
    src = open(WRAPPER).read()
    r = _compile(src, tmp_path, "FastWindingNumber.cpp")
    assert r.returncode == 0, r.stderr
    for inc in ("<lagrange/SurfaceMeshTypes.h>", "<lagrange/utils/Error.h>", "<lagrange/utils/assert.h>",
                "<lagrange/views.h>"):
        line = f"#include {inc}\n"
        assert line in src, inc
        r = _compile(src.replace(line, ""), tmp_path, "FastWindingNumber.cpp")
        assert r.returncode != 0, f"{inc} is not needed by the wrapper"
