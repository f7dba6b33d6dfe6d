"""Build recipe for the engine's shared library (hipcc, gfx950, in-tree).

``python -m lagrange_amd.build`` compiles every HIP source under ``csrc/`` into
``lagrange_amd/libfwn.so``.  Packed FP32 instructions are disabled for all device code
(``-target-feature -packed-fp32-ops``).
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfwn.so")
ARCH = os.environ.get("FWN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


# per-source additions: the surrogate's MFMA kernels keep their accumulators in VGPRs (the
# fused MLP's bias/ReLU/BF16 epilogue then reads them without 512 AGPR-to-VGPR moves per tile)
EXTRA_FLAGS = {
    "fwn_surrogate.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile each source to an object in parallel, then link libfwn.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc()] + HIP_FLAGS + EXTRA_FLAGS.get(os.path.basename(src), []) + inc + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


def _drain(procs):
    errs = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errs.append(f"{src}:\n{out}")
        elif out and "warning" in out and "packed-fp32-ops" not in out:
            sys.stderr.write(out)
    procs.clear()
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))


CXX_DIR = os.path.join(HERE, "cpp")
CXX_LIB = os.path.join(HERE, "liblagrange_winding.so")
IO_LIB = os.path.join(HERE, "libfwn_io.so")


def build_io(verbose: bool = False) -> str:
    """The OBJ / PLY reader and writer behind include/fwn_io.h (host code, g++)."""
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall",
           "-I", os.path.join(ROOT, "include"), "-o", IO_LIB + ".tmp",
           os.path.join(CXX_DIR, "src", "mesh_io.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout)
    os.replace(IO_LIB + ".tmp", IO_LIB)
    return IO_LIB


CXX_TEST = os.path.join(CXX_DIR, "build", "test_fast_winding_number")
CXX_EXAMPLES = ("fix_orientation", "sample_points_in_mesh")
CXX_BENCH = os.path.join(CXX_DIR, "build", "bench_point_calls")


def build_cxx(verbose: bool = False) -> str:
    """The lagrange::winding::FastWindingNumber drop-in (g++, links libfwn.so) and its test runner.

    Uses the compat Lagrange headers under cpp/compat; a build against real Lagrange puts its
    include directory first instead."""
    inc = ["-I", os.path.join(CXX_DIR, "include"), "-I", os.path.join(CXX_DIR, "compat"),
           "-I", os.path.join(ROOT, "include")]
    link = ["-L", HERE, "-lfwn", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    # compat/src/assert.cpp stands in for lagrange::core's assertion handler (a real Lagrange build
    # links lagrange::core instead)
    cmds = [
        ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-o", CXX_LIB,
         os.path.join(CXX_DIR, "src", "FastWindingNumber.cpp"),
         os.path.join(CXX_DIR, "compat", "src", "assert.cpp")] + inc + link,
    ]
    os.makedirs(os.path.dirname(CXX_TEST), exist_ok=True)
    cmds.append(["g++", "-std=c++17", "-O2", "-Wall", "-pthread", "-o", CXX_TEST,
                 os.path.join(CXX_DIR, "tests", "test_fast_winding_number.cpp")] + inc +
                ["-L", HERE, "-llagrange_winding", "-lfwn", "-Wl,-rpath," + HERE,
                 "-Wl,-rpath-link,/opt/rocm/lib"])
    # the reference's two winding examples (need libfwn_io.so for mesh input / output)
    if not os.path.exists(IO_LIB):
        build_io(verbose)
    for name in CXX_EXAMPLES:
        cmds.append(["g++", "-std=c++17", "-O2", "-Wall", "-o", os.path.join(CXX_DIR, "build", name),
                     os.path.join(CXX_DIR, "examples", name + ".cpp")] + inc +
                    ["-L", HERE, "-llagrange_winding", "-lfwn", "-lfwn_io", "-Wl,-rpath," + HERE,
                     "-Wl,-rpath-link,/opt/rocm/lib"])
    # the reference's per-point benchmark loop over the drop-in class
    cmds.append(["g++", "-std=c++17", "-O2", "-Wall", "-pthread", "-o", CXX_BENCH,
                 os.path.join(CXX_DIR, "bench", "bench_point_calls.cpp")] + inc +
                ["-L", HERE, "-llagrange_winding", "-lfwn", "-lfwn_io", "-Wl,-rpath," + HERE,
                 "-Wl,-rpath-link,/opt/rocm/lib"])
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("g++ failed:\n" + r.stdout)
    return CXX_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_io(verbose="-v" in sys.argv))
    print(build_cxx(verbose="-v" in sys.argv))
