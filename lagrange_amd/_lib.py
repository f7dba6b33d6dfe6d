"""ctypes binding of the engine's C ABI (include/fwn.h).

Loading fails loudly (RuntimeError) when ``libfwn.so`` has not been built: there is no CPU
fallback anywhere in the package.  torch is imported first when it is installed so that the
library binds to the same HIP runtime instance as torch (both use soname libamdhip64.so.7),
which lets torch device tensors be passed with FWN_DEVICE_POINTERS.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfwn.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "fwn.h")

FWN_OK = 0
FWN_INVALID_ARGUMENT = 1
FWN_OUT_OF_MEMORY = 2
FWN_NO_DEVICE = 3
FWN_HIP_ERROR = 4
FWN_LIMIT_EXCEEDED = 5
FWN_DEVICE_POINTERS = 1
FWN_LAUNCH_GRID = 0
FWN_LAUNCH_PERSISTENT = 1
FWN_TREE_LBVH = 0
FWN_TREE_SAH = 1
FWN_MOMENTS_LEVELS = 0
FWN_MOMENTS_CLIMB = 1

MAX_WIDTH = 8
SLOT_WORDS = 24
TRI_WORDS = 12

_STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "OUT_OF_MEMORY", 3: "NO_DEVICE", 4: "HIP_ERROR",
                 5: "LIMIT_EXCEEDED"}


class Options(C.Structure):
    _fields_ = [
        ("accuracy_scale", C.c_float),
        ("near_fp64_rel", C.c_float),
        ("leaf_max", C.c_int32),
        ("fp64_moment_min", C.c_int32),
        ("device", C.c_int32),
        ("waves_per_block", C.c_int32),
        ("tree_width", C.c_int32),
        ("launch_mode", C.c_int32),
        ("tree_builder", C.c_int32),
        ("moment_pass", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


class TreeInfo(C.Structure):
    _fields_ = [
        ("num_meshes", C.c_int64),
        ("num_nodes", C.c_int64),
        ("num_triangles", C.c_int64),
        ("width", C.c_int32),
        ("max_depth", C.c_int32),
        ("build_ms", C.c_double),
    ]


class FWNError(RuntimeError):
    """Raised on a non-OK status; mirrors lagrange::Error (a std::runtime_error)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{_STATUS_NAMES.get(status, status)}] {message}")
        self.status = status


_lib = None

VP = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32


def declared_functions():
    """Names of the functions include/fwn.h declares (parsed from the header text)."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fwn_[a-z0-9_]+)\s*\(", txt)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lagrange_amd.build` "
                           "(there is no CPU fallback)")
    try:  # bind to torch's HIP runtime when torch is present
        import torch  # noqa: F401
    except Exception:
        pass
    L = C.CDLL(LIB_PATH)
    sig = {
        "fwn_options_default": ([C.POINTER(Options)], I32),
        "fwn_create": ([C.POINTER(VP), VP, I64, VP, I64, C.POINTER(Options)], I32),
        "fwn_create_multi": ([C.POINTER(VP), I32, VP, VP, VP, VP, C.POINTER(Options)], I32),
        "fwn_destroy": ([VP], I32),
        "fwn_solid_angle": ([VP, VP, I64, VP, VP, I32], I32),
        "fwn_solid_angle_noalloc": ([VP, VP, I64, VP, VP], I32),
        "fwn_solid_angle_multi": ([VP, VP, VP, VP, VP, I32], I32),
        "fwn_solid_angle_point": ([VP, VP, VP], I32),
        "fwn_is_inside_point": ([VP, VP, VP], I32),
        "fwn_sign_voxel_box": ([VP, VP, VP, C.c_double, VP, VP, I32], I32),
        "fwn_sign_voxels": ([VP, VP, I64, C.c_double, VP, VP, I32], I32),
        "fwn_orientation_criterion": ([VP, VP, I64, VP, I64, C.c_float, VP, I32], I32),
        "fwn_sample_inside": ([VP, VP, VP, I64, C.c_uint64, VP, VP, I32], I32),
        "fwn_closest_point": ([VP, VP, I64, VP, VP, VP, I32], I32),
        "fwn_tree_info_get": ([VP, C.POINTER(TreeInfo)], I32),
        "fwn_export_tree": ([VP, VP, VP, VP, VP], I32),
        "fwn_query_stats": ([VP, VP, VP, I64, VP, VP, I32], I32),
        "fwn_bruteforce_f64": ([VP, VP, I64, VP, I32], I32),
        "fwn_last_timing": ([VP, VP, VP], I32),
        "fwn_create_tuned": ([C.POINTER(VP), I32, VP, VP, VP, VP, C.POINTER(Options), VP, VP, I64, I32, VP, I32, VP,
                              VP], I32),
        "fwn_set_launch": ([VP, I32, I32], I32),
        "fwn_autotune": ([VP, VP, VP, I64, I32, C.c_char_p, VP, VP, VP], I32),
        "fwn_last_error": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(status: int):
    if status != FWN_OK:
        msg = load().fwn_last_error()
        raise FWNError(status, msg.decode() if msg else "")


def default_options(**kw) -> Options:
    o = Options()
    check(load().fwn_options_default(C.byref(o)))
    for k, v in kw.items():
        if v is not None:
            setattr(o, k, v)
    return o
