"""ctypes binding of lib/libn2wos_amd.so (the C ABI declared in include/n2wos_amd.h).

The library is the product: there is no fallback.  If it has not been built,
loading fails with an error that says how to build it.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libn2wos_amd.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "n2wos_amd.h")

OK = 0
ERR_INVALID = -1
ERR_NULL = -2
ERR_BLOCK = -3
ERR_HIP = -4
ERR_EMPTY = -5
ERR_NOMEM = -6
ERR_UNSUPPORTED = -7
ERR_INTERNAL = -8

ERROR_NAMES = {
    ERR_INVALID: "N2WOS_ERR_INVALID", ERR_NULL: "N2WOS_ERR_NULL", ERR_BLOCK: "N2WOS_ERR_BLOCK",
    ERR_HIP: "N2WOS_ERR_HIP", ERR_EMPTY: "N2WOS_ERR_EMPTY", ERR_NOMEM: "N2WOS_ERR_NOMEM",
    ERR_UNSUPPORTED: "N2WOS_ERR_UNSUPPORTED", ERR_INTERNAL: "N2WOS_ERR_INTERNAL",
}


class N2wosError(RuntimeError):
    """A non-zero return code of the C ABI, with n2wos_last_error()'s text."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


class BvhStats(C.Structure):
    _fields_ = [
        ("triangle_count", C.c_int64), ("node_count", C.c_int64), ("index_count", C.c_int64),
        ("binary_nodes", C.c_int64), ("leaf_size", C.c_int), ("max_depth", C.c_int), ("depth4", C.c_int),
        ("prune_margin", C.c_float), ("build_milliseconds", C.c_double), ("device_bytes", C.c_int64),
        ("bounds_lo", C.c_float * 3), ("bounds_hi", C.c_float * 3), ("build_method", C.c_char * 32),
    ]


_lib = None

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f32 = C.c_float

_SIGNATURES = {
    "n2wos_last_error": ([], C.c_char_p),
    "n2wos_abi_version": ([], i32),
    "n2wos_set_tuning": ([C.c_char_p, i32], i32),
    "n2wos_device_summary": ([C.c_char_p, i32], i32),
    "n2wos_uv_sphere_counts": ([i32, i32, C.POINTER(i32), C.POINTER(i32)], i32),
    "n2wos_uv_sphere": ([i32, i32, i32, vp, i32, vp, vp], i32),
    "n2wos_normalize_to_unit_radius": ([vp, i32, vp, vp], i32),
    "n2wos_count_degenerate_triangles": ([vp, i32, vp, i32, f32], i64),
    "n2wos_mesh_load": ([C.c_char_p, C.c_char_p, C.POINTER(vp), C.POINTER(i32), C.POINTER(vp), C.POINTER(i32)], i32),
    "n2wos_mesh_free": ([vp, vp], i32),
    "n2wos_closest_point_bruteforce": ([vp, i32, vp, i32, vp, i32, vp, vp, vp, i32], i32),
    "n2wos_bvh_create": ([vp, i32, vp, i32, i32, C.c_char_p, C.POINTER(vp)], i32),
    "n2wos_bvh_destroy": ([vp], i32),
    "n2wos_bvh_get_stats": ([vp, C.POINTER(BvhStats)], i32),
    "n2wos_bvh_query_device": ([vp, vp, i32, vp, vp, vp, vp, i32, vp], i32),
    "n2wos_bvh_query_host": ([vp, vp, i32, vp, vp, vp, C.POINTER(i32), C.POINTER(f32), i32], i32),
    "n2wos_bvh_query_device_masked": ([vp, vp, vp, i32, vp, vp, vp, vp, i32, vp], i32),
    "n2wos_bvh_query_device_counted": ([vp, vp, i32, vp, vp, vp, vp, vp, vp, i32, vp], i32),
    "n2wos_bvh_build_host": ([vp, i32, vp, i32, i32, C.c_char_p, vp, C.POINTER(i64), vp, C.POINTER(f32)], i32),
    "n2wos_bvh_export": ([vp, vp, vp], i32),
    "n2wos_bvh_inside_device": ([vp, vp, i32, vp, vp], i32),
    "n2wos_wos_workspace_size": ([i32, i32], i64),
    "n2wos_wos_full_grid_blocks": ([i32, vp], i32),
    "n2wos_cache_create": ([vp, C.POINTER(vp)], i32),
    "n2wos_cache_destroy": ([vp], i32),
    "n2wos_cache_infer": ([vp, vp, i32, vp, vp], i32),
    "n2wos_cache_train_step": ([vp, vp, vp, i32, vp, vp], i32),
    "n2wos_cache_layout": ([vp, vp, vp, vp, vp, vp, vp, vp], i32),
    "n2wos_cache_set_output_offset": ([vp, f32], i32),
    "n2wos_cache_copy_params": ([vp, vp, vp], i32),
    "n2wos_cache_get_params": ([vp, vp, vp], i32),
    "n2wos_cache_set_params": ([vp, vp, vp], i32),
    "n2wos_cache_gradients": ([vp, vp, vp, i32, vp, vp, vp], i32),
    "n2wos_nc_workspace_size": ([i32, i32], i64),
    "n2wos_nc_walks": ([vp, vp, vp, vp, i32, i32, C.c_uint32, i32, i32, vp, vp, vp, vp, vp, vp, i64, vp], i32),
    "n2wos_nc_combine": ([i32, i32, i32, vp, vp, vp, vp, i32, vp, vp, vp, vp], i32),
    "n2wos_wos_solve": ([vp, vp, vp, vp, i32, i32, C.c_uint32, vp, vp, vp, vp, vp, i32, vp, i64, vp, vp, vp], i32),
    "n2wos_cv_solve": ([vp, vp, vp, vp, i32, i32, C.c_uint32, vp, vp, vp, vp, vp, vp, i32, vp, i64, vp, vp, vp, vp,
                        i32, vp], i32),
    "n2wos_wos_solve_noalloc": ([vp, vp, vp, vp, i32, i32, C.c_uint32, vp, vp, vp, vp, vp, i32, vp, vp, vp], i32),
}


def lib():
    """The loaded engine library (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc != OK:
        msg = lib().n2wos_last_error()
        raise N2wosError(rc, msg.decode() if msg else "")
    return rc


def header_functions() -> list[str]:
    """Names of every function declared in include/n2wos_amd.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(n2wos_[a-z0-9_]+)\s*\(", text)))
