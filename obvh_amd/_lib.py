"""ctypes binding of libobvh.so (the C ABI in include/obvh/obvh.h).

The library is the product path; there is no fallback.  If it is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.environ.get("OBVH_LIB_DIR") or HERE, "libobvh.so")

# every symbol the header declares (checked by tests/test_abi.py)
EXPORTS = [
    "kr_cuda_points_obb",
    "kr_points_obb",
    "kr_cuda_bvh_obb_tree",
    "obvh_dito_workspace_size",
    "obvh_dito_soa",
    "obvh_dito_soa_noalloc",
    "obvh_dito_batch_workspace_size",
    "obvh_dito_batch",
    "obvh_upload_sets",
    "obvh_gather_sets",
    "obvh_obb_tree_workspace_size",
    "obvh_obb_tree_async",
    "obvh_obb_tree_status_async",
    "obvh_obb_tree_status_string",
    "obvh_obb_tree_batch_workspace_size",
    "obvh_obb_tree_batch",
    "obvh_obb_tree_batch_status_async",
    "obvh_validate_tree_host",
    "obvh_lbvh_workspace_size",
    "obvh_lbvh_build",
    "obvh_atrbvh_workspace_size",
    "obvh_atrbvh_build",
    "kr_cuda_bvh_collapse",
    "obvh_collapse_workspace_size",
    "obvh_collapse_async",
    "kr_bvh_sah",
    "kr_cuda_obvh_persistent",
    "kr_cuda_soa_obvh",
    "kr_cuda_soa_obvh_intersect",
    "obvh_soa_obvh_free",
    "obvh_soa_pack_async",
    "obvh_soa_intersect_workspace_size",
    "obvh_soa_intersect_async",
    "obvh_records_pack_async",
    "obvh_records_intersect_async",
    "obvh_last_error",
    "obvh_version",
]


class KrVec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class KrObb3(C.Structure):
    _fields_ = [("mid", KrVec3), ("v0", KrVec3), ("v1", KrVec3), ("v2", KrVec3), ("ext", KrVec3)]


_lib = None


class ObvhError(RuntimeError):
    pass


def lib():
    """Load libobvh.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ObvhError(f"{LIB_PATH} is missing: build it with `python -m obvh_amd.build` "
                        "(or __graft_entry__.build()); there is no CPU fallback")
    # torch brings its own HIP runtime libraries: have them in the process
    # before libobvh.so pulls in the system ones, the order under which
    # device calls from both work in one process
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, sz, u32, u64 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint64
    L.kr_cuda_points_obb.argtypes = [vp, u64, C.POINTER(C.c_float)]
    L.kr_cuda_points_obb.restype = KrObb3
    L.kr_points_obb.argtypes = [vp, u64, C.POINTER(C.c_float)]
    L.kr_points_obb.restype = KrObb3
    L.kr_cuda_bvh_obb_tree.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float,
                                       u32, u32, u32]
    L.kr_cuda_bvh_obb_tree.restype = C.c_char_p
    L.obvh_dito_workspace_size.argtypes = [u64]
    L.obvh_dito_workspace_size.restype = sz
    L.obvh_dito_soa.argtypes = [vp, vp, vp, u64, vp, vp, vp, sz, vp]
    L.obvh_dito_soa.restype = C.c_char_p
    L.obvh_dito_soa_noalloc.argtypes = [vp, vp, vp, u64, vp, vp, vp]
    L.obvh_dito_soa_noalloc.restype = C.c_char_p
    L.obvh_dito_batch_workspace_size.argtypes = [u32, u64, u32]
    L.obvh_dito_batch_workspace_size.restype = sz
    L.obvh_dito_batch.argtypes = [vp, vp, vp, vp, u32, u64, vp, vp, u32, u32, vp, sz, vp]
    L.obvh_dito_batch.restype = C.c_char_p
    L.obvh_upload_sets.argtypes = [vp, vp, u32, vp, u64, vp]
    L.obvh_upload_sets.restype = C.c_char_p
    L.obvh_gather_sets.argtypes = [vp, vp, vp, vp, u32, vp, u64, vp]
    L.obvh_gather_sets.restype = C.c_char_p
    L.obvh_obb_tree_workspace_size.argtypes = [u32, u32]
    L.obvh_obb_tree_workspace_size.restype = sz
    L.obvh_obb_tree_async.argtypes = [vp, vp, vp, vp, vp, vp, vp, u32, u32, vp, sz, vp]
    L.obvh_obb_tree_async.restype = C.c_char_p
    L.obvh_obb_tree_status_async.argtypes = [vp, sz, u32, vp, vp]
    L.obvh_obb_tree_status_async.restype = C.c_char_p
    L.obvh_obb_tree_status_string.argtypes = [u32]
    L.obvh_obb_tree_status_string.restype = C.c_char_p
    L.obvh_obb_tree_batch_workspace_size.argtypes = [u32, u32]
    L.obvh_obb_tree_batch_workspace_size.restype = sz
    L.obvh_obb_tree_batch.argtypes = [vp, vp, vp, vp, vp, vp, u32, u32, vp, vp, vp, sz, vp]
    L.obvh_obb_tree_batch.restype = C.c_char_p
    L.obvh_obb_tree_batch_status_async.argtypes = [vp, sz, u32, vp, vp]
    L.obvh_obb_tree_batch_status_async.restype = C.c_char_p
    L.obvh_validate_tree_host.argtypes = [vp, vp, u32, u32]
    L.obvh_validate_tree_host.restype = C.c_char_p
    L.obvh_lbvh_workspace_size.argtypes = [u32]
    L.obvh_lbvh_workspace_size.restype = sz
    L.obvh_lbvh_build.argtypes = [vp, vp, u32, vp, vp, vp, vp, vp, vp, sz, vp]
    L.obvh_lbvh_build.restype = C.c_char_p
    L.obvh_atrbvh_workspace_size.argtypes = [u32]
    L.obvh_atrbvh_workspace_size.restype = sz
    L.obvh_atrbvh_build.argtypes = [vp, vp, u32, C.c_int, C.c_int, C.c_float, C.c_float, vp, vp, vp, vp, vp, sz,
                                    vp]
    L.obvh_atrbvh_build.restype = C.c_char_p
    L.kr_cuda_bvh_collapse.argtypes = [vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, u32, u32, vp]
    L.kr_cuda_bvh_collapse.restype = C.c_char_p
    L.obvh_collapse_workspace_size.argtypes = [u32]
    L.obvh_collapse_workspace_size.restype = sz
    L.obvh_collapse_async.argtypes = [vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, u32, vp, sz, vp]
    L.obvh_collapse_async.restype = C.c_char_p
    L.kr_bvh_sah.argtypes = [vp, C.c_int32, C.c_float, C.c_float, vp, C.c_char_p]
    L.kr_bvh_sah.restype = C.c_float
    L.kr_cuda_obvh_persistent.argtypes = [vp, vp, vp, vp, u32]
    L.kr_cuda_obvh_persistent.restype = C.c_char_p
    L.kr_cuda_soa_obvh.argtypes = [vp, vp, C.POINTER(C.c_void_p), u32, u32]
    L.kr_cuda_soa_obvh.restype = C.c_char_p
    L.kr_cuda_soa_obvh_intersect.argtypes = [vp, vp, vp, vp, vp, vp, u32, u32, u32, vp]
    L.kr_cuda_soa_obvh_intersect.restype = C.c_char_p
    L.obvh_soa_obvh_free.argtypes = [vp]
    L.obvh_soa_obvh_free.restype = C.c_char_p
    L.obvh_soa_pack_async.argtypes = [vp, vp, u32, u32, vp, vp, vp, vp]
    L.obvh_soa_pack_async.restype = C.c_char_p
    L.obvh_soa_intersect_workspace_size.argtypes = [u32]
    L.obvh_soa_intersect_workspace_size.restype = sz
    L.obvh_soa_intersect_async.argtypes = [vp, vp, vp, u32, u32, vp, u32, vp, vp, u32, vp, vp, u32, vp, sz, vp]
    L.obvh_soa_intersect_async.restype = C.c_char_p
    L.obvh_records_pack_async.argtypes = [vp, vp, vp, vp, u32, vp, u32, vp, vp, vp]
    L.obvh_records_pack_async.restype = C.c_char_p
    L.obvh_records_intersect_async.argtypes = [vp, vp, u32, vp, vp, u32, vp, sz, vp]
    L.obvh_records_intersect_async.restype = C.c_char_p
    L.obvh_last_error.argtypes = []
    L.obvh_last_error.restype = C.c_char_p
    L.obvh_version.argtypes = []
    L.obvh_version.restype = C.c_char_p
    _lib = L
    return L


def check(err) -> None:
    """Raise ObvhError if a kr_error (C string or None) reports a failure."""
    if err:
        raise ObvhError(err.decode() if isinstance(err, bytes) else str(err))
