"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker -- never by the product
path (obvh_amd), which must fail loudly when the HIP library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from obvh_amd.layouts import ISECT_DTYPE, NODE_DTYPE, RAY_DTYPE, TRACE_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
# OBVH_ORACLE_LIB: a differently built copy (the sanitizer run's)
_LIB_PATH = os.environ.get("OBVH_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no FMA contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.oracle_dito14.argtypes = [vp, C.c_int, vp, vp]
        L.oracle_points_obb.argtypes = [vp, C.c_int, vp, vp]
        L.oracle_points_obb_batch.argtypes = [vp, vp, C.c_int, vp, vp]
        L.oracle_points_obb_batch_mt.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int]
        L.oracle_obb_tree.argtypes = [vp, vp, vp, vp, C.c_uint32, vp, vp, vp, vp, vp]
        L.oracle_obb_tree.restype = C.c_int
        L.oracle_obb_tree_mt.argtypes = [vp, vp, vp, vp, C.c_uint32, vp, vp, vp, vp, vp, C.c_int]
        L.oracle_obb_tree_mt.restype = C.c_int
        L.oracle_lbvh_build.argtypes = [vp, vp, C.c_uint32, vp, vp, vp, vp, vp]
        L.oracle_lbvh_build.restype = C.c_int
        L.oracle_cpu_obb_tree.argtypes = [vp, vp, vp, vp, C.c_uint32, vp, vp, C.c_uint32, C.c_uint32]
        L.oracle_cpu_obb_tree.restype = C.c_int64
        L.oracle_atrbvh_build.argtypes = [vp, vp, C.c_uint32, C.c_int, C.c_int, C.c_float, C.c_float,
                                          vp, vp, vp, vp]
        L.oracle_atrbvh_build.restype = C.c_int
        L.oracle_collapse.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_uint32]
        L.oracle_collapse.restype = C.c_int64
        L.oracle_bvh_sah.argtypes = [vp, C.c_int32, C.c_float, C.c_float, vp, C.c_uint32]
        L.oracle_bvh_sah.restype = C.c_float
        L.oracle_obvh_pack.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp]
        L.oracle_obvh_pack.restype = C.c_int
        L.oracle_obvh_intersect.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_uint32, vp, vp]
        L.oracle_obvh_intersect.restype = C.c_int
        L.oracle_obvh_intersect_mt.argtypes = [vp, vp, vp, vp, vp, vp, vp, C.c_uint32, vp, vp, C.c_int]
        L.oracle_obvh_intersect_mt.restype = C.c_int
        L.oracle_ray_unit_box.argtypes = [vp, vp, vp]
        L.oracle_ray_triangle.argtypes = [vp, vp, vp, vp, vp, C.c_float, vp]
        L.oracle_minverse4.argtypes = [vp, vp]
        L.oracle_mmul4.argtypes = [vp, vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def dito14(points, trace: bool = False):
    """DiTO_14<float> (no extent floor).  Returns obb (15,) [, trace]."""
    p = _f32(points).reshape(-1, 3)
    out = np.zeros(15, dtype=np.float32)
    tr = np.zeros(1, dtype=TRACE_DTYPE)
    lib().oracle_dito14(_p(p), int(p.shape[0]), _p(out), _p(tr))
    return (out, tr[0]) if trace else out


def points_obb(points, trace: bool = False):
    """kr_points_obb: DiTO_14 + 0.001 extent floor.  Returns obb (15,) [, trace]."""
    p = _f32(points).reshape(-1, 3)
    out = np.zeros(15, dtype=np.float32)
    tr = np.zeros(1, dtype=TRACE_DTYPE)
    lib().oracle_points_obb(_p(p), int(p.shape[0]), _p(out), _p(tr))
    return (out, tr[0]) if trace else out


def points_obb_batch(points, offsets, threads: int = 1):
    """kr_points_obb of every set s = points[offsets[s]:offsets[s+1]];
    threads > 1 fits sets in parallel on host threads (same outputs)."""
    p = _f32(points).reshape(-1, 3)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = off.size - 1
    out = np.zeros((n, 15), dtype=np.float32)
    tr = np.zeros(n, dtype=TRACE_DTYPE)
    lib().oracle_points_obb_batch_mt(_p(p), _p(off), n, _p(out), _p(tr), int(threads))
    return out, tr


def lbvh_build(vertices, faces):
    """Returns (nodes, parents, counts, costs, mortons_sorted)."""
    V = _f32(vertices).reshape(-1, 3)
    F = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 4)
    n = F.shape[0]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    parents = np.zeros(2 * n - 1, dtype=np.uint32)
    counts = np.zeros(2 * n - 1, dtype=np.uint32)
    costs = np.zeros(2 * n - 1, dtype=np.float32)
    mortons = np.zeros(n, dtype=np.uint32)
    rc = lib().oracle_lbvh_build(_p(V), _p(F), n, _p(nodes), _p(parents), _p(counts), _p(costs), _p(mortons))
    if rc != 0:
        raise RuntimeError("oracle_lbvh_build failed")
    return nodes, parents, counts, costs, mortons


def host_threads() -> int:
    """Host threads for an all-cores oracle run: the cores this process may
    use, at most 16 (a one-GPU share of a GPU host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def obb_tree(nodes, parents, vertices, faces, with_ids: bool = False, with_ranges: bool = False,
             with_axes: bool = False, threads: int = 1):
    """The transform method.  Returns obbs (2N-1, 15), mats (2N-1, 16)
    [, ids (2N-1, 14)] [, ranges (N-1, 6)] [, candidate axes (N-1, 11):
    b0, b1, b2, degenerate flag, winning candidate id (-1: AABB axes)].
    threads > 1 runs it node-parallel on host threads (same outputs)."""
    V = _f32(vertices).reshape(-1, 3)
    F = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 4)
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    parents = np.ascontiguousarray(parents, dtype=np.uint32)
    total = nodes.shape[0]
    N = (total + 1) // 2
    obbs = np.zeros((total, 15), dtype=np.float32)
    mats = np.zeros((total, 16), dtype=np.float32)
    ids = np.zeros((total, 14), dtype=np.uint32) if with_ids else None
    rr = np.zeros((max(N - 1, 1), 6), dtype=np.float32) if with_ranges else None
    ca = np.zeros((max(N - 1, 1), 11), dtype=np.float32) if with_axes else None
    rc = lib().oracle_obb_tree_mt(
        _p(nodes), _p(parents), _p(V), _p(F), N, _p(obbs), _p(mats),
        _p(ids) if ids is not None else None, _p(rr) if rr is not None else None,
        _p(ca) if ca is not None else None, int(threads))
    if rc != 0:
        raise ValueError("oracle_obb_tree: malformed tree")
    res = [obbs, mats]
    if with_ids:
        res.append(ids)
    if with_ranges:
        res.append(rr)
    if with_axes:
        res.append(ca)
    return tuple(res)


def cpu_obb_tree(nodes, parents, vertices, faces, leaf_begin: int = 0, leaf_end: int | None = None):
    """Reference CPU transform (full DiTO per node).  Returns (obbs, mats, nodes_out, fitted)."""
    V = _f32(vertices).reshape(-1, 3)
    F = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 4)
    nodes = np.array(nodes, dtype=NODE_DTYPE, copy=True)
    parents = np.ascontiguousarray(parents, dtype=np.uint32)
    total = nodes.shape[0]
    N = (total + 1) // 2
    obbs = np.zeros((total, 15), dtype=np.float32)
    mats = np.zeros((total, 16), dtype=np.float32)
    le = N if leaf_end is None else leaf_end
    fitted = lib().oracle_cpu_obb_tree(_p(nodes), _p(parents), _p(V), _p(F), N, _p(obbs), _p(mats), leaf_begin, le)
    return obbs, mats, nodes, int(fitted)


def atrbvh_build(vertices, faces, treelet_size: int = 7, iterations: int = 2, ci: float = 1.2, ct: float = 1.0):
    """ATRBVH.  Returns (nodes, parents, counts, costs) like lbvh_build (no mortons)."""
    V = _f32(vertices).reshape(-1, 3)
    F = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 4)
    n = F.shape[0]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    parents = np.zeros(2 * n - 1, dtype=np.uint32)
    counts = np.zeros(2 * n - 1, dtype=np.uint32)
    costs = np.zeros(2 * n - 1, dtype=np.float32)
    rc = lib().oracle_atrbvh_build(_p(V), _p(F), n, treelet_size, iterations, C.c_float(ci), C.c_float(ct),
                                   _p(nodes), _p(parents), _p(counts), _p(costs))
    if rc != 0:
        raise ValueError("oracle_atrbvh_build failed")
    return nodes, parents, counts, costs


def collapse(nodes, parents, counts, ci: float = 1.2, ct: float = 1.0):
    """kr_cuda_bvh_collapse.  nodes: NODE_DTYPE (2N-1), parents uint32,
    counts uint32 (leaf entries used).  Returns dict(collapsed, counts, costs,
    primitives, table, used)."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    total = nodes.shape[0]
    N = (total + 1) // 2
    par = np.ascontiguousarray(parents, dtype=np.uint32)
    cnt = np.ascontiguousarray(counts, dtype=np.uint32).copy()
    costs = np.zeros(total, dtype=np.float32)
    out = np.zeros(total, dtype=NODE_DTYPE)
    prims = np.full(N, 0xFFFFFFFF, dtype=np.uint32)
    table = np.zeros(total, dtype=np.uint8)
    used = lib().oracle_collapse(_p(nodes), _p(par), _p(cnt), _p(costs), _p(out), _p(prims), _p(table),
                                 C.c_float(ci), C.c_float(ct), N)
    if used < 0:
        raise ValueError("malformed tree")
    return {"collapsed": out, "counts": cnt, "costs": costs, "primitives": prims[:used], "table": table,
            "used": int(used)}


def bvh_sah(nodes, root: int = 0, ci: float = 1.2, ct: float = 1.0, obbs=None) -> float:
    """kr_bvh_sah (AABB areas, or OBB areas when obbs (2N-1, 15) is given)."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    ob = None if obbs is None else _f32(obbs)
    return float(lib().oracle_bvh_sah(_p(nodes), root, C.c_float(ci), C.c_float(ct),
                                      None if ob is None else _p(ob), nodes.shape[0]))


def minverse4(m):
    m = _f32(m).reshape(16)
    out = np.zeros(16, dtype=np.float32)
    lib().oracle_minverse4(_p(m), _p(out))
    return out


def mmul4(a, b):
    a = _f32(a).reshape(16)
    b = _f32(b).reshape(16)
    out = np.zeros(16, dtype=np.float32)
    lib().oracle_mmul4(_p(a), _p(b), _p(out))
    return out


def obvh_pack(nodes, mats):
    """obvh_persistent_prepare + kr_cuda_soa_obvh.  nodes: NODE_DTYPE (2N-1,
    plain or collapsed), mats (2N-1, 16) float32.  Returns (left_T (N-1, 12),
    right_T (N-1, 12), meta (2N-1, 4) int32)."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    total = nodes.shape[0]
    internal = (total - 1) // 2
    m = _f32(mats).reshape(total, 16)
    L = np.zeros((internal, 12), dtype=np.float32)
    R = np.zeros((internal, 12), dtype=np.float32)
    meta = np.zeros((total, 4), dtype=np.int32)
    if lib().oracle_obvh_pack(_p(nodes), _p(m), total, internal, _p(L), _p(R), _p(meta)) != 0:
        raise ValueError("malformed tree")
    return L, R, meta


def obvh_intersect(L, R, meta, vertices, faces, primitives, rays, with_stats: bool = False, threads: int = 1):
    """soa_obvh_persistent_intersect for every ray (RAY_DTYPE array).
    Returns ISECT_DTYPE results (and per-ray stats: nodes visited, triangles
    tested, deepest stack slot).  threads > 1 traces rays in parallel on host
    threads (same outputs)."""
    L, R = _f32(L), _f32(R)
    meta = np.ascontiguousarray(meta, dtype=np.int32)
    V = _f32(vertices)
    F = np.ascontiguousarray(faces, dtype=np.uint32)
    P = np.ascontiguousarray(primitives if primitives is not None else np.zeros(1), dtype=np.uint32)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = rays.shape[0]
    out = np.zeros(n, dtype=ISECT_DTYPE)
    st = np.zeros((n, 3), dtype=np.uint32)
    lib().oracle_obvh_intersect_mt(_p(L), _p(R), _p(meta), _p(V), _p(F), _p(P), _p(rays), n, _p(out), _p(st),
                                   int(threads))
    return (out, st) if with_stats else out


def ray_unit_box(rd, ro):
    out = np.zeros(2, dtype=np.float32)
    lib().oracle_ray_unit_box(_p(_f32(rd)), _p(_f32(ro)), _p(out))
    return out


def ray_triangle(direction, origin, v1, v2, v3, tmax):
    out = np.zeros(3, dtype=np.float32)
    lib().oracle_ray_triangle(_p(_f32(direction)), _p(_f32(origin)), _p(_f32(v1)), _p(_f32(v2)), _p(_f32(v3)),
                              C.c_float(tmax), _p(out))
    return out
