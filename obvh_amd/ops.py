"""Python entry points over the HIP library.

Mirrors the reference operator interface for the OBB path:

* ``points_obb(points)``            kr_cuda_points_obb (dito.cuh:12-13): host
                                    AoS points in, one OBB (15 floats) out.
* ``bvh_obb_tree(...)``             kr_cuda_bvh_obb_tree (transforms.cuh:52-65):
                                    device tree in, obbs / transforms out.
* ``dito(x, y, z)``, ``dito_batch`` device-resident, stream-ordered DiTO.
* ``lbvh_build(V, F)``              device LBVH (lbvh.cu:1547-1692).
* ``obb_tree(...)``                 stream-ordered transform with workspace.
* ``obvh_persistent`` / ``soa_obvh`` / ``soa_obvh_intersect``
                                    the packing and traversal drop-ins
                                    (transforms.cuh:32-38,72-77,
                                    intersectors.cuh:81-91).
* ``soa_pack``, ``soa_intersect``   one-pass packing and the stream-ordered
                                    traversal with workspace.
* ``records_pack``, ``records_intersect``
                                    the native cache-line record layout and
                                    its traversal (same hits).

Device memory and streams come from torch (plumbing only); every compute step
runs in libobvh.so.  Errors raise ObvhError; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import ObvhError, check, lib
from .layouts import ISECT_DTYPE, NODE_DTYPE, RAY_DTYPE, TRACE_DTYPE


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise ObvhError("no GPU visible: the OBB engine runs on the HIP device only")
    return torch


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream=None) -> C.c_void_p:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _on(stream):
    """Context that makes `stream` torch's current stream, so that tensors
    allocated inside belong to its caching-allocator pool: they are not
    handed to another stream while this stream's kernels may still use
    them."""
    torch = _torch()
    return torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream())


def _workspace(nbytes: int):
    torch = _torch()
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device="cuda")


# ---------------------------------------------------------------- DiTO --

def points_obb(points, return_time: bool = False):
    """kr_cuda_points_obb on a host (n, 3) float32 array -> (15,) float32."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    t = C.c_float(0.0)
    o = lib().kr_cuda_points_obb(p.ctypes.data_as(C.c_void_p), p.shape[0], C.byref(t))
    out = np.array([o.mid.x, o.mid.y, o.mid.z, o.v0.x, o.v0.y, o.v0.z, o.v1.x, o.v1.y, o.v1.z,
                    o.v2.x, o.v2.y, o.v2.z, o.ext.x, o.ext.y, o.ext.z], dtype=np.float32)
    if np.isnan(out).all() and p.shape[0] > 0:
        err = lib().obvh_last_error()
        raise ObvhError(err.decode() if err else "kr_cuda_points_obb failed")
    return (out, t.value) if return_time else out


def dito(x, y, z, trace: bool = False, stream=None, workspace=None):
    """DiTO-14 (+0.001 extent floor) of one SoA set of device float32 tensors."""
    torch = _torch()
    n = int(x.numel())
    if not (y.numel() == n and z.numel() == n):
        raise ValueError("x, y, z must have the same length")
    for a in (x, y, z):
        if a.dtype != torch.float32 or not a.is_cuda or not a.is_contiguous():
            raise ValueError("points must be contiguous float32 CUDA tensors")
    with _on(stream):
        out = torch.empty(15, dtype=torch.float32, device="cuda")
        tr = torch.zeros(TRACE_DTYPE.itemsize, dtype=torch.uint8, device="cuda") if trace else None
        need = lib().obvh_dito_workspace_size(n)
        ws = workspace if workspace is not None else _workspace(need)
    check(lib().obvh_dito_soa(_ptr(x), _ptr(y), _ptr(z), n, _ptr(out), _ptr(tr), _ptr(ws),
                              ws.numel(), _stream(stream)))
    if trace:
        return out, tr
    return out


def dito_batch(x, y, z, offsets, tile_points: int = 0, workers: int = 0, trace: bool = False,
               stream=None, workspace=None, host_offsets=None):
    """Batched DiTO.  offsets: int64 CUDA tensor with nsets + 1 entries;
    host_offsets (optional): the same values on the host, validated here.
    Without them a set with decreasing offsets or offsets past the points
    gets a NaN box (trace degenerate = 4) on the device."""
    torch = _torch()
    nsets = int(offsets.numel()) - 1
    total = int(x.numel())
    if offsets.dtype != torch.int64 or not offsets.is_cuda:
        raise ValueError("offsets must be an int64 CUDA tensor")
    if host_offsets is not None:
        ho = np.asarray(host_offsets, dtype=np.int64)
        if ho.shape != (nsets + 1,) or (nsets > 0 and (np.any(np.diff(ho) < 0) or ho[0] < 0 or ho[-1] > total)):
            raise ValueError("offsets must be nsets + 1 non-decreasing values within [0, total points]")
    with _on(stream):
        out = torch.empty((max(nsets, 0), 15), dtype=torch.float32, device="cuda")
        tr = torch.zeros(max(nsets, 1) * TRACE_DTYPE.itemsize, dtype=torch.uint8, device="cuda") if trace else None
        need = lib().obvh_dito_batch_workspace_size(max(nsets, 1), total, tile_points)
        ws = workspace if workspace is not None else _workspace(need)
    if ws.numel() < need:
        raise ValueError(f"workspace needs {need} bytes")
    check(lib().obvh_dito_batch(_ptr(x), _ptr(y), _ptr(z), _ptr(offsets), nsets, total, _ptr(out), _ptr(tr),
                                tile_points, workers, _ptr(ws), ws.numel(), _stream(stream)))
    if trace:
        return out, tr
    return out


def trace_to_numpy(tr) -> np.ndarray:
    return tr.cpu().numpy().view(TRACE_DTYPE)


# ---------------------------------------------------------------- LBVH --

def lbvh_build(V, F, stream=None, out=None, workspace=None):
    """Device LBVH over (V,3) float32 vertices and (F,4) int32/uint32 faces.

    Returns dict(nodes=uint8 tensor (2F-1)*36 B, parents, counts (int32
    tensors, values are u32), costs (float32), mortons (int32, sorted)).
    out: optional dict of preallocated tensors with those keys (e.g. views
    into a batch's arrays)."""
    torch = _torch()
    nf = int(F.shape[0])
    total = 2 * nf - 1
    with _on(stream):
        if out is not None:
            nodes, parents, counts, costs, mortons = (out[k] for k in ("nodes", "parents", "counts", "costs",
                                                                      "mortons"))
            if (nodes.numel() != total * NODE_DTYPE.itemsize or parents.numel() != total or counts.numel() != total
                    or costs.numel() != total or mortons.numel() != nf):
                raise ValueError("lbvh_build: output tensors of the wrong size")
        else:
            nodes = torch.empty(total * NODE_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
            parents = torch.empty(total, dtype=torch.int32, device="cuda")
            counts = torch.empty(total, dtype=torch.int32, device="cuda")
            costs = torch.empty(total, dtype=torch.float32, device="cuda")
            mortons = torch.empty(nf, dtype=torch.int32, device="cuda")
        need = lib().obvh_lbvh_workspace_size(nf)
        ws = workspace if workspace is not None and workspace.numel() >= need else _workspace(need)
    check(lib().obvh_lbvh_build(_ptr(V), _ptr(F), nf, _ptr(nodes), _ptr(parents), _ptr(counts), _ptr(costs),
                                _ptr(mortons), _ptr(ws), ws.numel(), _stream(stream)))
    return {"nodes": nodes, "parents": parents, "counts": counts, "costs": costs, "mortons": mortons}


def atrbvh_build(V, F, treelet_size: int = 7, iterations: int = 2, ci: float = 1.2, ct: float = 1.0, stream=None,
                 workspace=None):
    """Device ATRBVH over (V,3) float32 vertices and (F,4) int32 faces.
    Returns dict(nodes, parents, counts, costs) like lbvh_build (no mortons)."""
    torch = _torch()
    nf = int(F.shape[0])
    total = 2 * nf - 1
    with _on(stream):
        nodes = torch.empty(total * NODE_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        parents = torch.empty(total, dtype=torch.int32, device="cuda")
        counts = torch.empty(total, dtype=torch.int32, device="cuda")
        costs = torch.empty(total, dtype=torch.float32, device="cuda")
        need = lib().obvh_atrbvh_workspace_size(nf)
        ws = workspace if workspace is not None else _workspace(need)
    if ws.numel() < need:
        raise ValueError(f"workspace needs {need} bytes")
    check(lib().obvh_atrbvh_build(_ptr(V), _ptr(F), nf, int(treelet_size), int(iterations), C.c_float(ci),
                                  C.c_float(ct), _ptr(nodes), _ptr(parents), _ptr(counts), _ptr(costs), _ptr(ws),
                                  ws.numel(), _stream(stream)))
    return {"nodes": nodes, "parents": parents, "counts": counts, "costs": costs}


def nodes_to_numpy(nodes) -> np.ndarray:
    return nodes.cpu().numpy().view(NODE_DTYPE)


# ----------------------------------------------------------- transform --

def obb_tree(nodes, parents, V, F, with_ids: bool = False, stream=None, workspace=None, out=None,
             status=None):
    """Stream-ordered transform.  Returns (obbs (2N-1,15), mats (2N-1,16)[, ids (2N-1,14)]).

    status: optional int32 tensor of one element (device or pinned host)
    that receives the call's device error word on the same stream (0 =
    success; tree_status_message() explains other values).  A malformed tree
    is only detected on the device."""
    torch = _torch()
    total = int(parents.numel())
    N = (total + 1) // 2
    with _on(stream):
        if out is None:
            obbs = torch.empty((total, 15), dtype=torch.float32, device="cuda")
            mats = torch.empty((total, 16), dtype=torch.float32, device="cuda")
        else:
            obbs, mats = out
        ids = torch.empty((total, 14), dtype=torch.int32, device="cuda") if with_ids else None
        need = lib().obvh_obb_tree_workspace_size(N, N - 1)
        ws = workspace if workspace is not None else _workspace(need)
    if ws.numel() < need:
        raise ValueError(f"workspace needs {need} bytes")
    check(lib().obvh_obb_tree_async(_ptr(nodes), _ptr(parents), _ptr(V), _ptr(F), _ptr(obbs), _ptr(mats),
                                    _ptr(ids), N, N - 1, _ptr(ws), ws.numel(), _stream(stream)))
    if status is not None:
        if status.dtype != torch.int32 or status.numel() < 1:
            raise ValueError("status must be an int32 tensor with one element")
        check(lib().obvh_obb_tree_status_async(_ptr(ws), ws.numel(), N, _ptr(status), _stream(stream)))
    return (obbs, mats, ids) if with_ids else (obbs, mats)


def pack_trees(trees):
    """Concatenate trees for obb_tree_batch.  trees: list of (nodes, parents,
    V, F) device tensors in the single-tree layout (nodes uint8 (2N-1)*36,
    parents int32 (2N-1), V float32 (nv, 3), F int32 (N, 4)).  Returns a dict
    with the batch arrays, the offsets (device int32 and host int64) and the
    primitive total."""
    torch = _torch()
    import numpy as _np

    if not trees:
        raise ValueError("no trees")
    prim = _np.zeros(len(trees) + 1, dtype=_np.int64)
    vert = _np.zeros(len(trees) + 1, dtype=_np.int64)
    for t, (nd, par, V, F) in enumerate(trees):
        n = int(F.shape[0])
        if n < 1 or int(par.numel()) != 2 * n - 1:
            raise ValueError(f"tree {t}: needs N >= 1 faces and 2N - 1 nodes")
        prim[t + 1] = prim[t] + n
        vert[t + 1] = vert[t] + int(V.shape[0])
    if prim[-1] > 0x7FFFFFFF:
        raise ValueError("more than 2^31 - 1 primitives in the batch")
    return {
        "nodes": torch.cat([nd.reshape(-1) for nd, _, _, _ in trees]),
        "parents": torch.cat([par.reshape(-1) for _, par, _, _ in trees]),
        "V": torch.cat([V.reshape(-1, 3) for _, _, V, _ in trees]),
        "F": torch.cat([F.reshape(-1, 4) for _, _, _, F in trees]),
        "prim_offsets": torch.from_numpy(prim.astype(_np.int32)).cuda(),
        "vert_offsets": torch.from_numpy(vert.astype(_np.int32)).cuda(),
        "prim_offsets_host": prim,
        "total": int(prim[-1]),
        "ntrees": len(trees),
    }


def obb_tree_batch(batch, stream=None, workspace=None, out=None, status=None):
    """obvh_obb_tree_batch over a pack_trees() dict.  Returns (obbs, mats) in
    the batch's node layout (tree t's nodes from 2 L_t - t, L_t its first
    primitive).  status: optional int32 tensor receiving the batch's error word."""
    torch = _torch()
    total, K = batch["total"], batch["ntrees"]
    nodes_total = 2 * total - K
    with _on(stream):
        if out is None:
            obbs = torch.empty((nodes_total, 15), dtype=torch.float32, device="cuda")
            mats = torch.empty((nodes_total, 16), dtype=torch.float32, device="cuda")
        else:
            obbs, mats = out
        need = lib().obvh_obb_tree_batch_workspace_size(K, total)
        ws = workspace if workspace is not None else _workspace(need)
    if ws.numel() < need:
        raise ValueError(f"workspace needs {need} bytes")
    check(lib().obvh_obb_tree_batch(_ptr(batch["nodes"]), _ptr(batch["parents"]), _ptr(batch["V"]), _ptr(batch["F"]),
                                    _ptr(batch["prim_offsets"]), _ptr(batch["vert_offsets"]), K, total, _ptr(obbs),
                                    _ptr(mats), _ptr(ws), ws.numel(), _stream(stream)))
    if status is not None:
        if status.dtype != torch.int32 or status.numel() < 1:
            raise ValueError("status must be an int32 tensor with one element")
        check(lib().obvh_obb_tree_batch_status_async(_ptr(ws), ws.numel(), total, _ptr(status), _stream(stream)))
    return obbs, mats


def tree_status_message(word: int):
    """Error message of a nonzero transform status word (None for 0)."""
    m = lib().obvh_obb_tree_status_string(int(word) & 0xFFFFFFFF)
    return m.decode() if m else None


def bvh_obb_tree(nodes, parents, V, F, obbs, mats, primitive_count: int, internal_count: int,
                 primitive_counts=None, costs=None, primitives=None, ci: float = 1.2, ct: float = 1.0,
                 obb_cost: float = 1.0, primitive_threshold: int = 0xFFFFFFFF) -> None:
    """kr_cuda_bvh_obb_tree (legacy, synchronous; all arguments device tensors)."""
    err = lib().kr_cuda_bvh_obb_tree(_ptr(nodes), _ptr(parents), _ptr(primitive_counts), _ptr(costs), _ptr(V),
                                     _ptr(F), _ptr(primitives), _ptr(obbs), _ptr(mats), ci, ct, obb_cost,
                                     primitive_count, internal_count, primitive_threshold)
    check(err)


# ------------------------------------------------------------ collapse --

class CollapseMetrics(C.Structure):
    _fields_ = [("collapse_mark_time", C.c_float), ("collapse_time", C.c_float)]


def collapse(nodes, parents, counts, ci: float = 1.2, ct: float = 1.0, stream=None, workspace=None):
    """Stream-ordered BVH collapse of a device tree (nodes uint8 (2N-1)*36 B,
    parents/counts int32 tensors; counts' leaf entries are read).

    Returns dict(collapsed (uint8 tensor, node layout), counts (int32, the
    updated copy), costs (float32), primitives (int32, N entries; the slices
    in use are those of the collapsed tree's leaves)).
    """
    torch = _torch()
    total = int(parents.numel())
    N = (total + 1) // 2
    with _on(stream):
        out = torch.empty(total * NODE_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        cnt = counts.clone()
        costs = torch.empty(total, dtype=torch.float32, device="cuda")
        prims = torch.full((N,), -1, dtype=torch.int32, device="cuda")
        need = lib().obvh_collapse_workspace_size(N)
        ws = workspace if workspace is not None else _workspace(need)
    if ws.numel() < need:
        raise ValueError(f"workspace needs {need} bytes")
    check(lib().obvh_collapse_async(_ptr(nodes), _ptr(parents), _ptr(out), _ptr(cnt), _ptr(prims), _ptr(costs),
                                    C.c_float(ci), C.c_float(ct), N, _ptr(ws), ws.numel(), _stream(stream)))
    return {"collapsed": out, "counts": cnt, "costs": costs, "primitives": prims}


def bvh_collapse(nodes, parents, collapsed, counts, primitives, costs, ci: float = 1.2, ct: float = 1.0):
    """The legacy synchronous kr_cuda_bvh_collapse on device tensors; returns the metrics (ms)."""
    total = int(parents.numel())
    N = (total + 1) // 2
    m = CollapseMetrics(0.0, 0.0)
    check(lib().kr_cuda_bvh_collapse(_ptr(nodes), _ptr(parents), _ptr(collapsed), _ptr(counts), _ptr(primitives),
                                     _ptr(costs), C.c_float(ci), C.c_float(ct), N, N - 1, C.byref(m)))
    return {"collapse_mark_time": m.collapse_mark_time, "collapse_time": m.collapse_time}


def bvh_sah(nodes: np.ndarray, root: int = 0, ci: float = 1.2, ct: float = 1.0, obbs=None,
            output_file: str | None = None) -> float:
    """kr_bvh_sah on host arrays (NODE_DTYPE nodes, optional (2N-1, 15) float32 OBBs)."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    ob = None if obbs is None else np.ascontiguousarray(obbs, dtype=np.float32)
    return float(lib().kr_bvh_sah(nodes.ctypes.data_as(C.c_void_p), int(root), C.c_float(ci), C.c_float(ct),
                                  None if ob is None else ob.ctypes.data_as(C.c_void_p),
                                  output_file.encode() if output_file else None))


def validate_tree(nodes: np.ndarray, parents: np.ndarray) -> None:
    """Host-side structural check of a tree (raises ObvhError)."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    parents = np.ascontiguousarray(parents).view(np.uint32)
    N = (nodes.shape[0] + 1) // 2
    check(lib().obvh_validate_tree_host(nodes.ctypes.data_as(C.c_void_p), parents.ctypes.data_as(C.c_void_p),
                                        N, N - 1))


# ---------------------------------------------------------------- OBVH --

def soa_pack(nodes, mats, stream=None):
    """One-pass OBVH packing of a device tree (uint8 node bytes, plain or
    collapsed) and its transforms ((2N-1, 16) float32).  Returns dict(left_T
    (N-1, 12), right_T (N-1, 12) float32, meta (2N-1, 4) int32)."""
    torch = _torch()
    total = nodes.numel() // NODE_DTYPE.itemsize
    internal = (total - 1) // 2
    with _on(stream):
        L = torch.empty((max(internal, 1), 12), dtype=torch.float32, device="cuda")
        R = torch.empty((max(internal, 1), 12), dtype=torch.float32, device="cuda")
        meta = torch.empty((total, 4), dtype=torch.int32, device="cuda")
    check(lib().obvh_soa_pack_async(_ptr(nodes), _ptr(mats), total, internal, _ptr(L), _ptr(R), _ptr(meta),
                                    _stream(stream)))
    return {"left_T": L[:internal], "right_T": R[:internal], "meta": meta}


def soa_intersect(packed, V, F, rays, primitives=None, stream=None, workspace=None, out=None):
    """Closest hits of device rays (uint8 tensor of kr_ray records, or a
    RAY_DTYPE numpy array) against a packed OBVH.  Returns a uint8 tensor of
    kr_intersection records (view the host copy with ISECT_DTYPE)."""
    torch = _torch()
    with _on(stream):
        if isinstance(rays, np.ndarray):
            rays = torch.from_numpy(np.ascontiguousarray(rays, dtype=RAY_DTYPE).view(np.uint8)).cuda()
        n = rays.numel() // RAY_DTYPE.itemsize
        if out is None:
            out = torch.empty(n * ISECT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        need = lib().obvh_soa_intersect_workspace_size(n)
        ws = workspace if workspace is not None else _workspace(need)
    if ws.numel() < need:
        raise ValueError(f"workspace needs {need} bytes")
    nf = F.numel() // 4
    check(lib().obvh_soa_intersect_async(_ptr(packed["left_T"]), _ptr(packed["right_T"]), _ptr(packed["meta"]),
                                         packed["left_T"].shape[0], packed["meta"].shape[0], _ptr(V),
                                         V.numel() // 3, _ptr(F), _ptr(primitives), nf, _ptr(rays), _ptr(out), n,
                                         _ptr(ws), ws.numel(), _stream(stream)))
    return out


def records_pack(nodes, mats, V, F, primitives=None, stream=None):
    """Native OBVH records of a device tree (plain or collapsed): dict(nodes
    (uint8, 128 B per internal node), tris (uint8, 48 B per slot, 2N slots),
    face_count)."""
    torch = _torch()
    total = nodes.numel() // NODE_DTYPE.itemsize
    N = (total + 1) // 2
    with _on(stream):
        rec = torch.zeros(max(N - 1, 1) * 128, dtype=torch.uint8, device="cuda")
        tri = torch.zeros(2 * N * 48, dtype=torch.uint8, device="cuda")
    check(lib().obvh_records_pack_async(_ptr(nodes), _ptr(mats), _ptr(primitives), _ptr(V), V.numel() // 3,
                                        _ptr(F), N, _ptr(rec), _ptr(tri), _stream(stream)))
    return {"nodes": rec, "tris": tri, "face_count": N}


def records_intersect(records, rays, stream=None, workspace=None, out=None):
    """Closest hits against the record layout; same output as soa_intersect."""
    torch = _torch()
    with _on(stream):
        if isinstance(rays, np.ndarray):
            rays = torch.from_numpy(np.ascontiguousarray(rays, dtype=RAY_DTYPE).view(np.uint8)).cuda()
        n = rays.numel() // RAY_DTYPE.itemsize
        if out is None:
            out = torch.empty(n * ISECT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        need = lib().obvh_soa_intersect_workspace_size(n)
        ws = workspace if workspace is not None else _workspace(need)
    if ws.numel() < need:
        raise ValueError(f"workspace needs {need} bytes")
    check(lib().obvh_records_intersect_async(_ptr(records["nodes"]), _ptr(records["tris"]), records["face_count"],
                                             _ptr(rays), _ptr(out), n, _ptr(ws), ws.numel(), _stream(stream)))
    return out


def obvh_persistent(nodes, mats):
    """kr_cuda_obvh_persistent (legacy, synchronous): AoS packed nodes (uint8,
    64 B each) and transformation pairs (uint8, 128 B per internal node).
    Slots the reference leaves untouched are zero here (fresh buffers)."""
    torch = _torch()
    total = nodes.numel() // NODE_DTYPE.itemsize
    internal = (total - 1) // 2
    packed = torch.zeros(total * 64, dtype=torch.uint8, device="cuda")
    pairs = torch.zeros(max(internal, 1) * 128, dtype=torch.uint8, device="cuda")
    check(lib().kr_cuda_obvh_persistent(_ptr(nodes), _ptr(mats), _ptr(packed), _ptr(pairs), total))
    return packed, pairs


def soa_obvh(packed, pairs, internal_count: int, node_count: int) -> int:
    """kr_cuda_soa_obvh (legacy): returns the device kr_SoA_OBVH pointer."""
    h = C.c_void_p(0)
    check(lib().kr_cuda_soa_obvh(_ptr(packed), _ptr(pairs), C.byref(h), internal_count, node_count))
    return int(h.value)


def soa_obvh_intersect(handle: int, V, F, primitives, rays, out) -> None:
    """kr_cuda_soa_obvh_intersect (legacy, default stream)."""
    n = rays.numel() // RAY_DTYPE.itemsize
    check(lib().kr_cuda_soa_obvh_intersect(C.c_void_p(handle), _ptr(V), _ptr(F), _ptr(primitives), _ptr(rays),
                                           _ptr(out), n, 128, 0, None))


def soa_obvh_free(handle: int) -> None:
    check(lib().obvh_soa_obvh_free(C.c_void_p(handle)))


def isects_to_numpy(isects) -> np.ndarray:
    return isects.cpu().numpy().view(ISECT_DTYPE)
