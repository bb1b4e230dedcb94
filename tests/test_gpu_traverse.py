"""GPU OBVH packing and traversal against the oracle: bit-exact packed
arrays and bit-exact hit records (primitive, barycentrics, normal,
instance) for LBVH and ATRBVH trees, plain and collapsed, camera and random
rays; the legacy three-call path (kr_cuda_obvh_persistent ->
kr_cuda_soa_obvh -> kr_cuda_soa_obvh_intersect) against the one-pass path;
stacks deeper than the LDS part, exactly full, and overflowing."""
import sys

import numpy as np
import pytest

from obvh_amd import ops, synth
from obvh_amd.layouts import ISECT_DTYPE, MISS, NODE_DTYPE, RAY_DTYPE, STACK_OVERFLOW
from test_oracle_traverse import chain_rays, chain_tree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same_isects(got, exp):
    """Bit-exact records; NaN barycentrics (degenerate hits) compare equal."""
    assert np.array_equal(got["primitive"], exp["primitive"])
    assert np.array_equal(got["instance"], exp["instance"])
    for f in ("normal", "barys"):
        a, b = got[f], exp[f]
        both_nan = np.isnan(a) & np.isnan(b)
        assert np.array_equal(np.where(both_nan, 0, a).view(np.uint32), np.where(both_nan, 0, b).view(np.uint32)), f


def device_tree(torch, V, F, builder):
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).cuda()
    t = ops.lbvh_build(Vd, Fd) if builder == "lbvh" else ops.atrbvh_build(Vd, Fd, 7, 2)
    obbs, mats = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd)
    return Vd, Fd, t, mats


@pytest.mark.parametrize("kind,builder,collapsed,rays_kind", [
    ("ico", "lbvh", False, "camera"),
    ("ico", "lbvh", True, "random"),
    ("hair", "lbvh", True, "camera"),
    ("hair", "atrbvh", False, "random"),
    ("hair", "atrbvh", True, "camera"),
])
def test_pack_and_intersect_match_oracle(torch, oracle_lib, kind, builder, collapsed, rays_kind):
    say(f"[traverse] {kind} {builder} collapsed={collapsed} rays={rays_kind}")
    V, F = synth.icosphere(40, seed=2) if kind == "ico" else synth.hairball(tubes=1500, seed=3)
    Vd, Fd, t, mats = device_tree(torch, V, F, builder)
    nodes, prims_d, prims_h = t["nodes"], None, None
    if collapsed:
        col = ops.collapse(t["nodes"], t["parents"], t["counts"])
        nodes, prims_d = col["collapsed"], col["primitives"]
    packed = ops.soa_pack(nodes, mats)
    rays = synth.camera_rays(V, 160, 120) if rays_kind == "camera" else synth.random_rays(V, 20000, seed=5)
    out = ops.soa_intersect(packed, Vd, Fd, rays, primitives=prims_d)
    torch.cuda.synchronize()

    nh = ops.nodes_to_numpy(nodes)
    eL, eR, em = oracle_lib.obvh_pack(nh, mats.cpu().numpy())
    assert np.array_equal(packed["left_T"].cpu().numpy().view(np.uint32), eL.view(np.uint32))
    assert np.array_equal(packed["right_T"].cpu().numpy().view(np.uint32), eR.view(np.uint32))
    assert np.array_equal(packed["meta"].cpu().numpy(), em)
    if prims_d is not None:
        prims_h = prims_d.cpu().numpy().view(np.uint32)
    exp = oracle_lib.obvh_intersect(eL, eR, em, V, F, prims_h, rays)
    got = ops.isects_to_numpy(out)
    same_isects(got, exp)
    assert np.count_nonzero(got["primitive"] != MISS) > rays.shape[0] // 10
    # the native record layout: same records, byte for byte
    recs = ops.records_pack(nodes, mats, Vd, Fd, primitives=prims_d)
    out_r = ops.records_intersect(recs, rays)
    torch.cuda.synchronize()
    assert np.array_equal(out_r.cpu().numpy(), out.cpu().numpy())
    check_records(recs, nh, eL, eR, em, V, F, prims_h)


def check_records(recs, nh, eL, eR, em, V, F, prims_h):
    """Decode the record layout of every internal node reachable from the
    root and compare it with the oracle's SoA packing.  (A collapsed tree
    also holds unreachable internal nodes below a collapsed leaf.)"""
    R = recs["nodes"].cpu().numpy().view(np.float32).reshape(-1, 32)
    T = recs["tris"].cpu().numpy().view(np.float32).reshape(-1, 12)
    stack = [0]
    while stack:
        i = stack.pop()
        if nh[i]["nprim"] != 0:
            continue
        stack += [int(nh[i]["a"]), int(nh[i]["b"])]
        assert np.array_equal(R[i, 0:12].view(np.uint32), eL[i].view(np.uint32))
        assert np.array_equal(R[i, 12:24].view(np.uint32), eR[i].view(np.uint32))
        for side, code in ((0, R[i, 24:25].view(np.int32)[0]), (1, R[i, 25:26].view(np.int32)[0])):
            exp_code = em[i, side]
            if exp_code >= 0:
                assert code == exp_code
                continue
            leaf = ~exp_code
            beg, end = em[leaf, 0], em[leaf, 1]
            pids = [beg] if end - beg == 1 else list(prims_h[beg:end])
            x = ~int(code)
            slot, cnt = x >> 4, (x & 15) + 1
            assert cnt == len(pids)
            for j, pid in enumerate(pids):
                rec = T[slot + j]
                assert rec[9:10].view(np.uint32)[0] == pid
                f = F[pid]
                assert np.array_equal(rec[0:9].reshape(3, 3), V[f[:3].astype(np.int64)])


def test_legacy_path_matches_one_pass(torch, oracle_lib):
    say("[traverse] legacy kr_cuda_obvh_persistent -> kr_cuda_soa_obvh -> intersect")
    V, F = synth.icosphere(24, seed=7)
    Vd, Fd, t, mats = device_tree(torch, V, F, "lbvh")
    col = ops.collapse(t["nodes"], t["parents"], t["counts"])
    total = t["parents"].numel()
    internal = (total - 1) // 2
    packed, pairs = ops.obvh_persistent(col["collapsed"], mats)
    torch.cuda.synchronize()
    # AoS contents: child boxes / metadata / mat4 pairs of every live node
    nh = ops.nodes_to_numpy(col["collapsed"])
    P = packed.cpu().numpy().view(np.uint32).reshape(total, 16)
    Pf, Pi = P.view(np.float32), P.view(np.int32)
    T = pairs.cpu().numpy().view(np.float32).reshape(-1, 32)
    m = mats.cpu().numpy()
    for i in range(total):
        n = nh[i]
        if n["nprim"] == 0xFFFF:
            assert (P[i] == 0).all()
            continue
        if n["nprim"] == 0:
            a, b = int(n["a"]), int(n["b"])
            assert np.array_equal(Pf[i, 0:3], nh[a]["bmin"]) and np.array_equal(Pf[i, 9:12], nh[b]["bmax"])
            assert Pi[i, 12] == (~a if nh[a]["nprim"] > 0 else a) and Pi[i, 13] == (~b if nh[b]["nprim"] > 0 else b)
            assert np.array_equal(T[i, :16], m[a]) and np.array_equal(T[i, 16:], m[b])
        else:
            assert np.array_equal(Pf[i, 0:3], n["bmin"])
            assert (Pi[i, 12], Pi[i, 13]) == (n["a"], n["a"] + n["nprim"])
        assert Pi[i, 14] == n["axis"] and Pi[i, 15] == 0
    handle = ops.soa_obvh(packed, pairs, internal, total)
    try:
        rays = synth.random_rays(V, 5000, seed=8)
        rd = torch.from_numpy(rays.view(np.uint8)).cuda()
        out_legacy = torch.empty(rays.shape[0] * ISECT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        ops.soa_obvh_intersect(handle, Vd, Fd, col["primitives"], rd, out_legacy)
        one = ops.soa_intersect(ops.soa_pack(col["collapsed"], mats), Vd, Fd, rd, primitives=col["primitives"])
        torch.cuda.synchronize()
        assert np.array_equal(out_legacy.cpu().numpy(), one.cpu().numpy())
    finally:
        ops.soa_obvh_free(handle)


@pytest.mark.parametrize("depth", [40, 63, 64])
def test_deep_stack(torch, oracle_lib, depth):
    say(f"[traverse] chain depth {depth}")
    L, R, meta, V, F = chain_tree(depth)
    packed = {"left_T": torch.from_numpy(L).cuda(), "right_T": torch.from_numpy(R).cuda(),
              "meta": torch.from_numpy(meta).cuda()}
    rays = chain_rays()
    out = ops.soa_intersect(packed, torch.from_numpy(V).cuda(), torch.from_numpy(F.view(np.int32)).cuda(), rays)
    torch.cuda.synchronize()
    got = ops.isects_to_numpy(out)
    exp = oracle_lib.obvh_intersect(L, R, meta, V, F, None, rays)
    same_isects(got, exp)
    if depth == 64:
        assert (got["instance"] == STACK_OVERFLOW).all()
    else:
        assert list(got["primitive"]) == [0, depth, MISS]


def test_bad_index_aborts_ray(torch, oracle_lib):
    """A child index past the node arrays stops the ray (no out-of-bounds read)."""
    say("[traverse] malformed metadata")
    L, R, meta, V, F = chain_tree(8)
    meta[3, 0] = 1000  # internal child outside [0, internal_count)
    packed = {"left_T": torch.from_numpy(L).cuda(), "right_T": torch.from_numpy(R).cuda(),
              "meta": torch.from_numpy(meta).cuda()}
    out = ops.soa_intersect(packed, torch.from_numpy(V).cuda(), torch.from_numpy(F.view(np.int32)).cuda(),
                            chain_rays())
    torch.cuda.synchronize()
    got = ops.isects_to_numpy(out)
    assert (got["instance"] == STACK_OVERFLOW).all()


def test_empty_and_all_miss(torch, oracle_lib):
    say("[traverse] empty batch, all-miss batch")
    V, F = synth.icosphere(8, seed=1)
    Vd, Fd, t, mats = device_tree(torch, V, F, "lbvh")
    packed = ops.soa_pack(t["nodes"], mats)
    empty = ops.soa_intersect(packed, Vd, Fd, np.zeros(0, dtype=RAY_DTYPE))
    assert empty.numel() == 0
    rays = synth.random_rays(V, 300, seed=2)
    rays["direction"] = -rays["direction"]  # pointing away from the mesh
    rays["origin"] *= 10
    out = ops.soa_intersect(packed, Vd, Fd, rays)
    torch.cuda.synchronize()
    got = ops.isects_to_numpy(out)
    assert (got["primitive"] == MISS).all()
    assert np.array_equal(got["normal"], rays["direction"])
