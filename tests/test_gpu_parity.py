"""GPU parity: every HIP entry point against the CPU oracle on the same
seeded inputs.  Integer/index results must be bit-exact; float results are
compared bit-exactly too (== on values, so -0 == +0), which is stricter than
the 1e-5 relative bound the path is specified to.  Full-size runs are
checked through size-independent properties (enclosure, determinism)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from obvh_amd import ops, synth  # noqa: E402
from obvh_amd.layouts import NODE_DTYPE, TRACE_DTYPE  # noqa: E402
from treeutil import contiguous_tree, doubled_vertices, flat_grid, slivers  # noqa: E402


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def to_dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def soa(torch, p):
    p = np.ascontiguousarray(p, np.float32)
    return [torch.from_numpy(np.ascontiguousarray(p[:, k])).cuda() for k in range(3)]


TRACE_INT_FIELDS = ["slab_min_idx", "slab_max_idx", "pair_k", "p2_sel", "q0_sel", "q1_sel", "q0_valid",
                    "q1_valid", "degenerate", "best", "result_aabb"]


def assert_trace_equal(got, exp):
    for f in TRACE_INT_FIELDS:
        assert np.array_equal(got[f], exp[f]), f


DITO_CASES = {
    "gaussian131072": lambda: synth.gaussian_points(131072, seed=1),
    "gaussian1000": lambda: synth.gaussian_points(1000, seed=2),
    "n1": lambda: np.array([[1.0, 2.0, 3.0]], np.float32),
    "n3": lambda: synth.gaussian_points(3, seed=3),
    "n14": lambda: synth.gaussian_points(14, seed=4),
    "n15": lambda: synth.gaussian_points(15, seed=5),
    "dups": lambda: (np.random.default_rng(6).integers(-8, 8, (5000, 3)) * 0.25).astype(np.float32),
    "line": lambda: (np.linspace(-1, 1, 777)[:, None] * np.array([0.3, -0.5, 0.8])).astype(np.float32),
    "point": lambda: np.tile(np.array([[0.5, 0.5, 0.5]], np.float32), (600, 1)),
    # coplanar points (a tilted plane through the origin, exact in float): no apex off the base triangle
    "plane": lambda: (np.random.default_rng(8).integers(-64, 64, (3000, 2)) @ np.array(
        [[0.5, 0.25, 0.0], [0.0, 0.5, 0.25]])).astype(np.float32),
    "ragged": lambda: synth.gaussian_points(512 * 7 + 13, seed=7),
}


@pytest.mark.parametrize("case", sorted(DITO_CASES))
def test_dito_soa_matches_oracle(torch, oracle_lib, case):
    say(f"[dito_soa] case={case}")
    p = DITO_CASES[case]()
    exp, etr = oracle_lib.points_obb(p, trace=True)
    x, y, z = soa(torch, p)
    out, tr = ops.dito(x, y, z, trace=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got, exp), (got, exp)
    assert_trace_equal(ops.trace_to_numpy(tr)[0], etr)


@pytest.mark.parametrize("case", ["gaussian131072", "n3", "dups", "point"])
def test_legacy_points_obb_matches_oracle(torch, oracle_lib, case):
    say(f"[kr_cuda_points_obb] case={case}")
    p = DITO_CASES[case]()
    got, ms = ops.points_obb(p, return_time=True)
    assert np.array_equal(got, oracle_lib.points_obb(p))
    assert ms >= 0.0


@pytest.mark.parametrize("tile", [0, 512, 4096])
def test_dito_batch_matches_oracle(torch, oracle_lib, tile):
    say(f"[dito_batch] tile={tile}")
    sizes = np.concatenate([[0, 1, 2, 14, 15, 64, 513], synth.batch_sizes(120, lo=64, hi=1 << 15, seed=9)])
    pts, off = synth.batch_sets(sizes, seed=10)
    exp, etr = oracle_lib.points_obb_batch(pts, off)
    x, y, z = soa(torch, pts)
    offs = torch.from_numpy(off.astype(np.int64)).cuda()
    out, tr = ops.dito_batch(x, y, z, offs, tile_points=tile, trace=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    gtr = ops.trace_to_numpy(tr)[: len(sizes)]
    for s in range(len(sizes)):
        assert np.array_equal(got[s], exp[s]), (s, sizes[s])
        assert_trace_equal(gtr[s], etr[s])


def build_gpu_tree(torch, V, F):
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(F.view(np.int32)).cuda()
    t = ops.lbvh_build(Vd, Fd)
    return Vd, Fd, t


MESHES = {
    "ico16": lambda: synth.icosphere(16, seed=1),
    "flat": lambda: flat_grid(60, seed=9),
    "dupverts": lambda: doubled_vertices(seed=10),
    "slivers": lambda: slivers(seed=11),
    "ico_n256": lambda: synth.icosphere(8, seed=2),  # 1280 faces = 5 * 256
    "hair200": lambda: synth.hairball(tubes=200, seed=3),
    "single": lambda: (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                       np.array([[0, 1, 2, 0]], np.uint32)),
    "two": lambda: (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32),
                    np.array([[0, 1, 2, 0], [0, 1, 3, 0]], np.uint32)),
}


@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_lbvh_matches_oracle(torch, oracle_lib, mesh):
    say(f"[lbvh] mesh={mesh}")
    V, F = MESHES[mesh]()
    Vd, Fd, t = build_gpu_tree(torch, V, F)
    torch.cuda.synchronize()
    en, ep, ec, eco, em = oracle_lib.lbvh_build(V, F)
    gn = ops.nodes_to_numpy(t["nodes"])
    assert np.array_equal(t["mortons"].cpu().numpy().view(np.uint32), em)
    assert np.array_equal(t["parents"].cpu().numpy().view(np.uint32), ep)
    assert np.array_equal(t["counts"].cpu().numpy().view(np.uint32), ec)
    assert np.array_equal(t["costs"].cpu().numpy(), eco)
    for f in ("bmin", "bmax", "a", "nprim", "axis"):
        assert np.array_equal(gn[f], en[f]), f
    I = len(F) - 1
    assert np.array_equal(gn["b"][:I], en["b"][:I])


@pytest.mark.parametrize("mesh", sorted(MESHES))
def test_obb_tree_matches_oracle(torch, oracle_lib, mesh):
    say(f"[obb_tree] mesh={mesh}")
    V, F = MESHES[mesh]()
    Vd, Fd, t = build_gpu_tree(torch, V, F)
    obbs, mats, ids = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd, with_ids=True)
    torch.cuda.synchronize()
    nodes = ops.nodes_to_numpy(t["nodes"])
    parents = t["parents"].cpu().numpy().view(np.uint32)
    eo, em, ei = oracle_lib.obb_tree(nodes, parents, V, F, with_ids=True)
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), ei)
    go, gm = obbs.cpu().numpy(), mats.cpu().numpy()
    # equal, or NaN in both (a zero-area leaf triangle's frame is NaN in the
    # reference's leaf OBB as well)
    bad = np.where(~np.all((go == eo) | (np.isnan(go) & np.isnan(eo)), axis=1))[0]
    assert bad.size == 0, (bad[:10], go[bad[:3]], eo[bad[:3]])
    assert np.all((gm == em) | (np.isnan(gm) & np.isnan(em)))
    if mesh != "slivers":
        assert not np.isnan(go).any()


def test_legacy_obb_tree_equals_async(torch, oracle_lib):
    say("[kr_cuda_bvh_obb_tree] ico16")
    V, F = synth.icosphere(16, seed=4)
    Vd, Fd, t = build_gpu_tree(torch, V, F)
    total = 2 * len(F) - 1
    o1 = torch.zeros((total, 15), dtype=torch.float32, device="cuda")
    m1 = torch.zeros((total, 16), dtype=torch.float32, device="cuda")
    ops.bvh_obb_tree(t["nodes"], t["parents"], Vd, Fd, o1, m1, len(F), len(F) - 1,
                     primitive_counts=t["counts"], costs=t["costs"])
    o2, m2 = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd)
    torch.cuda.synchronize()
    assert torch.equal(o1, o2) and torch.equal(m1, m2)


def test_legacy_obb_tree_rejects_malformed_parents(torch):
    say("[kr_cuda_bvh_obb_tree] malformed parents")
    V, F = synth.icosphere(4, seed=5)
    Vd, Fd, t = build_gpu_tree(torch, V, F)
    parents = t["parents"].clone()
    parents[7] = 3  # disagrees with the children of node 3 (bounds stay valid)
    total = 2 * len(F) - 1
    o = torch.zeros((total, 15), dtype=torch.float32, device="cuda")
    m = torch.zeros((total, 16), dtype=torch.float32, device="cuda")
    from obvh_amd._lib import ObvhError

    with pytest.raises(ObvhError):
        ops.bvh_obb_tree(t["nodes"], parents, Vd, Fd, o, m, len(F), len(F) - 1)


def test_full_size_transform_properties(torch, oracle_lib):
    # 1 003 520-triangle icosphere (f = 224): determinism across runs and
    # parity with the oracle on the whole tree (the oracle takes seconds).
    say("[obb_tree] ico224 full size")
    V, F = synth.icosphere(224, seed=2)
    Vd, Fd, t = build_gpu_tree(torch, V, F)
    o1, m1 = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd)
    o2, m2 = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd)
    torch.cuda.synchronize()
    assert torch.equal(o1, o2) and torch.equal(m1, m2)
    nodes = ops.nodes_to_numpy(t["nodes"])
    parents = t["parents"].cpu().numpy().view(np.uint32)
    en, ep, *_ = oracle_lib.lbvh_build(V, F)
    assert np.array_equal(parents, ep)
    eo, em = oracle_lib.obb_tree(nodes, parents, V, F)
    go = o1.cpu().numpy()
    mism = int((~np.all(go == eo, axis=1)).sum())
    assert mism == 0, mism
    assert np.array_equal(m1.cpu().numpy(), em)


def random_tree(n_leaves: int, seed: int):
    """Random binary tree over shuffled leaves: internal nodes numbered in
    breadth-first order (root 0), leaves I..2N-2 holding a random face
    permutation.  Sibling leaves are rarely adjacent, so the bottom-up
    kernels exchange most k-DOPs across workgroups."""
    rng = np.random.default_rng(seed)
    I = n_leaves - 1
    nodes = np.zeros(2 * n_leaves - 1, dtype=NODE_DTYPE)
    parents = np.zeros(2 * n_leaves - 1, dtype=np.uint32)
    parents[0] = 0xFFFFFFFF
    leaf_slots = rng.permutation(n_leaves)  # leaf slot of each face
    queue = [(0, np.arange(n_leaves))]
    next_internal = 1
    head = 0
    while head < len(queue):
        idx, faces = queue[head]
        head += 1
        cut = int(rng.integers(1, len(faces)))
        kids = []
        for part in (faces[:cut], faces[cut:]):
            if len(part) == 1:
                kids.append(I + int(leaf_slots[part[0]]))
            else:
                kids.append(next_internal)
                queue.append((next_internal, part))
                next_internal += 1
        nodes[idx]["a"], nodes[idx]["b"] = kids
        for k in kids:
            parents[k] = idx
    for f in range(n_leaves):
        L = nodes[I + leaf_slots[f]]
        L["a"] = f
        L["nprim"] = 1
    return nodes, parents


@pytest.mark.parametrize("n", [5000, 20000, 100000])
def test_obb_tree_random_topology(torch, oracle_lib, n):
    """n = 100 000 spans four 32768-leaf chunks: nearly every node has leaves
    in several workgroups, so the climb defers most parents across workgroups,
    folds split pairs into the global arrival words, overflows the last
    phase's LDS slots into global arrivals, and the refit's tiles have
    gapped ancestors everywhere."""
    say(f"[obb_tree] random topology n={n}")
    V, F = synth.hairball(tubes=n // 40 + 1, seed=6)
    F = F[:n]
    nodes, parents = random_tree(n, seed=n)
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).cuda()
    nd = torch.from_numpy(nodes.view(np.uint8)).cuda()
    pd = torch.from_numpy(parents.view(np.int32)).cuda()
    obbs, mats, ids = ops.obb_tree(nd, pd, Vd, Fd, with_ids=True)
    torch.cuda.synchronize()
    eo, em, ei = oracle_lib.obb_tree(nodes, parents, V, F, with_ids=True)
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), ei)
    assert np.array_equal(obbs.cpu().numpy(), eo)
    assert np.array_equal(mats.cpu().numpy(), em)


@pytest.mark.parametrize("n,spine", [(30000, 0), (81920, 512), (81920, 384), (40000, 16)])
def test_obb_tree_contiguous_topology(torch, oracle_lib, n, spine):
    """Trees whose subtrees hold contiguous leaf runs but whose internal nodes
    are numbered breadth-first (not Karras): the refit's LDS copy of the
    tile's index range mostly misses.  spine=512: the leftmost tile has 127
    ancestors holding it, more than the 64 listed per tile, so the rest are
    walked by every wave; spine=384: 171.  spine=16: a 2000-deep spine
    whose nodes mostly span more than one 32768-leaf chunk, so the
    top-of-tree climb phases walk long chains."""
    say(f"[obb_tree] contiguous topology n={n} spine={spine}")
    V, F = synth.hairball(tubes=n // 40 + 1, seed=7)
    F = F[:n]
    nodes, parents = contiguous_tree(n, seed=n, spine_block=spine)
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).cuda()
    nd = torch.from_numpy(nodes.view(np.uint8)).cuda()
    pd = torch.from_numpy(parents.view(np.int32)).cuda()
    obbs, mats, ids = ops.obb_tree(nd, pd, Vd, Fd, with_ids=True)
    torch.cuda.synchronize()
    eo, em, ei = oracle_lib.obb_tree(nodes, parents, V, F, with_ids=True)
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), ei)
    assert np.array_equal(obbs.cpu().numpy(), eo)
    assert np.array_equal(mats.cpu().numpy(), em)


def renumber(nodes, parents, perm, lperm):
    """The same tree with internal node i stored at perm[i] (perm[0] = 0) and
    the leaf at position j stored at position lperm[j]."""
    n = len(lperm)
    I = n - 1
    old2new = np.concatenate([perm, I + lperm]).astype(np.int64)
    new_nodes = np.zeros_like(nodes)
    new_parents = np.zeros_like(parents)
    new_nodes[old2new] = nodes
    for i in range(I):
        j = old2new[i]
        new_nodes[j]["a"] = old2new[nodes[i]["a"]]
        new_nodes[j]["b"] = old2new[nodes[i]["b"]]
    p = parents.astype(np.int64)
    new_parents[old2new] = np.where(p == 0xFFFFFFFF, 0xFFFFFFFF, old2new[np.minimum(p, 2 * n - 2)]).astype(np.uint32)
    return new_nodes, new_parents, old2new


@pytest.mark.parametrize("mesh,builder", [("ico", "lbvh"), ("hair", "atrbvh")])
def test_obb_tree_renumbering_canary(torch, oracle_lib, mesh, builder):
    """Determinism canary: the same tree stored with its internal nodes in a
    random order and its leaf slots reversed.  Every tile, chunk, workgroup
    and lane assignment of the climb, the candidate search and the refit
    changes (and the spans get gaps everywhere), so any result that
    depended on scheduling would move; every box and matrix, mapped back,
    must be bit-identical to the original tree's."""
    say(f"[obb_tree] renumbering canary {mesh} {builder}")
    V, F = synth.icosphere(48, seed=11) if mesh == "ico" else synth.hairball(tubes=1500, seed=12)
    if builder == "lbvh":
        nodes, parents, *_ = oracle_lib.lbvh_build(V, F)
    else:
        nodes, parents, *_ = oracle_lib.atrbvh_build(V, F, 7, 2)
    n = F.shape[0]
    rng = np.random.default_rng(n)
    perm = np.concatenate([[0], 1 + rng.permutation(n - 2)])
    lperm = np.arange(n)[::-1].copy()
    nn, npar, old2new = renumber(nodes, parents, perm, lperm)
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).cuda()
    outs = []
    for nd_, pr_ in ((nodes, parents), (nn, npar)):
        o, m = ops.obb_tree(torch.from_numpy(np.ascontiguousarray(nd_).view(np.uint8)).cuda(),
                            torch.from_numpy(np.ascontiguousarray(pr_).view(np.int32)).cuda(), Vd, Fd)
        torch.cuda.synchronize()
        outs.append((o.cpu().numpy(), m.cpu().numpy()))
    (o0, m0), (o1, m1) = outs
    assert np.array_equal(o1[old2new], o0)
    assert np.array_equal(m1[old2new], m0)
    eo, em = oracle_lib.obb_tree(nodes, parents, V, F)
    assert np.array_equal(o0, eo) and np.array_equal(m0, em)
