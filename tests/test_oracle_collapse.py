"""Pins of the collapse and SAH oracle (oracle/collapse.c): closed forms on
hand-built trees, subtree coverage of every collapsed leaf, the stack-walk
order, bottom-up cost == traversal SAH, and the library's host kr_bvh_sah
against the oracle."""
import numpy as np
import pytest

from obvh_amd import ops, synth
from obvh_amd.layouts import NODE_DTYPE
from treeutil import INV, fill_bounds, leaves_right_first, random_tree

CI, CT = 1.2, 1.0


def area(n):
    d = n["bmax"].astype(np.float64) - n["bmin"].astype(np.float64)
    return 2.0 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2])


def box_tree(boxes_leaves, boxes_internal, children):
    """Tree with given internal children lists (internal i -> (a, b)), leaf boxes
    and internal boxes; leaf k holds face k."""
    n = len(boxes_leaves)
    I = n - 1
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    parents = np.full(2 * n - 1, INV, dtype=np.uint32)
    for i, (a, b) in enumerate(children):
        nodes[i]["a"], nodes[i]["b"] = a, b
        parents[a] = parents[b] = i
        nodes[i]["bmin"], nodes[i]["bmax"] = boxes_internal[i]
    for k, (mn, mx) in enumerate(boxes_leaves):
        nodes[I + k]["bmin"], nodes[I + k]["bmax"] = mn, mx
        nodes[I + k]["a"] = k
        nodes[I + k]["nprim"] = 1
    counts = np.zeros(2 * n - 1, dtype=np.uint32)
    counts[I:] = 1
    return nodes, parents, counts


def test_two_equal_leaves_collapse(oracle_lib):
    """Parent box == child boxes (area 6): node cost 1.2*6 + 6 + 6 = 19.2 >
    leaf cost 2*6 = 12, so the parent becomes one leaf of 2 primitives."""
    unit = ((0, 0, 0), (1, 1, 1))
    nodes, parents, counts = box_tree([unit, unit], [unit], [(1, 2)])
    r = oracle_lib.collapse(nodes, parents, counts, CI, CT)
    assert r["costs"][0] == pytest.approx(12.0)
    assert list(r["table"]) == [1, 1, 1]
    # the root has no parent, so it stays internal; its children are marked
    assert r["collapsed"]["nprim"][1] == 0xFFFF and r["collapsed"]["nprim"][2] == 0xFFFF
    assert r["used"] == 0


def test_far_apart_leaves_stay(oracle_lib):
    """Tiny leaves far apart: the parent box is large, leaf cost 2A > 1.2A + eps."""
    a = ((0, 0, 0), (0.01, 0.01, 0.01))
    b = ((10, 10, 10), (10.01, 10.01, 10.01))
    nodes, parents, counts = box_tree([a, b], [((0, 0, 0), (10.01, 10.01, 10.01))], [(1, 2)])
    r = oracle_lib.collapse(nodes, parents, counts, CI, CT)
    A = area(nodes[0])
    assert r["costs"][0] == pytest.approx(CI * A + area(nodes[1]) + area(nodes[2]), rel=1e-6)
    assert not r["table"].any()


def test_collapsed_child_becomes_leaf(oracle_lib):
    """((l0, l1), far) : the pair collapses, the root does not, so node 1
    becomes a leaf of 2 primitives holding faces in right-first order."""
    unit = ((0, 0, 0), (1, 1, 1))
    far = ((50, 50, 50), (50.1, 50.1, 50.1))
    big = ((0, 0, 0), (50.1, 50.1, 50.1))
    # internal 0 = (1, leaf far), internal 1 = (leaf0, leaf1); leaves at 2, 3, 4
    nodes, parents, counts = box_tree([unit, unit, far], [big, unit], [(1, 4), (2, 3)])
    r = oracle_lib.collapse(nodes, parents, counts, CI, CT)
    c = r["collapsed"]
    assert c["nprim"][1] == 2 and c["a"][1] == 0
    assert list(r["primitives"]) == [1, 0]  # right child first
    assert c["nprim"][2] == 0xFFFF and c["nprim"][3] == 0xFFFF
    assert r["counts"][1] == 2 and r["used"] == 2


@pytest.mark.parametrize("kind", ["ico", "hair", "random"])
def test_collapse_invariants(oracle_lib, kind):
    if kind == "ico":
        V, F = synth.icosphere(12, seed=1)
        nodes, parents, counts, _, _ = oracle_lib.lbvh_build(V, F)
    elif kind == "hair":
        V, F = synth.hairball(tubes=120, seed=2)
        nodes, parents, counts, _, _ = oracle_lib.lbvh_build(V, F)
    else:
        V, F = synth.hairball(tubes=80, seed=5)
        nodes, parents = random_tree(F.shape[0], seed=7)
        nodes = fill_bounds(nodes, V, F)
        counts = np.zeros(nodes.shape[0], dtype=np.uint32)
        counts[(nodes.shape[0] - 1) // 2:] = 1
    total = nodes.shape[0]
    I = (total - 1) // 2
    r = oracle_lib.collapse(nodes, parents, counts, CI, CT)
    c, prims, table = r["collapsed"], r["primitives"], r["table"]
    # walk the collapsed tree: every face reachable exactly once
    seen = np.zeros(I + 1, dtype=np.int64)
    stack = [0]
    while stack:
        v = stack.pop()
        n = c[v]
        assert n["nprim"] != 0xFFFF
        if n["nprim"] > 0:
            if v >= I:
                seen[n["a"]] += 1
            else:
                k = int(n["nprim"])
                assert 2 <= k <= 8
                got = prims[n["a"]:n["a"] + k]
                want = [int(nodes[u]["a"]) for u in leaves_right_first(nodes, v, I)]
                assert list(got) == want
                seen[got] += 1
        else:
            stack += [int(n["a"]), int(n["b"])]
    assert (seen == 1).all()
    # slices in node-index order, back to back
    roots = [v for v in range(I) if c[v]["nprim"] not in (0, 0xFFFF)]
    assert [int(c[v]["a"]) for v in roots] == list(np.cumsum([0] + [int(c[v]["nprim"]) for v in roots])[:-1])
    # the bottom-up cost of the root equals the traversal SAH numerator
    sah = oracle_lib.bvh_sah(c, 0, CI, CT)
    assert r["costs"][0] / np.float32(area(nodes[0])) == pytest.approx(sah, rel=2e-5)
    # collapsing never raises the AABB SAH, and every decision follows its children
    assert sah <= oracle_lib.bvh_sah(nodes, 0, CI, CT) * (1 + 1e-6)
    for v in range(I):
        a, b = int(nodes[v]["a"]), int(nodes[v]["b"])
        A = area(nodes[v])
        node_cost = CI * A + r["costs"][a] + r["costs"][b]
        leaf_cost = CT * A * r["counts"][v] if c[v]["nprim"] in (0, 0xFFFF) else None
        if table[v] and (v == 0 or not table[parents[v]]) and v != 0:
            assert c[v]["nprim"] not in (0, 0xFFFF)
        if not table[v] and leaf_cost is not None and int(r["counts"][v]) <= 8:
            assert leaf_cost >= node_cost * (1 - 1e-5)
    # the library's host SAH equals the oracle's, with and without OBBs
    assert ops.bvh_sah(c, 0, CI, CT) == sah
    obbs, _ = oracle_lib.obb_tree(nodes, parents, V, F)
    assert ops.bvh_sah(c, 0, CI, CT, obbs=obbs) == oracle_lib.bvh_sah(c, 0, CI, CT, obbs=obbs)


def test_sah_closed_form(oracle_lib):
    """Root with two leaves: (ci*A + ct*A_l + ct*A_r) / A; with OBBs the root
    and internal areas come from 2*ext, one-primitive leaves keep AABBs."""
    a = ((0, 0, 0), (1, 2, 3))
    b = ((1, 0, 0), (2, 1, 1))
    nodes, parents, counts = box_tree([a, b], [((0, 0, 0), (2, 2, 3))], [(1, 2)])
    A, Al, Ar = area(nodes[0]), area(nodes[1]), area(nodes[2])
    assert oracle_lib.bvh_sah(nodes, 0, CI, CT) == pytest.approx((CI * A + CT * Al + CT * Ar) / A, rel=1e-6)
    obbs = np.zeros((3, 15), dtype=np.float32)
    obbs[:, 12:15] = [[0.5, 0.75, 1.25], [9, 9, 9], [9, 9, 9]]
    Ao = 2.0 * (1.0 * 1.5 + 1.0 * 2.5 + 1.5 * 2.5)
    assert oracle_lib.bvh_sah(nodes, 0, CI, CT, obbs=obbs) == pytest.approx((CI * Ao + Al + Ar) / Ao, rel=1e-6)
    assert ops.bvh_sah(nodes, 0, CI, CT, obbs=obbs) == oracle_lib.bvh_sah(nodes, 0, CI, CT, obbs=obbs)


def test_sah_output_file(tmp_path):
    unit = ((0, 0, 0), (1, 1, 1))
    nodes, _, _ = box_tree([unit, unit], [unit], [(1, 2)])
    f = tmp_path / "sah.csv"
    q1 = ops.bvh_sah(nodes, 0, CI, CT, output_file=str(f))
    q2 = ops.bvh_sah(nodes, 0, CI, CT, output_file=str(f))
    assert q1 == q2
    assert f.read_text() == f"{q1:f},{q2:f},"
