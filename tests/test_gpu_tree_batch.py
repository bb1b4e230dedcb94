"""obvh_obb_tree_batch: many trees in one transform.  Every tree's boxes and
matrices must equal the single-tree call's (obvh_obb_tree_async) and the
oracle's, bit for bit, for batches of uneven LBVH and ATRBVH trees
(including one- and two-triangle trees), of random topologies (every span
gapped) and of degenerate meshes (flat grids, doubled vertices); the call
replays from a HIP graph;
a malformed tree is reported through the batch status word."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from obvh_amd import ops, synth  # noqa: E402
from obvh_amd.layouts import NODE_DTYPE  # noqa: E402
from treeutil import doubled_vertices, flat_grid, random_tree  # noqa: E402


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def meshes(seed):
    rng = np.random.default_rng(seed)
    out = [(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2, 0]], np.uint32)),
           (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]], np.float32),
            np.array([[0, 1, 2, 0], [1, 3, 2, 0]], np.uint32))]
    for k in range(28):
        if k % 2:
            out.append(synth.icosphere(int(rng.integers(1, 24)), seed=int(rng.integers(1 << 30))))
        else:
            out.append(synth.hairball(tubes=int(rng.integers(1, 300)), seed=int(rng.integers(1 << 30))))
    return out


def build_trees(oracle_lib, ms, builder):
    trees = []
    for k, (V, F) in enumerate(ms):
        if builder in ("lbvh", "degenerate") or F.shape[0] < 2:
            nodes, parents, *_ = oracle_lib.lbvh_build(V, F)
        elif builder == "random":  # shuffled leaves, breadth-first internal nodes
            nodes, parents = random_tree(F.shape[0], seed=1000 + k)
        else:
            nodes, parents, *_ = oracle_lib.atrbvh_build(V, F, 7, 2)
        trees.append((nodes, parents, V, F))
    return trees


def to_dev(torch, trees):
    dev = []
    for nodes, parents, V, F in trees:
        dev.append((torch.from_numpy(np.ascontiguousarray(nodes).view(np.uint8)).cuda(),
                    torch.from_numpy(np.ascontiguousarray(parents).view(np.int32)).cuda(),
                    torch.from_numpy(np.ascontiguousarray(V)).cuda(),
                    torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).cuda()))
    return dev


def degenerate_meshes():
    """Flat grids (every node coplanar) and icospheres with doubled vertices,
    between one- and two-triangle meshes."""
    ms = meshes(14)[:2]
    for k in range(10):
        ms.append(flat_grid(8 + 7 * k, seed=40 + k) if k % 2 == 0 else doubled_vertices(seed=40 + k))
    return ms


@pytest.mark.parametrize("builder", ["lbvh", "atrbvh", "random", "degenerate"])
def test_batch_equals_single_tree_calls_and_oracle(torch, oracle_lib, builder):
    ms = degenerate_meshes() if builder == "degenerate" else meshes({"lbvh": 7, "atrbvh": 8, "random": 13}[builder])
    say(f"[tree batch] {builder}: {len(ms)} trees, {sum(F.shape[0] for _, F in ms)} triangles")
    trees = build_trees(oracle_lib, ms, builder)
    dev = to_dev(torch, trees)
    b = ops.pack_trees(dev)
    status = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    obbs, mats = ops.obb_tree_batch(b, status=status)
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    go, gm = obbs.cpu().numpy(), mats.cpu().numpy()
    off = b["prim_offsets_host"]
    for t, ((nodes, parents, V, F), (nd, pr, Vd, Fd)) in enumerate(zip(trees, dev)):
        n = int(F.shape[0])
        s = 2 * int(off[t]) - t
        so, sm = ops.obb_tree(nd, pr, Vd, Fd)
        eo, em = oracle_lib.obb_tree(nodes, parents, V, F)
        torch.cuda.synchronize()
        assert np.array_equal(go[s:s + 2 * n - 1], so.cpu().numpy()), t
        assert np.array_equal(gm[s:s + 2 * n - 1], sm.cpu().numpy()), t
        assert np.array_equal(go[s:s + 2 * n - 1], eo), t
        assert np.array_equal(gm[s:s + 2 * n - 1], em), t


def test_batch_replays_from_a_graph(torch, oracle_lib):
    say("[tree batch] graph capture, 3 replays")
    trees = build_trees(oracle_lib, meshes(9)[:12], "lbvh")
    b = ops.pack_trees(to_dev(torch, trees))
    eager = [a.clone() for a in ops.obb_tree_batch(b)]
    torch.cuda.synchronize()
    ws = torch.empty(ops.lib().obvh_obb_tree_batch_workspace_size(b["ntrees"], b["total"]), dtype=torch.uint8,
                     device="cuda")
    out = (torch.empty_like(eager[0]), torch.empty_like(eager[1]))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ops.obb_tree_batch(b, stream=s, workspace=ws, out=out)  # warm-up outside the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        ops.obb_tree_batch(b, stream=s, workspace=ws, out=out)
    for k in range(3):
        out[0].fill_(0.0)
        out[1].fill_(0.0)
        junk = torch.randn(1 << 18, device="cuda").sum()  # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out[0], eager[0]) and torch.equal(out[1], eager[1]), k
        del junk


def test_malformed_tree_sets_the_batch_status(torch, oracle_lib):
    say("[tree batch] a tree with a bad child index")
    trees = build_trees(oracle_lib, meshes(10)[:6], "lbvh")
    nodes = trees[3][0].copy()
    nodes[0]["b"] = 10 ** 6  # right child of the root outside the tree
    trees[3] = (nodes,) + trees[3][1:]
    b = ops.pack_trees(to_dev(torch, trees))
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    ops.obb_tree_batch(b, status=status)
    torch.cuda.synchronize()
    word = int(status.item())
    assert word != 0 and ops.tree_status_message(word) is not None
