"""GPU collapse against the oracle: bit-exact collapsed nodes, counts, costs
and primitive slices; legacy synchronous entry equals the async one."""
import sys

import numpy as np
import pytest

from obvh_amd import ops, synth
from obvh_amd.layouts import NODE_DTYPE
from treeutil import fill_bounds, random_tree

pytestmark = pytest.mark.gpu

CI, CT = 1.2, 1.0


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def device_tree(torch, kind):
    if kind == "random":
        V, F = synth.hairball(tubes=300, seed=8)
        nodes, parents = random_tree(F.shape[0], seed=9)
        nodes = fill_bounds(nodes, V, F)
        counts = np.zeros(nodes.shape[0], dtype=np.uint32)
        counts[(nodes.shape[0] - 1) // 2:] = 1
        return (torch.from_numpy(nodes.view(np.uint8)).cuda(), torch.from_numpy(parents.view(np.int32)).cuda(),
                torch.from_numpy(counts.view(np.int32)).cuda())
    if kind == "ico":
        V, F = synth.icosphere(40, seed=3)
    elif kind == "ico224":
        V, F = synth.icosphere(224, seed=2)
    else:
        V, F = synth.hairball(tubes=2000, seed=4)
    t = ops.lbvh_build(torch.from_numpy(V).cuda(), torch.from_numpy(F.view(np.int32)).cuda())
    return t["nodes"], t["parents"], t["counts"]


@pytest.mark.parametrize("kind", ["ico", "hair", "random", "ico224"])
def test_collapse_matches_oracle(torch, oracle_lib, kind):
    say(f"[collapse] {kind}")
    nodes, parents, counts = device_tree(torch, kind)
    r = ops.collapse(nodes, parents, counts, CI, CT)
    torch.cuda.synchronize()
    nh = ops.nodes_to_numpy(nodes)
    ph = parents.cpu().numpy().view(np.uint32)
    ch = counts.cpu().numpy().view(np.uint32)
    e = oracle_lib.collapse(nh, ph, ch, CI, CT)
    got = r["collapsed"].cpu().numpy()
    assert np.array_equal(got, e["collapsed"].view(np.uint8)), "collapsed nodes differ"
    assert np.array_equal(r["counts"].cpu().numpy().view(np.uint32), e["counts"])
    assert np.array_equal(r["costs"].cpu().numpy(), e["costs"])
    used = e["used"]
    assert np.array_equal(r["primitives"].cpu().numpy().view(np.uint32)[:used], e["primitives"])
    cn = got.view(NODE_DTYPE)
    assert ops.bvh_sah(cn, 0, CI, CT) == oracle_lib.bvh_sah(e["collapsed"], 0, CI, CT)


def test_legacy_collapse_equals_async(torch):
    say("[collapse] legacy")
    nodes, parents, counts = device_tree(torch, "hair")
    r = ops.collapse(nodes, parents, counts, CI, CT)
    total = int(parents.numel())
    out = torch.empty_like(nodes)
    cnt = counts.clone()
    prims = torch.full(((total + 1) // 2,), -1, dtype=torch.int32, device="cuda")
    costs = torch.empty(total, dtype=torch.float32, device="cuda")
    m = ops.bvh_collapse(nodes, parents, out, cnt, prims, costs, CI, CT)
    assert m["collapse_mark_time"] > 0 and m["collapse_time"] > 0
    assert torch.equal(out, r["collapsed"]) and torch.equal(cnt, r["counts"])
    assert torch.equal(costs, r["costs"]) and torch.equal(prims, r["primitives"])
