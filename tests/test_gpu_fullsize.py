"""Full-size workloads on the GPU against the threaded oracle, bit-exact:

* the batched DiTO of 8192 log-uniform sets (64 .. 2^20 points, about
  885 M points) at the tile the bench runs (32 768-point tiles, one
  workgroup per tile), boxes and traces;
* the 2M-triangle hairball (50 000 tubes) through the ATRBVH builder
  (treelet 7, 2 passes), the OBB-tree transform (boxes, matrices and the 14
  extremal vertex ids of every node) and the collapse (ci 1.2, ct 1.0).

The oracle runs node- or set-parallel on the host's cores (its outputs do
not depend on the thread count)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from obvh_amd import batch, ops, synth  # noqa: E402

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


TRACE_INT_FIELDS = ["slab_min_idx", "slab_max_idx", "pair_k", "p2_sel", "q0_sel", "q1_sel", "q0_valid",
                    "q1_valid", "degenerate", "best", "result_aabb"]


def test_batched_dito_8192_sets_full_size(torch, oracle_lib):
    sizes = synth.batch_sizes(8192, lo=64, hi=1 << 20, seed=4)
    say(f"[fullsize dito] 8192 sets, {int(sizes.sum())} points, tile 32768, one workgroup per tile")
    x, y, z, off = batch.device_sets(sizes, seed=5)
    out, tr = ops.dito_batch(x, y, z, off, tile_points=32768, workers=0, trace=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    got_tr = ops.trace_to_numpy(tr)
    pts = np.stack([x.cpu().numpy(), y.cpu().numpy(), z.cpu().numpy()], 1)
    del x, y, z
    exp, etr = oracle_lib.points_obb_batch(pts, off.cpu().numpy().astype(np.uint64), threads=oracle_lib.host_threads())
    del pts
    assert np.array_equal(got, exp), np.nonzero(~(got == exp).all(axis=1))[0][:10]
    for f in TRACE_INT_FIELDS:
        assert np.array_equal(got_tr[f], etr[f]), f


def test_hairball_2m_atrbvh_transform_collapse(torch, oracle_lib):
    V, F = synth.hairball(seed=3)
    say(f"[fullsize hairball] {F.shape[0]} triangles: ATRBVH (treelet 7, 2 passes), transform, collapse")
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).cuda()
    t = ops.atrbvh_build(Vd, Fd, 7, 2, CI, CT)
    torch.cuda.synchronize()
    en, ep, ec, eco = oracle_lib.atrbvh_build(V, F, 7, 2, CI, CT)
    nodes = ops.nodes_to_numpy(t["nodes"])
    parents = t["parents"].cpu().numpy().view(np.uint32)
    counts = t["counts"].cpu().numpy().view(np.uint32)
    assert np.array_equal(parents, ep), "parents differ"
    assert np.array_equal(nodes.view(np.uint8), en.view(np.uint8)), "nodes differ"
    assert np.array_equal(counts, ec)
    assert np.array_equal(t["costs"].cpu().numpy(), eco)
    say("[fullsize hairball] transform")
    obbs, mats, ids = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd, with_ids=True)
    col = ops.collapse(t["nodes"], t["parents"], t["counts"], CI, CT)
    torch.cuda.synchronize()
    eo, em, eid = oracle_lib.obb_tree(nodes, parents, V, F, with_ids=True, threads=oracle_lib.host_threads())
    assert np.array_equal(ids.cpu().numpy().view(np.uint32).reshape(-1, 14), eid), "extremal vertex ids differ"
    go, gm = obbs.cpu().numpy(), mats.cpu().numpy()
    assert np.array_equal(go, eo), np.nonzero(~(go == eo).all(axis=1))[0][:10]
    assert np.array_equal(gm, em), np.nonzero(~(gm == em).all(axis=1))[0][:10]
    say("[fullsize hairball] collapse")
    e = oracle_lib.collapse(nodes, parents, counts, CI, CT)
    assert np.array_equal(col["collapsed"].cpu().numpy(), e["collapsed"].view(np.uint8))
    assert np.array_equal(col["primitives"].cpu().numpy().view(np.uint32)[: e["used"]], e["primitives"])
