"""GPU parity on non-finite input, and the device status of the stream-ordered
transform.

DiTO (dito.cpp:116-173, :193-206) starts every running extreme at point 0
and compares strictly: a NaN never wins, a NaN at point 0 is never replaced.
The transform keeps the first vertex of a node in DFS order in the same way
(transforms.cu:216-411) and reduces each triangle from its vertex 0 before
the refit's strict range update (:855-897).  The oracle restates those rules
(pinned in test_oracle_dito.py / test_oracle_tree.py); here the kernels must
give the same boxes, traces and ids, bit for bit (NaN where the oracle has
NaN)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from obvh_amd import ops, synth  # noqa: E402


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def soa(torch, p):
    p = np.ascontiguousarray(p, np.float32)
    return [torch.from_numpy(np.ascontiguousarray(p[:, k])).cuda() for k in range(3)]


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


TRACE_INT_FIELDS = ["slab_min_idx", "slab_max_idx", "pair_k", "p2_sel", "q0_sel", "q1_sel", "q0_valid",
                    "q1_valid", "degenerate", "best", "result_aabb"]


def _nan_at(n, rows, cols=(0, 1, 2), seed=0, value=np.nan):
    p = synth.gaussian_points(n, seed=seed)
    for r in rows:
        p[r, list(cols)] = value
    return p


CASES = {
    "nan_point0": lambda: _nan_at(5000, [0], seed=1),
    "nan_point0_x": lambda: _nan_at(5000, [0], cols=(0,), seed=2),
    "nan_mid": lambda: _nan_at(5000, [1234, 4999], seed=3),
    "nan_mid_z": lambda: _nan_at(131072, [70000], cols=(2,), seed=4),
    "pos_inf": lambda: _nan_at(3000, [17], cols=(0,), seed=5, value=np.inf),
    "neg_inf": lambda: _nan_at(3000, [40], cols=(1,), seed=6, value=-np.inf),
    "inf_point0": lambda: _nan_at(3000, [0], cols=(2,), seed=7, value=np.inf),
    "mixed_inf": lambda: _nan_at(3000, [5, 9], cols=(0, 1), seed=8, value=np.inf),
    "all_nan": lambda: np.full((700, 3), np.nan, np.float32),
    "small_nan0": lambda: _nan_at(10, [0], seed=9),
    "small_nan_mid": lambda: _nan_at(12, [6], seed=10),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_dito_soa_nonfinite(torch, oracle_lib, case):
    say(f"[dito_soa nonfinite] case={case}")
    p = CASES[case]()
    exp, etr = oracle_lib.points_obb(p, trace=True)
    x, y, z = soa(torch, p)
    out, tr = ops.dito(x, y, z, trace=True)
    torch.cuda.synchronize()
    assert same(out.cpu().numpy(), exp), (out.cpu().numpy(), exp)
    g = ops.trace_to_numpy(tr)[0]
    for f in TRACE_INT_FIELDS:
        assert np.array_equal(g[f], etr[f]), f


@pytest.mark.parametrize("case", ["nan_point0", "nan_mid", "pos_inf", "all_nan"])
def test_legacy_points_obb_nonfinite(torch, oracle_lib, case):
    say(f"[kr_cuda_points_obb nonfinite] case={case}")
    p = CASES[case]()
    exp = oracle_lib.points_obb(p)
    assert not np.isnan(exp).all()  # identity axes: never the legacy entry's all-NaN error value
    o = ops.points_obb(p)
    assert same(o, exp), (o, exp)


@pytest.mark.parametrize("tile", [0, 512])
def test_dito_batch_nonfinite_sets_mid_batch(torch, oracle_lib, tile):
    say(f"[dito_batch nonfinite] tile={tile}")
    sets = [synth.gaussian_points(n, seed=20 + n) for n in (3000, 513, 64, 20000)]
    sets.insert(1, CASES["nan_point0"]())
    sets.insert(3, CASES["nan_mid"]())
    sets.insert(4, CASES["all_nan"]())
    sets.insert(6, CASES["pos_inf"]())
    sets.append(CASES["small_nan_mid"]())
    pts = np.concatenate(sets).astype(np.float32)
    off = np.zeros(len(sets) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in sets])
    exp, etr = oracle_lib.points_obb_batch(pts, off.astype(np.uint64))
    x, y, z = soa(torch, pts)
    offs = torch.from_numpy(off).cuda()
    out, tr = ops.dito_batch(x, y, z, offs, tile_points=tile, trace=True, host_offsets=off)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    gtr = ops.trace_to_numpy(tr)[: len(sets)]
    for s in range(len(sets)):
        assert same(got[s], exp[s]), s
        for f in TRACE_INT_FIELDS:
            assert np.array_equal(gtr[s][f], etr[s][f]), (s, f)


def test_dito_batch_invalid_offsets_give_nan_boxes(torch, oracle_lib):
    # offsets that decrease or run past the points: that set gets a NaN box
    # and degenerate = 4; the other sets are unaffected
    say("[dito_batch] invalid offsets")
    pts = synth.gaussian_points(4000, seed=30)
    off = np.array([0, 1000, 600, 4000, 9000], np.int64)  # set 1 decreases, set 3 runs past
    x, y, z = soa(torch, pts)
    out, tr = ops.dito_batch(x, y, z, torch.from_numpy(off).cuda(), trace=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    g = ops.trace_to_numpy(tr)
    assert np.isnan(got[1]).all() and np.isnan(got[3]).all()
    assert int(g[1]["degenerate"]) == 4 and int(g[3]["degenerate"]) == 4
    assert np.array_equal(got[0], oracle_lib.points_obb(pts[:1000]))
    assert np.array_equal(got[2], oracle_lib.points_obb(pts[600:4000]))
    with pytest.raises(ValueError):
        ops.dito_batch(x, y, z, torch.from_numpy(off).cuda(), host_offsets=off)


@pytest.mark.parametrize("kind", ["nan", "inf"])
@pytest.mark.parametrize("f", [16, 60])
def test_obb_tree_nonfinite_vertex(torch, oracle_lib, kind, f):
    # tree built on the finite mesh, then one vertex made non-finite; f = 60
    # gives 72 000 triangles: the tile holding the vertex takes the
    # per-triangle rule up to the root, the other tiles their tile-wide pass
    say(f"[obb_tree nonfinite] kind={kind} f={f}")
    V, F = synth.icosphere(f, seed=3)
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(F.view(np.int32)).cuda()
    t = ops.lbvh_build(Vd, Fd)
    W = V.copy()
    w = int(F[len(F) // 3, 1])
    W[w] = np.nan if kind == "nan" else np.array([np.inf, 0.25, -np.inf], np.float32)
    Wd = torch.from_numpy(W).cuda()
    st = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    obbs, mats, ids = ops.obb_tree(t["nodes"], t["parents"], Wd, Fd, with_ids=True, status=st)
    torch.cuda.synchronize()
    assert int(st.item()) == 0
    nodes = ops.nodes_to_numpy(t["nodes"])
    parents = t["parents"].cpu().numpy().view(np.uint32)
    eo, em, ei = oracle_lib.obb_tree(nodes, parents, W, F, with_ids=True)
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), ei)
    go, gm = obbs.cpu().numpy(), mats.cpu().numpy()
    bad = [v for v in range(len(go)) if not (same(go[v], eo[v]) and same(gm[v], em[v]))]
    assert not bad, (bad[:10], go[bad[0]], eo[bad[0]])


def test_obb_tree_status_reports_malformed_tree(torch):
    say("[obb_tree status] malformed parents")
    V, F = synth.icosphere(4, seed=5)
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(F.view(np.int32)).cuda()
    t = ops.lbvh_build(Vd, Fd)
    host = torch.full((1,), -1, dtype=torch.int32).pin_memory()
    ops.obb_tree(t["nodes"], t["parents"], Vd, Fd, status=host)
    torch.cuda.synchronize()
    assert int(host.item()) == 0 and ops.tree_status_message(0) is None
    parents = t["parents"].clone()
    parents[7] = 3  # disagrees with the children of node 3 (indices stay in range)
    ops.obb_tree(t["nodes"], parents, Vd, Fd, status=host)
    torch.cuda.synchronize()
    word = int(host.item())
    assert word != 0
    assert "parent/child" in ops.tree_status_message(word) or "never completed" in ops.tree_status_message(word)


@pytest.mark.parametrize("kind", ["all", "one"])
def test_obb_tree_huge_coordinates(torch, oracle_lib, kind):
    """Finite coordinates beyond 2^61 in magnitude: the candidate search takes
    the reference's compare-and-select mode (the separate redo launch) for
    every node that holds such a point, all nodes ("all": the mesh scaled by
    2^62) or the ancestors of one far vertex ("one"), the rest their
    distinct-point search."""
    say(f"[obb_tree huge] kind={kind}")
    V, F = synth.icosphere(8, seed=6)
    Vd0 = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(F.view(np.int32)).cuda()
    t = ops.lbvh_build(Vd0, Fd)
    W = V.copy()
    if kind == "all":
        W *= np.float32(2.0 ** 62)
    else:
        W[int(F[len(F) // 2, 0])] = np.array([3.0e18, -2.5e18, 1.0e18], np.float32)
    Wd = torch.from_numpy(W).cuda()
    obbs, mats, ids = ops.obb_tree(t["nodes"], t["parents"], Wd, Fd, with_ids=True)
    torch.cuda.synchronize()
    nodes = ops.nodes_to_numpy(t["nodes"])
    parents = t["parents"].cpu().numpy().view(np.uint32)
    eo, em, ei = oracle_lib.obb_tree(nodes, parents, W, F, with_ids=True)
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), ei)
    go, gm = obbs.cpu().numpy(), mats.cpu().numpy()
    bad = [v for v in range(len(go)) if not (same(go[v], eo[v]) and same(gm[v], em[v]))]
    assert not bad, (bad[:10], go[bad[0]], eo[bad[0]])
