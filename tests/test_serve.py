"""Serving loop: admission by HBM budget (CPU) and alone-vs-mix equality (GPU)."""
import numpy as np
import pytest

from obvh_amd import serve


class _Shape:
    def __init__(self, *shape):
        self.shape = shape


def test_admission_respects_budget_and_order():
    srv = serve.Server(hbm_budget=serve.fit_bytes(1000) * 3 + 10)
    reqs = [serve.Request(i, "fit", 0.0, points=_Shape(3, 1000)) for i in range(5)]
    for r in reqs:
        srv.submit(r)
    fits, trees = srv._admit()
    assert [r.rid for r in fits] == [0, 1, 2] and trees == []
    assert srv.in_use == 3 * serve.fit_bytes(1000)
    assert [r.rid for r in srv.waiting] == [3, 4]
    srv.in_use -= reqs[0].nbytes  # one result came back
    fits, _ = srv._admit()
    assert [r.rid for r in fits] == [3]


def test_request_larger_than_budget_is_rejected():
    srv = serve.Server(hbm_budget=1000)
    with pytest.raises(ValueError):
        srv.submit(serve.Request(0, "fit", 0.0, points=_Shape(3, 10 ** 6)))


def test_batch_point_cap_splits_batches():
    srv = serve.Server(hbm_budget=1e12, max_batch_points=2500)
    for i in range(4):
        srv.submit(serve.Request(i, "fit", 0.0, points=_Shape(3, 1000)))
    first, _ = srv._admit()
    second, _ = srv._admit()
    assert [r.rid for r in first] == [0, 1] and [r.rid for r in second] == [2, 3]


def test_poisson_schedule_rate_and_mix():
    s = serve.poisson_schedule(2000.0, 5.0, seed=1)
    assert abs(len(s) - 10000) < 400
    assert all(a[0] < b[0] for a, b in zip(s, s[1:]))
    share = sum(1 for x in s if x[1] == "fit") / len(s)
    assert abs(share - 0.7) < 0.02


@pytest.mark.gpu
def test_results_alone_equal_results_in_the_mix(oracle_lib):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("[serve] alone vs mix", flush=True)
    fit_pool, tree_pool = serve.request_pool(seed=3, fits=12, trees=3, fit_range=(1000, 60000),
                                             tree_range=(2000, 40000))
    sched = serve.poisson_schedule(5000.0, 0.03, seed=4)  # ~150 requests, overlapping
    mixed = {r.rid: r for r in serve.run_open_loop(serve.Server(keep_trees=True), sched, fit_pool, tree_pool)}
    assert len(mixed) == len(sched)
    alone_srv = serve.Server(keep_trees=True)
    for rid, (at, kind, pick) in enumerate(sched):
        if kind == "fit":
            r = serve.Request(rid, "fit", 0.0, points=fit_pool[pick % len(fit_pool)])
        else:
            V, F = tree_pool[pick % len(tree_pool)]
            r = serve.Request(rid, "tree", 0.0, vertices=V, faces=F)
        alone_srv.submit(r)
        while not alone_srv.idle():
            alone_srv.step(0.0)
    alone = {r.rid: r for r in alone_srv.finished}
    for rid, r in mixed.items():
        a = alone[rid]
        if r.kind == "fit":
            assert np.array_equal(r.result, a.result)
            assert np.array_equal(r.result, oracle_lib.points_obb(r.points.numpy().T))
        else:
            assert torch.equal(r.result["obbs"], a.result["obbs"]) and torch.equal(r.result["mats"], a.result["mats"])
            assert np.array_equal(r.result["root"], r.result["obbs"][0].cpu().numpy())
    # the mixed run really overlapped work: more than one fit per batch on average
    assert any(r.kind == "tree" for r in mixed.values())
    # resident trees stay charged to the budget until released
    srv = serve.Server(keep_trees=True)
    held = [r for r in mixed.values() if r.kind == "tree"]
    for r in held:
        srv.in_use += r.nbytes
        srv.release(r)
    assert srv.in_use == 0 and all(set(r.result) == {"root"} for r in held)


def test_waiting_request_times_out():
    srv = serve.Server(hbm_budget=serve.fit_bytes(1000) * 2, request_timeout=0.5)
    r = serve.Request(0, "fit", 0.0, points=_Shape(3, 1000))
    srv.submit(r)
    srv.in_use = srv.budget  # no room: the request keeps waiting
    srv.step(0.4)
    assert r.error is None and list(srv.waiting) == [r]
    srv.step(0.6)
    assert r.error and "timed out" in r.error and not srv.waiting and srv.finished == [r]
    assert srv.timeouts == 1


class _Event:
    def __init__(self):
        self.ready = False

    def query(self):
        return self.ready


def test_inflight_request_is_answered_at_its_timeout():
    """A request past its timeout on the device is answered with an error at
    once; its bytes return to the budget only when the device work ends."""
    srv = serve.Server(hbm_budget=1e9, request_timeout=1.0)
    r = serve.Request(0, "fit", 0.0, points=_Shape(3, 1000))
    r.nbytes = 100
    srv.in_use = 100
    ev, calls = _Event(), []

    def finish(now, r=r):
        calls.append(now)
        srv.in_use -= r.nbytes
        if not r.abandoned:
            r.result = "box"
            srv.finished.append(r)

    srv.inflight.append((ev, finish, [r]))
    srv.poll(0.5)
    assert r.error is None and srv.finished == []
    srv.poll(1.5)
    assert r.abandoned and "timed out" in r.error and srv.finished == [r] and srv.in_use == 100
    ev.ready = True
    srv.poll(2.0)
    assert len(calls) == 1 and calls[0] >= 2.0 and srv.in_use == 0 and r.result is None and srv.finished == [r]


@pytest.mark.gpu
def test_tree_burst_runs_as_one_batch_and_matches_alone():
    """Tree requests admitted together go through one batched transform;
    every tree equals the same request served alone (single-tree path)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("[serve] tree burst as one batch", flush=True)
    _, tree_pool = serve.request_pool(seed=5, fits=1, trees=6, tree_range=(500, 30000))
    srv = serve.Server(keep_trees=True)
    for i, (V, F) in enumerate(tree_pool):
        srv.submit(serve.Request(i, "tree", 0.0, vertices=V, faces=F))
    while not srv.idle():
        srv.step(0.0)
    assert srv.tree_batches == 1 and len(srv.finished) == len(tree_pool)
    for r in srv.finished:
        alone = serve.Server(keep_trees=True)
        a = serve.Request(r.rid, "tree", 0.0, vertices=r.vertices, faces=r.faces)
        alone.submit(a)
        while not alone.idle():
            alone.step(0.0)
        assert r.error is None and a.error is None
        assert torch.equal(r.result["obbs"], a.result["obbs"]) and torch.equal(r.result["mats"], a.result["mats"])
        assert np.array_equal(r.result["root"], a.result["root"])
