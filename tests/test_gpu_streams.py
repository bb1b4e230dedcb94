"""The transform's second stream (leaf boxes forked off the caller's stream
and joined back inside the call) under the two uses the C-ABI promises:
capture into a HIP graph, replayed several times with eager work and
synchronisations between the replays, and calls made at the same time from
two host threads on two streams.  Every result is compared bit for bit with
a plain eager call, which is itself checked against the oracle."""
import sys
import threading

import numpy as np
import pytest

from obvh_amd import ops, synth

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


def device_tree(torch, V, F):
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(F.view(np.int32)).cuda()
    return Vd, Fd, ops.lbvh_build(Vd, Fd)


def bits(x):
    """float32 tensor -> its int32 bit patterns (NaN compares by bits)."""
    import torch

    return x.contiguous().view(torch.int32)


def checked_eager(torch, oracle_lib, V, F, Vd, Fd, t):
    obbs, mats = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd)
    torch.cuda.synchronize()
    eo, em = oracle_lib.obb_tree(ops.nodes_to_numpy(t["nodes"]), t["parents"].cpu().numpy().view(np.uint32), V, F)
    assert np.array_equal(obbs.cpu().numpy().view(np.uint32), eo.view(np.uint32))
    assert np.array_equal(mats.cpu().numpy().view(np.uint32), em.view(np.uint32))
    return obbs, mats


def outputs(torch, N):
    total = 2 * N - 1
    return (torch.full((total, 15), float("nan"), dtype=torch.float32, device="cuda"),
            torch.full((total, 16), float("nan"), dtype=torch.float32, device="cuda"))


def workspace(torch, N):
    return torch.empty(ops.lib().obvh_obb_tree_workspace_size(N, N - 1), dtype=torch.uint8, device="cuda")


def test_transform_captured_in_a_graph(torch, oracle_lib):
    say("[streams] obvh_obb_tree_async captured in a HIP graph, 4 replays")
    V, F = synth.icosphere(64, seed=11)  # 81920 faces: every climb phase runs
    Vd, Fd, t = device_tree(torch, V, F)
    eo, em = checked_eager(torch, oracle_lib, V, F, Vd, Fd, t)
    V2, F2 = synth.hairball(tubes=400, seed=12)
    Vd2, Fd2, t2 = device_tree(torch, V2, F2)
    e2o, e2m = checked_eager(torch, oracle_lib, V2, F2, Vd2, Fd2, t2)

    N = len(F)
    ws = workspace(torch, N)
    obbs, mats = outputs(torch, N)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up outside capture
        ops.obb_tree(t["nodes"], t["parents"], Vd, Fd, workspace=ws, out=(obbs, mats))
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ops.obb_tree(t["nodes"], t["parents"], Vd, Fd, workspace=ws, out=(obbs, mats))
    torch.cuda.synchronize()

    for rep in range(4):
        obbs.fill_(float("nan"))
        mats.fill_(float("nan"))
        # an eager call on another tree between replays shares the side stream
        o2, m2 = ops.obb_tree(t2["nodes"], t2["parents"], Vd2, Fd2)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        say(f"[streams]   replay {rep}")
        assert torch.equal(bits(obbs), bits(eo)), rep
        assert torch.equal(bits(mats), bits(em)), rep
        assert torch.equal(bits(o2), bits(e2o)) and torch.equal(bits(m2), bits(e2m)), rep


def test_concurrent_calls_from_two_threads(torch, oracle_lib):
    say("[streams] two host threads, two streams, 6 calls each")
    meshes = [synth.icosphere(64, seed=13), synth.hairball(tubes=3000, seed=14)]
    jobs = []
    for V, F in meshes:
        Vd, Fd, t = device_tree(torch, V, F)
        eo, em = checked_eager(torch, oracle_lib, V, F, Vd, Fd, t)
        N = len(F)
        jobs.append({"args": (t["nodes"], t["parents"], Vd, Fd), "expect": (eo, em), "ws": workspace(torch, N),
                     "outs": [outputs(torch, N) for _ in range(6)], "stream": torch.cuda.Stream()})
    torch.cuda.synchronize()
    start = threading.Barrier(len(jobs))
    errors = []

    def run(job):
        try:
            start.wait()
            for out in job["outs"]:
                ops.obb_tree(*job["args"], stream=job["stream"], workspace=job["ws"], out=out)
        except Exception as e:  # reported from the main thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(j,)) for j in jobs]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for k, job in enumerate(jobs):
        eo, em = job["expect"]
        for i, (o, m) in enumerate(job["outs"]):
            assert torch.equal(bits(o), bits(eo)) and torch.equal(bits(m), bits(em)), (k, i)
