"""Batched DiTO on the GPU: graph capture, autotuner, device-generated sets."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from obvh_amd import batch, ops, synth  # noqa: E402


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_graph_replays_match_eager(torch, oracle_lib):
    say("[batch graph] 300 sets, 5 replays with eager work between")
    sizes = np.concatenate([[0, 1, 14, 15], synth.batch_sizes(296, lo=64, hi=1 << 14, seed=21)])
    x, y, z, off = batch.device_sets(sizes, seed=22)
    eager = ops.dito_batch(x, y, z, off, tile_points=2048, workers=512)
    torch.cuda.synchronize()
    eager_h = eager.cpu().numpy()
    # the eager result equals the oracle on the same bytes
    pts = torch.stack([x, y, z], 1).cpu().numpy()
    exp, _ = oracle_lib.points_obb_batch(pts, off.cpu().numpy().astype(np.uint64))
    assert np.array_equal(eager_h, exp)
    g = batch.BatchGraph(x, y, z, off, tile_points=2048, workers=512)
    for k in range(5):
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        out = g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), eager_h), k
        del junk


def test_autotuner_picks_a_measured_config(torch, tmp_path):
    say("[autotune] small sweep")
    sizes = synth.batch_sizes(64, lo=64, hi=1 << 15, seed=23)
    x, y, z, off = batch.device_sets(sizes, seed=24)
    tuner = batch.Autotuner(path=str(tmp_path / "tune.json"))
    entry, results = tuner.tune(x, y, z, off, sizes, reps=1, tiles=(2048, 8192), workers=(256, 1024))
    assert (entry["tile_points"], entry["workers"]) in {(t, w) for _, t, w in results}
    assert entry["ms"] == pytest.approx(min(r[0] for r in results), rel=1e-3, abs=1e-3)
    again = batch.Autotuner(path=str(tmp_path / "tune.json"))
    assert again.lookup(sizes)["tile_points"] == entry["tile_points"]
    # every candidate gives the same boxes
    ref = ops.dito_batch(x, y, z, off, tile_points=2048, workers=256)
    for _, t, w in results:
        assert torch.equal(ops.dito_batch(x, y, z, off, tile_points=t, workers=w), ref)


def test_set_upload_and_gather(torch):
    """obvh_upload_sets (2-D copies) and obvh_gather_sets (one kernel reading
    pinned host memory) place every set at its offset, empty sets included;
    a pageable set is refused before any launch."""
    import ctypes as C

    from obvh_amd._lib import ObvhError, check, lib

    say("[upload] 2-D copies and the gather kernel")
    sizes = [0, 5, 1, 0, 1000, 77777, 3, 0]
    sets = [torch.arange(3 * n, dtype=torch.float32).reshape(3, n).add_(1000.0 * k).pin_memory()
            for k, n in enumerate(sizes)]
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    row = total + 13  # rows padded: the row stride differs from the total
    expect = torch.zeros((3, row), dtype=torch.float32)
    for k, (s, n) in enumerate(zip(sets, sizes)):
        expect[:, int(off[k]):int(off[k + 1])] = s
    ptrs = (C.c_void_p * len(sets))(*[s.data_ptr() for s in sets])
    stream = torch.cuda.current_stream()
    a = torch.zeros((3, row), dtype=torch.float32, device="cuda")
    check(lib().obvh_upload_sets(ptrs, off.ctypes.data_as(C.c_void_p), len(sets), C.c_void_p(a.data_ptr()), row,
                                 C.c_void_p(stream.cuda_stream)))
    b = torch.zeros((3, row), dtype=torch.float32, device="cuda")
    ptr_d = torch.tensor([s.data_ptr() for s in sets], dtype=torch.int64).cuda()
    off_d = torch.from_numpy(off).cuda()
    check(lib().obvh_gather_sets(ptrs, C.c_void_p(ptr_d.data_ptr()), off.ctypes.data_as(C.c_void_p),
                                 C.c_void_p(off_d.data_ptr()), len(sets), C.c_void_p(b.data_ptr()), row,
                                 C.c_void_p(stream.cuda_stream)))
    torch.cuda.synchronize()
    assert torch.equal(a.cpu(), expect) and torch.equal(b.cpu(), expect)
    # a pageable (not pinned) set: refused on the host, nothing launched
    pageable = torch.ones((3, 10), dtype=torch.float32)
    ptrs2 = (C.c_void_p * 1)(pageable.data_ptr())
    off2 = np.array([0, 10], dtype=np.int64)
    with pytest.raises(ObvhError):
        check(lib().obvh_gather_sets(ptrs2, C.c_void_p(ptr_d.data_ptr()), off2.ctypes.data_as(C.c_void_p),
                                     C.c_void_p(off_d.data_ptr()), 1, C.c_void_p(b.data_ptr()), row,
                                     C.c_void_p(stream.cuda_stream)))
