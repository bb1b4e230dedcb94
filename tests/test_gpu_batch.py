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
