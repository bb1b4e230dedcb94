"""obvh_dito_soa_noalloc (static device scratch, no allocation) against
obvh_dito_soa (caller workspace), both called through the C ABI: boxes and
traces bit-identical on two calls in a row, the second with fewer points, so
that the second call runs on scratch the first one left behind."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from obvh_amd import _lib, synth  # noqa: E402
from obvh_amd.layouts import TRACE_DTYPE  # noqa: E402


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def _p(t):
    return C.c_void_p(t.data_ptr())


def _fit_both(torch, pts):
    L = _lib.lib()
    n = pts.shape[0]
    x, y, z = (torch.from_numpy(np.ascontiguousarray(pts[:, k])).cuda() for k in range(3))
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ws = torch.empty(max(int(L.obvh_dito_workspace_size(n)), 256), dtype=torch.uint8, device="cuda")
    out_ws = torch.full((15,), -1.0, dtype=torch.float32, device="cuda")
    out_na = torch.full((15,), -2.0, dtype=torch.float32, device="cuda")
    tr_ws = torch.zeros(TRACE_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    tr_na = torch.full((TRACE_DTYPE.itemsize,), 0x5A, dtype=torch.uint8, device="cuda")
    _lib.check(L.obvh_dito_soa(_p(x), _p(y), _p(z), n, _p(out_ws), _p(tr_ws), _p(ws), ws.numel(), stream))
    _lib.check(L.obvh_dito_soa_noalloc(_p(x), _p(y), _p(z), n, _p(out_na), _p(tr_na), stream))
    torch.cuda.synchronize()
    return (out_ws.cpu().numpy(), tr_ws.cpu().numpy().view(TRACE_DTYPE)[0],
            out_na.cpu().numpy(), tr_na.cpu().numpy().view(TRACE_DTYPE)[0])


@pytest.mark.parametrize("sizes", [(131072, 5000), (3_000_000, 14)])
def test_noalloc_equals_workspace_entry(torch, sizes):
    for i, n in enumerate(sizes):
        print(f"[dito_soa_noalloc] call {i} n={n}", flush=True)
        sys.stdout.flush()
        pts = synth.gaussian_points(n, seed=40 + i)
        a, ta, b, tb = _fit_both(torch, pts)
        assert a.view(np.uint32).tolist() == b.view(np.uint32).tolist(), (n, a, b)
        for f in TRACE_DTYPE.names:
            if f != "pad":
                assert np.array_equal(ta[f], tb[f]), (n, f)


def test_noalloc_rejects_oversized_count(torch):
    L = _lib.lib()
    out = torch.empty(15, dtype=torch.float32, device="cuda")
    x = torch.zeros(1, dtype=torch.float32, device="cuda")
    err = L.obvh_dito_soa_noalloc(_p(x), _p(x), _p(x), (1 << 28) + 1, _p(out), None, None)
    assert err and b"OBVH_DITO_NOALLOC_MAX_POINTS" in err
