"""The C-ABI library loads without a GPU and exports exactly what
include/obvh/obvh.h declares; host-only entry points work on the CPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from obvh_amd import _lib
from obvh_amd.layouts import NODE_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "obvh", "obvh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:kr|obvh)_[a-z0-9_]+)\s*\(", text)))


def test_header_and_export_list_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    L = _lib.lib()
    for name in _lib.EXPORTS:
        assert hasattr(L, name), name
    assert L.obvh_version().startswith(b"obvh")


def test_struct_sizes():
    assert C.sizeof(_lib.KrObb3) == 60
    assert NODE_DTYPE.itemsize == 36


def test_workspace_sizes_are_positive():
    L = _lib.lib()
    assert L.obvh_dito_workspace_size(131072) > 0
    assert L.obvh_dito_batch_workspace_size(8192, 10 ** 8, 16384) > L.obvh_dito_batch_workspace_size(8192, 10 ** 6, 16384)
    assert L.obvh_obb_tree_workspace_size(1 << 20, (1 << 20) - 1) > 112 * (1 << 20)


def test_host_tree_validation(oracle_lib):
    from obvh_amd import ops, synth

    V, F = synth.icosphere(4)
    nodes, parents, *_ = oracle_lib.lbvh_build(V, F)
    ops.validate_tree(nodes, parents)
    bad = parents.copy()
    bad[3] = 9
    with pytest.raises(_lib.ObvhError):
        ops.validate_tree(nodes, bad)
    cyc = nodes.copy()
    cyc[1]["a"] = 0  # child points back at the root
    with pytest.raises(_lib.ObvhError):
        ops.validate_tree(cyc, parents)


def test_product_path_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from obvh_amd import ops

    with pytest.raises(_lib.ObvhError):
        ops.dito(torch.zeros(4), torch.zeros(4), torch.zeros(4))


def test_noalloc_dito_scratch_covers_its_limit():
    """obvh_dito_soa_noalloc's static 1 MiB scratch holds obvh_dito_soa's
    workspace for every count up to OBVH_DITO_NOALLOC_MAX_POINTS; larger
    counts are rejected before any device call."""
    L = _lib.lib()
    limit = 1 << 28
    ns = np.unique(np.concatenate([np.arange(0, 20000, 7), np.logspace(4, np.log10(limit), 4000).astype(np.int64),
                                   [limit - 1, limit]]))
    assert max(L.obvh_dito_workspace_size(int(n)) for n in ns) <= 1 << 20
    one = C.c_void_p(16)
    err = L.obvh_dito_soa_noalloc(one, one, one, limit + 1, one, None, None)
    assert err and b"OBVH_DITO_NOALLOC_MAX_POINTS" in err
