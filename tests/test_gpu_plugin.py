"""The intersector plugins end to end on the GPU, driven like the
reference's host drives lib<name>.so (CREATE, INIT, COMMIT of a kr_scene
with host mesh arrays, INTERSECTION_QUERY with device rays and hits,
DESTROY): the hits equal, byte for byte, those of the same build, transform,
collapse, packing and traversal called entry point by entry point, and the
benchmark file gets the reference's SAH figures and timing line."""
import ctypes as C
import sys

import numpy as np
import pytest

import korangar_abi as K
from obvh_amd import ops, synth
from obvh_amd.layouts import ISECT_DTYPE, MISS

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


def expected_hits(torch, V, F, rays_d, builder, collapse):
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(np.ascontiguousarray(F).view(np.int32)).cuda()
    t = ops.lbvh_build(Vd, Fd) if builder == "lbvh_gpu" else ops.atrbvh_build(Vd, Fd, 9, 2, ci=1.2, ct=1.0)
    obbs, mats = ops.obb_tree(t["nodes"], t["parents"], Vd, Fd)
    nodes, prims = t["nodes"], None
    if collapse:
        col = ops.collapse(t["nodes"], t["parents"], t["counts"], ci=1.2, ct=1.0)
        nodes, prims = col["collapsed"], col["primitives"]
    out = ops.soa_intersect(ops.soa_pack(nodes, mats), Vd, Fd, rays_d, primitives=prims)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("name,collapse", [("lbvh_gpu", True), ("lbvh_gpu", False), ("atrbvh_gpu", True)])
def test_plugin_commit_and_query(torch, tmp_path, name, collapse):
    say(f"[plugin] {name} collapse={collapse}")
    _, fn = K.load(name)
    V, F = synth.hairball(tubes=600, seed=21) if name == "atrbvh_gpu" else synth.icosphere(30, seed=22)
    rays = synth.random_rays(V, 30000, seed=23)
    rays_d = torch.from_numpy(rays.view(np.uint8)).cuda()
    bench = tmp_path / "bench.csv"

    ads = K.Ads()
    ads.context = fn(None, None, K.ACTION["CREATE"])
    assert ads.context
    cont, keep = K.settings(collapse="y" if collapse else "n", use_obbs="y", use_persistent_kernel="y",
                            ct="1.0", ci="1.2")
    assert fn(C.byref(ads), C.byref(cont), K.ACTION["INIT"]) is None
    scene, keep2 = K.mesh_scene(V, F, benchmark_file=str(bench))
    err = fn(C.byref(ads), C.byref(scene), K.ACTION["COMMIT"])
    assert err is None, K.message(err)

    isects = torch.full((rays.shape[0] * ISECT_DTYPE.itemsize,), 0xAB, dtype=torch.uint8, device="cuda")
    q = K.Query(K.QUERY_TYPE_CUDA, rays_d.data_ptr(), isects.data_ptr(), rays.shape[0])
    err = fn(C.byref(ads), C.byref(q), K.ACTION["INTERSECTION_QUERY"])
    assert err is None, K.message(err)
    torch.cuda.synchronize()
    got = isects.cpu().numpy()
    exp = expected_hits(torch, V, F, rays_d, name, collapse)
    assert np.array_equal(got, exp)
    hits = got.view(ISECT_DTYPE)
    assert np.count_nonzero(hits["primitive"] != MISS) > rays.shape[0] // 10

    # a second query on the same context gives the same records
    isects.fill_(0)
    assert fn(C.byref(ads), C.byref(q), K.ACTION["INTERSECTION_QUERY"]) is None
    torch.cuda.synchronize()
    assert np.array_equal(isects.cpu().numpy(), exp)

    # kr_bvh_sah appends "%f," per stage (LBVH plugin only: AABB, OBB, then
    # after a collapse both again), the timing line follows on the same line
    text = bench.read_text()
    assert text.endswith("\n") and text.count("\n") == 1
    vals = [float(x) for x in text.split(",")]
    nsah = (4 if collapse else 2) if name == "lbvh_gpu" else 0
    assert len(vals) == nsah + 14
    sah, timing = vals[:nsah], vals[nsah:]
    assert timing[10] > 0 and timing[11] > 0  # build and transform totals
    if collapse:
        assert timing[12] > 0
    if sah:
        assert all(s > 0 for s in sah)
        assert sah[1] < sah[0]  # OBBs bound tighter than AABBs
    assert fn(C.byref(ads), None, K.ACTION["DESTROY"]) is None


def test_plugin_recommit_replaces_the_tree(torch, tmp_path):
    say("[plugin] lbvh_gpu COMMIT twice on one context")
    _, fn = K.load("lbvh_gpu")
    ads = K.Ads()
    ads.context = fn(None, None, K.ACTION["CREATE"])
    cont, keep = K.settings(collapse="y", use_obbs="y", use_persistent_kernel="y", ct="1.0", ci="1.2")
    assert fn(C.byref(ads), C.byref(cont), K.ACTION["INIT"]) is None
    V1, F1 = synth.icosphere(12, seed=31)
    V2, F2 = synth.icosphere(20, seed=32)
    rays = synth.random_rays(V2, 5000, seed=33)
    rays_d = torch.from_numpy(rays.view(np.uint8)).cuda()
    isects = torch.zeros(rays.shape[0] * ISECT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    q = K.Query(K.QUERY_TYPE_CUDA, rays_d.data_ptr(), isects.data_ptr(), rays.shape[0])
    for V, F in ((V1, F1), (V2, F2)):
        scene, keep2 = K.mesh_scene(V, F)
        assert fn(C.byref(ads), C.byref(scene), K.ACTION["COMMIT"]) is None
    assert fn(C.byref(ads), C.byref(q), K.ACTION["INTERSECTION_QUERY"]) is None
    torch.cuda.synchronize()
    assert np.array_equal(isects.cpu().numpy(), expected_hits(torch, V2, F2, rays_d, "lbvh_gpu", True))
    assert fn(C.byref(ads), None, K.ACTION["DESTROY"]) is None
