"""Benchmark of the OBB path on one MI355X.

Workload (default): AABB-BVH -> OBB-tree transform of a 30-bit-Morton LBVH over
a seeded geodesic icosphere of frequency 224 (1 003 520 triangles, 501 762
vertices, radius 1 + U(-0.05, 0.05)).  The tree is built on the GPU by the
library's LBVH builder before timing; one step is one transform of the whole
resident tree (obbs + matrices for all 2 007 039 nodes).

value = internal nodes / wall time per step (Mnodes/s), inputs resident in HBM.
roofline = algorithmic bytes of one transform call
           (2N-1)(36 node + 4 parent + 60 obb + 64 mat4) + 16 N + 12 V
           over its device time (hipEvents on the launch stream).
cpu_baseline = the CPU oracle (a C port of the same method) on the same tree,
           node-parallel on the host's cores (at most 16 threads), timed on
           this host; cpu_baseline_1core = the same oracle on one thread.
           Both outputs are compared bit-exactly with the GPU output
           ("parity").
cpu_reference_transform = the reference's own CPU transform (kr_bvh_obb_tree,
           a full DiTO of every subtree; a different method) restated in C,
           single thread, on the whole tree.
roofline_c3 = the same roofline for the 2M-triangle hairball ATRBVH transform.
host = the host's core count, CPU model and the threads the CPU legs used.
Further legs (reported in the same JSON line): collapse + SAH of the tree,
OBVH packing and closest-hit traversal of 1M camera and 1M random rays
through the collapsed OBB tree, 2M-triangle transforms (LBVH and ATRBVH
inputs), batched DiTO, and the serving loop.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def pmc_traffic(profile: str):
    """HBM bytes per call from a committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes, calibrated on a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_call:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        return None
    return None


# The kernels of one transform call; the roofline's device time is the call's,
# on the launch stream.
CALL_KERNELS = ("obb_tree call: tree_clear, climb_leaves, climb_phase, candidates x 2 (the first on a second "
                "stream beside climb_phase), cand_redo, refit, finalize")


def pmc_value(profile: str, key: str):
    """A `key: value` line of a committed profile summary, or None."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1].strip())
    except OSError:
        return None
    return None


def host_info() -> dict:
    """Core count, CPU model and the threads the CPU legs use."""
    import oracle

    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"nproc": os.cpu_count(), "usable_cores": affinity, "cpu_model": model,
            "cpu_threads_used": oracle.host_threads()}


def copy_probe() -> int:
    """Two device copies of 1 GiB (torch), the known-size kernel the HBM
    counters are calibrated against."""
    import torch

    n = 1 << 28
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy probe: 2 copies of {n * 4} bytes", flush=True)
    return 0


def b_alg_transform(N: int, V: int) -> int:
    return (2 * N - 1) * (36 + 4 + 60 + 64) + 16 * N + 12 * V


def dito_legs(args) -> dict:
    """Batched DiTO over log-uniform sets (64 .. 2^20 points), the same sets fitted
    one call per set, and one 131 072-point fit (device resident and legacy)."""
    import numpy as np
    import torch

    from obvh_amd import batch, ops, synth

    out = {}
    sizes = synth.batch_sizes(args.sets, lo=64, hi=1 << 20, seed=4)
    t0 = time.perf_counter()
    x, y, z, off = batch.device_sets(sizes, seed=5)
    torch.cuda.synchronize()
    total = int(sizes.sum())
    print(f"[bench] dito batch: {args.sets} sets, {total} points generated in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    tuner = batch.Autotuner()
    entry = tuner.lookup(sizes)
    if entry is None:
        entry, _ = tuner.tune(x, y, z, off, sizes, reps=2)
    g = batch.BatchGraph(x, y, z, off, entry["tile_points"], entry["workers"])
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - w0) / reps
    dev = e0.elapsed_time(e1) / reps / 1e3
    batched = g.out.clone()
    out["batched"] = {
        "sets": args.sets, "points": total, "tile_points": entry["tile_points"], "workers": entry["workers"],
        "ms": round(dev * 1e3, 4), "wall_ms": round(wall * 1e3, 4),
        "Gpts_per_s": round(total / dev / 1e9, 2), "sets_per_s": round(args.sets / dev, 1),
        "achieved_GBs": round(24.0 * total / dev / 1e9, 1),
        "roofline_frac": round(24.0 * total / dev / 1e9 / HBM_PEAK_GBS, 4),
    }
    print(f"[bench] dito batched: {dev*1e3:.3f} ms, {total/dev/1e9:.1f} Gpts/s", file=sys.stderr, flush=True)
    offs = off.cpu().numpy()
    if args.no_plain:
        return out
    # plain per-request launches: one obvh_dito_soa call per set
    ws = torch.empty(ops.lib().obvh_dito_workspace_size(int(sizes.max())), dtype=torch.uint8, device="cuda")
    res = []
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for s in range(args.sets):
        a, b = int(offs[s]), int(offs[s + 1])
        res.append(ops.dito(x[a:b], y[a:b], z[a:b], workspace=ws))
    torch.cuda.synchronize()
    plain = time.perf_counter() - w0
    same = bool(torch.equal(torch.stack(res), batched))
    out["plain_per_set"] = {"s": round(plain, 4), "sets_per_s": round(args.sets / plain, 1),
                            "Gpts_per_s": round(total / plain / 1e9, 3), "equal_to_batched": same}
    out["batched_over_plain"] = round((args.sets / dev) / (args.sets / plain), 2)
    print(f"[bench] dito plain per set: {plain*1e3:.1f} ms ({out['batched_over_plain']}x slower), equal={same}",
          file=sys.stderr, flush=True)
    # one 131 072-point set: back-to-back calls of the C entry points (ctypes,
    # no allocation per call), so that the host does not set the pace
    import ctypes
    from obvh_amd import _lib
    p = synth.gaussian_points(131072, seed=1)
    xs, ys, zs = (torch.from_numpy(np.ascontiguousarray(p[:, k])).cuda() for k in range(3))
    ws1 = torch.empty(ops.lib().obvh_dito_workspace_size(131072), dtype=torch.uint8, device="cuda")
    o1 = torch.empty(15, dtype=torch.float32, device="cuda")
    L = _lib.lib()
    cs = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    fits = {
        "soa": lambda: _lib.check(L.obvh_dito_soa(vp(xs), vp(ys), vp(zs), 131072, vp(o1), None, vp(ws1),
                                                  ws1.numel(), cs)),
        "noalloc": lambda: _lib.check(L.obvh_dito_soa_noalloc(vp(xs), vp(ys), vp(zs), 131072, vp(o1), None, cs)),
    }
    fit_us = {}
    for name, fit in fits.items():
        for _ in range(10):
            fit()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(200):
            fit()
        e1.record()
        torch.cuda.synchronize()
        fit_us[name] = e0.elapsed_time(e1) / 200 * 1e3
    single = fit_us["noalloc"] / 1e6
    legacy = []
    for _ in range(5):
        w0 = time.perf_counter()
        r = ops.points_obb(p)
        legacy.append(time.perf_counter() - w0)
    out["single_131072"] = {"device_us": round(single * 1e6, 2), "Gpts_per_s": round(131072 / single / 1e9, 3),
                            "entry": "obvh_dito_soa_noalloc (two launches)",
                            "soa_device_us": round(fit_us["soa"], 2),
                            "legacy_wall_us": round(min(legacy) * 1e6, 1),
                            "equal_to_legacy": bool(np.array_equal(o1.cpu().numpy(),
                                                                   np.asarray(r, dtype=np.float32).reshape(-1)))}
    print(f"[bench] dito single 131072: {single*1e6:.1f} us device, legacy {min(legacy)*1e6:.0f} us wall",
          file=sys.stderr, flush=True)
    del o1, r
    if not args.no_cpu:
        import oracle

        # CPU DiTO on a bounded sample of the batch (sets 0..K until ~10 s of work)
        xh, yh, zh = x.cpu().numpy(), y.cpu().numpy(), z.cpu().numpy()
        done, pts, t_cpu = 0, 0, 0.0
        for s in range(args.sets):
            a, b = int(offs[s]), int(offs[s + 1])
            P = np.stack([xh[a:b], yh[a:b], zh[a:b]], 1)
            w0 = time.perf_counter()
            oracle.points_obb(P)
            t_cpu += time.perf_counter() - w0
            done += 1
            pts += b - a
            if t_cpu > 8.0:
                break
        out["cpu_baseline_1core"] = {"Gpts_per_s": round(pts / t_cpu / 1e9, 4),
                                     "sets_per_s": round(done / t_cpu, 1), "cores": 1, "kind": "port",
                                     "sample": f"first {done} sets of the batch ({pts} points, {t_cpu:.1f} s)"}
        # all host cores, sets in parallel, on the whole batch; compared with
        # the GPU's boxes set by set
        thr = oracle.host_threads()
        P = np.stack([xh, yh, zh], 1)
        del xh, yh, zh
        tc = time.perf_counter()
        exp, _ = oracle.points_obb_batch(P, offs, threads=thr)
        t_all = time.perf_counter() - tc
        del P
        out["parity"] = bool(np.array_equal(exp, batched.cpu().numpy()))
        out["cpu_baseline"] = {"Gpts_per_s": round(total / t_all / 1e9, 4), "sets_per_s": round(args.sets / t_all, 1),
                               "cores": thr, "kind": "port",
                               "sample": f"whole batch: {args.sets} sets ({total} points, {t_all:.2f} s, "
                                         f"{thr} threads, sets in parallel)"}
        print(f"[bench] cpu DiTO: whole batch in {t_all:.2f}s on {thr} threads, parity={out['parity']}",
              file=sys.stderr, flush=True)
    return out


def serving_leg(args) -> dict:
    """Open-loop Poisson stream (70% DiTO fits of 1k-256k points, 30% mesh ->
    LBVH -> OBB tree of 10k-2M triangles) under a 200 GB HBM budget, then a
    closed-loop burst of fits served vs one plain call per request."""
    import numpy as np
    import torch

    from obvh_amd import ops, serve

    torch.cuda.empty_cache()  # hand back what the DiTO legs cached
    t0 = time.perf_counter()
    fit_pool, tree_pool = serve.request_pool()
    print(f"[bench] serving pool ready in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    # warm-up run: every kernel loaded, the allocator's pools grown
    serve.run_open_loop(serve.Server(hbm_budget=200e9), serve.poisson_schedule(args.serve_rate, 2.0, seed=99),
                        fit_pool, tree_pool)
    sched = serve.poisson_schedule(args.serve_rate, args.serve_seconds)
    srv = serve.Server(hbm_budget=200e9, request_timeout=1.0)
    w0 = time.perf_counter()
    done = serve.run_open_loop(srv, sched, fit_pool, tree_pool)
    wall = time.perf_counter() - w0
    out = {"open_loop": serve.latency_stats(done)}
    out["open_loop"].update({"rate": args.serve_rate, "seconds": args.serve_seconds,
                             "req_per_s": round(len(done) / wall, 1), "fit_batches": srv.batches,
                             "mean_fits_per_batch": round(float(np.mean(srv.batch_sizes)), 2),
                             "tree_batches": srv.tree_batches, "request_timeout_s": 1.0,
                             "timeouts": srv.timeouts})
    print(f"[bench] serving open loop: {json.dumps(out['open_loop'])}", file=sys.stderr, flush=True)
    # closed loop: a burst of fits, all arrived at t = 0
    burst = 4096
    srv = serve.Server(hbm_budget=200e9)
    for i in range(burst):
        srv.submit(serve.Request(i, "fit", 0.0, points=fit_pool[i % len(fit_pool)]))
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    while not srv.idle():
        srv.step(time.perf_counter() - w0)
    served = time.perf_counter() - w0
    ws = torch.empty(ops.lib().obvh_dito_workspace_size(1 << 18), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for i in range(burst):
        pts = fit_pool[i % len(fit_pool)].to("cuda", non_blocking=True)
        box = ops.dito(pts[0], pts[1], pts[2], workspace=ws).cpu()
    plain = time.perf_counter() - w0
    out["closed_loop_fits"] = {"requests": burst, "served_sets_per_s": round(burst / served, 1),
                               "plain_sets_per_s": round(burst / plain, 1),
                               "served_over_plain": round(plain / served, 2)}
    print(f"[bench] serving closed loop: {json.dumps(out['closed_loop_fits'])}", file=sys.stderr, flush=True)
    del box
    return out


def traversal_leg(V, F, Vd, Fd, nodes, mats, prims, label: str, steps: int, warmup: int, with_cpu: bool,
                  width: int = 1024) -> dict:
    """Pack the (collapsed) OBB tree and trace closest-hit queries through it:
    a width x width pinhole camera (coherent) and as many random rays from a
    sphere around the mesh (incoherent).  Per batch: device time and Mrays/s
    of the native record layout (the headline) and of the reference's SoA
    layout (the drop-in), whose outputs must be identical; the CPU oracle
    re-traces a sample of every 16th ray (bit-exact parity and single-thread
    CPU Mrays/s)."""
    import numpy as np
    import torch

    from obvh_amd import ops, synth

    ops.soa_pack(nodes, mats)
    ops.records_pack(nodes, mats, Vd, Fd, primitives=prims)
    torch.cuda.synchronize()
    p0, p1, p2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    p0.record()
    packed = ops.soa_pack(nodes, mats)
    p1.record()
    recs = ops.records_pack(nodes, mats, Vd, Fd, primitives=prims)
    p2.record()
    torch.cuda.synchronize()
    out = {"pack_ms": round(p0.elapsed_time(p1), 5), "records_pack_ms": round(p1.elapsed_time(p2), 5),
           "batches": {}}
    nh = L = R = M = ph = None
    for kind in ("camera", "random"):
        rays = synth.camera_rays(V, width, width) if kind == "camera" else synth.random_rays(V, width * width)
        n = rays.shape[0]
        rd = torch.from_numpy(rays.view(np.uint8)).cuda()
        res = torch.empty(n * 28, dtype=torch.uint8, device="cuda")
        res_r = torch.empty(n * 28, dtype=torch.uint8, device="cuda")
        ws = torch.empty(ops.lib().obvh_soa_intersect_workspace_size(n), dtype=torch.uint8, device="cuda")

        def timed(fn):
            for _ in range(warmup):
                fn()
            torch.cuda.synchronize()
            evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                   for _ in range(steps)]
            t0 = time.perf_counter()
            for k in range(steps):
                evs[k][0].record()
                fn()
                evs[k][1].record()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / steps, sum(a.elapsed_time(b) for a, b in evs) / steps

        wall_r, dev_r = timed(lambda: ops.records_intersect(recs, rd, workspace=ws, out=res_r))
        wall, dev = timed(lambda: ops.soa_intersect(packed, Vd, Fd, rd, primitives=prims, workspace=ws, out=res))
        got = ops.isects_to_numpy(res)
        rec = {"rays": n, "ms_per_batch": round(wall_r, 5), "device_ms": round(dev_r, 5),
               "Mrays_per_s": round(n / (wall_r * 1e-3) / 1e6, 2),
               "soa": {"ms_per_batch": round(wall, 5), "device_ms": round(dev, 5),
                       "Mrays_per_s": round(n / (wall * 1e-3) / 1e6, 2)},
               "records_equal_soa": bool(torch.equal(res, res_r)),
               "hit_fraction": round(float(np.count_nonzero(got["primitive"] != 0xFFFFFFFF)) / n, 4),
               "aborted": int(np.count_nonzero(got["instance"] != 0))}
        if with_cpu:
            import oracle

            if nh is None:
                nh = ops.nodes_to_numpy(nodes)
                L, R, M = oracle.obvh_pack(nh, mats.cpu().numpy())
                ph = None if prims is None else prims.cpu().numpy().view(np.uint32)
            sample = rays[::16]
            tc = time.perf_counter()
            exp, st = oracle.obvh_intersect(L, R, M, V, F, ph, sample, with_stats=True)
            cpu_s = time.perf_counter() - tc
            g = np.ascontiguousarray(got[::16])
            rec["parity"] = bool(np.array_equal(g.view(np.uint8), np.ascontiguousarray(exp).view(np.uint8)))
            rec["nodes_per_ray"] = round(float(st[:, 0].mean()), 2)
            rec["triangles_per_ray"] = round(float(st[:, 1].mean()), 2)
            rec["max_stack"] = int(st[:, 2].max())
            rec["cpu_baseline_1core"] = {"value": round(sample.shape[0] / cpu_s / 1e6, 4), "unit": "Mrays/s",
                                         "cores": 1, "kind": "port",
                                         "sample": f"every 16th ray ({sample.shape[0]} rays, {cpu_s:.2f} s)"}
            # all host cores on the whole batch, compared ray by ray with the GPU
            thr = oracle.host_threads()
            tc = time.perf_counter()
            exp_all = oracle.obvh_intersect(L, R, M, V, F, ph, rays, threads=thr)
            cpu_all = time.perf_counter() - tc
            rec["parity"] = rec["parity"] and bool(
                np.array_equal(np.ascontiguousarray(got).view(np.uint8), np.ascontiguousarray(exp_all).view(np.uint8)))
            rec["cpu_baseline"] = {"value": round(n / cpu_all / 1e6, 4), "unit": "Mrays/s", "cores": thr,
                                   "kind": "port",
                                   "sample": f"whole batch ({n} rays, {cpu_all:.2f} s, {thr} threads, rays in parallel)"}
        print(f"[bench] trace {label} {kind}: {n} rays, records {dev_r:.4f} ms device "
              f"({rec['Mrays_per_s']} Mrays/s), SoA {dev:.4f} ms ({rec['soa']['Mrays_per_s']} Mrays/s), "
              f"records == SoA {rec['records_equal_soa']}, hits {rec['hit_fraction']}, "
              f"oracle parity {rec.get('parity')}", file=sys.stderr, flush=True)
        out["batches"][kind] = rec
        del rd, res, res_r, ws
    return out


def transform_leg(V, F, label: str, steps: int, warmup: int, with_cpu: bool, builder: str = "lbvh",
                  trace: bool = False) -> dict:
    """LBVH (or ATRBVH: treelet 7, 2 passes) of the mesh built on the GPU and
    its build timed (mean of 5 LBVH builds, one ATRBVH build), then `steps`
    timed transforms of the resident tree, then the collapse; optionally
    the CPU oracle on the same tree (bit-exact comparison and single-thread
    CPU throughput)."""
    import numpy as np
    import torch

    from obvh_amd import ops

    N, NV = F.shape[0], V.shape[0]
    Vd = torch.from_numpy(V).cuda()
    Fd = torch.from_numpy(F.view(np.int32)).cuda()
    build_ms = None
    if builder == "atrbvh":
        ops.atrbvh_build(Vd, Fd, 7, 2)  # warm
        torch.cuda.synchronize()
        b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        b0.record()
        tree = ops.atrbvh_build(Vd, Fd, 7, 2)
        b1.record()
        torch.cuda.synchronize()
        build_ms = b0.elapsed_time(b1)
        print(f"[bench] atrbvh build {label}: {build_ms:.3f} ms device", file=sys.stderr, flush=True)
    else:
        ops.lbvh_build(Vd, Fd)  # warm
        torch.cuda.synchronize()
        b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        b0.record()
        for _ in range(5):
            tree = ops.lbvh_build(Vd, Fd)
        b1.record()
        torch.cuda.synchronize()
        build_ms = b0.elapsed_time(b1) / 5
        print(f"[bench] lbvh build {label}: {build_ms:.3f} ms device", file=sys.stderr, flush=True)
    total = 2 * N - 1
    obbs = torch.empty((total, 15), dtype=torch.float32, device="cuda")
    mats = torch.empty((total, 16), dtype=torch.float32, device="cuda")
    ws = torch.empty(ops.lib().obvh_obb_tree_workspace_size(N, N - 1), dtype=torch.uint8, device="cuda")

    def step():
        ops.obb_tree(tree["nodes"], tree["parents"], Vd, Fd, workspace=ws, out=(obbs, mats))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record()
        step()
        evs[k][1].record()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / steps
    dev_ms = sorted(a.elapsed_time(b) for a, b in evs)
    dev_ms_mean = sum(dev_ms) / len(dev_ms)
    internal = N - 1
    value = internal / (ms_step * 1e-3) / 1e6
    balg = b_alg_transform(N, NV)
    achieved = balg / (dev_ms_mean * 1e-3) / 1e9
    print(f"[bench] transform {label}: {N} triangles, {ms_step:.4f} ms/step wall, {dev_ms_mean:.4f} ms device "
          f"(min {dev_ms[0]:.4f}), {value:.1f} Mnodes/s, {achieved:.1f} GB/s", file=sys.stderr, flush=True)
    # collapse after the transform (ci = 1.2, ct = 1.0) and the tree quality
    cws = torch.empty(ops.lib().obvh_collapse_workspace_size(N), dtype=torch.uint8, device="cuda")
    col = ops.collapse(tree["nodes"], tree["parents"], tree["counts"], workspace=cws)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(5):
        col = ops.collapse(tree["nodes"], tree["parents"], tree["counts"], workspace=cws)
    c1.record()
    torch.cuda.synchronize()
    collapse_ms = c0.elapsed_time(c1) / 5
    nodes_h = ops.nodes_to_numpy(tree["nodes"])
    coll_h = ops.nodes_to_numpy(col["collapsed"])
    obbs_h = obbs.cpu().numpy()
    sah = {"aabb": round(ops.bvh_sah(nodes_h), 4), "aabb_collapsed": round(ops.bvh_sah(coll_h), 4),
           "obb_collapsed": round(ops.bvh_sah(coll_h, obbs=obbs_h), 4)}
    print(f"[bench] collapse {label}: {collapse_ms:.4f} ms device; SAH {sah}", file=sys.stderr, flush=True)
    out = {"builder": builder, "build_ms": None if build_ms is None else round(build_ms, 4),
           "triangles": N, "vertices": NV, "internal_nodes": internal, "steps": steps,
           "collapse_ms": round(collapse_ms, 5), "sah": sah,
           "ms_per_step": round(ms_step, 5), "device_ms": round(dev_ms_mean, 5),
           "Mnodes_per_s": round(value, 2), "bytes_per_call": balg, "achieved_GBs": round(achieved, 1),
           "roofline_frac": round(achieved / HBM_PEAK_GBS, 4), "cpu_baseline": None, "parity": None,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_call": balg,
                        "device_ms": round(dev_ms_mean, 5)}}
    if with_cpu:
        import oracle

        parents_h = tree["parents"].cpu().numpy().view(np.uint32)
        tc = time.perf_counter()
        eo, em = oracle.obb_tree(nodes_h, parents_h, V, F)
        cpu_s = time.perf_counter() - tc
        thr = oracle.host_threads()
        tc = time.perf_counter()
        eo_mt, em_mt = oracle.obb_tree(nodes_h, parents_h, V, F, threads=thr)
        cpu_mt = time.perf_counter() - tc
        go, gm = obbs.cpu().numpy(), mats.cpu().numpy()
        out["parity"] = bool(np.array_equal(go, eo) and np.array_equal(gm, em)
                             and np.array_equal(go, eo_mt) and np.array_equal(gm, em_mt))
        out["cpu_baseline"] = {
            "value": round(internal / cpu_mt / 1e6, 4), "unit": "Mnodes/s", "cores": thr, "kind": "port",
            "sample": f"whole workload: one transform of the {N}-triangle tree ({cpu_mt:.2f} s, {thr} threads, "
                      f"node-parallel)",
        }
        out["cpu_baseline_1core"] = {
            "value": round(internal / cpu_s / 1e6, 4), "unit": "Mnodes/s", "cores": 1, "kind": "port",
            "sample": f"whole workload: one transform of the {N}-triangle tree ({cpu_s:.2f} s, single thread)",
        }
        print(f"[bench] cpu oracle {label}: {cpu_s:.2f}s single thread, {cpu_mt:.2f}s on {thr} threads, "
              f"parity={out['parity']}", file=sys.stderr, flush=True)
        # the reference's own CPU transform (kr_bvh_obb_tree: a full DiTO of
        # every subtree), a different method, on the whole tree
        tc = time.perf_counter()
        _, _, _, fitted = oracle.cpu_obb_tree(nodes_h, parents_h, V, F)
        ref_s = time.perf_counter() - tc
        out["cpu_reference_transform"] = {
            "value": round(fitted / ref_s / 1e6, 4), "unit": "Mnodes/s", "cores": 1, "kind": "port",
            "sample": f"whole workload: kr_bvh_obb_tree restated, all {fitted} internal nodes "
                      f"({ref_s:.2f} s, single thread)",
        }
        print(f"[bench] reference CPU transform {label}: {fitted} nodes in {ref_s:.2f}s", file=sys.stderr,
              flush=True)
    if trace:
        try:
            out["traversal"] = traversal_leg(V, F, Vd, Fd, col["collapsed"], mats, col["primitives"], label,
                                             max(5, steps // 4), 2, with_cpu)
        except Exception as e:
            print(f"[bench] traversal leg failed: {e!r}", file=sys.stderr, flush=True)
            out["traversal"] = {"error": repr(e)}
    return out


def tree_batch_leg(args) -> dict:
    """1000 LBVH trees of 10k..200k triangles (log-uniform; noisy icospheres of
    frequency round(sqrt(n / 20)), one mesh per frequency) through one
    obvh_obb_tree_batch call, against 1000 single-tree calls
    (obvh_obb_tree_async, one after another on one stream, shared workspace).
    Every box and matrix of the batch is compared with the single-tree calls'
    ones, and one tree per frequency with the oracle."""
    import numpy as np
    import torch

    from obvh_amd import ops, synth

    rng = np.random.default_rng(11)
    n = np.exp(rng.uniform(np.log(10_000), np.log(200_000), size=args.trees))
    freqs = np.clip(np.round(np.sqrt(n / 20.0)).astype(int), 1, None)
    meshes, trees_by_f = {}, {}
    for f in sorted(set(freqs.tolist())):
        V, F = synth.icosphere(int(f), seed=int(f))
        Vd = torch.from_numpy(V).cuda()
        Fd = torch.from_numpy(F.view(np.int32)).cuda()
        t = ops.lbvh_build(Vd, Fd)
        meshes[f] = (V, F)
        trees_by_f[f] = (t["nodes"], t["parents"], Vd, Fd)
    trees = [trees_by_f[int(f)] for f in freqs]
    b = ops.pack_trees(trees)
    total, K = b["total"], b["ntrees"]
    ws = torch.empty(ops.lib().obvh_obb_tree_batch_workspace_size(K, total), dtype=torch.uint8, device="cuda")
    obbs = torch.empty((2 * total - K, 15), dtype=torch.float32, device="cuda")
    mats = torch.empty((2 * total - K, 16), dtype=torch.float32, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    ops.obb_tree_batch(b, workspace=ws, out=(obbs, mats), status=status)
    torch.cuda.synchronize()
    reps = 3
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record()
    for _ in range(reps):
        ops.obb_tree_batch(b, workspace=ws, out=(obbs, mats))
    e1.record()
    torch.cuda.synchronize()
    wall_b = (time.perf_counter() - w0) / reps
    dev_b = e0.elapsed_time(e1) / reps / 1e3
    # plain: one single-tree call per tree, outputs into the same layout
    off = b["prim_offsets_host"]
    nmax = int(max(F.shape[0] for F in (m[1] for m in meshes.values())))
    ws1 = torch.empty(ops.lib().obvh_obb_tree_workspace_size(nmax, nmax - 1), dtype=torch.uint8, device="cuda")
    po = torch.empty_like(obbs)
    pm = torch.empty_like(mats)

    def plain():
        for t, (nd, par, Vd, Fd) in enumerate(trees):
            s, e = 2 * int(off[t]) - t, 2 * int(off[t + 1]) - t - 1
            ops.obb_tree(nd, par, Vd, Fd, workspace=ws1, out=(po[s:e], pm[s:e]))

    plain()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record()
    plain()
    e1.record()
    torch.cuda.synchronize()
    wall_p = time.perf_counter() - w0
    dev_p = e0.elapsed_time(e1) / 1e3
    same = bool(torch.equal(po, obbs) and torch.equal(pm, mats))
    # one tree per frequency against the oracle
    import oracle

    thr = oracle.host_threads()
    first = {}
    for t, f in enumerate(freqs.tolist()):
        first.setdefault(int(f), t)
    oracle_ok = True
    go, gm = obbs.cpu().numpy(), mats.cpu().numpy()
    for f, t in first.items():
        V, F = meshes[f]
        nd, par, _, _ = trees_by_f[f]
        eo, em = oracle.obb_tree(ops.nodes_to_numpy(nd), par.cpu().numpy().view(np.uint32), V, F, threads=thr)
        s, e = 2 * int(off[t]) - t, 2 * int(off[t + 1]) - t - 1
        oracle_ok &= bool(np.array_equal(go[s:e], eo) and np.array_equal(gm[s:e], em))
    out = {"trees": K, "triangles": total, "internal_nodes": total - K,
           "batched": {"ms": round(dev_b * 1e3, 3), "wall_ms": round(wall_b * 1e3, 3),
                       "trees_per_s": round(K / dev_b, 1), "Mnodes_per_s": round((total - K) / dev_b / 1e6, 1)},
           "plain_per_tree": {"ms": round(dev_p * 1e3, 3), "wall_ms": round(wall_p * 1e3, 3),
                              "trees_per_s": round(K / max(dev_p, wall_p), 1)},
           "batched_over_plain": round(max(dev_p, wall_p) / max(dev_b, wall_b), 2),
           "status": int(status.item()), "equal_to_plain": same, "oracle_parity": oracle_ok,
           "distinct_meshes": len(first)}
    print(f"[bench] tree batch: {K} trees, {total} triangles: batched {dev_b*1e3:.2f} ms, plain "
          f"{max(dev_p, wall_p)*1e3:.2f} ms ({out['batched_over_plain']}x), equal={same}, oracle={oracle_ok}",
          file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--freq", type=int, default=224, help="icosphere frequency (224 -> 1 003 520 triangles)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-dito", action="store_true", help="skip the DiTO legs")
    ap.add_argument("--sets", type=int, default=8192, help="batched DiTO: number of sets")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-call-per-set DiTO comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-2m", action="store_true", help="skip the 2M-triangle transform legs")
    ap.add_argument("--no-trace", action="store_true", help="skip the OBVH packing + traversal leg")
    ap.add_argument("--trees", type=int, default=1000, help="batched transform: number of trees")
    ap.add_argument("--no-tree-batch", action="store_true", help="skip the batched tree transform leg")
    ap.add_argument("--only", choices=["c2", "c3", "trees"], default=None,
                    help="time only one transform workload and print its leg as JSON (profiling runs): "
                         "c2 = the icosphere LBVH, c3 = the 2M-triangle hairball ATRBVH")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run two 1 GiB device copies (calibration of the HBM byte counters)")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="open-loop serving duration (s)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="open-loop arrival rate (req/s)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the OBB path runs on a single MI355X; "
              "sharding over several GPUs is not launched", file=sys.stderr)
        return 2

    import torch

    from obvh_amd import synth

    torch.cuda.set_device(0)
    dev = torch.cuda.get_device_name(0)
    if args.only == "trees":
        print(json.dumps({"only": "trees", **tree_batch_leg(args)}))
        return 0
    if args.only:
        V, F = synth.icosphere(args.freq, seed=2) if args.only == "c2" else synth.hairball(seed=3)
        leg = transform_leg(V, F, args.only, args.steps, args.warmup, False, "lbvh" if args.only == "c2" else "atrbvh")
        print(json.dumps({"only": args.only, **leg}))
        return 0
    t_gen = time.perf_counter()
    V, F = synth.icosphere(args.freq, seed=2)
    print(f"[bench] device={dev}", file=sys.stderr, flush=True)
    main_leg = transform_leg(V, F, f"icosphere f={args.freq}", args.steps, args.warmup, not args.no_cpu,
                             trace=not args.no_trace)
    print(f"[bench] main workload ready and timed in {time.perf_counter() - t_gen:.1f}s", file=sys.stderr, flush=True)
    N, NV, internal = main_leg["triangles"], main_leg["vertices"], main_leg["internal_nodes"]
    ms_step, dev_ms_mean, value, achieved = (main_leg["ms_per_step"], main_leg["device_ms"],
                                              main_leg["Mnodes_per_s"], main_leg["achieved_GBs"])
    balg, cpu, parity = main_leg["bytes_per_call"], main_leg["cpu_baseline"], main_leg["parity"]

    extra = None
    if not args.no_2m:
        extra = {}
        for name, gen, builder in (("icosphere_f316_lbvh", lambda: synth.icosphere(316, seed=2), "lbvh"),
                                   ("hairball_2M_atrbvh", lambda: synth.hairball(seed=3), "atrbvh")):
            try:
                V2, F2 = gen()
                extra[name] = transform_leg(V2, F2, name, max(5, args.steps // 4), 2, not args.no_cpu, builder)
                del V2, F2
            except Exception as e:
                print(f"[bench] {name} leg failed: {e!r}", file=sys.stderr, flush=True)
                extra[name] = {"error": repr(e)}

    tree_batch = None
    if not args.no_tree_batch:
        try:
            tree_batch = tree_batch_leg(args)
        except Exception as e:
            print(f"[bench] tree batch leg failed: {e!r}", file=sys.stderr, flush=True)
            tree_batch = {"error": repr(e)}

    dito = None
    if not args.no_dito:
        try:
            dito = dito_legs(args)
        except Exception as e:  # the transform line must still be reported
            print(f"[bench] DiTO legs failed: {e!r}", file=sys.stderr, flush=True)
            dito = {"error": repr(e)}

    served = None
    if not args.no_serve:
        try:
            served = serving_leg(args)
        except Exception as e:
            print(f"[bench] serving leg failed: {e!r}", file=sys.stderr, flush=True)
            served = {"error": repr(e)}

    roofline_c3 = None
    c3 = (extra or {}).get("hairball_2M_atrbvh")
    if c3 and "roofline" in c3:
        roofline_c3 = dict(c3["roofline"])
        roofline_c3["traffic"] = pmc_traffic("transform_hairball2m_pmc.txt")
        roofline_c3["workload"] = (f"ATRBVH (treelet 7, 2 passes) -> OBB-tree transform, hairball of 50000 tubes: "
                                   f"{c3['triangles']} triangles, {c3['internal_nodes']} internal nodes")
        roofline_c3["kernel"] = CALL_KERNELS
    # The whole call against the VALU issue peak: the profiled VALU wave
    # instructions of one call over this run's device time per call.
    issue_peak = 1024 * 2.4e9 / 4
    valu_call = {}
    for key, prof, ms in (("c2", "transform_icosphere224_valu.txt", dev_ms_mean),
                          ("c3", "transform_hairball2m_valu.txt",
                           (c3 or {}).get("roofline", {}).get("device_ms"))):
        vw = pmc_value(prof, "call_valu_wave_instructions")
        if vw and ms and (key != "c2" or args.freq == 224):
            valu_call[key] = {"bound": "valu_issue", "achieved": round(vw / (ms * 1e-3) / 1e9, 1),
                              "peak": round(issue_peak / 1e9, 1), "unit": "G wave-instructions/s",
                              "frac": round(vw / (ms * 1e-3) / issue_peak, 4), "wave_instructions": int(vw),
                              "source": "profiles/" + prof}
    valu = None
    vi, vus = (pmc_value("transform_icosphere224_valu.txt", "candidates_valu_wave_instructions_per_call"),
               pmc_value("transform_icosphere224_valu.txt", "candidates_avg_us"))
    if vi and vus:
        # 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 VALU instruction
        peak = 1024 * 2.4e9 / 4
        valu = {"bound": "valu_issue", "kernel": "candidates (the launch over the tile-completed nodes)", "achieved": round(vi / (vus * 1e-6) / 1e9, 1),
                "peak": round(peak / 1e9, 1), "unit": "G wave-instructions/s",
                "frac": round(vi / (vus * 1e-6) / peak, 4), "source": "profiles/transform_icosphere224_valu.txt"}

    line = {
        "metric": "OBB-tree transform Mnodes/s; batched DiTO Gpts/s; served p50/p99 ms",
        "value": round(value, 2),
        "unit": "Mnodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded geodesic icosphere, radial noise)",
        "config": {
            "workload": f"LBVH->OBB-tree transform, icosphere f={args.freq}: {N} triangles, "
                        f"{internal} internal nodes",
            "triangles": N,
            "vertices": NV,
            "internal_nodes": internal,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic("transform_icosphere224_pmc.txt") if args.freq == 224 else None,
            "kernel": CALL_KERNELS,
            "bytes_per_call": balg,
            "device_ms": round(dev_ms_mean, 5),
        },
        "roofline_c3": roofline_c3,
        "valu_roofline_candidates": valu,
        "valu_roofline_call": valu_call or None,
        "cpu_baseline": cpu,
        "cpu_reference_transform": main_leg.get("cpu_reference_transform"),
        "host": host_info(),
        "parity": parity,
        "lbvh_build_ms": main_leg["build_ms"],
        "collapse_ms": main_leg["collapse_ms"],
        "sah": main_leg["sah"],
        "traversal": main_leg.get("traversal"),
        "transform_2m": extra,
        "tree_batch": tree_batch,
        "dito": dito,
        "serving": served,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
