"""Serving loop: continuous batching of OBB requests on one GPU.

Two request kinds:
* ``fit``   DiTO-14 box of one point set (host float32 (n, 3)).
* ``tree``  mesh (host vertices (V, 3), faces (F, 4)) -> LBVH -> OBB tree;
            the host receives the root box.  With ``keep_trees`` the tree
            (obbs (2F-1, 15), matrices (2F-1, 16), nodes) stays resident in
            HBM, still charged to the budget, until ``release(request)``.

Each scheduler iteration
  1. fails requests that waited longer than the request timeout;
  2. admits arrived requests in arrival order while their device footprint
     fits the HBM budget (inputs, outputs and workspaces stay resident until
     the result is back on the host);
  3. packs every admitted fit into ONE batched DiTO launch (one gather
     kernel reads every set from pinned host memory into the batch's SoA
     array, then obvh_dito_batch, download) on one of two alternating
     streams, so a new batch starts while the previous one still runs;
  4. packs every admitted tree request into ONE batched transform on the
     next of several streams: each mesh is uploaded into the batch's arrays
     and gets its LBVH built there (obvh_lbvh_build), then
     obvh_obb_tree_batch transforms all of them at once and the root boxes
     come back to the host;
  5. polls completion events and returns results; a request in flight past
     its timeout is answered with an error at once (its device work still
     finishes and returns its bytes to the budget).
Latency is arrival -> result on the host.  The kernels are deterministic, so
a request's result is the same alone and in any mix.  Each request records
when it was admitted and when its launch was enqueued, so a slow request
can be attributed to queueing, host-side launch work or the device.
"""
from __future__ import annotations

import collections
import ctypes
import dataclasses
import math
import time

import numpy as np

from . import ops
from ._lib import check, lib
from .layouts import NODE_DTYPE


@dataclasses.dataclass
class Request:
    rid: int
    kind: str                 # "fit" or "tree"
    arrival: float            # seconds, on the server clock
    points: object = None     # pinned torch tensor (3, n) for fits
    vertices: object = None   # pinned torch tensor (V, 3)
    faces: object = None      # pinned torch tensor (F, 4) int32
    done: float = math.nan
    result: object = None
    error: str | None = None  # set instead of result when the request failed
    nbytes: int = 0
    admitted: float = math.nan   # server clock when admitted
    launched: float = math.nan   # server clock when its launch was enqueued
    abandoned: bool = False      # answered (timeout) while its device work still ran


def fit_bytes(n: int) -> int:
    return 12 * n + 15 * 4 + 96


def tree_bytes(nv: int, nf: int) -> int:
    """Device bytes of one tree request in a batch: mesh, tree, boxes and
    matrices, LBVH scratch, and its share of the batched transform's
    workspace (the forest copy of its tree plus the transform workspace)."""
    total = 2 * nf - 1
    return (12 * nv + 16 * nf + total * (36 + 4 + 4 + 4 + 60 + 64) + nf * 4
            + int(lib().obvh_obb_tree_batch_workspace_size(1, nf)) + int(lib().obvh_lbvh_workspace_size(nf)))


class Server:
    # A batch of fits holds at most max_batch_points points (16 M = 201 MB of
    # SoA by default).  Bounded batches let the next batch's host packing and
    # PCIe gather overlap the previous batch on the other fit stream: a burst
    # of 4096 fits (2.8 GB) measured 72 K sets/s at 2^23..2^25 against
    # 60-69 K with 2^27 (two batches, the first gather waiting for the host
    # to pack half the burst), before the gather's 16-byte reads (77 K).
    def __init__(self, hbm_budget: float = 200e9, tree_streams: int = 4, max_batch_points: int = 1 << 24,
                 keep_trees: bool = False, request_timeout: float | None = None):
        import torch

        self.torch = torch
        self.keep_trees = keep_trees
        self.request_timeout = request_timeout
        self.timeouts = 0
        self.tree_batches = 0
        self.budget = int(hbm_budget)
        self.in_use = 0
        self.max_batch_points = max_batch_points
        self.n_tree_streams = tree_streams
        self.fit_streams: list = []
        self.tree_streams: list = []
        self._fs = 0
        self._ts = 0
        self.waiting: collections.deque[Request] = collections.deque()
        self.inflight: list = []  # (event, callback, requests)
        self._base = (0.0, time.perf_counter())  # (server clock, perf_counter) at the current step
        self.finished: list[Request] = []
        self.batches = 0
        self.batch_sizes: list[int] = []

    # -------------------------------------------------------- admission --
    def submit(self, req: Request) -> None:
        if req.kind == "fit":
            req.nbytes = fit_bytes(int(req.points.shape[1]))
        else:
            req.nbytes = tree_bytes(int(req.vertices.shape[0]), int(req.faces.shape[0]))
        if req.nbytes > self.budget:
            raise ValueError(f"request {req.rid} needs {req.nbytes} B, more than the HBM budget")
        self.waiting.append(req)

    def _expire(self, now: float) -> None:
        """Fail waiting requests older than the request timeout."""
        if self.request_timeout is None or not self.waiting:
            return
        keep = collections.deque()
        for r in self.waiting:
            if now - r.arrival > self.request_timeout:
                r.error = "timed out waiting for admission"
                r.done = now
                self.timeouts += 1
                self.finished.append(r)
            else:
                keep.append(r)
        self.waiting = keep

    def _admit(self, now: float = math.nan) -> tuple[list[Request], list[Request]]:
        fits, trees, pts = [], [], 0
        while self.waiting:
            r = self.waiting[0]
            if self.in_use + r.nbytes > self.budget:
                break
            if r.kind == "fit":
                n = int(r.points.shape[1])
                if fits and pts + n > self.max_batch_points:
                    break
                pts += n
                fits.append(r)
            else:
                trees.append(r)
            self.in_use += r.nbytes
            r.admitted = now
            self.waiting.popleft()
        return fits, trees

    # ----------------------------------------------------------- launch --
    def _streams(self) -> None:
        if not self.fit_streams:  # created on first use (admission works without a GPU)
            self.fit_streams = [self.torch.cuda.Stream(), self.torch.cuda.Stream()]
            self.tree_streams = [self.torch.cuda.Stream() for _ in range(self.n_tree_streams)]

    def _launch_fits(self, fits: list[Request]) -> None:
        torch = self.torch
        self._streams()
        stream = self.fit_streams[self._fs]
        self._fs ^= 1
        for r in fits:
            p = r.points
            # pinned host or device memory is checked per set by obvh_gather_sets
            # (one pointer query each; a second one here cost as much again)
            if p.dim() != 2 or p.shape[0] != 3 or p.dtype != torch.float32 or not p.is_contiguous():
                raise ValueError(f"request {r.rid}: points must be a contiguous float32 (3, n) tensor "
                                 "in pinned host or device memory")
        sizes = np.array([int(r.points.shape[1]) for r in fits], dtype=np.int64)
        off = np.zeros(len(fits) + 1, dtype=np.int64)
        off[1:] = np.cumsum(sizes)
        total = int(off[-1])
        with torch.cuda.stream(stream):
            soa = torch.empty((3, max(total, 1)), dtype=torch.float32, device="cuda")
            addrs = [r.points.data_ptr() for r in fits]
            ptrs = (ctypes.c_void_p * len(fits))(*addrs)
            ptrs_h = torch.tensor(addrs, dtype=torch.int64).pin_memory()
            offs_h = torch.from_numpy(off).pin_memory()
            ptrs_d = ptrs_h.to("cuda", non_blocking=True)
            offs = offs_h.to("cuda", non_blocking=True)
            check(lib().obvh_gather_sets(ptrs, ctypes.c_void_p(ptrs_d.data_ptr()), off.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(offs.data_ptr()), len(fits),
                                         ctypes.c_void_p(soa.data_ptr()), max(total, 1),
                                         ctypes.c_void_p(stream.cuda_stream)))
            out = ops.dito_batch(soa[0], soa[1], soa[2], offs, stream=stream)
            host = torch.empty(out.shape, dtype=torch.float32, pin_memory=True)
            host.copy_(out, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        keep = (soa, offs, offs_h, ptrs_h, ptrs_d, out)

        def finish(now, fits=fits, host=host, keep=keep):
            res = host.numpy()
            for k, r in enumerate(fits):
                self.in_use -= r.nbytes
                if r.abandoned:
                    continue
                r.result = res[k].copy()
                r.done = now
                self.finished.append(r)

        self.inflight.append((ev, finish, fits))
        self.batches += 1
        self.batch_sizes.append(len(fits))

    def _launch_trees(self, trees: list[Request]) -> None:
        """Meshes -> LBVHs -> one batched OBB-tree transform; the host
        receives each root box, which marks completion.  With ``keep_trees``
        each result also holds its tree's device tensors (views into the
        batch's arrays) and stays charged until ``release``."""
        torch = self.torch
        self._streams()
        stream = self.tree_streams[self._ts]
        self._ts = (self._ts + 1) % len(self.tree_streams)
        if len(trees) == 1:  # alone: the single-tree transform, no packing
            self._launch_tree(trees[0], stream)
            return
        nf = [int(r.faces.shape[0]) for r in trees]
        nv = [int(r.vertices.shape[0]) for r in trees]
        for r, n in zip(trees, nf):
            if n < 1 or r.faces.dim() != 2 or r.faces.shape[1] != 4:
                raise ValueError(f"request {r.rid}: faces must be an (F >= 1, 4) int32 tensor")
        prim = np.zeros(len(trees) + 1, dtype=np.int64)
        vert = np.zeros(len(trees) + 1, dtype=np.int64)
        prim[1:] = np.cumsum(nf)
        vert[1:] = np.cumsum(nv)
        total, K = int(prim[-1]), len(trees)
        nodes_total = 2 * total - K
        with torch.cuda.stream(stream):
            V = torch.empty((int(vert[-1]), 3), dtype=torch.float32, device="cuda")
            F = torch.empty((total, 4), dtype=torch.int32, device="cuda")
            nodes = torch.empty(nodes_total * NODE_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
            parents = torch.empty(nodes_total, dtype=torch.int32, device="cuda")
            counts = torch.empty(nodes_total, dtype=torch.int32, device="cuda")
            costs = torch.empty(nodes_total, dtype=torch.float32, device="cuda")
            mortons = torch.empty(total, dtype=torch.int32, device="cuda")
            lws = torch.empty(max(int(lib().obvh_lbvh_workspace_size(max(nf))), 256), dtype=torch.uint8,
                              device="cuda")
            offs_h = torch.from_numpy(np.stack([prim, vert]).astype(np.int32)).pin_memory()
            offs = offs_h.to("cuda", non_blocking=True)
            spans = []
            for t, r in enumerate(trees):
                s0, s1 = 2 * int(prim[t]) - t, 2 * int(prim[t + 1]) - t - 1
                v0, v1 = int(vert[t]), int(vert[t + 1])
                f0, f1 = int(prim[t]), int(prim[t + 1])
                V[v0:v1].copy_(r.vertices, non_blocking=True)
                F[f0:f1].copy_(r.faces, non_blocking=True)
                ops.lbvh_build(V[v0:v1], F[f0:f1], stream=stream, workspace=lws,
                               out={"nodes": nodes[s0 * NODE_DTYPE.itemsize:s1 * NODE_DTYPE.itemsize],
                                    "parents": parents[s0:s1], "counts": counts[s0:s1], "costs": costs[s0:s1],
                                    "mortons": mortons[f0:f1]})
                spans.append((s0, s1))
            batch = {"nodes": nodes, "parents": parents, "V": V, "F": F, "prim_offsets": offs[0],
                     "vert_offsets": offs[1], "prim_offsets_host": prim, "total": total, "ntrees": K}
            status = torch.full((1,), -1, dtype=torch.int32, pin_memory=True)
            obbs, mats = ops.obb_tree_batch(batch, stream=stream, status=status)
            roots_idx = torch.tensor([s0 for s0, _ in spans], dtype=torch.int64).pin_memory().to(
                "cuda", non_blocking=True)
            roots = torch.empty((K, 15), dtype=torch.float32, pin_memory=True)
            roots.copy_(obbs.index_select(0, roots_idx), non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        keep = (V, F, nodes, parents, counts, costs, mortons, lws, offs, offs_h, roots_idx)
        now = self._clock()
        for r in trees:
            r.launched = now

        def finish(now, trees=trees, roots=roots, obbs=obbs, mats=mats, keep=keep, stream=stream, status=status,
                   spans=spans):
            word = int(status.item())
            res = roots.numpy()
            for t, r in enumerate(trees):
                if r.abandoned:
                    self.in_use -= r.nbytes
                    continue
                r.done = now
                if word != 0:  # the device found a malformed tree in the batch: fail its requests
                    r.error = ops.tree_status_message(word)
                    r.result = None
                    self.in_use -= r.nbytes
                else:
                    r.result = {"root": res[t].copy()}
                    if self.keep_trees:
                        s0, s1 = spans[t]
                        r.result.update({"obbs": obbs[s0:s1], "mats": mats[s0:s1],
                                         "nodes": keep[2][s0 * NODE_DTYPE.itemsize:s1 * NODE_DTYPE.itemsize],
                                         "stream": stream})
                    else:
                        self.in_use -= r.nbytes
                self.finished.append(r)

        self.inflight.append((ev, finish, trees))
        self.tree_batches += 1

    def _launch_tree(self, r: Request, stream) -> None:
        torch = self.torch
        with torch.cuda.stream(stream):
            V = r.vertices.to("cuda", non_blocking=True)
            F = r.faces.to("cuda", non_blocking=True)
            t = ops.lbvh_build(V, F, stream=stream)
            status = torch.full((1,), -1, dtype=torch.int32, pin_memory=True)
            obbs, mats = ops.obb_tree(t["nodes"], t["parents"], V, F, stream=stream, status=status)
            root = torch.empty(15, dtype=torch.float32, pin_memory=True)
            root.copy_(obbs[0], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        keep = (V, F, t)
        r.launched = self._clock()

        def finish(now, r=r, root=root, obbs=obbs, mats=mats, keep=keep, stream=stream, status=status):
            word = int(status.item())
            if r.abandoned:
                self.in_use -= r.nbytes
                return
            r.done = now
            if word != 0:  # the device found a malformed tree: fail the request
                r.error = ops.tree_status_message(word)
                r.result = None
                self.in_use -= r.nbytes
            else:
                r.result = {"root": root.numpy().copy()}
                if self.keep_trees:
                    r.result.update({"obbs": obbs, "mats": mats, "nodes": keep[2]["nodes"], "stream": stream})
                else:
                    self.in_use -= r.nbytes
            self.finished.append(r)

        self.inflight.append((ev, finish, [r]))

    def release(self, r: Request) -> None:
        """Drop a resident tree (``keep_trees``) and return its bytes to the budget."""
        if r.kind == "tree" and self.keep_trees and r.result is not None and "obbs" in r.result:
            for k in ("obbs", "mats", "nodes", "stream"):
                r.result.pop(k)
            self.in_use -= r.nbytes

    # ------------------------------------------------------------- loop --
    def poll(self, now: float) -> int:
        if now >= self._base[0]:
            self._base = (now, time.perf_counter())
        still, n = [], 0
        for ev, cb, reqs in self.inflight:
            if ev.query():
                cb(max(now, self._clock()))
                n += 1
            else:
                if self.request_timeout is not None:
                    for r in reqs:  # answer late requests now; their bytes return when the work ends
                        if not r.abandoned and now - r.arrival > self.request_timeout:
                            r.abandoned = True
                            r.error = "timed out on the device"
                            r.done = now
                            self.timeouts += 1
                            self.finished.append(r)
                still.append((ev, cb, reqs))
        self.inflight = still
        return n

    def _clock(self) -> float:
        now, base = self._base
        return now + (time.perf_counter() - base)

    def step(self, now: float) -> None:
        self._base = (now, time.perf_counter())
        self._expire(now)
        fits, trees = self._admit(now)
        if fits:
            self._launch_fits(fits)
            t = self._clock()
            for r in fits:
                r.launched = t
        if trees:
            self._launch_trees(trees)
        self.poll(self._clock())

    def idle(self) -> bool:
        return not self.waiting and not self.inflight


# ------------------------------------------------------------ workload --

def request_pool(seed: int = 7, fits: int = 64, trees: int = 12, fit_range=(1000, 262144),
                 tree_range=(10000, 2000000)):
    """Pinned host payloads reused by the generated requests: point sets with
    log-uniform sizes and noisy icospheres with log-uniform triangle counts."""
    import torch

    from . import synth

    rng = np.random.default_rng(seed)
    fit_pool = []
    for k in range(fits):
        n = int(np.exp(rng.uniform(np.log(fit_range[0]), np.log(fit_range[1]))))
        p = synth.gaussian_points(n, std=tuple(np.sort(rng.uniform(1, 20, 3))[::-1]), seed=seed * 1000 + k)
        fit_pool.append(torch.from_numpy(np.ascontiguousarray(p.T)).pin_memory())
    tree_pool = []
    for k in range(trees):
        nf = float(np.exp(rng.uniform(np.log(tree_range[0]), np.log(tree_range[1]))))
        f = max(1, int(round(math.sqrt(nf / 20.0))))
        V, F = synth.icosphere(f, seed=seed * 1000 + 500 + k)
        tree_pool.append((torch.from_numpy(V).pin_memory(), torch.from_numpy(F.view(np.int32)).pin_memory()))
    return fit_pool, tree_pool


def poisson_schedule(rate: float, duration: float, fit_share: float = 0.7, seed: int = 8):
    """[(arrival_time, kind, pool_index)] for a Poisson stream."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append((t, "fit" if rng.random() < fit_share else "tree", int(rng.integers(1 << 30))))


def run_open_loop(server: Server, schedule, fit_pool, tree_pool, timeout: float = 600.0, pause_gc: bool = True):
    """Feed the schedule in real time; returns the finished requests.
    pause_gc: no cyclic garbage collection during the run (a full collection
    over the run's request objects stalls the loop for tens of ms); it runs
    once afterwards."""
    import gc

    was = gc.isenabled()
    if pause_gc:
        gc.collect()
        gc.disable()
    try:
        return _run_open_loop(server, schedule, fit_pool, tree_pool, timeout)
    finally:
        if pause_gc and was:
            gc.enable()


def _run_open_loop(server: Server, schedule, fit_pool, tree_pool, timeout: float):
    t0 = time.perf_counter()
    k = 0
    rid = 0
    while True:
        now = time.perf_counter() - t0
        while k < len(schedule) and schedule[k][0] <= now:
            at, kind, pick = schedule[k]
            if kind == "fit":
                r = Request(rid, "fit", at, points=fit_pool[pick % len(fit_pool)])
            else:
                V, F = tree_pool[pick % len(tree_pool)]
                r = Request(rid, "tree", at, vertices=V, faces=F)
            server.submit(r)
            rid += 1
            k += 1
        server.step(time.perf_counter() - t0)
        if k >= len(schedule) and server.idle():
            break
        if now > timeout:
            raise TimeoutError("serving run exceeded its time limit")
        if k < len(schedule) and not server.inflight and not server.waiting:
            time.sleep(max(0.0, min(schedule[k][0] - now, 0.0005)))
    return server.finished


def latency_stats(done: list[Request]) -> dict:
    """p50 / p99 / p99.9 / max latency (ms) of the answered requests, per kind,
    errors, and the phases of the five slowest requests: queue (arrival ->
    admission), launch (admission -> launch enqueued) and device (launch ->
    result on the host, including the poll interval)."""
    ok = [r for r in done if r.error is None]
    lat = np.array([r.done - r.arrival for r in ok]) * 1e3
    out = {"requests": len(done), "errors": len(done) - len(ok),
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "p99_9_ms": round(float(np.percentile(lat, 99.9)), 3), "max_ms": round(float(lat.max()), 3)}
    slow = sorted(ok, key=lambda r: r.done - r.arrival)[-5:][::-1]
    out["slowest"] = [{"kind": r.kind, "ms": round((r.done - r.arrival) * 1e3, 3),
                       "queue_ms": round((r.admitted - r.arrival) * 1e3, 3),
                       "launch_ms": round((r.launched - r.admitted) * 1e3, 3),
                       "device_ms": round((r.done - r.launched) * 1e3, 3)} for r in slow]
    done = ok
    for kind in ("fit", "tree"):
        l2 = np.array([r.done - r.arrival for r in done if r.kind == kind]) * 1e3
        if l2.size:
            out[kind] = {"n": int(l2.size), "p50_ms": round(float(np.percentile(l2, 50)), 3),
                         "p99_ms": round(float(np.percentile(l2, 99)), 3)}
    return out


def main(argv=None) -> int:
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Open-loop Poisson serving run")
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--timeout", type=float, default=None, help="per-request timeout (s)")
    a = ap.parse_args(argv)
    fit_pool, tree_pool = request_pool()
    run_open_loop(Server(hbm_budget=a.budget_gb * 1e9), poisson_schedule(a.rate, 1.0, seed=99), fit_pool,
                  tree_pool)  # warm-up: kernels loaded, allocator pools grown
    srv = Server(hbm_budget=a.budget_gb * 1e9, request_timeout=a.timeout)
    sched = poisson_schedule(a.rate, a.duration)
    t0 = time.perf_counter()
    done = run_open_loop(srv, sched, fit_pool, tree_pool)
    wall = time.perf_counter() - t0
    st = latency_stats(done)
    st.update({"req_per_s": round(len(done) / wall, 1), "batches": srv.batches, "tree_batches": srv.tree_batches,
               "timeouts": srv.timeouts,
               "mean_fits_per_batch": round(float(np.mean(srv.batch_sizes)), 2) if srv.batch_sizes else 0})
    print(json.dumps(st))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
