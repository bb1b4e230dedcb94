"""Serving loop: continuous batching of OBB requests on one GPU.

Two request kinds:
* ``fit``   DiTO-14 box of one point set (host float32 (n, 3)).
* ``tree``  mesh (host vertices (V, 3), faces (F, 4)) -> LBVH -> OBB tree;
            the host receives the root box.  With ``keep_trees`` the tree
            (obbs (2F-1, 15), matrices (2F-1, 16), nodes) stays resident in
            HBM, still charged to the budget, until ``release(request)``.

Each scheduler iteration
  1. admits arrived requests in arrival order while their device footprint
     fits the HBM budget (inputs, outputs and workspaces stay resident until
     the result is back on the host);
  2. packs every admitted fit into ONE batched DiTO launch (one gather
     kernel reads every set from pinned host memory into the batch's SoA
     array, then obvh_dito_batch, download) on one of two alternating
     streams, so a new batch starts while the previous one still runs;
  3. starts each admitted tree request on the next of several streams
     (upload, obvh_lbvh_build, obvh_obb_tree_async, root box download);
  4. polls completion events and returns results.
Latency is arrival -> result on the host.  The kernels are deterministic, so
a request's result is the same alone and in any mix.
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


def fit_bytes(n: int) -> int:
    return 12 * n + 15 * 4 + 96


def tree_bytes(nv: int, nf: int) -> int:
    total = 2 * nf - 1
    return (12 * nv + 16 * nf + total * (36 + 4 + 4 + 4 + 60 + 64) + nf * 4
            + int(lib().obvh_obb_tree_workspace_size(nf, nf - 1)) + int(lib().obvh_lbvh_workspace_size(nf)))


class Server:
    # A batch of fits holds at most max_batch_points points (16 M = 201 MB of
    # SoA by default).  Bounded batches let the next batch's host packing and
    # PCIe gather overlap the previous batch on the other fit stream: a burst
    # of 4096 fits (2.8 GB) measured 72 K sets/s at 2^23..2^25 against
    # 60-69 K with 2^27 (two batches, the first gather waiting for the host
    # to pack half the burst), before the gather's 16-byte reads (77 K).
    def __init__(self, hbm_budget: float = 200e9, tree_streams: int = 4, max_batch_points: int = 1 << 24,
                 keep_trees: bool = False):
        import torch

        self.torch = torch
        self.keep_trees = keep_trees
        self.budget = int(hbm_budget)
        self.in_use = 0
        self.max_batch_points = max_batch_points
        self.n_tree_streams = tree_streams
        self.fit_streams: list = []
        self.tree_streams: list = []
        self._fs = 0
        self._ts = 0
        self.waiting: collections.deque[Request] = collections.deque()
        self.inflight: list = []  # (event, callback)
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

    def _admit(self) -> tuple[list[Request], list[Request]]:
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
                r.result = res[k].copy()
                r.done = now
                self.in_use -= r.nbytes
                self.finished.append(r)

        self.inflight.append((ev, finish))
        self.batches += 1
        self.batch_sizes.append(len(fits))

    def _launch_tree(self, r: Request) -> None:
        """Mesh -> LBVH -> OBB tree; the host receives the root box, which
        marks completion.  With ``keep_trees`` the result also holds the
        device tensors and the request stays charged until ``release``."""
        torch = self.torch
        self._streams()
        stream = self.tree_streams[self._ts]
        self._ts = (self._ts + 1) % len(self.tree_streams)
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

        def finish(now, r=r, root=root, obbs=obbs, mats=mats, keep=keep, stream=stream, status=status):
            word = int(status.item())
            if word != 0:  # the device found a malformed tree: fail the request
                r.error = ops.tree_status_message(word)
                r.result = None
                r.done = now
                self.in_use -= r.nbytes
                self.finished.append(r)
                return
            r.result = {"root": root.numpy().copy()}
            if self.keep_trees:
                r.result.update({"obbs": obbs, "mats": mats, "nodes": keep[2]["nodes"], "stream": stream})
            else:
                self.in_use -= r.nbytes
            r.done = now
            self.finished.append(r)

        self.inflight.append((ev, finish))

    def release(self, r: Request) -> None:
        """Drop a resident tree (``keep_trees``) and return its bytes to the budget."""
        if r.kind == "tree" and self.keep_trees and r.result is not None and "obbs" in r.result:
            for k in ("obbs", "mats", "nodes", "stream"):
                r.result.pop(k)
            self.in_use -= r.nbytes

    # ------------------------------------------------------------- loop --
    def poll(self, now: float) -> int:
        still, n = [], 0
        for ev, cb in self.inflight:
            if ev.query():
                cb(now)
                n += 1
            else:
                still.append((ev, cb))
        self.inflight = still
        return n

    def step(self, now: float) -> None:
        fits, trees = self._admit()
        if fits:
            self._launch_fits(fits)
        for r in trees:
            self._launch_tree(r)
        self.poll(now)

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


def run_open_loop(server: Server, schedule, fit_pool, tree_pool, timeout: float = 600.0):
    """Feed the schedule in real time; returns the finished requests."""
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
    lat = np.array([r.done - r.arrival for r in done]) * 1e3
    out = {"requests": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3), "max_ms": round(float(lat.max()), 3)}
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
    a = ap.parse_args(argv)
    fit_pool, tree_pool = request_pool()
    srv = Server(hbm_budget=a.budget_gb * 1e9)
    sched = poisson_schedule(a.rate, a.duration)
    t0 = time.perf_counter()
    done = run_open_loop(srv, sched, fit_pool, tree_pool)
    wall = time.perf_counter() - t0
    st = latency_stats(done)
    st.update({"req_per_s": round(len(done) / wall, 1), "batches": srv.batches,
               "mean_fits_per_batch": round(float(np.mean(srv.batch_sizes)), 2) if srv.batch_sizes else 0})
    print(json.dumps(st))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
