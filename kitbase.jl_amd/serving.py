"""Serving loop: many solve requests on one GPU, resident in HBM, stepped together.

Requests arrive over time (Poisson arrivals in the benchmark).  The loop
  * admits waiting requests first come, first served while their device footprint
    (kbg_solve_bytes) fits the memory budget; an admitted request is uploaded once and stays
    resident until it finishes;
  * steps every admitted request together with kbg_batch_step, in chunks: a chunk is at most
    `chunk_steps` steps and never more than the fewest steps any admitted request still needs, so
    a finished request leaves and waiting requests join between chunks (continuous batching);
  * downloads a finished request's state, frees it and records its time to solution (arrival to
    completion).
Each request's arithmetic is the same alone or in any mix (solves never share data; their sums have
a fixed order), so results do not depend on what else is being served.  The FP32 sweep advances a
1D solve two steps per pass where it can, and a two-step pass may round differently (by an ulp, in
rare FP32 Shakhov cases) from two one-step passes; so requests advance in step pairs counted from
their own first step, as a solo kbg_batch_step does: a chunk has an even number of steps, and a
request down to its last odd step takes it in a chunk of its own.

The loop talks to the device through a backend object (DeviceBackend here; tests use a fake one).
"""
from __future__ import annotations

import dataclasses
import threading
import time
from dataclasses import dataclass, field
from typing import Any

import numpy as np


@dataclass
class Request:
    rid: int
    cfg: dict                 # KitBase config dict (SolverSet input)
    nsteps: int | None        # time steps to run; None: to maxTime (KitBase's solve!: nt = floor(T / dt) + 1)
    arrival: float            # seconds after the start of the run
    precision: str = "fp32"
    setup: Any = None         # optional callable(ks, ctr, face) that overrides the initial state
    # filled by the server
    admitted: float = -1.0
    done: float = -1.0
    result: Any = None        # host ControlVolumes after the last step
    points: int = 0           # cell x velocity points of one step

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class DeviceBackend:
    """Solve lifecycle on libkbgpu.so through the host mirror (kitbase_amd).

    sweep_vw: most velocities per wavefront of the fused 1D sweep, set on the context
    (kbg_ctx_set_sweep_vw).  Serving batches are small and latency-bound, and 16 velocities per
    wavefront instead of 64 cut the bench mix's p50 / p99 from about 90 / 180 ms to 69 / 140 ms; a
    request's results do not depend on the batch it runs in either way."""

    def __init__(self, kb, ctx=None, sweep_vw: int = 16):
        self.kb = kb
        self.ctx = ctx or kb.Context(0)
        self.ctx.set_sweep_vw(sweep_vw)

    def prepare(self, req: Request):
        ks = self.kb.SolverSet(req.cfg)
        ctr, face = self.kb.init_fvm(ks)
        if req.setup is not None:
            req.setup(ks, ctr, face)
        dt = self.kb.timestep(ks, ctr)
        if req.nsteps is None:
            req.nsteps = int(np.floor(ks.set.maxTime / dt)) + 1
        nx = ks.ps.nx * (ks.ps.ny if ks.dim == 2 else 1)
        req.points = int(nx * ks.vs.nv)
        return ks, ctr, face, dt

    def footprint(self, prepared, req: Request) -> int:
        return self.kb.solve_bytes(prepared[0], req.precision)

    def create(self, prepared, req: Request):
        ks, ctr, face, dt = prepared
        # the face arrays hold no state the fused step reads (boundary fluxes are recomputed every
        # step), so only the cells are uploaded and no phase-level buffers are allocated
        return self.kb.DeviceSolve(ks, ctr, None, precision=req.precision, ctx=self.ctx), dt

    def step(self, handles, dts, nsteps: int):
        self.kb.batch_step(handles, dts, nsteps, ctx=self.ctx)

    def download(self, handle, prepared):
        # fresh cell arrays for the result (the slope arrays of the prepared state stay zero and
        # are shared: a download does not touch them)
        p = prepared[1]
        ctr = dataclasses.replace(p, w=np.empty_like(p.w), prim=np.empty_like(p.prim), f=np.empty_like(p.f))
        handle.download(ctr)
        return ctr

    def release(self, handle):
        handle.close()

    def finish(self, handle, prepared):
        ctr = self.download(handle, prepared)
        self.release(handle)
        return ctr


@dataclass
class _Live:
    req: Request
    handle: Any
    dt: float
    left: int
    nbytes: int
    prepared: Any


@dataclass
class ServeStats:
    latencies_ms: list = field(default_factory=list)
    updates: float = 0.0
    wall_s: float = 0.0
    step_s: float = 0.0       # time inside backend.step (device work plus launch overhead)
    host_s: float = 0.0       # time preparing, uploading and downloading requests
    chunks: int = 0
    max_live: int = 0
    peak_bytes: int = 0

    def summary(self) -> dict:
        lat = np.asarray(self.latencies_ms)
        return {
            "requests": int(lat.size),
            "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
            "max_ms": float(lat.max()) if lat.size else None,
            "updates_per_s": self.updates / self.wall_s if self.wall_s > 0 else 0.0,
            "chunks": self.chunks,
            "step_s": self.step_s,
            "host_s": self.host_s,
            "max_live": self.max_live,
            "peak_bytes": self.peak_bytes,
        }


class Server:
    """overlap: while a chunk runs on the device (backend.step on a worker thread; the library call
    releases the interpreter lock), the calling thread prepares the requests that arrive meanwhile
    (host-side set-up only: no device calls), so host set-up does not stall the live requests.

    transfer (with overlap): a second backend whose solves live on another context of the same
    device (kbg_batch_step steps solves of any context of its device).  While a chunk runs, the
    calling thread also uploads, through it, the waiting requests that fit (they join the next
    chunk, the one they would have joined anyway) and downloads the requests the previous chunk
    finished; their device memory is released between chunks.  The device then no longer idles
    during transfers."""

    def __init__(self, backend, budget_bytes: int, chunk_steps: int = 25, clock=time.perf_counter, sleep=time.sleep,
                 overlap: bool = False, transfer=None):
        if chunk_steps < 1:
            raise ValueError("chunk_steps must be >= 1")
        self.backend = backend
        self.budget = int(budget_bytes)
        self.chunk = int(chunk_steps)
        self.clock = clock
        self.sleep = sleep
        self.overlap = overlap
        self.transfer = transfer

    def run(self, requests: list[Request]) -> ServeStats:
        if self.overlap and self.transfer is not None:
            return self._run_transfer(requests)
        return self._run(requests)

    def _chunk(self, live):
        """Steps of the next chunk and the requests it takes (pairs from each request's own start)."""
        k = min(self.chunk, min(l.left for l in live))
        if k == 1 and self.chunk > 1:
            return 1, [l for l in live if l.left == 1]  # last odd steps, on their own
        if k > 1:
            k -= k & 1
        return k, list(live)

    def _run_transfer(self, requests: list[Request]) -> ServeStats:
        stats = ServeStats()
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        waiting: list[tuple[Request, Any, int]] = []
        live: list[_Live] = []
        ready: list[_Live] = []      # uploaded while a chunk ran: they join the next one
        finished: list[_Live] = []   # done stepping, not yet downloaded
        spent: list[_Live] = []      # downloaded, memory released between chunks
        used = 0
        xf = self.transfer
        t0 = self.clock()
        now = lambda: self.clock() - t0

        def prepare_arrived():
            while pending and pending[0].arrival <= now():
                r = pending.pop(0)
                h0 = self.clock()
                prep = xf.prepare(r)
                nb = xf.footprint(prep, r)
                stats.host_s += self.clock() - h0
                if nb > self.budget:
                    raise ValueError(f"request {r.rid} needs {nb} bytes, more than the budget {self.budget}")
                waiting.append((r, prep, nb))

        def admit_one(into) -> bool:
            nonlocal used
            if not (waiting and used + waiting[0][2] <= self.budget):
                return False
            r, prep, nb = waiting.pop(0)
            h0 = self.clock()
            h, dt = xf.create(prep, r)
            stats.host_s += self.clock() - h0
            r.admitted = now()
            into.append(_Live(r, h, dt, r.nsteps, nb, prep))
            used += nb
            return True

        def download_one():
            l = finished.pop(0)
            h0 = self.clock()
            l.req.result = xf.download(l.handle, l.prepared)
            stats.host_s += self.clock() - h0
            l.req.done = now()
            stats.latencies_ms.append(1e3 * l.req.latency)
            spent.append(l)

        while pending or waiting or live or ready or finished or spent:
            prepare_arrived()
            live.extend(ready)
            ready.clear()
            for l in spent:
                xf.release(l.handle)
                used -= l.nbytes
            spent.clear()
            while admit_one(live):  # the device is between chunks: what fits joins the next one now
                pass
            stats.max_live = max(stats.max_live, len(live))
            stats.peak_bytes = max(stats.peak_bytes, used)
            if not live:
                while finished:
                    download_one()
                if pending and not spent and not waiting:
                    self.sleep(max(0.0, pending[0].arrival - now()))
                continue
            k, batch = self._chunk(live)
            err = []

            def run():
                try:
                    self.backend.step([l.handle for l in batch], [l.dt for l in batch], k)
                except BaseException as e:  # re-raised on the calling thread
                    err.append(e)

            h0 = self.clock()
            th = threading.Thread(target=run)
            th.start()
            while th.is_alive():
                if pending and pending[0].arrival <= now():
                    prepare_arrived()
                elif finished:
                    download_one()
                elif not admit_one(ready):
                    th.join(timeout=5e-4)
            th.join()
            if err:
                raise err[0]
            stats.step_s += self.clock() - h0
            stats.chunks += 1
            for l in batch:
                l.left -= k
                stats.updates += float(k) * l.req.points
            for l in [l for l in live if l.left == 0]:
                live.remove(l)
                finished.append(l)
            stats.peak_bytes = max(stats.peak_bytes, used)
        stats.wall_s = now()
        return stats

    def _run(self, requests: list[Request]) -> ServeStats:
        stats = ServeStats()
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        waiting: list[tuple[Request, Any, int]] = []
        live: list[_Live] = []
        used = 0
        t0 = self.clock()
        now = lambda: self.clock() - t0

        def prepare_arrived():
            t = now()
            while pending and pending[0].arrival <= t:
                r = pending.pop(0)
                h0 = self.clock()
                prep = self.backend.prepare(r)
                nb = self.backend.footprint(prep, r)
                stats.host_s += self.clock() - h0
                if nb > self.budget:
                    raise ValueError(f"request {r.rid} needs {nb} bytes, more than the budget {self.budget}")
                waiting.append((r, prep, nb))

        while pending or waiting or live:
            prepare_arrived()
            # first come, first served while the footprint fits
            while waiting and used + waiting[0][2] <= self.budget:
                r, prep, nb = waiting.pop(0)
                h0 = self.clock()
                h, dt = self.backend.create(prep, r)
                stats.host_s += self.clock() - h0
                r.admitted = now()
                live.append(_Live(r, h, dt, r.nsteps, nb, prep))
                used += nb
            stats.max_live = max(stats.max_live, len(live))
            stats.peak_bytes = max(stats.peak_bytes, used)
            if not live:
                if pending:
                    self.sleep(max(0.0, pending[0].arrival - now()))
                continue
            k = min(self.chunk, min(l.left for l in live))
            if k == 1 and self.chunk > 1:
                batch = [l for l in live if l.left == 1]  # last odd steps, on their own
            else:
                if k > 1:
                    k -= k & 1
                batch = live
            h0 = self.clock()
            if self.overlap:
                err = []

                def run():
                    try:
                        self.backend.step([l.handle for l in batch], [l.dt for l in batch], k)
                    except BaseException as e:  # re-raised on the calling thread
                        err.append(e)

                th = threading.Thread(target=run)
                th.start()
                while th.is_alive():
                    if pending and pending[0].arrival <= now():
                        prepare_arrived()
                    else:
                        th.join(timeout=5e-4)
                th.join()
                if err:
                    raise err[0]
            else:
                self.backend.step([l.handle for l in batch], [l.dt for l in batch], k)
            stats.step_s += self.clock() - h0
            stats.chunks += 1
            for l in batch:
                l.left -= k
                stats.updates += float(k) * l.req.points
            for l in [l for l in live if l.left == 0]:
                h0 = self.clock()
                l.req.result = self.backend.finish(l.handle, l.prepared)
                stats.host_s += self.clock() - h0
                l.req.done = now()
                stats.latencies_ms.append(1e3 * l.req.latency)
                used -= l.nbytes
                live.remove(l)
        stats.wall_s = now()
        return stats
