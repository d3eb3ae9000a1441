"""Serving loop over the batch engine: continuous batching of solve requests.

A request carries its mesh tables (built by the client beforehand, as the
reference's model builder writes a .mesh file before the solver runs), a
step count and an arrival time.  The loop

  * enqueues K steps of every resident solve (asynchronous, one HIP-graph
    replay over all of them),
  * meanwhile admits arrived requests in arrival order while they fit the
    HBM budget (their device tables are built while the GPU runs),
  * collects the solves that finished, reads their probe series and DFT and
    frees them, recording latency = results ready - arrival.

run_sequential() is the baseline: the same stream served one request at a
time with the single-solve path (tlm_step), each to completion.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import capi
from .problem import Tables


@dataclass
class Request:
    rid: int
    tables: Tables
    steps: int
    arrival: float                 # seconds after the start of the stream
    finish: float = float("nan")   # results ready, seconds after the start
    outputs: Optional[np.ndarray] = None
    dft: Optional[np.ndarray] = None
    handle: Optional[object] = field(default=None, repr=False)

    @property
    def cells(self) -> int:
        return self.tables.nx * self.tables.ny * self.tables.nz

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def request_from_files(rid: int, inp_path: str, arrival: float = 0.0, mesh_path: Optional[str] = None,
                       material_dir: Optional[str] = None, steps: Optional[int] = None) -> Request:
    """A request given in the reference's own formats: the `.inp` file (and
    the model builder's `.mesh`, when the geometry needs it), read by
    ingest.load; steps defaults to the deck's own step count."""
    from . import ingest
    from .problem import build_tables

    t = build_tables(ingest.load(inp_path, mesh_path, material_dir).problem)
    return Request(rid, t, int(steps or t.n_timesteps), arrival)


def _results(req: Request, h):
    out, _ = h.read_cell_outputs(req.steps)
    req.outputs = capi.probe_series(out, h.t) if len(out) else out
    req.dft = h.read_dft() if len(h.t.dft_frequencies) else None


def run_batched(requests: List[Request], form=capi.FORM_PRECISION_FP32, budget_bytes: int = 200 << 30,
                chunk: int = 64, lib=None, log=None) -> dict:
    """Serve the stream with continuous batching; returns statistics.

    Each round enqueues `chunk` steps of the resident solves (asynchronous),
    admits the requests that arrived meanwhile (their set-up runs on the host
    while the GPU steps), then collects the solves that finished."""
    lib = lib or capi.engine_library()
    b = capi.Batch(form, budget_bytes, lib)
    b.set_graph_steps(chunk)
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    waiting: deque = deque()
    inflight = {}
    t0 = time.perf_counter()
    peak_resident = 0

    def admit() -> int:
        n = 0
        while waiting:
            r = waiting[0]
            if r.handle is None:
                r.handle = capi.Handle(lib, r.tables, form, output_capacity=r.steps)
                r.handle.release_tables()
            if not b.add(r.handle, 1, r.steps):
                if not inflight:
                    raise RuntimeError(f"request {r.rid} does not fit the HBM budget on its own")
                break
            inflight[r.handle.h.value] = r
            waiting.popleft()
            n += 1
        return n

    while pending or waiting or inflight:
        now = time.perf_counter() - t0
        while pending and pending[0].arrival <= now:
            waiting.append(pending.popleft())
        running = bool(inflight)
        if running:
            b.run(chunk)
            peak_resident = max(peak_resident, b.info()[2])
        admitted = admit()
        if not running:
            if not admitted and pending:
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
            continue
        for h in b.collect():
            r = inflight.pop(h.h.value)
            _results(r, h)
            r.finish = time.perf_counter() - t0
            h.close()
            r.handle = None
            if log:
                log(f"[serve] request {r.rid} n={r.tables.nx} steps={r.steps} latency {r.latency:.3f} s")
    b.close()
    st = stats(requests)
    st["peak_resident_bytes"] = int(peak_resident)
    return st


def run_sequential(requests: List[Request], form=capi.FORM_PLAIN_FP64, lib=None, log=None) -> dict:
    """Serve the stream one request at a time (single-solve path)."""
    lib = lib or capi.engine_library()
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda q: q.arrival):
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        h = capi.Handle(lib, r.tables, form, output_capacity=r.steps)
        h.release_tables()
        h.step(1, r.steps)
        h.synchronize()
        _results(r, h)
        r.finish = time.perf_counter() - t0
        h.close()
        if log:
            log(f"[serial] request {r.rid} n={r.tables.nx} steps={r.steps} latency {r.latency:.3f} s")
    return stats(requests)


def stats(requests: List[Request]) -> dict:
    lat = np.array([r.latency for r in requests])
    first = min(r.arrival for r in requests)
    last = max(r.finish for r in requests)
    work = float(sum(r.cells * r.steps for r in requests))
    return {
        "requests": len(requests),
        "makespan_s": round(last - first, 3),
        "solves_per_s": round(len(requests) / (last - first), 3),
        "gcell_updates_per_s": round(work / (last - first) / 1e9, 3),
        "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
        "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
        "mean_latency_s": round(float(lat.mean()), 3),
    }


def poisson_arrivals(n: int, rate: float, rng) -> np.ndarray:
    """Arrival times (s) of a Poisson process of the given rate."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n)) - 0.0
