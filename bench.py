#!/usr/bin/env python3
"""Benchmark of the SCN hot path on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): 256^3
cells, dl = 5 mm, 8-cell PML on all six faces (R_pml = 1e-4, order 2,
PEC behind it), a dielectric sphere of radius 60 cells (eps_r = 4,
sigma = 0.01 S/m) at a seeded offset from the centre, an x-directed
Gaussian plane wave with E along z injected by a Huygens box two cells
inside the PML, 16 seeded probes; 4000 timed steps in the FP32-link
precision form (FP64 arithmetic, FP64 PML and filter states).

A step is one complete loop body (excitation, scatter, cell output,
connect, outer boundary) over the whole grid with inputs resident in HBM.
value = cells * timed steps / wall time of the timed region (Gcell-updates/s).

The roofline object prices the dominant kernel (the kernel class with
the longest average launch: the pure free-space line kernel or the
mixed-line kernel on this workload) at its algorithmic bytes per launch
(tlm_kernel_bytes: 24 link words per cell, state and table words) over its
average launch duration, measured with HIP events on the engine's own
stream in a separate profiled pass.  traffic is the HBM bytes per launch
of the same kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this
workload (profiles/traffic_<workload>.json, counters scaled by a copy of
known size), or null when no such file matches the kernel's algorithmic
bytes.

cpu_baseline times the CPU oracle (a single-threaded FP64 restatement of
the reference's loop, built with gcc -O2) on a bounded sample of the same
workload: the full 256^3 grid for a few steps.

Usage: python bench.py [--gpus 1] [--steps 4000] [--warmup 50]
       [--form precision|plain] [--no-cpu] [--profile-steps 200]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ggi_tlm_amd import capi, workloads  # noqa: E402
from ggi_tlm_amd.problem import build_tables  # noqa: E402

HBM_PEAK_GBS = 8000.0


def workload(n=256, npml=8, radius=60, steps=4000, seed=2024):
    return workloads.single_large(n=n, npml=npml, radius=radius, steps=steps, seed=seed)


def measured_traffic(a, kernel, alg_bytes):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload, if it was measured on the same kernel and tables."""
    name = f"{a.n}cube" if a.workload == "large" else a.workload
    path = os.path.join(ROOT, "profiles", f"traffic_{name}_{a.form}.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    if rec.get("algorithmic_bytes") != alg_bytes:
        return None
    return rec.get("traffic_bytes")


def host_cpu():
    """CPU model and the cores this process may use (lscpu / affinity)."""
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
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "usable_cores": usable, "os_cpu_count": os.cpu_count()}


def _oracle_rate(tables, steps, threads):
    """Gcell-updates/s of the oracle on `tables`: one untimed step (first
    touch), then `steps` timed steps."""
    from oracle import oracle

    h = oracle.handle(tables, output_capacity=steps + 2, threads=threads)
    h.step(1, 1)
    t0 = time.perf_counter()
    h.step(2, steps)
    t1 = time.perf_counter()
    h.close()
    N = tables.nx * tables.ny * tables.nz
    return N * steps / (t1 - t0) / 1e9, t1 - t0


def cpu_baseline(tables, steps, steps_all=None):
    """The CPU oracle (FP64 C restatement of the reference's loop, gcc -O2)
    on the same grid: single-threaded (the reference's sequential build)
    and over all usable cores as z-slabs (its MPI build's partition, as
    OpenMP threads).  value is the all-core rate."""
    from oracle import oracle

    host = host_cpu()
    threads = max(1, min(oracle.max_threads(), host["usable_cores"] or 1))
    r1, t1 = _oracle_rate(tables, steps, 1)
    steps_all = steps_all or max(steps, int(steps * min(threads, 8)))
    rn, tn = _oracle_rate(tables, steps_all, threads) if threads > 1 else (r1, t1)
    grid = f"{tables.nx}x{tables.ny}x{tables.nz}"
    return {"value": rn, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{grid} grid of the same workload: {steps_all} steps over {threads} z-slab threads "
                      f"({tn:.1f} s) and {steps} steps single-threaded ({t1:.1f} s); C oracle, gcc -O2, FP64",
            "sequential": {"value": r1, "cores": 1},
            "host": host}


def dtype_of(form):
    """Arithmetic and storage of a numerical form."""
    return {capi.FORM_PLAIN_FP64: "f64", capi.FORM_PRECISION_FP32: "f32-links/f64-arith",
            capi.FORM_EXTENDED_FP32: "f32+i8-links/f64-arith"}[form]


FORM_NAMES = {"plain": capi.FORM_PLAIN_FP64, "precision": capi.FORM_PRECISION_FP32,
              "extended": capi.FORM_EXTENDED_FP32}


def _build(problem):
    return build_tables(problem)


def build_all(problems, workers=8):
    """Tables of many decks, built in worker processes (CPU only, before any
    GPU call of this process)."""
    import concurrent.futures as cf
    import multiprocessing as mp

    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork")) as ex:
        return list(ex.map(_build, problems, chunksize=4))


def main_batch(a):
    """Batch of enclosure solves (shielding effectiveness), all resident and
    stepped together; beside it, the same solves one at a time in the plain
    FP64 single-solve path on a subset, as solves/s."""
    from ggi_tlm_amd import serving

    t_setup = time.perf_counter()
    probs = [workloads.enclosure(s, steps=a.batch_steps) for s in range(a.batch_solves)]
    tables = build_all(probs)
    cells = [t.nx * t.ny * t.nz for t in tables]
    print(f"[bench] batch: {len(tables)} enclosures, {sum(cells) / 1e6:.1f} M cells, tables in "
          f"{time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)
    lib = capi.engine_library()
    b = capi.Batch(capi.FORM_PRECISION_FP32, 250 << 30, lib)
    handles = []
    for t in tables:
        h = capi.Handle(lib, t, capi.FORM_PRECISION_FP32, output_capacity=a.batch_steps)
        h.release_tables()
        assert b.add(h, 1, a.batch_steps)
        handles.append(h)
    resident = b.info()[2]
    b.run(1)
    b.synchronize()
    print(f"[bench] batch admitted, {resident / 2**30:.1f} GiB resident; set-up {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    k, tuned = b.autotune(64)
    b.run(max(0, a.batch_steps - 1 - tuned))
    done = b.collect()
    elapsed = time.perf_counter() - t0
    assert len(done) == len(handles)
    work = sum(cells) * (a.batch_steps - 1)
    value = len(handles) / elapsed * (a.batch_steps - 1) / a.batch_steps
    # shielding effectiveness at the centre probe from one solve's DFT
    se = None
    h0 = handles[0]
    dft = h0.read_dft()
    if dft is not None and dft.size:
        se = float(np.median(np.abs(dft[0])))
    b.close()
    for h in handles:
        h.close()
    base = None
    if a.baseline_solves:
        sub = [serving.Request(i, tables[i], a.batch_steps, 0.0) for i in range(a.baseline_solves)]
        st = serving.run_sequential(sub, capi.FORM_PLAIN_FP64, lib)
        base = {"solves_per_s": st["solves_per_s"], "gcell_updates_per_s": st["gcell_updates_per_s"],
                "solves": len(sub), "form": "plain FP64, one solve at a time (tlm_step)"}
    line = {
        "metric": "solves/s", "value": round(value, 3), "unit": "solves/s", "n_gpus": 1,
        "steps": a.batch_steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype_of(capi.FORM_PRECISION_FP32), "data": "synthetic",
        "config": {"workload": f"{len(tables)} shielding enclosures (PEC box, one aperture, 6-cell PML, "
                               f"Huygens plane wave, centre probe DFT at 100 f), {a.batch_steps} steps each",
                   "form": "precision: FP32 link planes", "graph_steps": k, "cells": int(sum(cells)),
                   "gcell_updates_per_s": round(work / elapsed / 1e9, 3), "elapsed_s": round(elapsed, 3),
                   "resident_GiB": round(resident / 2**30, 2), "median_abs_dft_solve0": se},
        "baseline_one_at_a_time": base,
    }
    print(json.dumps(line), flush=True)


def timed_solve(t, form, steps, warmup, tune, profile_steps):
    """Autotune (untimed), warm up, time `steps` steps between two
    synchronisations, then one profiled pass of per-class kernel durations
    (HIP events on the engine's stream).  Returns the elapsed seconds, the
    per-class figures, the step's algorithmic bytes, the schedule and the
    x-segments."""
    h = capi.Handle(capi.engine_library(), t, form, output_capacity=1024)
    first = 1
    if tune:
        first += h.autotune(1, tune)
    h.step(first, warmup)
    first += warmup
    h.synchronize()
    t0 = time.perf_counter()
    h.step(first, steps)
    first += steps
    h.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"[bench] form {form}: timed {steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    h.set_profiling(True)
    h.step(first, profile_steps)
    kt = h.read_kernel_times()
    h.set_profiling(False)
    kb = h.kernel_bytes()
    per_class = {}
    step_bytes = 0
    for k, (n_launch, ms) in kt.items():
        cells, b = kb[k]
        avg_ms = ms / n_launch if n_launch else 0.0
        step_bytes += b
        per_class[k] = {"cells": cells, "bytes_per_launch": b, "avg_ms": round(avg_ms, 5),
                        "GB_s": round(b / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 else None}
    print(f"[bench] kernels {json.dumps(per_class)}", file=sys.stderr, flush=True)
    sched, segs = h.get_schedule(), h.layout_segments()
    h.read_cell_outputs()
    h.close()
    return elapsed, per_class, step_bytes, sched, segs


def main_accuracy(a):
    """Precision and extended forms against FP64 on the full-size workloads:
    relative L2 of every probe series and of every probe's DFT (transfer
    function), each against its own FP64 norm.  The FP64 side is the plain
    form of the engine, which equals the CPU oracle bit for bit on every
    parity deck (the oracle would take hours here)."""
    out = []
    for name in a.acc_workloads:
        if name == "large":
            p = workload(n=a.n, steps=a.acc_steps_large)
        else:
            p = workloads.dispersive_load(steps=a.acc_steps_dispersive)
        t = build_tables(p)
        steps = t.n_timesteps
        res = {}
        for form in (capi.FORM_PLAIN_FP64, capi.FORM_PRECISION_FP32, capi.FORM_EXTENDED_FP32):
            h = capi.Handle(capi.engine_library(), t, form, output_capacity=steps)
            t0 = time.perf_counter()
            h.step(1, steps)
            h.synchronize()
            el = time.perf_counter() - t0
            series = capi.probe_series(h.read_cell_outputs()[0], t)
            dft = h.read_dft() if len(t.dft_frequencies) else None
            h.close()
            res[form] = (series, dft)
            print(f"[bench] accuracy {name}: form {form}, {steps} steps in {el:.1f} s", file=sys.stderr, flush=True)
        s64, d64 = res[capi.FORM_PLAIN_FP64]
        norms = np.array([np.linalg.norm(s64[:, k]) for k in range(s64.shape[1])])
        rec = {"workload": name, "steps": steps,
               "probe_norm_rel_to_peak": [round(float(v), 4) for v in norms / norms.max()]}
        if d64 is not None:
            dn = np.array([np.linalg.norm(d64[k]) for k in range(d64.shape[0])])
            rec["dft_frequencies"] = len(t.dft_frequencies)
            rec["dft_norm_rel_to_peak"] = [round(float(v), 4) for v in dn / dn.max()]
        for fname, form in (("precision", capi.FORM_PRECISION_FP32), ("extended", capi.FORM_EXTENDED_FP32)):
            s32, d32 = res[form]
            rel = [float(np.linalg.norm(s32[:, k] - s64[:, k]) / norms[k]) for k in range(s64.shape[1])]
            f = {"probe_rel_l2": [float(f"{e:.3e}") for e in rel], "max_probe_rel_l2": float(f"{max(rel):.3e}"),
                 "max_rel_l2_probes_above_5pct_of_peak":
                     float(f"{max(e for e, n in zip(rel, norms) if n >= 0.05 * norms.max()):.3e}")}
            if d64 is not None:
                drel = [float(np.linalg.norm(d32[k] - d64[k]) / dn[k]) for k in range(d64.shape[0])]
                f["dft_rel_l2"] = [float(f"{e:.3e}") for e in drel]
                f["max_dft_rel_l2"] = float(f"{max(drel):.3e}")
            rec[fname] = f
        out.append(rec)
        print(f"[bench] accuracy {json.dumps(rec)}", file=sys.stderr, flush=True)
    line = {"metric": "largest per-probe relative L2 vs FP64 (extended form; precision form in config)",
            "value": max(r["extended"]["max_probe_rel_l2"] for r in out),
            "unit": "relative L2", "n_gpus": 1, "higher_is_better": False, "dtype": dtype_of(capi.FORM_EXTENDED_FP32),
            "data": "synthetic", "config": {"workloads": out,
                                            "precision_form_max_probe_rel_l2":
                                                max(r["precision"]["max_probe_rel_l2"] for r in out)}}
    print(json.dumps(line), flush=True)


def main_serve_batch(a):
    """The serving stream's requests (same seed) admitted at once and stepped
    to the end as one resident batch in the form --form names: the GPU-bound
    part of the saturated stream, for kernel profiles (one graph capture per
    chunk set, no arrivals).  The timed work counts each request's own
    steps after the untimed first chunk; the algorithmic bytes are each
    solve's per-step class bytes over those steps."""
    rng = np.random.default_rng(a.seed)
    probs = [workloads.serving_request(rng) for _ in range(a.serve_requests)]
    tables = build_all(probs)
    lib = capi.engine_library()
    form = FORM_NAMES[a.form]
    b = capi.Batch(form, 200 << 30, lib)
    chunk = 512
    b.set_graph_steps(chunk)
    hs = []
    step_bytes = []
    for t, p in zip(tables, probs):
        h = capi.Handle(lib, t, form, output_capacity=p.n_timesteps)
        h.release_tables()
        step_bytes.append(sum(v[1] for v in h.kernel_bytes().values()))
        if not b.add(h, 1, p.n_timesteps):
            raise RuntimeError("the serving requests do not fit the batch budget")
        hs.append(h)
    steps = max(p.n_timesteps for p in probs)
    b.run(chunk)  # first chunk (graph capture) untimed
    b.synchronize()
    t0 = time.perf_counter()
    done = chunk
    while done < steps:
        k = min(chunk, steps - done)
        b.run(k)
        done += k
    b.synchronize()
    dt = time.perf_counter() - t0
    n = 0
    while n < len(hs):
        got = b.collect()
        n += len(got)
        for h in got:
            h.close()
    b.close()
    timed_cells = sum(t.nx * t.ny * t.nz * max(0, p.n_timesteps - chunk) for t, p in zip(tables, probs))
    timed_bytes = sum(sb * max(0, p.n_timesteps - chunk) for sb, p in zip(step_bytes, probs))
    print(json.dumps({"metric": "Gcell-updates/s", "value": round(timed_cells / dt / 1e9, 3),
                      "unit": "Gcell-updates/s", "n_gpus": 1, "steps": steps - chunk, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": dtype_of(form), "data": "synthetic",
                      "config": {"workload": f"{len(tables)} serving requests (seed {a.seed}) stepped as one "
                                             "resident batch", "form": a.form, "seconds": round(dt, 3),
                                 "timed_cell_updates": int(timed_cells),
                                 "algorithmic_GB_s": round(timed_bytes / dt / 1e9, 1),
                                 "frac_of_hbm": round(timed_bytes / dt / 1e9 / HBM_PEAK_GBS, 4)}}))


def request_errors(test, ref):
    """Per request: the largest relative L2 over its probes of `test`
    against `ref` (each probe by its own norm), and the same error over the
    request's largest probe norm."""
    own, peak = [], []
    for a_, r in zip(test, ref):
        so, sr = a_.outputs, r.outputs
        norms = np.array([np.linalg.norm(sr[:, k]) for k in range(sr.shape[1])])
        errs = np.array([np.linalg.norm(so[:, k] - sr[:, k]) for k in range(sr.shape[1])])
        own.append(float(np.max(errs / np.where(norms > 0, norms, 1.0))))
        peak.append(float(np.max(errs) / max(norms.max(), 1e-300)))
    return own, peak


def main_serve(a):
    """Poisson stream of mixed requests served six ways on the same arrivals:
    continuous batching in the precision and extended forms, one request at
    a time in plain FP64 (the baseline), and the same-form baselines (one at
    a time in the precision and extended forms, batched in plain FP64) that
    separate the gain of batching from the gain of narrower links.  p50/p99
    latency and solves/s for each; per-request relative L2 of the batched
    precision and extended series against the one-at-a-time FP64 series."""
    from ggi_tlm_amd import serving

    rng = np.random.default_rng(a.seed)
    probs = [workloads.serving_request(rng) for _ in range(a.serve_requests)]
    t_setup = time.perf_counter()
    tables = build_all(probs)
    print(f"[bench] serve: {len(tables)} requests, tables in {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)
    lib = capi.engine_library()
    P32, P64, PX = capi.FORM_PRECISION_FP32, capi.FORM_PLAIN_FP64, capi.FORM_EXTENDED_FP32
    out = {}
    for rate in a.rates:
        arr = serving.poisson_arrivals(len(tables), rate, np.random.default_rng(a.seed + int(rate)))

        def stream():
            return [serving.Request(i, t, p.n_timesteps, float(arr[i])) for i, (t, p) in
                    enumerate(zip(tables, probs))]

        runs = {}
        for key, form, batched in (("batched_precision", P32, True), ("one_at_a_time_fp64", P64, False),
                                   ("one_at_a_time_precision", P32, False), ("batched_fp64", P64, True),
                                   ("batched_extended", PX, True), ("one_at_a_time_extended", PX, False)):
            reqs = stream()
            if batched:
                st = serving.run_batched(reqs, form, chunk=a.chunk, lib=lib)
            else:
                st = serving.run_sequential(reqs, form, lib)
            runs[key] = (st, reqs)
            print(f"[bench] rate {rate}: {key} {st}", file=sys.stderr, flush=True)
        ref = runs["one_at_a_time_fp64"][1]
        same64 = all(np.array_equal(x.outputs, y.outputs) for x, y in zip(runs["batched_fp64"][1], ref))
        sps = {k: v[0]["solves_per_s"] for k, v in runs.items()}

        def errors(key):
            own, peak = request_errors(runs[key][1], ref)
            return {"own_norm_max": float(f"{max(own):.3e}"), "own_norm_median": float(f"{np.median(own):.3e}"),
                    "requests_within_1e-5": int(sum(e <= 1e-5 for e in own)),
                    "peak_norm_max": float(f"{max(peak):.3e}"),
                    "requests_within_1e-5_of_peak": int(sum(e <= 1e-5 for e in peak)),
                    "own_norm": [float(f"{e:.2e}") for e in own]}

        out[str(rate)] = {
            **{k: v[0] for k, v in runs.items()},
            "speedup_vs_one_at_a_time_fp64": round(sps["batched_precision"] / sps["one_at_a_time_fp64"], 2),
            "speedup_extended_vs_one_at_a_time_fp64": round(sps["batched_extended"] / sps["one_at_a_time_fp64"], 2),
            "speedup_same_form_precision": round(sps["batched_precision"] / sps["one_at_a_time_precision"], 2),
            "speedup_same_form_extended": round(sps["batched_extended"] / sps["one_at_a_time_extended"], 2),
            "speedup_same_form_fp64": round(sps["batched_fp64"] / sps["one_at_a_time_fp64"], 2),
            "batched_fp64_bitwise_equal_to_one_at_a_time": bool(same64),
            "precision_vs_fp64_rel_l2_per_request": errors("batched_precision"),
            "extended_vs_fp64_rel_l2_per_request": errors("batched_extended"),
        }
    top = out[str(max(a.rates))]
    line = {
        "metric": "solves/s", "value": top["batched_precision"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype_of(P32),
        "data": "synthetic",
        "config": {"workload": f"{len(tables)} Poisson-arrival requests, cubic n log-U[24,160], steps U[1000,8000], "
                               "40% free / 35% sphere / 25% 2-pole blocks, 6-cell PML, 1-8 probes",
                   "rates_req_per_s": a.rates, "chunk_steps": a.chunk, "by_rate": out},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--form", choices=tuple(FORM_NAMES), default="precision")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extended", action="store_true", help="skip the extended-form figures of the default line")
    ap.add_argument("--mode", choices=("single", "batch", "serve", "serve-batch", "accuracy"), default="single")
    ap.add_argument("--acc-workloads", nargs="+", default=["large", "dispersive"])
    ap.add_argument("--acc-steps-large", type=int, default=4000)
    ap.add_argument("--acc-steps-dispersive", type=int, default=6000)
    ap.add_argument("--workload", choices=("large", "dispersive"), default="large")
    ap.add_argument("--tune-steps", type=int, default=4)
    ap.add_argument("--batch-solves", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=4096)
    ap.add_argument("--baseline-solves", type=int, default=8)
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--rates", type=float, nargs="+", default=[5.0, 200.0])
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--seed", type=int, default=11)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs one solve per GPU on a single MI355X; "
                                   "multi-GPU z-slab sharding is designed but not built (--gpus 1 only)"}))
        sys.exit(2)

    if a.mode == "batch":
        return main_batch(a)
    if a.mode == "serve-batch":
        return main_serve_batch(a)
    if a.mode == "serve":
        return main_serve(a)
    if a.mode == "accuracy":
        return main_accuracy(a)
    tune_steps = 14 * (a.tune_steps + 1) if a.tune_steps else 0  # at most 14 candidates
    total = tune_steps + a.warmup + a.steps + a.profile_steps
    t_setup = time.perf_counter()
    if a.workload == "dispersive":
        p = workloads.dispersive_load(steps=total)
        desc = (f"{p.nx}x{p.ny}x{p.nz} cells, 10-cell PML, 40 two-pole Debye/Lorentz blocks, probe DFT at "
                f"200 frequencies, 16 probes, {a.steps} timed steps")
    else:
        p = workload(n=a.n, steps=total)
        desc = (f"{p.nx}^3 cells, 8-cell PML, eps_r=4 sigma=0.01 sphere r=60, Huygens plane wave, "
                f"16 probes, {a.steps} timed steps")
    t = build_tables(p)
    form = FORM_NAMES[a.form]
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] set-up {setup_s:.1f} s; cells {t.nx * t.ny * t.nz}, PML {t.n_pml_cells}, "
          f"special {t.n_special_cells}, special faces {t.n_special_faces}", file=sys.stderr, flush=True)
    N = t.nx * t.ny * t.nz
    elapsed, per_class, step_bytes, sched, segs = timed_solve(t, form, a.steps, a.warmup, a.tune_steps,
                                                              a.profile_steps)
    gcells = N * a.steps / elapsed / 1e9
    dom = max((k for k in per_class if k != "dft" and per_class[k]["cells"]),
              key=lambda k: per_class[k]["avg_ms"])
    achieved = per_class[dom]["GB_s"]
    traffic = measured_traffic(a, dom, per_class[dom]["bytes_per_launch"])
    # the same workload in the extended form (every probe within 1e-5 of FP64)
    ext = None
    if form == capi.FORM_PRECISION_FP32 and not a.no_extended:
        ex_steps = min(a.steps, 2000)
        e_el, e_pc, e_bytes, e_sched, _ = timed_solve(t, capi.FORM_EXTENDED_FP32, ex_steps, a.warmup,
                                                       a.tune_steps, a.profile_steps)
        ext = {"value": round(N * ex_steps / e_el / 1e9, 3), "ms_per_step": round(e_el / ex_steps * 1e3, 4),
               "steps": ex_steps, "dtype": dtype_of(capi.FORM_EXTENDED_FP32), "link_bytes": 5,
               "bulk_GB_s": e_pc["bulk"]["GB_s"],
               "bulk_frac_of_hbm": round(e_pc["bulk"]["GB_s"] / HBM_PEAK_GBS, 4) if e_pc["bulk"]["GB_s"] else None,
               "whole_step_GB_s": round(e_bytes / (e_el / ex_steps) / 1e9, 1), "kernels": e_pc,
               "schedule": e_sched}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(t, a.cpu_steps)
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)

    ms_per_step = elapsed / a.steps * 1e3
    line = {
        "metric": "Gcell-updates/s",
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_of(form),
        "data": "synthetic",
        "config": {
            "workload": desc,
            "form": {capi.FORM_PLAIN_FP64: "plain FP64",
                     capi.FORM_PRECISION_FP32: "precision: FP32 link planes, FP64 arithmetic and states",
                     capi.FORM_EXTENDED_FP32: "extended: FP32 link planes plus an 8-bit residual plane (32-bit "
                                              "significand), FP64 arithmetic and states"}[form],
            "link_bytes": {capi.FORM_PLAIN_FP64: 8, capi.FORM_PRECISION_FP32: 4, capi.FORM_EXTENDED_FP32: 5}[form],
            "cells": N,
            "pml_cells": t.n_pml_cells,
            "special_cells": t.n_special_cells,
            "whole_step_GB_s": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
            "whole_step_frac_of_hbm": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": per_class,
            "schedule": sched,
            "x_segments": segs,
            "extended_form": ext,
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
