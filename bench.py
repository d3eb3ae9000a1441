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

from ggi_tlm_amd import capi  # noqa: E402
from ggi_tlm_amd.filters import Sfilter  # noqa: E402
from ggi_tlm_amd.problem import (PML, Box, ExcitationFunction, HuygensBox, OutputPoint,  # noqa: E402
                                 Problem, Sphere, VolumeMaterial, build_tables)

HBM_PEAK_GBS = 8000.0


def workload(n=256, npml=8, radius=60, steps=4000, seed=2024):
    rng = np.random.default_rng(seed)
    dl = 5e-3
    dt = dl / (2 * 2.99792458e8)
    off = rng.integers(-40, 41, size=3)
    cx, cy, cz = (n / 2 + float(o) for o in off)
    b = npml + 3  # Huygens box; source faces sit on the cells just outside it
    probes = []
    comps = ("ez", "ez", "ez", "ez", "ey", "hy", "hx", "ex")
    while len(probes) < 16:
        p = rng.integers(b + 1, n - b, size=3)
        if (p[0] - 0.5 - cx) ** 2 + (p[1] - 0.5 - cy) ** 2 + (p[2] - 0.5 - cz) ** 2 <= (radius + 2) ** 2:
            continue
        probes.append(OutputPoint(tuple(int(v) for v in p), comps[len(probes) % len(comps)]))
    return Problem(
        nx=n, ny=n, nz=n, dl=dl, n_timesteps=steps, reflection=(-1.0,) * 6,
        pml=PML((npml,) * 6, (1e-4,) * 6, 2),
        materials=[VolumeMaterial("dispersive", Sfilter.constant(4.0), Sfilter.constant(1.0), sigma_e=0.01)],
        material_regions=[(Sphere(cx, cy, cz, radius), 1)],
        functions=[ExcitationFunction("gaussian", (1.0, 40 * dt, 160 * dt))],
        huygens=HuygensBox(Box(b, n - b + 1, b, n - b + 1, b, n - b + 1), 1, 90.0, 0.0, -1.0, 0.0, side=1),
        outputs=probes)


def measured_traffic(a, kernel, alg_bytes):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload, if it was measured on the same kernel and tables."""
    path = os.path.join(ROOT, "profiles", f"traffic_{a.n}cube_{a.form}.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    if rec.get("algorithmic_bytes") != alg_bytes:
        return None
    return rec.get("traffic_bytes")


def cpu_baseline(tables, steps):
    """Oracle on the same grid for `steps` steps; per-step time from the
    difference of two step counts removes set-up."""
    from oracle import oracle

    h = oracle.handle(tables, output_capacity=2 * steps + 2)
    h.step(1, 1)  # touch everything once
    t0 = time.perf_counter()
    h.step(2, steps)
    t1 = time.perf_counter()
    h.close()
    N = tables.nx * tables.ny * tables.nz
    rate = N * steps / (t1 - t0) / 1e9
    return {"value": rate, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{tables.nx}x{tables.ny}x{tables.nz} grid of the same workload, {steps} steps "
                      f"({t1 - t0:.1f} s), single-threaded C oracle (gcc -O2, FP64)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--form", choices=("precision", "plain"), default="precision")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs one solve per GPU on a single MI355X; "
                                   "multi-GPU z-slab sharding is designed but not built (--gpus 1 only)"}))
        sys.exit(2)

    total = a.warmup + a.steps + a.profile_steps
    t_setup = time.perf_counter()
    p = workload(n=a.n, steps=total)
    t = build_tables(p)
    form = capi.FORM_PRECISION_FP32 if a.form == "precision" else capi.FORM_PLAIN_FP64
    h = capi.Handle(capi.engine_library(), t, form, output_capacity=1024)
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] set-up {setup_s:.1f} s; cells {t.nx * t.ny * t.nz}, PML {t.n_pml_cells}, "
          f"special {t.n_special_cells}, special faces {t.n_special_faces}", file=sys.stderr, flush=True)

    # warm-up, then the timed region
    h.step(1, a.warmup)
    h.synchronize()
    t0 = time.perf_counter()
    h.step(a.warmup + 1, a.steps)
    h.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    N = t.nx * t.ny * t.nz
    gcells = N * a.steps / elapsed / 1e9
    print(f"[bench] timed {a.steps} steps in {elapsed:.3f} s: {gcells:.2f} Gcell/s", file=sys.stderr, flush=True)

    # profiled pass: per-class kernel durations with HIP events on the engine stream
    h.set_profiling(True)
    h.step(a.warmup + a.steps + 1, a.profile_steps)
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
    dom = max((k for k in per_class if k != "dft" and per_class[k]["cells"]),
              key=lambda k: per_class[k]["avg_ms"])
    achieved = per_class[dom]["GB_s"]
    traffic = measured_traffic(a, dom, per_class[dom]["bytes_per_launch"])
    h.read_cell_outputs()
    h.close()

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
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{t.nx}^3 cells, 8-cell PML, eps_r=4 sigma=0.01 sphere r=60, Huygens plane wave, "
                        f"16 probes, {a.steps} timed steps",
            "form": "precision: FP32 link planes, FP64 arithmetic and states" if form else "plain FP64",
            "link_bytes": 4 if form else 8,
            "cells": N,
            "pml_cells": t.n_pml_cells,
            "special_cells": t.n_special_cells,
            "whole_step_GB_s": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
            "whole_step_frac_of_hbm": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": per_class,
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
