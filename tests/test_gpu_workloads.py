"""The measured workloads against the oracle, unmodified.

* Shielding enclosures (the batch workload) stepped together through the
  batch engine (tlm_batch: K-step graph replays, class block tables, PML
  recompute rotation) against the oracle solve by solve.
* Serving-stream requests, decks unmodified (materials, PML, probes as
  generated), served with continuous batching (serving.run_batched) against
  the oracle; only the step count is cut so that the oracle finishes in
  seconds.
* The 256^3 single large solve and the 384x384x192 dispersive load at full
  size for a few steps, with their x-segment layout and autotuned schedule,
  against the oracle's final link state and probe rows.

Plain FP64: bit-identical.  Precision form (FP32 links) and extended form
(FP32 links plus a residual byte): relative L2 per probe against the FP64
oracle below 1e-5 (a probe whose FP64 signal is negligible is measured
against 1e-3 of the solve's largest probe norm), and the link state within
1e-5 relative L2; on the serving requests the precision form is held to
2e-4 and the extended form to 1e-5.
"""
import os

import numpy as np
import pytest

from ggi_tlm_amd import capi, serving, workloads
from ggi_tlm_amd.capi import FORM_EXTENDED_FP32, FORM_PLAIN_FP64, FORM_PRECISION_FP32, probe_series
from ggi_tlm_amd.problem import build_tables

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
TOL = 1e-5
# Point-source requests ringing in small PML boxes with dielectric spheres
# (eps_r up to 8) or two-pole blocks accumulate the FP32 rounding of the
# links on their weak probes: measured up to 9.2e-5 after 400 steps on this
# stream (request 11, three two-pole blocks).  The extended form rounds each
# link 256 times less and is held to the 1e-5 target.
TOL_SERVE = {FORM_PRECISION_FP32: 2e-4, FORM_EXTENDED_FP32: 1e-5}
FORMS = [FORM_PLAIN_FP64, FORM_PRECISION_FP32, FORM_EXTENDED_FP32]


def oracle_solve(t, steps):
    from oracle import oracle

    h = oracle.handle(t, output_capacity=steps, threads=THREADS)
    h.step(1, steps)
    out = h.read_cell_outputs()[0]
    dft = h.read_dft() if len(t.dft_frequencies) and len(t.output_rows) else None
    V = h.download_links()[:, 1:-1, 1:-1]
    h.close()
    return out, dft, V


def check(form, t, got, want, label):
    (og, dg, Vg), (oo, do, Vo) = got, want
    assert og.shape == oo.shape, label
    if form == FORM_PLAIN_FP64:
        assert np.array_equal(og, oo), (label, np.max(np.abs(og - oo)))
        assert np.array_equal(Vg, Vo), label
        if do is not None:
            assert np.max(np.abs(dg - do)) <= 1e-12 * max(np.max(np.abs(do)), 1e-300), label
        return
    sg, so = probe_series(og, t), probe_series(oo, t)
    norms = [np.linalg.norm(so[:, k]) for k in range(so.shape[1])]
    floor = 1e-3 * max(norms)
    worst = 0.0
    for k in range(so.shape[1]):
        rel = np.linalg.norm(sg[:, k] - so[:, k]) / max(norms[k], floor)
        worst = max(worst, rel)
    print(f"[{label}] probes {so.shape[1]}, worst relative L2 {worst:.2e}", flush=True)
    assert worst < TOL, (label, worst)
    vrel = np.linalg.norm(Vg - Vo) / np.linalg.norm(Vo)
    assert vrel < TOL, (label, vrel)
    if do is not None:
        dn = [np.linalg.norm(do[p]) for p in range(do.shape[0])]
        if max(dn) == 0.0:
            assert not np.any(dg), label
            return
        for p in range(do.shape[0]):
            rel = np.linalg.norm(dg[p] - do[p]) / max(dn[p], 1e-3 * max(dn))
            assert rel < TOL, (label, "dft", p, rel)


def engine_results(h):
    out = h.read_cell_outputs()[0]
    dft = h.read_dft() if len(h.t.dft_frequencies) and len(h.t.output_rows) else None
    return out, dft, h.download_links()[:, 1:-1, 1:-1]


@pytest.mark.parametrize("form", FORMS)
def test_enclosure_batch_vs_oracle(form):
    steps = 500
    seeds = (0, 1, 2, 3)
    tables = [build_tables(workloads.enclosure(s)) for s in seeds]
    lib = capi.engine_library()
    b = capi.Batch(form, 64 << 30, lib)
    b.set_graph_steps(16)
    handles = []
    for t in tables:
        h = capi.Handle(lib, t, form, output_capacity=steps)
        assert b.add(h, 1, steps)
        handles.append(h)
    print(f"[enclosures] form={form} grids={[(t.nx, t.ny, t.nz) for t in tables]}", flush=True)
    done = []
    for _ in range(steps):
        b.run(64)
        done += b.collect()
        if len(done) == len(handles):
            break
    assert len(done) == len(handles)
    b.close()
    for s, t, h in zip(seeds, tables, handles):
        check(form, t, engine_results(h), oracle_solve(t, steps), f"enclosure {s}")
        h.close()


def serving_decks(seed=11, count=12, steps=400, max_n=90):
    """Requests of the serving stream as generated (seed 11, the stream the
    serving figures use), those with n <= max_n, each cut to `steps`."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        p = workloads.serving_request(rng)
        if p.nx <= max_n:
            out.append((i, build_tables(p), min(steps, p.n_timesteps), len(p.materials)))
    return out


@pytest.mark.parametrize("form", FORMS)
def test_serving_requests_vs_oracle(form):
    decks = serving_decks()
    assert any(m > 1 for *_, m in decks) and any(m == 0 for *_, m in decks)
    arr = serving.poisson_arrivals(len(decks), 200.0, np.random.default_rng(3))
    reqs = [serving.Request(i, t, s, float(arr[k])) for k, (i, t, s, _) in enumerate(decks)]
    st = serving.run_batched(reqs, form, budget_bytes=32 << 30, chunk=32)
    print(f"[serve] form={form} n={[t.nx for _, t, _, _ in decks]} {st}", flush=True)
    for r in reqs:
        t = r.tables
        oo, _, _ = oracle_solve(t, r.steps)
        so = probe_series(oo, t)
        sg = r.outputs
        assert sg.shape == so.shape, r.rid
        if form == FORM_PLAIN_FP64:
            assert np.array_equal(sg, so), r.rid
        else:
            norms = [np.linalg.norm(so[:, k]) for k in range(so.shape[1])]
            floor = 1e-3 * max(norms)
            rel = max(np.linalg.norm(sg[:, k] - so[:, k]) / max(norms[k], floor) for k in range(so.shape[1]))
            print(f"[serve] request {r.rid} n={t.nx} worst relative L2 {rel:.2e}", flush=True)
            assert rel < TOL_SERVE[form], (r.rid, rel)


@pytest.mark.parametrize("name", ["single_large", "dispersive_load"])
@pytest.mark.parametrize("form", FORMS)
def test_full_size_vs_oracle(name, form):
    # 16 autotuner steps (8 candidates on the x-segment layout), then enough
    # steps for the sources' fields to reach the probes
    steps = 120
    p = workloads.single_large(steps=steps) if name == "single_large" else workloads.dispersive_load(steps=steps)
    t = build_tables(p)
    h = capi.Handle(capi.engine_library(), t, form, output_capacity=steps)
    print(f"[full] {name} {t.nx}x{t.ny}x{t.nz} form={form} segments={h.layout_segments()}", flush=True)
    done = h.autotune(1, 1)
    assert 0 < done < steps
    print(f"[full] schedule {h.get_schedule()} after {done} tuning steps", flush=True)
    h.step(1 + done, steps - done)
    got = engine_results(h)
    h.close()
    want = oracle_solve(t, steps)
    assert np.abs(probe_series(want[0], t)).max() > 0, "no probe sees the field yet"
    check(form, t, got, want, name)
