"""Batch engine (include/tlmbatch.h): every solve gives the same bits alone
and in a mix of uneven solves admitted at different times, in both
numerical forms; admission respects the HBM budget."""
import numpy as np
import pytest

import problems
from ggi_tlm_amd import capi
from ggi_tlm_amd.capi import FORM_EXTENDED_FP32, FORM_PLAIN_FP64, FORM_PRECISION_FP32
from ggi_tlm_amd.problem import build_tables

pytestmark = pytest.mark.gpu

DECKS = [
    ("free_cube", lambda: problems.free_cube(16, 60, dft=np.linspace(1e8, 3e9, 5)), 40),
    ("pml_sphere", problems.pml_sphere, 60),
    ("special_mix", problems.special_mix, 33),
    ("dispersive_blocks", problems.dispersive_blocks, 50),
    ("pml_material", lambda: problems.pml_material(steps=120), 57),
    ("free_cube_odd", lambda: problems.free_cube(13, 41, r=0.3), 41),
    ("pml_sphere_64", lambda: problems.pml_sphere(64, 50, 8, radius=20.0), 21),
    ("face_outputs", problems.face_outputs, 45),
    ("huygens_face_outputs", problems.huygens_face_outputs, 70),
]


def results(h):
    out, first = h.read_cell_outputs()
    fo, ffirst = h.read_face_outputs()
    return out, first, h.read_dft(), h.download_links()[:, 1:-1, 1:-1], fo, ffirst


@pytest.mark.parametrize("kernels", [(-1, 2), (1, 1), (0, 2)])
@pytest.mark.parametrize("form", [FORM_PLAIN_FP64, FORM_PRECISION_FP32, FORM_EXTENDED_FP32])
def test_batch_matches_alone(form, kernels):
    lib = capi.engine_library()
    tables = [(name, build_tables(f()), n) for name, f, n in DECKS]
    alone = {}
    for name, t, n in tables:
        h = capi.Handle(lib, t, form, output_capacity=n)
        h.step(1, n)
        alone[name] = results(h)
        h.close()
    b = capi.Batch(form, 16 << 30)
    b.set_graph_steps(8)
    if form == FORM_EXTENDED_FP32 and kernels[0] == 0:
        with pytest.raises(capi.TlmError, match="LDS-staged"):
            b.set_kernels(*kernels)     # the extended form has the LDS-staged mixed kernel only
        b.close()
        return
    b.set_kernels(*kernels)
    handles = {name: capi.Handle(lib, t, form, output_capacity=n) for name, t, n in tables}
    steps = {name: n for name, _, n in tables}
    done = {}
    # three solves first, the rest admitted while those run
    for name in list(handles)[:3]:
        assert b.add(handles[name], 1, steps[name])
    b.run(7)
    for name in list(handles)[3:]:
        assert b.add(handles[name], 1, steps[name])
    print(f"[batch] form={form} kernels={kernels} solves={b.info()}", flush=True)
    for _ in range(20):
        b.run(13)
        for h in b.collect():
            name = next(k for k, v in handles.items() if v is h)
            done[name] = results(h)
        if len(done) == len(handles):
            break
    assert set(done) == set(handles)
    for name in handles:
        ao, af, ad, av, afo, aff = alone[name]
        bo, bf, bd, bv, bfo, bff = done[name]
        assert af == bf == 1 and ao.shape == bo.shape, name
        assert np.array_equal(ao, bo), name
        assert np.array_equal(ad, bd), name
        assert np.array_equal(av, bv), name
        # face outputs of every connect, the last one applied after collection
        assert aff == bff and afo.shape == bfo.shape, name
        assert np.array_equal(afo, bfo), name
        if "face" in name:
            assert afo.shape[0] == steps[name] and np.abs(afo).max() > 0, name
    b.close()
    for h in handles.values():
        h.close()


@pytest.mark.parametrize("form", [FORM_PLAIN_FP64, FORM_PRECISION_FP32, FORM_EXTENDED_FP32])
def test_batch_flush_and_continue_alone(form):
    """The small-value flush inside a batch (PML cells recompute the previous
    step's incident voltages with that step's flush), and a collected solve
    continued on the single-solve path (the link-buffer rotation and the
    face-output series handed back): the same bits as stepping alone
    throughout."""
    lib = capi.engine_library()
    decks = [("pml_sphere", build_tables(problems.pml_sphere(24, 60)), 30, 60),
             ("face_outputs", build_tables(problems.face_outputs(14, 40)), 17, 40)]
    for name, t, n_batch, n_total in decks:
        a = capi.Handle(lib, t, form, output_capacity=n_total)
        a.set_small_to_zero(3, 1e-5)
        a.step(1, n_total)
        ref = results(a)
        a.close()
        h = capi.Handle(lib, t, form, output_capacity=n_total)
        h.set_small_to_zero(3, 1e-5)
        b = capi.Batch(form, 16 << 30)
        b.set_graph_steps(4)
        assert b.add(h, 1, n_batch)
        got = []
        for _ in range(n_batch):
            b.run(1)
            got += b.collect()
            if got:
                break
        assert got and got[0] is h, name
        b.close()
        h.step(n_batch + 1, n_total - n_batch)
        out, first, dft, links, fo, ffirst = results(h)
        assert first == 1 and np.array_equal(out, ref[0]), name
        assert np.array_equal(links, ref[3]), name
        # face outputs: the batch's connects, then the single-solve path's
        assert ffirst == ref[5] and np.array_equal(fo, ref[4]), name
        h.close()


def test_batch_admission_by_budget():
    lib = capi.engine_library()
    t = build_tables(problems.pml_sphere(24, 40))
    a = capi.Handle(lib, t, FORM_PRECISION_FP32, output_capacity=40)
    c = capi.Handle(lib, t, FORM_PRECISION_FP32, output_capacity=40)
    b = capi.Batch(FORM_PRECISION_FP32, 1)
    assert not b.add(a, 1, 40)            # nothing fits a 1-byte budget
    b.close()
    a.kernel_bytes()                      # builds the device tables: exact footprint
    c.kernel_bytes()
    dev = a.device_bytes()
    assert dev > 0 and c.device_bytes() == dev
    b = capi.Batch(FORM_PRECISION_FP32, dev * 3 // 2)
    assert b.add(a, 1, 40)
    assert not b.add(c, 1, 40)            # the second does not fit beside the first
    b.run(40)
    assert [h is a for h in b.collect()] == [True]
    assert b.add(c, 1, 40)                # room again after collection
    b.run(40)
    assert [h is c for h in b.collect()] == [True]
    oa, _ = a.read_cell_outputs()
    oc, _ = c.read_cell_outputs()
    assert np.array_equal(oa, oc)
    b.close()
    a.close()
    c.close()
