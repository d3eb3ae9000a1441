"""tlm_download_links_noalloc, the allocation-free download: through the C
ABI it returns exactly what tlm_download_links returns (bit for bit, the
caller's halo kept), on two calls in a row with the second on a smaller
grid, in both forms; it refuses a handle that is not set up and a z-plane
larger than its staging buffer."""
import numpy as np
import pytest

import problems
from ggi_tlm_amd import capi
from ggi_tlm_amd.capi import FORM_EXTENDED_FP32, FORM_PLAIN_FP64, FORM_PRECISION_FP32
from ggi_tlm_amd.problem import ExcitationFunction, OutputPoint, PointSource, Problem, build_tables

pytestmark = pytest.mark.gpu

STAGING = 4 * 1024 * 1024  # TLM_NOALLOC_STAGING_DOUBLES


def marked(t, seed):
    """A Fortran-layout link array whose every entry (halo included) holds
    a marker value, so that a kept halo and a written interior both show."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((12, t.nx + 2, t.ny + 2, t.nz)))


def both(h, t, seed):
    V = marked(t, seed)
    a = h.download_links(V.copy(order="F"))
    b = h.download_links_noalloc(V.copy(order="F"))
    return V, a, b


@pytest.mark.parametrize("form", [FORM_PLAIN_FP64, FORM_PRECISION_FP32, FORM_EXTENDED_FP32])
def test_noalloc_download_equals_download(form):
    lib = capi.engine_library()
    # first call: 80^3 (12*82*82*80 doubles, more than one staging slab);
    # second call: a smaller grid
    decks = [problems.pml_sphere(80, 30, 6, radius=20.0), problems.special_mix(14, 30)]
    for k, p in enumerate(decks):
        t = build_tables(p)
        plane = 12 * (t.nx + 2) * (t.ny + 2)
        print(f"[noalloc] call {k + 1}: {t.nx}x{t.ny}x{t.nz} form={form} slabs={-(-t.nz // (STAGING // plane))}",
              flush=True)
        h = capi.Handle(lib, t, form, output_capacity=30)
        h.step(1, 29)               # the last connect pending: the download applies it
        V, a, b = both(h, t, 11 + k)
        assert np.array_equal(a, b)
        assert np.array_equal(a[:, 0, :, :], V[:, 0, :, :])      # halo kept
        assert not np.array_equal(a[:, 1:-1, 1:-1, :], V[:, 1:-1, 1:-1, :])
        h.scatter(30)               # scattered state, no connect pending
        _, a, b = both(h, t, 21 + k)
        assert np.array_equal(a, b)
        h.close()


def test_noalloc_download_refusals():
    lib = capi.engine_library()
    t = build_tables(problems.free_cube(8, 4))
    h = capi.Handle(lib, t, FORM_PLAIN_FP64, output_capacity=4)
    with pytest.raises(capi.TlmError, match="set up"):
        h.download_links_noalloc()
    h.close()
    # a z-plane of 12*(nx+2)*(ny+2) > 4 Mi doubles
    dl = 0.01
    p = Problem(nx=600, ny=600, nz=2, dl=dl, n_timesteps=4, reflection=(-1.0,) * 6,
                functions=[ExcitationFunction("gaussian", (1.0, 4 * dl / 6e8, 12 * dl / 6e8))],
                sources=[PointSource((300, 300, 1), "ez", 1)], outputs=[OutputPoint((301, 300, 1), "ez")])
    t = build_tables(p)
    assert 12 * (t.nx + 2) * (t.ny + 2) > STAGING
    h = capi.Handle(lib, t, FORM_PRECISION_FP32, output_capacity=4)
    h.step(1, 2)
    with pytest.raises(capi.TlmError, match="staging buffer"):
        h.download_links_noalloc()
    V = h.download_links()         # the allocating entry point still serves it
    assert np.isfinite(V).all()
    h.close()
