"""The oracle against the reference's own golden results (TEST_DATA/*_ref,
extracted into tests/golden/reference_goldens.npz by make_golden.py).

The reference prints probe series with (E16.7,I5,E16.7), so these pin the
oracle to about 1e-7 of the peak.  Each check rounds the oracle's series
the same way the reference's writer does (values below 1e-30 printed as 0,
output_points.F90:597) and, for .fout files, applies the post-processor's
Fourier transform (POST_PROCESSOR/fourier_transform.F90:34-200: linear
frequency list fmin + (k-1) fstep, dt taken from the first two printed
times, sum of v(t_n) exp(-j w t_n) dt).
"""
import os

import numpy as np
import pytest

import problems
from ggi_tlm_amd.capi import probe_series
from ggi_tlm_amd.filters import Sfilter
from ggi_tlm_amd.problem import build_tables

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.npz"))


def e16_7(x):
    return float("%.6e" % x)


def run_oracle(p, steps=None):
    from oracle import oracle

    t = build_tables(p)
    n = t.n_timesteps if steps is None else steps
    h = oracle.handle(t, output_capacity=n)
    h.step(1, n)
    out, first = h.read_cell_outputs()
    assert first == 1 and out.shape[0] == n
    return t, probe_series(out, t)


def printed(series, dt):
    times = np.array([e16_7(i * dt) for i in range(len(series))])
    vals = np.array([0.0 if abs(v) < 1e-30 else e16_7(v) for v in series])
    return times, vals


def post_process_ft(times, vals, fmin, fmax, n):
    fstep = (fmax - fmin) / float(n - 1)
    f = np.array([fmin + k * fstep for k in range(n)])
    dt = times[1] - times[0]
    return np.array([np.sum(vals * np.exp(-1j * 2 * np.pi * fk * times) * dt) for fk in f])


@pytest.mark.parametrize("key,eps,mu", [
    ("cavity_free_space_fout", None, None),
    ("cavity_epsr2_fout", Sfilter(1e8, [2.0], [1.0]), None),
    ("cavity_mur2_fout", None, Sfilter(1e8, [2.0], [1.0])),
])
def test_cavity_fourier_transform(oracle_lib, key, eps, mu):
    t, s = run_oracle(problems.cavity(eps, mu))
    assert t.n_timesteps == 5996
    times, vals = printed(s[:, 0], t.dt)
    X = post_process_ft(times, vals, 1e6, 3e8, 1000)
    ref = GOLD[key]
    err = np.max(np.abs(X - (ref[:, 2] + 1j * ref[:, 3]))) / np.max(ref[:, 4])
    assert err < 5e-7, err


def test_material_test2_lorentz(oracle_lib):
    ref = GOLD["material_test2_tout"]
    steps = 2000
    t, s = run_oracle(problems.material_test2(), steps)
    assert t.n_timesteps == len(ref)
    r = ref[:steps, 2]
    assert np.allclose(ref[:steps, 0], [e16_7(i * t.dt) for i in range(steps)], rtol=1e-6, atol=0)
    err = np.max(np.abs(s[:, 0] - r)) / np.max(np.abs(r))
    assert err < 5e-7, err


def test_huygens_surface_centre_fields(oracle_lib):
    t, s = run_oracle(problems.huygens_surface())
    assert t.n_timesteps == 120
    for i, c in enumerate(("ex", "ey", "ez", "hx", "hy", "hz")):
        r = GOLD[f"huygens_{c}_centre_tout"][:, 2]
        err = np.max(np.abs(s[:, i] - r)) / np.max(np.abs(r))
        assert err < 1e-6, (c, err)


def test_tem_pml_reflection(oracle_lib):
    t, s = run_oracle(problems.tem_pml())
    r = GOLD["tem_pml_ey_tout"][:, 2]
    assert t.n_timesteps == len(r) == 1799
    err = np.max(np.abs(s[:, 0] - r)) / np.max(np.abs(r))
    assert err < 2e-7, err
