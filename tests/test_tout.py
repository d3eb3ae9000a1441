"""The .tout probe-file writer against the reference's own output file text
(tests/golden/tem_pml_ey_head.tout: the first 400 lines of
TEM_PML_TEST/Ey.tout_ref, written by the reference)."""
import os

import numpy as np

from ggi_tlm_amd import tout

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "tem_pml_ey_head.tout")


def test_rewrite_reproduces_reference_bytes(tmp_path):
    times, series = tout.read_time_domain(FIX)
    out = tmp_path / "Ey.tout"
    tout.write_time_domain(str(out), times, series, n_timesteps=1799)
    assert out.read_text() == open(FIX).read()


def test_fortran_e_edit_descriptor():
    assert tout.fortran_e(0.0) == "   0.0000000E+00"
    assert tout.fortran_e(1.0) == "   0.1000000E+01"
    assert tout.fortran_e(-0.7840943e-07) == "  -0.7840943E-07"
    assert tout.fortran_e(9.99999996e-5) == "   0.1000000E-03"   # carry into the exponent
    assert tout.fortran_e(1.5e-120) == "   0.1500000-119"       # three-digit exponent drops the E


def test_tiny_values_print_as_zero(tmp_path):
    out = tmp_path / "p.tout"
    tout.write_time_domain(str(out), [0.0, 1e-9], np.array([[1e-31, -2e-31], [3e-30, 0.5]]))
    lines = out.read_text().splitlines()
    assert lines[3] == "   0.0000000E+00    1   0.0000000E+00"
    assert lines[4] == "   0.0000000E+00    2   0.0000000E+00"
    assert lines[5] == "   0.1000000E-08    1   0.3000000E-29"
