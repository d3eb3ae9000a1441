"""The C ABI boundary: libraries load and export every declared symbol."""
import ctypes
import os
import re

import pytest

from ggi_tlm_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "tlmscn.h")).read()
    return sorted(set(re.findall(r"\b(tlm_[a-z_]+)\s*\(", txt)))


def batch_symbols():
    txt = open(os.path.join(ROOT, "include", "tlmbatch.h")).read()
    return sorted(set(re.findall(r"\b(tlm_[a-z_]+)\s*\(", txt)))


def test_header_matches_binding():
    names = declared_symbols()
    assert set(names) == set(capi.EXPORTED), set(names) ^ set(capi.EXPORTED)
    assert set(batch_symbols()) == set(capi.BATCH_EXPORTED), set(batch_symbols()) ^ set(capi.BATCH_EXPORTED)


def test_fortran_module_binds_every_single_solve_entry_point():
    """tlm_gpu_iface.F90 declares every tlmscn.h entry point under its C name
    (the batch ABI, tlmbatch.h, is for C and Python serving hosts)."""
    txt = open(os.path.join(ROOT, "ggi_tlm_amd", "fortran", "tlm_gpu_iface.F90")).read()
    bound = set(re.findall(r"name='(tlm_[a-z_]+)'", txt))
    assert set(declared_symbols()) <= bound, set(declared_symbols()) - bound


def test_engine_exports_batch_symbols():
    if not os.path.exists(capi.ENGINE_SO):
        pytest.fail("libtlmscn.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(capi.ENGINE_SO)
    for n in batch_symbols():
        assert hasattr(lib, n), n


def test_engine_exports_every_symbol():
    if not os.path.exists(capi.ENGINE_SO):
        pytest.fail("libtlmscn.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(capi.ENGINE_SO)
    for n in declared_symbols():
        assert hasattr(lib, n), n


def test_oracle_exports_same_abi(oracle_lib):
    for n in declared_symbols():
        assert hasattr(oracle_lib.dll, "oracle_" + n[len("tlm_"):]), n


def test_missing_engine_fails_loudly(tmp_path):
    with pytest.raises(FileNotFoundError):
        capi.Library(str(tmp_path / "libtlmscn.so"))


def test_struct_sizes_match_header():
    from ggi_tlm_amd.problem import HUYGENS_FACE_DTYPE, PML_DTYPE, POINT_SOURCE_DTYPE, VMAT_DTYPE
    # sizes implied by include/tlmscn.h on LP64
    assert ctypes.sizeof(capi.TlmGrid) == 4 * 4 + 9 * 8
    assert VMAT_DTYPE.itemsize == 6 * 4 + 4 * 8 + 4 * 9 * 8
    assert PML_DTYPE.itemsize == 18 * 8
    assert POINT_SOURCE_DTYPE.itemsize == 5 * 4
    assert HUYGENS_FACE_DTYPE.itemsize == 4 * 4 + 8
