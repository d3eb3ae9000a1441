"""Python access to the CPU oracle (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU-baseline leg use this.  The
oracle library (liboracle_scn.so, built from scn_oracle.c by the Makefile
here) exports the engine's C ABI under the prefix ``oracle_``, so the same
host binding drives both.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_scn.so")

_lib = None


def library():
    from ggi_tlm_amd.capi import Library

    global _lib
    if _lib is None:
        _lib = Library(ORACLE_SO, "oracle_",
                       extra=[("zfilter_sequence", (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32,
                                                              C.c_void_p, C.c_int32, C.c_void_p])),
                              ("set_threads", (C.c_int, [C.c_void_p, C.c_int32])),
                              ("max_threads", (C.c_int, []))])
    return _lib


def handle(tables, form=0, output_capacity=None, threads=1):
    """An oracle solve; threads > 1 runs scatter and connect over that many
    z-slabs (identical results)."""
    from ggi_tlm_amd.capi import Handle

    h = Handle(library(), tables, form, output_capacity)
    if threads != 1:
        rc = library().fn["set_threads"](h.h, int(threads))
        if rc:
            raise ValueError(f"oracle_set_threads({threads}) failed")
    return h


def max_threads() -> int:
    """Threads the oracle's OpenMP runtime would use by default."""
    return int(library().fn["max_threads"]())


def build():
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def zfilter_sequence(a, b, G):
    """Outputs of the oracle's filter recursion for the input sequence G."""
    import numpy as np

    order = len(a) - 1
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    G = np.ascontiguousarray(G, dtype=np.float64)
    out = np.zeros(len(G))
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    rc = library().fn["zfilter_sequence"](p(a), p(b), order, p(G), len(G), p(out))
    assert rc == 0
    return out


def face_outputs(h):
    """face_output_field(row, side, 1:6) after the last connect (consumes
    the recorded blocks)."""
    blocks, _ = h.read_face_outputs()
    return blocks[-1]
