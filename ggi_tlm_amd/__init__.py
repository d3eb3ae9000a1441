"""MI355X (gfx950) engine for GGI_TLM's SCN scatter/connect hot path.

The engine is the C-ABI library libtlmscn.so (include/tlmscn.h for one
solve, include/tlmbatch.h for many resident solves) built from csrc/ by
build.py.  This package holds the host side: reference-layout table set-up
(problem.py, filters.py), the ctypes binding (capi.py), the measured
workloads (workloads.py), the continuous-batching request loop
(serving.py), the .tout probe-file writer (tout.py) and the Fortran
iso_c_binding module (fortran/).
"""
__all__ = ["capi", "problem", "filters", "constants", "workloads", "serving", "tout"]
