"""MI355X (gfx950) engine for GGI_TLM's SCN scatter/connect hot path.

The engine is the C-ABI library libtlmscn.so (include/tlmscn.h) built from
csrc/; this package holds the host side: reference-layout table set-up
(problem.py, filters.py), the ctypes binding (capi.py) and the solver
driver mirroring run_TLM (solver.py).
"""
__all__ = ["capi", "problem", "filters", "constants"]
