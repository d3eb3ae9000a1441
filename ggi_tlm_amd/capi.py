"""ctypes binding of the C ABI declared in include/tlmscn.h.

The same signature table binds any library exporting the ABI under a
prefix ("tlm_" for the HIP engine, libtlmscn.so).  Python is only the host
shell here; every hot-path operation runs in the library.  There is no
fallback: if the engine library is missing, :func:`engine_library` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .problem import Tables

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_SO = os.path.join(HERE, "libtlmscn.so")

# tlm_frequency_output (include/tlmscn.h)
FREQUENCY_OUTPUT_DTYPE = np.dtype([("frequency", np.float64), ("kind", np.int32), ("field_component", np.int32),
                                   ("first_timestep", np.int32), ("last_timestep", np.int32),
                                   ("timestep_interval", np.int32), ("n_items", np.int32)])

# tlm_far_field_surface and tlm_rcs_surface (include/tlmscn.h)
FAR_FIELD_DTYPE = np.dtype([("frequency", np.float64), ("theta_min", np.float64), ("theta_step", np.float64),
                            ("phi_min", np.float64), ("phi_step", np.float64), ("n_theta", np.int32),
                            ("n_phi", np.int32), ("dim", np.int32), ("n_faces", np.int32)])
RCS_DTYPE = np.dtype([("theta", np.float64), ("phi", np.float64), ("n_faces", np.int32), ("n_points", np.int32),
                      ("point_offset", np.int32), ("points_per_cell", np.int32)])

# tlm_field_output (include/tlmscn.h)
FIELD_OUTPUT_DTYPE = np.dtype([("kind", np.int32), ("field_component", np.int32), ("both_sides", np.int32),
                               ("first_timestep", np.int32), ("last_timestep", np.int32),
                               ("timestep_interval", np.int32), ("n_items", np.int32)])

FORM_PLAIN_FP64 = 0
FORM_PRECISION_FP32 = 1
FORM_EXTENDED_FP32 = 2   # FP32 links plus an 8-bit residual plane (32-bit significand)


class TlmGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("n_timesteps", C.c_int32), ("dl", C.c_double), ("dt", C.c_double),
                ("reduced_c_factor", C.c_double),
                ("r_xmin", C.c_double), ("r_xmax", C.c_double), ("r_ymin", C.c_double),
                ("r_ymax", C.c_double), ("r_zmin", C.c_double), ("r_zmax", C.c_double)]


_p = C.c_void_p
_i32 = C.c_int32
_h = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "create": (C.c_int, [C.POINTER(TlmGrid), _i32, C.POINTER(C.c_void_p)]),
    "destroy": (None, [_h]),
    "last_error": (C.c_char_p, [_h]),
    "device_bytes": (C.c_int64, [_h]),
    "set_cell_codes": (C.c_int, [_h, _p, _i32, _p, _p, _p, _p, _p]),
    "set_face_codes": (C.c_int, [_h, _p, _i32, _p, _p, _p, _p]),
    "set_volume_materials": (C.c_int, [_h, _i32, _p]),
    "set_surface_materials": (C.c_int, [_h, _i32, _p]),
    "set_pml": (C.c_int, [_h, _i32, _p, _i32, _p]),
    "set_excitation_functions": (C.c_int, [_h, _i32, _p, _p]),
    "set_point_sources": (C.c_int, [_h, _i32, _p, _i32, _i32]),
    "set_huygens_surface": (C.c_int, [_h, _i32, C.c_double, _p, _p, _i32, _p]),
    "set_output_points": (C.c_int, [_h, _i32, _p, _p]),
    "set_dft_frequencies": (C.c_int, [_h, _i32, _p]),
    "upload_links": (C.c_int, [_h, _p]),
    "download_links": (C.c_int, [_h, _p]),
    "download_links_noalloc": (C.c_int, [_h, _p]),
    "excitation": (C.c_int, [_h, _i32]),
    "scatter": (C.c_int, [_h, _i32]),
    "connect": (C.c_int, [_h, _i32]),
    "outer_boundary": (C.c_int, [_h]),
    "step": (C.c_int, [_h, _i32, _i32]),
    "synchronize": (C.c_int, [_h]),
    "set_output_capacity": (C.c_int, [_h, _i32]),
    "read_cell_outputs": (C.c_int, [_h, _p, _i32, C.POINTER(_i32), C.POINTER(_i32)]),
    "read_dft": (C.c_int, [_h, _p]),
    "set_frequency_outputs": (C.c_int, [_h, _i32, _p, _p, _p, _p]),
    "read_frequency_output": (C.c_int, [_h, _i32, _p]),
    "set_field_outputs": (C.c_int, [_h, _i32, _p, _p, _p, _p]),
    "read_field_output": (C.c_int, [_h, _i32, _p, _i32, _p, _p]),
    "set_far_field_surfaces": (C.c_int, [_h, _i32, _p, _p, _p, _p, _p]),
    "read_far_field": (C.c_int, [_h, _i32, _p]),
    "set_rcs_surface": (C.c_int, [_h, _p, _p, _p, _p, _p]),
    "read_rcs": (C.c_int, [_h, _p]),
    "set_schedule": (C.c_int, [_h, _i32, _i32, _i32, _i32, _i32]),
    "set_profiling": (C.c_int, [_h, _i32]),
    "set_small_to_zero": (C.c_int, [_h, _i32, C.c_double]),
    "autotune": (C.c_int, [_h, _i32, _i32, C.POINTER(_i32)]),
    "read_face_outputs": (C.c_int, [_h, _p, _i32, _p, _p]),
    "read_kernel_times": (C.c_int, [_h, _p, _p]),
    "kernel_bytes": (C.c_int, [_h, _p, _p]),
    "set_layout": (C.c_int, [_h, _i32]),
    "layout_segments": (C.c_int, [_h, C.POINTER(_i32), _p]),
    "set_streams": (C.c_int, [_h, _i32]),
    "get_schedule": (C.c_int, [_h, _p, _p, _p, _p, _p, _p]),
    "host_cells": (C.c_int, [_h, C.POINTER(_i32), _p]),
    "gather_host_cells": (C.c_int, [_h, _p]),
    "scatter_host_cells": (C.c_int, [_h, _p]),
    "host_faces": (C.c_int, [_h, C.POINTER(_i32), _p]),
    "gather_host_faces": (C.c_int, [_h, _p]),
    "scatter_host_faces": (C.c_int, [_h, _p]),
}

LAYOUT_X_SEGMENTS = 0
LAYOUT_ROWS = 1
MAX_SEGMENTS = 16

KERNEL_CLASSES = ("bulk", "pml", "generic", "dft", "material", "mixed", "pml_material")

EXPORTED = tuple("tlm_" + n for n in SIGNATURES)

# include/tlmbatch.h (engine only; the oracle has no batch)
BATCH_SIGNATURES = {
    "batch_create": (C.c_int, [_i32, C.c_int64, C.POINTER(C.c_void_p)]),
    "batch_destroy": (None, [_h]),
    "batch_last_error": (C.c_char_p, [_h]),
    "batch_add": (C.c_int, [_h, _h, _i32, _i32]),
    "batch_run": (C.c_int, [_h, _i32]),
    "batch_synchronize": (C.c_int, [_h]),
    "batch_collect": (C.c_int, [_h, _p, _i32, C.POINTER(_i32)]),
    "batch_info": (C.c_int, [_h, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(C.c_int64)]),
    "batch_set_graph_steps": (C.c_int, [_h, _i32]),
    "batch_set_kernels": (C.c_int, [_h, _i32, _i32]),
    "batch_autotune": (C.c_int, [_h, _i32, C.POINTER(_i32), C.POINTER(_i32)]),
}

BATCH_EXPORTED = tuple("tlm_" + n for n in BATCH_SIGNATURES)
ERR_BUDGET = 6


class TlmError(RuntimeError):
    pass


class Library:
    """A loaded library exporting the ABI under ``prefix``."""

    def __init__(self, path: str, prefix: str = "tlm_", extra=()):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        self.path = path
        self.prefix = prefix
        self.dll = C.CDLL(path)
        self.fn = {}
        for name, (res, args) in list(SIGNATURES.items()) + list(extra):
            f = getattr(self.dll, prefix + name)
            f.restype = res
            f.argtypes = args
            self.fn[name] = f


_engine: Optional[Library] = None


def engine_library() -> Library:
    """The HIP engine (in-tree libtlmscn.so)."""
    global _engine
    if _engine is None:
        _engine = Library(ENGINE_SO, "tlm_", extra=list(BATCH_SIGNATURES.items()))
    return _engine


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def grid_of(t: Tables) -> TlmGrid:
    r = t.reflection
    return TlmGrid(t.nx, t.ny, t.nz, t.n_timesteps, t.dl, t.dt, t.reduced_c_factor,
                   r[0], r[1], r[2], r[3], r[4], r[5])


class Handle:
    """One solve loaded into a library (engine or oracle), mirroring the
    reference's per-step call sequence (run_TLM.F90:191-295)."""

    def __init__(self, lib: Library, tables: Tables, form: int = FORM_PLAIN_FP64,
                 output_capacity: Optional[int] = None):
        self.lib = lib
        self.t = tables
        self._keep = []
        h = C.c_void_p()
        g = grid_of(tables)
        self._check(lib.fn["create"](C.byref(g), form, C.byref(h)), None)
        self.h = h
        try:
            self._load(tables, output_capacity)
        except Exception:
            self.close()
            raise

    def _check(self, rc, h="self"):
        if rc != 0:
            msg = ""
            hh = self.h if h == "self" else h
            if hh:
                m = self.lib.fn["last_error"](hh)
                msg = m.decode() if m else ""
            raise TlmError(f"{self.lib.prefix}call failed (status {rc}): {msg}")

    def _arr(self, a, dtype):
        a = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F")) if a is not None else None
        self._keep.append(a)
        return a

    def _load(self, t: Tables, cap):
        f = self.lib.fn
        i32 = np.int32
        if cap is not None:
            self._check(f["set_output_capacity"](self.h, int(cap)))
        self._check(f["set_volume_materials"](self.h, len(t.volume_materials),
                                               _ptr(np.ascontiguousarray(t.volume_materials))))
        self._check(f["set_surface_materials"](self.h, len(t.surface_materials),
                                                _ptr(self._arr(t.surface_materials, i32))))
        self._check(f["set_cell_codes"](
            self.h, _ptr(self._arr(t.cell_code, i32)), t.n_special_cells,
            _ptr(self._arr(t.cell_to_material, i32)), _ptr(self._arr(t.cell_to_pml, i32)),
            _ptr(self._arr(t.cell_to_excitation, i32)), _ptr(self._arr(t.cell_to_output, i32)),
            _ptr(self._arr(t.cell_to_cable, i32))))
        self._check(f["set_face_codes"](
            self.h, _ptr(self._arr(t.face_code, i32)), t.n_special_faces,
            _ptr(self._arr(t.face_to_material, i32)), _ptr(self._arr(t.face_to_excitation, i32)),
            _ptr(self._arr(t.face_to_output, i32)), _ptr(self._arr(t.face_to_cable, i32))))
        # materials again so that filter storage sees the cell codes
        self._check(f["set_volume_materials"](self.h, len(t.volume_materials),
                                               _ptr(np.ascontiguousarray(t.volume_materials))))
        self._check(f["set_pml"](self.h, len(t.pml_params), _ptr(np.ascontiguousarray(t.pml_params)),
                                 len(t.pml_cell_param), _ptr(self._arr(t.pml_cell_param, i32))))
        nf = t.value.shape[1] if t.value.ndim == 2 else 0
        self._check(f["set_excitation_functions"](self.h, nf, _ptr(self._arr(t.value, np.float64)),
                                                   _ptr(self._arr(t.value_face, np.float64))))
        self._check(f["set_point_sources"](self.h, len(t.point_sources),
                                            _ptr(np.ascontiguousarray(t.point_sources)),
                                            t.n_cell_excitation_rows, t.n_face_excitation_rows))
        if len(t.huygens_faces):
            self._check(f["set_huygens_surface"](
                self.h, t.huygens_function, t.huygens_offset_min,
                _ptr(self._arr(t.huygens_Ei, np.float64)), _ptr(self._arr(t.huygens_Hi, np.float64)),
                len(t.huygens_faces), _ptr(np.ascontiguousarray(t.huygens_faces))))
        self._check(f["set_dft_frequencies"](self.h, len(t.dft_frequencies),
                                              _ptr(self._arr(t.dft_frequencies, np.float64))))
        self._check(f["set_output_points"](self.h, len(t.output_rows),
                                            _ptr(self._arr(t.output_rows, i32)),
                                            _ptr(self._arr(t.output_components, i32))))
        fos = getattr(t, "freq_outputs", None) or []
        if fos:
            outs = np.zeros(len(fos), dtype=FREQUENCY_OUTPUT_DTYPE)
            for k, e in enumerate(fos):
                outs[k] = (e["frequency"], e["kind"], e["component"], e["first_timestep"], e["last_timestep"],
                           e["timestep_interval"], len(e["rows"]))
            cat = lambda key: self._arr(np.concatenate([e[key] for e in fos]).astype(i32), i32)
            self._check(f["set_frequency_outputs"](self.h, len(fos), _ptr(outs), _ptr(cat("rows")),
                                                    _ptr(cat("sides")), _ptr(cat("normals"))))
        fds = getattr(t, "field_outputs", None) or []
        if fds:
            outs = np.zeros(len(fds), dtype=FIELD_OUTPUT_DTYPE)
            for k, e in enumerate(fds):
                outs[k] = (e["kind"], e["component"], e["both_sides"], e["first_timestep"], e["last_timestep"],
                           e["timestep_interval"], len(e["rows"]))
            cat = lambda key: self._arr(np.concatenate([np.asarray(e[key], dtype=i32) for e in fds]), i32)
            self._check(f["set_field_outputs"](self.h, len(fds), _ptr(self._arr(outs, FIELD_OUTPUT_DTYPE)),
                                                _ptr(cat("rows")), _ptr(cat("sides")), _ptr(cat("normals"))))
        ffs = getattr(t, "far_fields", None) or []
        if ffs:
            surf = np.zeros(len(ffs), dtype=FAR_FIELD_DTYPE)
            for k, e in enumerate(ffs):
                surf[k] = (e["frequency"], e["theta_min"], e["theta_step"], e["phi_min"], e["phi_step"],
                           e["n_theta"], e["n_phi"], e["dim"], len(e["rows"]))
            cat = lambda key, dt: self._arr(np.concatenate([np.asarray(e[key], dtype=dt).ravel() for e in ffs]), dt)
            self._check(f["set_far_field_surfaces"](self.h, len(ffs), _ptr(self._arr(surf, FAR_FIELD_DTYPE)),
                                                     _ptr(cat("rows", i32)), _ptr(cat("sides", i32)),
                                                     _ptr(cat("normals", i32)), _ptr(cat("centres", np.float64))))
        rcs = getattr(t, "rcs", None)
        if rcs is not None:
            s = np.zeros(1, dtype=RCS_DTYPE)
            s[0] = (rcs["theta"], rcs["phi"], len(rcs["rows"]), rcs["n_points"], rcs["point_offset"],
                    rcs["points_per_cell"])
            self._check(f["set_rcs_surface"](self.h, _ptr(self._arr(s, RCS_DTYPE)),
                                              _ptr(self._arr(rcs["rows"], i32)), _ptr(self._arr(rcs["sides"], i32)),
                                              _ptr(self._arr(rcs["normals"], i32)),
                                              _ptr(self._arr(np.asarray(rcs["centres"]).ravel(), np.float64))))

    # ---- reference call sequence ---------------------------------------
    def excitation(self, timestep):
        self._check(self.lib.fn["excitation"](self.h, int(timestep)))

    def scatter(self, timestep):
        self._check(self.lib.fn["scatter"](self.h, int(timestep)))

    def connect(self, timestep):
        self._check(self.lib.fn["connect"](self.h, int(timestep)))

    def outer_boundary(self):
        self._check(self.lib.fn["outer_boundary"](self.h))

    def step(self, first_timestep, n_steps):
        self._check(self.lib.fn["step"](self.h, int(first_timestep), int(n_steps)))

    def synchronize(self):
        self._check(self.lib.fn["synchronize"](self.h))

    def device_bytes(self) -> int:
        return int(self.lib.fn["device_bytes"](self.h))

    def upload_links(self, V):
        V = np.asfortranarray(V, dtype=np.float64)
        self._check(self.lib.fn["upload_links"](self.h, _ptr(V)))

    def download_links(self, V=None):
        t = self.t
        if V is None:
            V = np.zeros((12, t.nx + 2, t.ny + 2, t.nz), dtype=np.float64, order="F")
        self._check(self.lib.fn["download_links"](self.h, _ptr(V)))
        return V

    def download_links_noalloc(self, V=None):
        """download_links through the library's fixed staging buffer (no
        device allocation); see tlm_download_links_noalloc."""
        t = self.t
        if V is None:
            V = np.zeros((12, t.nx + 2, t.ny + 2, t.nz), dtype=np.float64, order="F")
        self._check(self.lib.fn["download_links_noalloc"](self.h, _ptr(V)))
        return V

    def read_cell_outputs(self, max_steps=None):
        """Recorded cell_output_field rows: array (steps, n_output_cells, 6)
        and the first timestep."""
        rows = self.t.n_output_cells
        cap = max_steps if max_steps is not None else self.t.n_timesteps
        out = np.zeros((max(cap, 1), max(rows, 1), 6), dtype=np.float64)
        n = C.c_int32(0)
        first = C.c_int32(0)
        self._check(self.lib.fn["read_cell_outputs"](self.h, _ptr(out), int(cap), C.byref(n), C.byref(first)))
        return out[: n.value, :rows], first.value

    def read_face_outputs(self, max_steps=None):
        """Recorded face_output_field rows: array (connects, n_output_faces,
        2 sides, 6) and the timestep of the first connect."""
        rows = self.t.n_output_faces
        cap = max_steps if max_steps is not None else self.t.n_timesteps
        out = np.zeros((max(cap, 1), max(rows, 1), 2, 6), dtype=np.float64)
        n = C.c_int32(0)
        first = C.c_int32(0)
        self._check(self.lib.fn["read_face_outputs"](self.h, _ptr(out), int(cap), C.byref(n), C.byref(first)))
        return out[: n.value, :rows], first.value

    def read_dft(self):
        npnt, nfreq = len(self.t.output_rows), len(self.t.dft_frequencies)
        out = np.zeros((max(npnt, 1), max(nfreq, 1), 2), dtype=np.float64)
        self._check(self.lib.fn["read_dft"](self.h, _ptr(out)))
        return out[:npnt, :nfreq, 0] + 1j * out[:npnt, :nfreq, 1]

    def read_frequency_output(self, k: int):
        """Accumulators of frequency output k (complex, one per item)."""
        n = len(self.t.freq_outputs[k]["rows"])
        out = np.zeros((max(n, 1), 2), dtype=np.float64)
        self._check(self.lib.fn["read_frequency_output"](self.h, int(k), _ptr(out)))
        return out[:n, 0] + 1j * out[:n, 1]

    def read_field_output(self, k: int, max_frames: int = None):
        """The unread frames of field output k: (values (frames, items),
        timestep of the first frame); reading consumes them."""
        n = len(self.t.field_outputs[k]["rows"])
        cap = int(max_frames if max_frames is not None else self.t.n_timesteps)
        out = np.zeros((max(cap, 1), max(n, 1)), dtype=np.float64)
        nf, first = C.c_int32(0), C.c_int32(0)
        self._check(self.lib.fn["read_field_output"](self.h, int(k), _ptr(out), cap, C.byref(nf), C.byref(first)))
        return out[: nf.value, :n], first.value

    def read_far_field(self, k: int):
        """Far field of far-field surface k at r = 1 m: complex E_theta and
        E_phi, each of shape (n_theta, n_phi)."""
        e = self.t.far_fields[k]
        out = np.zeros((e["n_theta"] * e["n_phi"], 4), dtype=np.float64)
        self._check(self.lib.fn["read_far_field"](self.h, int(k), _ptr(out)))
        shape = (e["n_theta"], e["n_phi"])
        return (out[:, 0] + 1j * out[:, 1]).reshape(shape), (out[:, 2] + 1j * out[:, 3]).reshape(shape)

    def read_rcs(self):
        """The RCS surface's far-field time series: array (n_points, 4) of
        E_theta, E_phi, H_theta, H_phi, row p-1 for point p."""
        out = np.zeros((self.t.rcs["n_points"], 4), dtype=np.float64)
        self._check(self.lib.fn["read_rcs"](self.h, _ptr(out)))
        return out

    def set_layout(self, layout: int):
        """Storage order of the link planes (before the first step)."""
        self._check(self.lib.fn["set_layout"](self.h, int(layout)))

    def layout_segments(self) -> list:
        """x-segment boundaries x0[0..n] of the layout in use."""
        n = C.c_int32(0)
        x0 = np.zeros(MAX_SEGMENTS + 1, dtype=np.int32)
        self._check(self.lib.fn["layout_segments"](self.h, C.byref(n), _ptr(x0)))
        return [int(v) for v in x0[: n.value + 1]]

    def get_schedule(self) -> dict:
        """The schedule in use (after autotune: the one it kept)."""
        v = np.zeros(6, dtype=np.int32)
        ptrs = [C.c_void_p(v.ctypes.data + 4 * k) for k in range(6)]
        self._check(self.lib.fn["get_schedule"](self.h, *ptrs))
        return dict(zip(("variant", "rows", "zc", "xcd_remap", "mixed", "streams"), (int(x) for x in v)))

    def host_cells(self) -> np.ndarray:
        """Host-path (cable) cells, (n, 3) 1-based (x, y, z) in scan order."""
        n = C.c_int32(0)
        self._check(self.lib.fn["host_cells"](self.h, C.byref(n), None))
        cells = np.zeros((n.value, 3), dtype=np.int32)
        if n.value:
            self._check(self.lib.fn["host_cells"](self.h, C.byref(n), _ptr(cells)))
        return cells

    def gather_host_cells(self, n_host: int) -> np.ndarray:
        """Incident voltages (n_host, 12) of the host-path cells; between scatter and connect."""
        V = np.zeros((n_host, 12), dtype=np.float64)
        self._check(self.lib.fn["gather_host_cells"](self.h, _ptr(V)))
        return V

    def scatter_host_cells(self, V: np.ndarray):
        """Write the host's scattered voltages (n_host, 12) of the host-path cells back."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        self._check(self.lib.fn["scatter_host_cells"](self.h, _ptr(V)))

    def host_faces(self) -> np.ndarray:
        """Host-path faces (thin layer, diode, spice, switch), (n, 4): the
        1-based (x, y, z) of the cell on the face's + side and the axis 1..3
        of its normal, in the connect's scan order."""
        n = C.c_int32(0)
        self._check(self.lib.fn["host_faces"](self.h, C.byref(n), None))
        faces = np.zeros((n.value, 4), dtype=np.int32)
        if n.value:
            self._check(self.lib.fn["host_faces"](self.h, C.byref(n), _ptr(faces)))
        return faces

    def gather_host_faces(self, n_faces: int) -> np.ndarray:
        """Scattered pulses (n, 4) beside each host-path face: a on the + side,
        a on the - side, b on the + side, b on the - side; between scatter and
        connect."""
        V = np.zeros((n_faces, 4), dtype=np.float64)
        self._check(self.lib.fn["gather_host_faces"](self.h, _ptr(V)))
        return V

    def scatter_host_faces(self, T: np.ndarray):
        """The host model's face voltages (n, 4) = (Va_min, Va_max, Vb_min,
        Vb_max) of each host-path face, for the coming connect."""
        T = np.ascontiguousarray(T, dtype=np.float64)
        self._check(self.lib.fn["scatter_host_faces"](self.h, _ptr(T)))

    def set_streams(self, streams: int):
        """1: a step's kernels in turn on one stream; 2: class kernels concurrent."""
        self._check(self.lib.fn["set_streams"](self.h, int(streams)))

    def set_schedule(self, variant: int, rows: int, zc: int, xcd_remap: int = 0, mixed: int = -1):
        self._check(self.lib.fn["set_schedule"](self.h, int(variant), int(rows), int(zc), int(xcd_remap),
                                                int(mixed)))

    def autotune(self, first_timestep: int, steps_per_candidate: int) -> int:
        """Run candidate bulk schedules on the next steps and keep the
        fastest; returns the number of steps run."""
        done = _i32(0)
        self._check(self.lib.fn["autotune"](self.h, int(first_timestep), int(steps_per_candidate), C.byref(done)))
        return done.value

    def set_small_to_zero(self, n_small_to_zero: int, small_to_zero_value: float):
        """TLM_set_small_to_zero every n steps (0: off)."""
        self._check(self.lib.fn["set_small_to_zero"](self.h, int(n_small_to_zero), float(small_to_zero_value)))

    def set_profiling(self, on: bool):
        self._check(self.lib.fn["set_profiling"](self.h, 1 if on else 0))

    def read_kernel_times(self):
        """{class: (launches, total_ms)} since the previous read."""
        n = np.zeros(len(KERNEL_CLASSES), dtype=np.int64)
        ms = np.zeros(len(KERNEL_CLASSES), dtype=np.float64)
        self._check(self.lib.fn["read_kernel_times"](self.h, _ptr(n), _ptr(ms)))
        return {k: (int(n[i]), float(ms[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def kernel_bytes(self):
        """{class: (cells, algorithmic bytes per launch)}."""
        c = np.zeros(len(KERNEL_CLASSES), dtype=np.int64)
        b = np.zeros(len(KERNEL_CLASSES), dtype=np.int64)
        self._check(self.lib.fn["kernel_bytes"](self.h, _ptr(c), _ptr(b)))
        return {k: (int(c[i]), int(b[i])) for i, k in enumerate(KERNEL_CLASSES)}

    def release_tables(self):
        """Drop the host copies of the tables (the library holds its own),
        keeping the few fields the readers need."""
        from types import SimpleNamespace
        t = self.t
        self.t = SimpleNamespace(n_output_cells=t.n_output_cells, n_output_faces=t.n_output_faces,
                                 n_timesteps=t.n_timesteps, output_rows=t.output_rows,
                                 output_components=t.output_components, dft_frequencies=t.dft_frequencies,
                                 nx=t.nx, ny=t.ny, nz=t.nz, freq_outputs=getattr(t, "freq_outputs", []),
                                 far_fields=getattr(t, "far_fields", []), rcs=getattr(t, "rcs", None),
                                 field_outputs=getattr(t, "field_outputs", []))
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.fn["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def probe_series(cell_outputs: np.ndarray, tables: Tables) -> np.ndarray:
    """(steps, n_points): the output-point values picked from recorded
    cell_output_field rows (cell_output_point, output_points.F90:545-552)."""
    rows = tables.output_rows - 1
    comps = tables.output_components - 1
    return cell_outputs[:, rows, comps]


class Batch:
    """Resident solves stepped together (include/tlmbatch.h).

    add() returns False when the solve does not fit the HBM budget left (the
    caller keeps it queued); run(n) advances every unfinished solve by n
    steps; collect() hands back the finished Handles, which then read like
    after Handle.step()."""

    def __init__(self, form: int, budget_bytes: int, lib: Optional[Library] = None):
        self.lib = lib or engine_library()
        self.b = C.c_void_p()
        rc = self.lib.fn["batch_create"](int(form), int(budget_bytes), C.byref(self.b))
        if rc != 0:
            raise TlmError(f"tlm_batch_create failed (status {rc})")
        self.members = {}

    def _check(self, rc):
        if rc != 0:
            m = self.lib.fn["batch_last_error"](self.b)
            raise TlmError(f"batch call failed (status {rc}): {m.decode() if m else ''}")

    def add(self, handle: "Handle", first_timestep: int, n_steps: int) -> bool:
        rc = self.lib.fn["batch_add"](self.b, handle.h, int(first_timestep), int(n_steps))
        if rc == ERR_BUDGET:
            return False
        self._check(rc)
        self.members[handle.h.value] = handle
        return True

    def run(self, n_steps: int):
        """Enqueue n_steps steps (returns before the device is done)."""
        self._check(self.lib.fn["batch_run"](self.b, int(n_steps)))

    def synchronize(self):
        self._check(self.lib.fn["batch_synchronize"](self.b))

    def collect(self):
        n = max(len(self.members), 1)
        arr = (C.c_void_p * n)()
        got = _i32(0)
        self._check(self.lib.fn["batch_collect"](self.b, arr, n, C.byref(got)))
        return [self.members.pop(arr[i]) for i in range(got.value)]

    def info(self):
        n, u, r = _i32(0), _i32(0), C.c_int64(0)
        self._check(self.lib.fn["batch_info"](self.b, C.byref(n), C.byref(u), C.byref(r)))
        return n.value, u.value, r.value

    def set_graph_steps(self, k: int):
        self._check(self.lib.fn["batch_set_graph_steps"](self.b, int(k)))

    def set_kernels(self, mixed: int = -1, streams: int = 2):
        self._check(self.lib.fn["batch_set_kernels"](self.b, int(mixed), int(streams)))

    def autotune(self, steps_per_candidate: int):
        k, done = _i32(0), _i32(0)
        self._check(self.lib.fn["batch_autotune"](self.b, int(steps_per_candidate), C.byref(k), C.byref(done)))
        return k.value, done.value

    def close(self):
        if self.b:
            self.lib.fn["batch_destroy"](self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
