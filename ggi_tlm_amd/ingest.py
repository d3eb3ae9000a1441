"""Standalone ingest of the reference's problem files: the solver packets
of a `.inp` file, the model builder's `.mesh` file and `.vmat` volume
material files, into a Problem (problem.py) whose tables the engine takes.

The serving layer can then take requests in the reference's own formats
without the Fortran set-up.  Sources followed:

* packet dispatch: SRC/PROCESS_INPUT_FILE/read_input_file_solver.F90:59-203
  (lower-cased packet names, one per line);
* list-directed reads (Fortran `read(unit,*)`): values separated by blanks
  or commas across lines, the rest of the last line ignored; reals may omit
  the exponent letter ("3.0-3" is 3.0e-3);
* `.mesh`: SRC/MESH/read_mesh.F90 / write_mesh.F90:279-310;
* `.vmat`: read_volume_material_list.F90:81-178 with read_Sfilter
  (FILTER/filter_subroutines.F90:54-78);
* mesh parameters: MESH/set_mesh_parameters.F90 (dl = mean of the three
  cell sizes when the mesh is given in cells).

Geometry comes from the `.mesh` file when one is given.  Without one, the
axis-aligned shapes the reference's own test decks use are meshed here as
the model builder meshes them: rectangular_block volumes (cells whose
centre lies inside), rectangular_block, rectangle and x/y/zplane surfaces
(the cell faces on them), and points (point_to_cell,
general_mesh_subroutines.F90:345-383); rotated or other shapes need the
model builder's `.mesh`.  Surface and volume field outputs, frequency
output surfaces and volumes, far-field surfaces and the RCS surface become
device-recorded outputs.  Packets the engine has no path for (modes,
cables, volume averages and peaks) are listed in `Deck.unsupported` and left
out.
"""
from __future__ import annotations

import gzip
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .constants import CENTRE, FACE_XMAX, FACE_XMIN, FACE_YMAX, FACE_YMIN, FACE_ZMAX, FACE_ZMIN
from .filters import Sfilter
from .problem import (PML, Box, ExcitationFunction, FarFieldSurface, FieldOutput, FrequencyOutput, HuygensBox,
                      OutputPoint, PointSource, Problem, RCSSurface, VolumeMaterial, point_to_cell)

_LOCATIONS = {"xmin": FACE_XMIN, "xmax": FACE_XMAX, "ymin": FACE_YMIN, "ymax": FACE_YMAX, "zmin": FACE_ZMIN,
              "zmax": FACE_ZMAX, "centre": CENTRE}
_COMPONENTS = ("ex", "ey", "ez", "hx", "hy", "hz", "jx", "jy", "jz")
# excitation function types and their parameter counts
# (read_excitation_function_list.F90; excitation.F90:155-559)
_FUNCTIONS = {"impulse": 1, "gaussian": 3, "gaussian_step": 3, "step": 2, "sinusoid": 3,
              "gaussian_sinusoid": 5}

_REAL = re.compile(r"([+-]?(?:\d+\.?\d*|\.\d+))(?:[eEdDqQ]([+-]?\d+)|([+-]\d+))?")


class IngestError(ValueError):
    pass


def fortran_real(tok: str) -> float:
    """A real as Fortran's list-directed input reads it."""
    m = _REAL.fullmatch(tok)
    if not m:
        raise IngestError(f"not a real number: {tok!r}")
    e = m.group(2) if m.group(2) is not None else m.group(3)
    return float(m.group(1) + ("e" + e if e else ""))


def fortran_int(tok: str) -> int:
    try:
        return int(tok)
    except ValueError:
        raise IngestError(f"not an integer: {tok!r}")


class Records:
    """Sequential records of a text file with Fortran-style reads."""

    def __init__(self, text: str, name: str = "input"):
        self.lines = text.splitlines()
        self.i = 0
        self.name = name

    def eof(self) -> bool:
        return self.i >= len(self.lines)

    def record(self) -> str:
        """read(unit,'(A)'): the next record as it is."""
        if self.eof():
            raise IngestError(f"{self.name}: unexpected end of file")
        s = self.lines[self.i]
        self.i += 1
        return s

    def backspace(self):
        self.i -= 1

    def values(self, n: int, conv) -> list:
        """read(unit,*) of n values: across records, the rest of the last
        record ignored."""
        out = []
        while len(out) < n:
            toks = [t for t in re.split(r"[\s,]+", self.record().strip()) if t]
            for t in toks:
                if len(out) == n:
                    break
                out.append(conv(t))
        return out

    def reals(self, n: int) -> List[float]:
        return self.values(n, fortran_real)

    def ints(self, n: int) -> List[int]:
        return self.values(n, fortran_int)

    def word(self) -> str:
        """The first word of the next record, lower case (the reference's
        convert_to_lower_case on an '(A)' read)."""
        s = self.record().strip().lower()
        return s.split()[0] if s else ""


@dataclass
class Shape:
    kind: str
    params: List[float]
    rotation_deg: Tuple[float, float, float]
    translation: Tuple[float, float, float]
    roughness: Optional[Tuple[float, ...]] = None   # surface_roughness parameters of a surface


@dataclass
class Deck:
    """The packets of a `.inp` file."""

    mesh_min: Optional[Tuple[float, float, float]] = None
    mesh_max: Optional[Tuple[float, float, float]] = None
    cells: Optional[Tuple[int, int, int]] = None
    dl: Optional[float] = None
    reflection: Tuple[float, ...] = (-1.0,) * 6
    pml: Optional[Tuple[Tuple[float, ...], Tuple[float, ...], int]] = None
    volumes: Dict[int, Shape] = field(default_factory=dict)
    surfaces: Dict[int, Shape] = field(default_factory=dict)
    points: Dict[int, Shape] = field(default_factory=dict)
    volume_materials: List[dict] = field(default_factory=list)
    surface_materials: List[dict] = field(default_factory=list)
    functions: List[ExcitationFunction] = field(default_factory=list)
    excitation_points: List[dict] = field(default_factory=list)
    huygens: Optional[dict] = None
    output_points: List[dict] = field(default_factory=list)
    simulation_time: Optional[float] = None
    reduced_c_factor: float = 1.0
    small_to_zero: Optional[Tuple[int, float]] = None
    frequency_outputs: List[dict] = field(default_factory=list)
    far_field_surfaces: List[dict] = field(default_factory=list)
    rcs_surface: Optional[dict] = None
    field_outputs: List[dict] = field(default_factory=list)
    unsupported: List[str] = field(default_factory=list)


def _shapes(r: Records, kinds: Dict[str, int], what: str) -> Dict[int, Shape]:
    """volume_list / surface_list / point_list entries
    (read_volume_list.F90, read_surface_list.F90, read_point_list.F90)."""
    n = r.ints(1)[0]
    out = {}
    for k in range(1, n + 1):
        num = r.ints(1)[0]
        if num != k:
            raise IngestError(f"{what} numbers must run 1..n in order (got {num} for {k})")
        if kinds:
            kind = r.word()
            if kind not in kinds:
                raise IngestError(f"unknown {what} type {kind!r}")
            if kind in _FILE_KINDS:
                r.record()                    # the shape's file name
            params = r.reals(kinds[kind])
        else:
            kind, params = "point", r.reals(3)
        rough = None
        if what == "surface":
            # an optional surface_roughness record and its parameters (one
            # for a sphere, two otherwise), read_surface_list.F90:242-262
            if r.record()[:17].strip().lower() == "surface_roughness":
                rough = tuple(r.reals(1 if kind == "sphere" else 2))
            else:
                r.backspace()
        rot = tuple(r.reals(3))
        trans = tuple(r.reals(3))
        out[num] = Shape(kind, params, rot, trans, rough)
    return out


# shape types and their parameter counts (read_volume_list.F90,
# read_surface_list.F90); the file-based ones read a file name first.  Only
# the axis-aligned ones are meshed here; the others need the `.mesh` file.
_VOLUME_KINDS = {"rectangular_block2": 6, "rectangular_block": 3, "cylinder": 2, "sphere": 1, "tet": 12, "hex": 24,
                 "pyramid_ram": 3, "pyramid": 2, "tet_volume_mesh": 1, "fill_surface": 1}
_SURFACE_KINDS = {"rectangular_block2": 6, "rectangular_block": 3, "cylinder": 2, "sphere": 1, "hemisphere": 1,
                  "rectangle": 2, "circle": 1, "quad": 12, "xplane": 6, "yplane": 6, "zplane": 6, "triangle": 9,
                  "annulus": 2, "split_ring": 3, "helix": 6, "triangulated_surface": 2,
                  "vtk_triangulated_surface": 2, "stl_triangulated_surface": 2}
_FILE_KINDS = ("tet_volume_mesh", "triangulated_surface", "vtk_triangulated_surface", "stl_triangulated_surface")


def _time_information(r: Records) -> dict:
    """read_output_time_information.F90: optional timestep or time window."""
    w = r.record().strip().lower()
    if w == "output_timestep_information":
        a, b, c = r.ints(1)[0], r.ints(1)[0], r.ints(1)[0]
        return {"first_timestep": a, "last_timestep": b, "timestep_interval": c}
    if w == "output_time_information":
        a, b, c = r.reals(1)[0], r.reals(1)[0], r.reals(1)[0]
        return {"first_time": a, "last_time": b, "time_interval": c}
    r.backspace()
    return {"first_timestep": 0, "last_timestep": 1000000000, "timestep_interval": 1}


def read_inp(path: str) -> Deck:
    """Read the packets of a `.inp` file (read_input_file_solver.F90:59-203,
    plus the geometry packets a deck without a `.mesh` needs)."""
    with open(path, "rb") as f:
        text = f.read().decode("latin-1")
    return read_inp_text(text, os.path.basename(path))


def read_inp_text(text: str, name: str = "input") -> Deck:
    """read_inp of a `.inp` file's text."""
    r = Records(text, name)
    d = Deck()
    while not r.eof():
        line = r.record().strip().lower()
        if not line or line.startswith("#"):
            continue
        if line == "mesh_outer_boundary_dimension":
            v = r.reals(6)
            d.mesh_min, d.mesh_max = (v[0], v[2], v[4]), (v[1], v[3], v[5])
        elif line == "mesh_dimensions_in_cells":
            d.cells = tuple(r.ints(3))
        elif line == "mesh_cell_dimension":
            d.dl = r.reals(1)[0]
        elif line == "outer_boundary_reflection_coefficient":
            d.reflection = tuple(r.reals(6))
        elif line == "pml":
            # read_pml.F90: six thicknesses (m) or one, six reflections or one, order
            def six(s):
                toks = [t for t in re.split(r"[\s,]+", s.strip()) if t]
                try:
                    return tuple(fortran_real(t) for t in toks[:6])
                except IngestError:
                    return (fortran_real(toks[0]),) * 6
            t = six(r.record())
            rr = six(r.record())
            d.pml = (t, rr, r.ints(1)[0])
        elif line == "volume_list":
            d.volumes = _shapes(r, _VOLUME_KINDS, "volume")
        elif line == "surface_list":
            d.surfaces = _shapes(r, _SURFACE_KINDS, "surface")
        elif line == "point_list":
            d.points = _shapes(r, {}, "point")
        elif line == "volume_material_list":
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("volume materials must be numbered 1..n in order")
                kind = r.word()
                m = {"kind": kind}
                if kind == "dispersive":
                    m["file"] = r.record().strip()
                elif kind not in ("pec", "pmc"):
                    raise IngestError(f"volume material type {kind!r} has no path here")
                nv = r.ints(1)[0]
                m["volumes"] = r.ints(nv) if nv else []
                d.volume_materials.append(m)
        elif line == "surface_material_list":
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("surface materials must be numbered 1..n in order")
                kind = r.word()
                m = {"kind": kind}
                if kind not in ("pec", "pmc", "free_space"):
                    raise IngestError(f"surface material type {kind!r} stays on the host path")
                ns = r.ints(1)[0]
                m["surfaces"] = r.ints(ns) if ns else []
                if ns:
                    m["orientation"] = r.ints(ns)
                d.surface_materials.append(m)
        elif line == "excitation_function_list":
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("excitation functions must be numbered 1..n in order")
                kind = r.word()
                if kind not in _FUNCTIONS:
                    raise IngestError(f"excitation function {kind!r} has no path here")
                d.functions.append(ExcitationFunction(kind, tuple(r.reals(_FUNCTIONS[kind]))))
        elif line == "excitation_point_list":
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("excitation points must be numbered 1..n in order")
                fn, pt = r.ints(1)[0], r.ints(1)[0]
                comp, loc, soft = r.word(), r.word(), r.word()
                if comp not in _COMPONENTS or loc not in _LOCATIONS or soft not in ("soft", "hard"):
                    raise IngestError(f"excitation point {k}: {comp!r} {loc!r} {soft!r}")
                d.excitation_points.append({"function": fn, "point": pt, "component": comp, "location": loc,
                                            "soft": soft == "soft"})
        elif line == "huygens_surface":
            # read_huygens_surface.F90: surface, side, optional wave impedance,
            # function, wave vector angles, polarisation
            surf = r.ints(1)[0]
            side = r.ints(1)[0]
            z0 = None
            if r.word() == "wave_impedance":
                z0 = r.reals(1)[0]
            else:
                r.backspace()
            fn = r.ints(1)[0]
            for what in ("wave vector", "polarisation"):
                if r.record().strip()[:1] in ("r", "R"):
                    raise IngestError(f"a random Huygens {what} is not reproducible here")
                r.backspace()
                if what == "wave vector":
                    theta, phi = r.reals(2)
                else:
                    et, ep = r.reals(2)
            d.huygens = {"surface": surf, "side": side, "wave_impedance": z0, "function": fn, "theta": theta,
                         "phi": phi, "etheta": et, "ephi": ep}
        elif line == "output_point_list":
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("output points must be numbered 1..n in order")
                pt = r.ints(1)[0]
                comp, loc = r.word(), r.word()
                if comp not in _COMPONENTS or loc not in _LOCATIONS:
                    raise IngestError(f"output point {k}: {comp!r} {loc!r}")
                d.output_points.append({"point": pt, "component": comp, "location": loc, **_time_information(r)})
        elif line == "frequency_output_surface_list":
            # read_frequency_output_surface.F90: surface, side, frequency,
            # component (field or current), time window
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("frequency output surfaces must be numbered 1..n in order")
                surf, side = r.ints(1)[0], r.ints(1)[0]
                freq = r.reals(1)[0]
                comp = r.word()
                if comp not in _COMPONENTS:
                    raise IngestError(f"frequency output surface {k}: component {comp!r} has no path here")
                d.frequency_outputs.append({"kind": "surface", "number": surf, "side": side, "frequency": freq,
                                            "component": comp, **_time_information(r)})
        elif line == "frequency_output_volume_list":
            # read_frequency_output_volume.F90 (normal format): volume,
            # frequency, field component, time window
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("frequency output volumes must be numbered 1..n in order")
                vol = r.ints(1)[0]
                freq = r.reals(1)[0]
                comp = r.word()
                if comp not in _COMPONENTS[:6]:
                    raise IngestError(f"frequency output volume {k}: component {comp!r}")
                d.frequency_outputs.append({"kind": "volume", "number": vol, "frequency": freq, "component": comp,
                                            **_time_information(r)})
        elif line == "output_surface_list":
            # read_output_surface_list.F90:74-131: surface, side (+1, -1, or
            # 0 for both sides), component (its first two letters), window
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("output surfaces must be numbered 1..n in order")
                surf, side = r.ints(1)[0], r.ints(1)[0]
                if side not in (1, -1, 0):
                    raise IngestError("output surface side must be +1, -1 or 0")
                comp = r.record().strip()[:2].lower()
                if comp not in _COMPONENTS + ("em", "hm", "jm"):
                    raise IngestError(f"output surface {k}: component {comp!r}")
                if side == 0 and comp not in ("jx", "jy", "jz", "jm"):
                    raise IngestError(f"output surface {k}: both sides take Jx, Jy, Jz or Jmagnitude")
                d.field_outputs.append({"kind": "surface", "number": surf, "side": side, "component": comp,
                                        **_time_information(r)})
        elif line == "output_volume_list":
            # read_output_volume_list.F90:70-90: volume, field component, window
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("output volumes must be numbered 1..n in order")
                vol = r.ints(1)[0]
                comp = r.word()
                if comp not in _COMPONENTS[:6]:
                    raise IngestError(f"output volume {k}: component {comp!r}")
                d.field_outputs.append({"kind": "volume", "number": vol, "component": comp, **_time_information(r)})
        elif line == "far_field_surface_list":
            # read_far_field_surface.F90:77-143: surface, side, frequency,
            # theta and phi ranges (degrees), then an optional '2D' record
            # and its axis (z only)
            n = r.ints(1)[0]
            for k in range(1, n + 1):
                if r.ints(1)[0] != k:
                    raise IngestError("far field surfaces must be numbered 1..n in order")
                surf, side = r.ints(1)[0], r.ints(1)[0]
                if side not in (1, -1):
                    raise IngestError("far field surface side must be +1 or -1")
                freq = r.reals(1)[0]
                theta, phi = tuple(r.reals(3)), tuple(r.reals(3))
                dim = 3
                if not r.eof():
                    if r.record()[:2] in ("2D", "2d"):
                        if r.record()[:1] not in ("z", "Z"):
                            raise IngestError("the 2D far-field transform takes the z axis only")
                        dim = 2
                    else:
                        r.backspace()
                d.far_field_surfaces.append({"number": surf, "side": side, "frequency": freq, "theta": theta,
                                             "phi": phi, "dim": dim})
        elif line == "rcs_surface":
            # read_rcs_surface.F90:56-79: surface, side, fmin fmax fstep,
            # theta, phi (degrees)
            surf, side = r.ints(1)[0], r.ints(1)[0]
            if side not in (1, -1):
                raise IngestError("RCS surface side must be +1 or -1")
            fmin, fmax, fstep = r.reals(3)
            theta, phi = r.reals(1)[0], r.reals(1)[0]
            d.rcs_surface = {"number": surf, "side": side, "fmin": fmin, "fmax": fmax, "fstep": fstep,
                             "theta": theta, "phi": phi}
        elif line == "simulation_time":
            d.simulation_time = r.reals(1)[0]
        elif line == "reduced_c_factor":
            d.reduced_c_factor = r.reals(1)[0]
        elif line == "set_small_to_zero":
            d.small_to_zero = (r.ints(1)[0], r.reals(1)[0])
        elif re.fullmatch(r"[a-z_0-9]+", line):
            d.unsupported.append(line)
            _skip_packet(r)
    return d


def _skip_packet(r: Records):
    """Skip a packet this ingest leaves out: up to the next blank line."""
    while not r.eof():
        if not r.record().strip():
            return


# ---- .vmat ------------------------------------------------------------------------

def _read_sfilter(r: Records) -> Sfilter:
    """read_Sfilter (FILTER/filter_subroutines.F90:54-78): wnorm, a order,
    a coefficients, b order, b coefficients."""
    wnorm = r.reals(1)[0]
    na = r.ints(1)[0]
    a = r.reals(na + 1)
    nb = r.ints(1)[0]
    b = r.reals(nb + 1)
    return Sfilter(wnorm, a, b)


def read_vmat(path: str) -> VolumeMaterial:
    """A dispersive volume material file (read_volume_material_list.F90:
    81-178): label, frequency range, comment, permittivity filter, sigma_e,
    comment, permeability filter, sigma_m."""
    with open(path, "rb") as f:
        r = Records(f.read().decode("latin-1"), os.path.basename(path))
    r.record()                # label
    r.reals(2)                # fmin, fmax
    r.record()                # comment
    eps = _read_sfilter(r)
    sigma_e = r.reals(1)[0]
    r.record()                # comment
    mu = _read_sfilter(r)
    sigma_m = r.reals(1)[0]
    return VolumeMaterial("dispersive", eps, mu, sigma_e=sigma_e, sigma_m=sigma_m)


# ---- .mesh ------------------------------------------------------------------------

@dataclass
class MeshFile:
    """The model builder's mesh (write_mesh.F90): cell lists of volumes,
    face lists of surfaces, cell-segment lists of lines, points, and the
    PML volumes with their parameters."""

    dl: float
    n: Tuple[int, int, int]
    mesh_min: Tuple[float, float, float]
    mesh_max: Tuple[float, float, float]
    volumes: List[np.ndarray]                  # (N, 4): i, j, k, point
    surfaces: List[dict]                       # bounds, cell bounds, faces (N, 4)
    lines: List[np.ndarray]                    # (N, 8)
    points: List[dict]                         # xyz, cell, face (i, j, k, point)
    pml_volumes: List[np.ndarray] = field(default_factory=list)
    pml_volume_to_face: Tuple[int, ...] = (0,) * 6
    pml_thickness: Tuple[float, ...] = (0.0,) * 6
    pml_reflection: Tuple[float, ...] = (0.0,) * 6
    pml_order: int = 0


def read_mesh(path: str) -> MeshFile:
    """Read a `.mesh` file (read_mesh.F90), plain or gzip-compressed."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        r = Records(f.read().decode("latin-1"), os.path.basename(path))
    dl = r.reals(1)[0]
    n = tuple(r.ints(3))
    x = r.reals(2)
    y = r.reals(2)
    z = r.reals(2)

    def rows(count, width):
        a = np.array([r.ints(width) for _ in range(count)], dtype=np.int64)
        return a.reshape(count, width)

    volumes = [rows(r.ints(1)[0], 4) for _ in range(r.ints(1)[0])]
    surfaces = []
    for _ in range(r.ints(1)[0]):
        nf = r.ints(1)[0]
        b = [r.reals(2), r.reals(2), r.reals(2)]
        cb = [r.ints(2), r.ints(2), r.ints(2)]
        surfaces.append({"bounds": b, "cell_bounds": cb, "faces": rows(nf, 4)})
    lines = [rows(r.ints(1)[0], 8) for _ in range(r.ints(1)[0])]
    points = []
    for _ in range(r.ints(1)[0]):
        xyz = tuple(r.reals(3))
        cell = tuple(r.ints(3))
        face = tuple(r.ints(4))
        points.append({"xyz": xyz, "cell": cell, "face": face})
    m = MeshFile(dl, n, (x[0], y[0], z[0]), (x[1], y[1], z[1]), volumes, surfaces, lines, points)
    if r.eof():
        return m
    m.pml_volumes = [rows(r.ints(1)[0], 4) for _ in range(r.ints(1)[0])]
    r.record()
    m.pml_volume_to_face = tuple(r.ints(6))
    r.record()
    t = r.reals(2) + r.reals(2) + r.reals(2)
    r.record()
    rr = r.reals(2) + r.reals(2) + r.reals(2)
    r.record()
    m.pml_thickness, m.pml_reflection, m.pml_order = tuple(t), tuple(rr), r.ints(1)[0]
    return m


def _r(v) -> str:
    return repr(float(v))


def write_mesh(m: MeshFile, path: str):
    """Write a `.mesh` file in write_mesh.F90's layout (round trip with
    read_mesh)."""
    L = [f"{_r(m.dl)} dl", f"{m.n[0]} {m.n[1]} {m.n[2]} nx ny nz",
         f"{_r(m.mesh_min[0])} {_r(m.mesh_max[0])} mesh_xmin,mesh_xmax",
         f"{_r(m.mesh_min[1])} {_r(m.mesh_max[1])} mesh_ymin,mesh_ymax",
         f"{_r(m.mesh_min[2])} {_r(m.mesh_max[2])} mesh_zmin,mesh_zmax", f"{len(m.volumes)} n_volumes"]
    for k, v in enumerate(m.volumes, 1):
        L.append(f"{len(v)} number_of_cells in volume= {k}")
        L += [" ".join(str(int(t)) for t in row) for row in v]
    L.append(f"{len(m.surfaces)} n_surfaces")
    for s in m.surfaces:
        L.append(f"{len(s['faces'])} number_of_faces")
        L += [f"{_r(a)} {_r(b)}" for a, b in s["bounds"]]
        L += [f"{a} {b}" for a, b in s["cell_bounds"]]
        L += [" ".join(str(int(t)) for t in row) for row in s["faces"]]
    L.append(f"{len(m.lines)} n_lines")
    for ln in m.lines:
        L.append(f"{len(ln)} number_of_cell_segments")
        L += [" ".join(str(int(t)) for t in row) for row in ln]
    L.append(f"{len(m.points)} n_points")
    for p in m.points:
        L += [" ".join(_r(t) for t in p["xyz"]), " ".join(str(int(t)) for t in p["cell"]),
              " ".join(str(int(t)) for t in p["face"])]
    L.append(f"{len(m.pml_volumes)} n_pml_volumes")
    for v in m.pml_volumes:
        L.append(f"{len(v)} number_of_cells in PML volume")
        L += [" ".join(str(int(t)) for t in row) for row in v]
    L += ["pml_volume_to_face", " ".join(str(int(t)) for t in m.pml_volume_to_face),
          "pml thicknesses on xmin, xmax, ymin, ymax, zmin and zmax boundaries"]
    t = m.pml_thickness
    L += [f"{_r(t[0])} {_r(t[1])}", f"{_r(t[2])} {_r(t[3])}", f"{_r(t[4])} {_r(t[5])}", "pml_r"]
    rr = m.pml_reflection
    L += [f"{_r(rr[0])} {_r(rr[1])}", f"{_r(rr[2])} {_r(rr[3])}", f"{_r(rr[4])} {_r(rr[5])}", "pml_order",
          str(m.pml_order)]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


# ---- meshing of the axis-aligned shapes -------------------------------------------

def _extent(s: Shape) -> Tuple[np.ndarray, np.ndarray]:
    """Axis-aligned extent (lo, hi) of a block, rectangle or plane after its
    translation (rotations are not meshed here)."""
    if any(a != 0.0 for a in s.rotation_deg):
        raise IngestError(f"rotated {s.kind} is not meshed here: give the model builder's .mesh file")
    if s.roughness is not None:
        raise IngestError(f"a rough {s.kind} (random surface points) is not meshed here: "
                          "give the model builder's .mesh file")
    t = np.array(s.translation)
    p = s.params
    if s.kind == "rectangular_block":
        h = np.array(p[:3]) / 2.0
        return t - h, t + h
    if s.kind == "rectangular_block2":      # corners (x1, y1, z1), (x2, y2, z2)
        a, b = np.array(p[:3]), np.array(p[3:6])
        return t + np.minimum(a, b), t + np.maximum(a, b)
    if s.kind == "rectangle":           # in the z = 0 plane of its frame
        h = np.array([p[0] / 2.0, p[1] / 2.0, 0.0])
        return t - h, t + h
    if s.kind in ("xplane", "yplane", "zplane"):  # corners (x1, y1, z1), (x2, y2, z2)
        a, b = np.array(p[:3]), np.array(p[3:6])
        return t + np.minimum(a, b), t + np.maximum(a, b)
    raise IngestError(f"a {s.kind} is not meshed here: give the model builder's .mesh file")


def _cells_inside(lo, hi, mesh_min, dl, n):
    """Index ranges (1-based, inclusive) of the cells whose centre lies
    inside [lo, hi] on each axis."""
    out = []
    for a in range(3):
        c = mesh_min[a] + (np.arange(1, n[a] + 1) - 0.5) * dl
        idx = np.nonzero((c >= lo[a]) & (c <= hi[a]))[0]
        out.append((int(idx[0]) + 1, int(idx[-1]) + 1) if len(idx) else (1, 0))
    return out


def _grid_index(v, m0, dl):
    """The grid line nearest v, ties away from zero (Fortran NINT)."""
    x = (v - m0) / dl
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


@dataclass
class CellList:
    """A material region given as a list of cells (i, j, k), 1-based."""

    cells: np.ndarray

    def mask(self, nx, ny, nz):
        m = np.zeros((nz, ny, nx), dtype=bool)
        c = np.asarray(self.cells, dtype=np.int64).reshape(-1, 3)
        m[c[:, 2] - 1, c[:, 1] - 1, c[:, 0] - 1] = True
        return m


def _plane_faces(lo, hi, mesh_min, dl, n, normal_sign):
    """Cell faces on an axis-aligned rectangle (one extent of zero width) on
    its nearest grid line, over the cells whose centre lines meet it.  The
    model builder's mesher (mesh_triangle.F90:155-168) keeps each face on the
    side its triangles' normal points away from: a normal along +axis gives
    the *max face of the cell on the low side, along -axis the *min face of
    the cell on the high side."""
    axis = [a for a in range(3) if hi[a] == lo[a]]
    if len(axis) != 1:
        raise IngestError("a rectangle or plane must be flat along exactly one axis")
    a = axis[0]
    g = _grid_index(lo[a], mesh_min[a], dl)
    rng = _cells_inside(lo, hi, mesh_min, dl, n)
    cell, face = (g, (FACE_XMAX, FACE_YMAX, FACE_ZMAX)[a]) if normal_sign > 0 else \
        (g + 1, (FACE_XMIN, FACE_YMIN, FACE_ZMIN)[a])
    faces = []
    others = [b for b in range(3) if b != a]
    for u in range(rng[others[0]][0], rng[others[0]][1] + 1):
        for v in range(rng[others[1]][0], rng[others[1]][1] + 1):
            c = [0, 0, 0]
            c[a], c[others[0]], c[others[1]] = cell, u, v
            faces.append((c[0], c[1], c[2], face))
    return faces


def _plane_normal_sign(s: Shape) -> int:
    """Sign of the triangle normal along the plane's axis:
    triangle_normal (functions.F90:241-278) is (v3 - v1) x (v2 - v1) of the
    triangles the geometry builders make (build_surface_xplane.F90,
    _yplane, _zplane: (p1, p3, p2) over the corners p1 = (x1, y1, z1)
    .. p4; build_surface_rectangle.F90: (p1, p3, p2) over a grid in x, y)."""
    p = s.params
    if s.kind == "rectangle":
        return -1
    x1, y1, z1, x2, y2, z2 = p[:6]
    if s.kind == "xplane":
        v = (y2 - y1) * (z2 - z1)
    elif s.kind == "yplane":
        v = -((x2 - x1) * (z2 - z1))
    else:
        v = (x2 - x1) * (y2 - y1)
    return 1 if v > 0 else -1


def mesh_surface(s: Shape, mesh_min, dl: float, n) -> List[Tuple[int, int, int, int]]:
    """Faces (i, j, k, face) of an unrotated rectangular block, rectangle
    or x/y/z plane surface, as the model builder's mesher lists them."""
    lo, hi = _extent(s)
    if s.kind.startswith("rectangular_block"):
        # walls on their nearest grid lines g0 (low) and g1 (high), over
        # the cells whose centre lines meet them; the builder keeps each
        # wall face on the outside cell, facing the block
        # (build_surface_rectangular_block.F90:106-: its triangles'
        # normals, by triangle_normal, point into the block)
        (i0, i1), (j0, j1), (k0, k1) = _cells_inside(lo, hi, mesh_min, dl, n)
        gx = [_grid_index(lo[0], mesh_min[0], dl), _grid_index(hi[0], mesh_min[0], dl)]
        gy = [_grid_index(lo[1], mesh_min[1], dl), _grid_index(hi[1], mesh_min[1], dl)]
        gz = [_grid_index(lo[2], mesh_min[2], dl), _grid_index(hi[2], mesh_min[2], dl)]
        f = []
        for k in range(k0, k1 + 1):
            for j in range(j0, j1 + 1):
                f += [(gx[0], j, k, FACE_XMAX), (gx[1] + 1, j, k, FACE_XMIN)]
        for k in range(k0, k1 + 1):
            for i in range(i0, i1 + 1):
                f += [(i, gy[0], k, FACE_YMAX), (i, gy[1] + 1, k, FACE_YMIN)]
        for j in range(j0, j1 + 1):
            for i in range(i0, i1 + 1):
                f += [(i, j, gz[0], FACE_ZMAX), (i, j, gz[1] + 1, FACE_ZMIN)]
    else:
        f = _plane_faces(lo, hi, mesh_min, dl, n, _plane_normal_sign(s))
    # faces whose cell lies outside the mesh are dropped
    # (cell_point_inside_mesh, mesh_triangle.F90:170)
    return [c for c in f if all(1 <= c[a] <= n[a] for a in range(3))]


# ---- deck -> problem --------------------------------------------------------------

@dataclass
class Ingested:
    problem: Problem
    unsupported: List[str]                  # packets left out
    output_time_information: List[dict]     # per output point, in order
    boundary_points: List[int] = field(default_factory=list)  # points on a cell boundary


def problem_from(deck: Deck, mesh: Optional[MeshFile] = None, material_dir: Optional[str] = None,
                 base_dir: str = ".", point_cells: Optional[Dict[int, Tuple[int, int, int]]] = None) -> Ingested:
    """Compose a Problem from the packets of a `.inp` file and, when given,
    the model builder's `.mesh` file.  Without a `.mesh`, point_cells may
    give the cells of points as the model builder placed them (its `.mesh`
    point records): a point exactly on a cell boundary lands on either side
    by the last bit of the builder's arithmetic, and such points are listed
    in Ingested.boundary_points."""
    if mesh is not None:
        dl, n, mesh_min = mesh.dl, mesh.n, mesh.mesh_min
    else:
        if deck.mesh_min is None:
            raise IngestError("no Mesh_outer_boundary_dimension packet and no .mesh file")
        mesh_min, mesh_max = deck.mesh_min, deck.mesh_max
        if deck.dl is None:
            if deck.cells is None:
                raise IngestError("neither Mesh_dimensions_in_cells nor Mesh_cell_dimension")
            # set_mesh_parameters.F90: dl = mean of the three cell sizes.  The
            # reference builds with -O4 -funsafe-math-optimizations
            # (make_configuration_data:39), under which the division by 3
            # is a product with its reciprocal: 0.1 for three cells of 0.1,
            # where the division gives 0.10000000000000002 (and a point on
            # a cell boundary then lands in the other cell).
            d = [(mesh_max[a] - mesh_min[a]) / float(deck.cells[a]) for a in range(3)]
            dl = (d[0] + d[1] + d[2]) * (1.0 / 3.0)
        else:
            dl = deck.dl
        n = tuple(max(1, int(round((mesh_max[a] - mesh_min[a]) / dl))) for a in range(3))
    nx, ny, nz = n

    boundary = []

    def point_cell(num):
        if mesh is not None:
            return tuple(int(v) for v in mesh.points[num - 1]["cell"])
        if point_cells and num in point_cells:
            return tuple(point_cells[num])
        s = deck.points[num]
        if any(a != 0.0 for a in s.rotation_deg):
            raise IngestError("rotated points are not meshed here: give the model builder's .mesh file")
        xyz = tuple(s.params[a] + s.translation[a] for a in range(3))
        q = [(xyz[a] - mesh_min[a]) / dl for a in range(3)]
        if any(abs(v - round(v)) <= 1e-9 * max(1.0, abs(v)) for v in q) and num not in boundary:
            boundary.append(num)
        return point_to_cell(mesh_min, dl, xyz)

    def volume_region(num):
        if mesh is not None:
            return CellList(mesh.volumes[num - 1][:, :3])
        lo, hi = _extent(deck.volumes[num])
        (i0, i1), (j0, j1), (k0, k1) = _cells_inside(lo, hi, mesh_min, dl, n)
        return Box(i0, i1, j0, j1, k0, k1)

    def surface_faces(num):
        if mesh is not None:
            return [tuple(int(v) for v in f) for f in mesh.surfaces[num - 1]["faces"]]
        return mesh_surface(deck.surfaces[num], mesh_min, dl, n)

    def surface_box(num):
        """The surface's extent as a cell box (grid lines i0-1 .. i1)."""
        if mesh is not None:
            b = mesh.surfaces[num - 1]["bounds"]
            lo, hi = [b[a][0] for a in range(3)], [b[a][1] for a in range(3)]
        else:
            lo, hi = _extent(deck.surfaces[num])
        g0 = [_grid_index(lo[a], mesh_min[a], dl) for a in range(3)]
        g1 = [_grid_index(hi[a], mesh_min[a], dl) for a in range(3)]
        return Box(g0[0] + 1, g1[0], g0[1] + 1, g1[1], g0[2] + 1, g1[2])

    materials, regions = [], []
    for k, m in enumerate(deck.volume_materials, 1):
        if m["kind"] == "dispersive":
            f = m["file"]
            if material_dir is not None and f.startswith("MATERIAL_DIRECTORY/"):
                f = os.path.join(material_dir, f[len("MATERIAL_DIRECTORY/"):])
            elif not os.path.isabs(f):
                f = os.path.join(base_dir, f)
            materials.append(read_vmat(f + ".vmat"))
        else:
            materials.append(VolumeMaterial(m["kind"]))
        for v in m["volumes"]:
            regions.append((volume_region(v), k))
    smats, sfaces = [], []
    for k, m in enumerate(deck.surface_materials, 1):
        smats.append(m["kind"])
        for s in m["surfaces"]:
            sfaces.append((np.array(surface_faces(s), dtype=np.int64), k))
    sources = []
    for e in deck.excitation_points:
        sources.append(PointSource(point_cell(e["point"]), e["component"], e["function"], soft=e["soft"],
                                   point=e["location"]))
    outputs = [OutputPoint(point_cell(o["point"]), o["component"], point=o["location"]) for o in deck.output_points]
    huygens = None
    if deck.huygens is not None:
        h = deck.huygens
        kw = {} if h["wave_impedance"] is None else {"wave_impedance": h["wave_impedance"]}
        s = h["surface"]
        box = surface_box(s)
        explicit = None
        kind = deck.surfaces[s].kind if (mesh is None and s in deck.surfaces) else "mesh"
        if kind == "mesh" or not kind.startswith("rectangular_block"):
            explicit = surface_faces(s)
        else:
            (i0, i1), (j0, j1), (k0, k1) = _cells_inside(*_extent(deck.surfaces[s]), mesh_min, dl, n)
            box = Box(i0, i1, j0, j1, k0, k1)
        huygens = HuygensBox(box, h["function"], h["theta"], h["phi"], h["etheta"], h["ephi"], side=h["side"],
                             faces=explicit, **kw)
    pml = None
    if mesh is not None and mesh.pml_order > 0:
        pml = PML(tuple(int(round(t / dl)) for t in mesh.pml_thickness), tuple(mesh.pml_reflection), mesh.pml_order)
    elif deck.pml is not None:
        t, rr, order = deck.pml
        pml = PML(tuple(int(round(x / dl)) for x in t), tuple(rr), order)
    freq_outputs = []
    for fo in deck.frequency_outputs:
        if "first_time" in fo:
            raise IngestError("frequency output windows given in time (output_time_information) have no path here")
        window = {"first_timestep": fo["first_timestep"], "last_timestep": fo["last_timestep"],
                  "timestep_interval": fo["timestep_interval"]}
        if fo["kind"] == "volume":
            freq_outputs.append(FrequencyOutput("volume", fo["frequency"], fo["component"],
                                                region=volume_region(fo["number"]), **window))
        else:
            freq_outputs.append(FrequencyOutput("surface", fo["frequency"], fo["component"],
                                                faces=surface_faces(fo["number"]), side=fo["side"], **window))
    field_outs = []
    for fo in deck.field_outputs:
        if "first_time" in fo:
            raise IngestError("field output windows given in time (output_time_information) have no path here")
        window = {"first_timestep": fo["first_timestep"], "last_timestep": fo["last_timestep"],
                  "timestep_interval": fo["timestep_interval"]}
        if fo["kind"] == "volume":
            field_outs.append(FieldOutput("volume", fo["component"], region=volume_region(fo["number"]), **window))
        else:
            field_outs.append(FieldOutput("surface", fo["component"], faces=surface_faces(fo["number"]),
                                          side=fo["side"], **window))
    far_fields = [FarFieldSurface(surface_faces(f["number"]), f["frequency"], theta=f["theta"], phi=f["phi"],
                                  side=f["side"], dim=f["dim"]) for f in deck.far_field_surfaces]
    rcs = None
    if deck.rcs_surface is not None:
        q = deck.rcs_surface
        rcs = RCSSurface(surface_faces(q["number"]), q["fmin"], q["fmax"], q["fstep"], theta=q["theta"],
                         phi=q["phi"], side=q["side"])
    p = Problem(nx=nx, ny=ny, nz=nz, dl=dl, simulation_time=deck.simulation_time, mesh_min=tuple(mesh_min),
                reflection=tuple(deck.reflection), materials=materials, material_regions=regions,
                surface_materials=smats, surface_faces=sfaces, functions=list(deck.functions), sources=sources,
                huygens=huygens, outputs=outputs, pml=pml, reduced_c_factor=deck.reduced_c_factor,
                frequency_outputs=freq_outputs, far_field_surfaces=far_fields, rcs_surface=rcs,
                field_outputs=field_outs)
    info = [{k: v for k, v in o.items() if k not in ("point", "component", "location")} for o in deck.output_points]
    return Ingested(p, list(deck.unsupported), info, sorted(boundary))


def load(inp_path: str, mesh_path: Optional[str] = None, material_dir: Optional[str] = None,
         point_cells: Optional[Dict[int, Tuple[int, int, int]]] = None) -> Ingested:
    """Read a `.inp` file (and the model builder's `.mesh` file, if given)
    into a Problem."""
    deck = read_inp(inp_path)
    mesh = read_mesh(mesh_path) if mesh_path else None
    return problem_from(deck, mesh, material_dir, os.path.dirname(os.path.abspath(inp_path)), point_cells)
