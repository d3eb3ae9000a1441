"""Host-side problem description and the reference-layout tables built from it.

A :class:`Problem` describes one TLM solve: the mesh, outer boundary, PML,
volume materials placed on cell regions, surface materials on cell faces,
excitation functions, point sources, a Huygens plane-wave box, probes and
the probe-DFT frequency list.  :func:`build_tables` turns it into exactly
the module tables GGI_TLM's solver holds after its set-up phase, in the
reference's conventions (1-based rows, scan order with x fastest, Fortran
column-major for 2-D tables), which is what the C ABI (include/tlmscn.h)
takes.  Each part cites the set-up routine it restates:

  set_cell_update_codes        SRC/TLM_SOLVER/set_update_codes.F90:38-363
  set_face_update_codes        SRC/TLM_SOLVER/set_update_codes.F90:383-791
  allocate_volume_material_filter_data
                               SRC/MATERIALS/TLM_volume_materials.F90:401-518
  set_pml_volume_material_mesh SRC/MATERIALS/TLM_PML_volume_materials.F90:41-428
  calc_pml_dist / calc_sigma_0 / PML_conductivity  (same file, :446-581)
  build_PML_geometry           SRC/GEOMETRY/build_PML_geometry.F90:41-110
  initialise_excitation_functions SRC/EXCITATION/excitation.F90:155-559
  set_huygens_surface_data     SRC/EXCITATION/huygens_surface.F90:247-715
  point_to_cell / get_cell_point_coordinate
                               SRC/MESH/general_mesh_subroutines.F90:46-383
  set_TLM_solver_parameters    SRC/TLM_SOLVER/set_TLM_solver_parameters.F90:94-96

Geometry rasterisation is reduced to what the hot path needs: axis-aligned
cell boxes and spheres tested at cell centres (the reference's tet-mesh
model builder is out of scope).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .constants import (C0, CENTRE, COMPONENTS, EPS0, FACE_XMAX, FACE_XMIN,
                        FACE_YMAX, FACE_YMIN, FACE_ZMAX, FACE_ZMIN, FACES,
                        MAX_FILTER_ORDER, PI, SMAT_FREE_SPACE, SMAT_PEC,
                        SMAT_PMC, SOURCE_HARD, SOURCE_SOFT, SURFACE_MATERIALS, VMAT_DISPERSIVE, Z0)
from .filters import Sfilter, volume_material_row

# ---------------------------------------------------------------------------
# problem description
# ---------------------------------------------------------------------------


@dataclass
class Box:
    """Inclusive 1-based cell ranges."""

    i0: int
    i1: int
    j0: int
    j1: int
    k0: int
    k1: int

    def mask(self, nx, ny, nz):
        m = np.zeros((nz, ny, nx), dtype=bool)
        m[self.k0 - 1:self.k1, self.j0 - 1:self.j1, self.i0 - 1:self.i1] = True
        return m


@dataclass
class Sphere:
    """Cells whose centre lies within radius (cells) of centre (cell
    coordinates, cell i spans [i-1, i])."""

    cx: float
    cy: float
    cz: float
    radius: float

    def mask(self, nx, ny, nz):
        x = (np.arange(1, nx + 1) - 0.5 - self.cx) ** 2
        y = (np.arange(1, ny + 1) - 0.5 - self.cy) ** 2
        z = (np.arange(1, nz + 1) - 0.5 - self.cz) ** 2
        return (x[None, None, :] + y[None, :, None] + z[:, None, None]) <= self.radius ** 2


@dataclass
class VolumeMaterial:
    kind: str = "dispersive"  # 'dispersive' | 'pec' | 'pmc'
    eps: Sfilter = field(default_factory=lambda: Sfilter.constant(1.0))
    mu: Sfilter = field(default_factory=lambda: Sfilter.constant(1.0))
    sigma_e: float = 0.0
    sigma_m: float = 0.0


@dataclass
class ExcitationFunction:
    """Excitation function (SRC/EXCITATION/excitation.F90:155-559).
    kind: impulse(amp) | gaussian(amp,width,delay) |
    gaussian_step(amp,width,delay) | step(amp,delay) |
    sinusoid(amp,freq,phase_deg) | gaussian_sinusoid(amp,width,delay,freq,phase_deg)."""

    kind: str
    params: Tuple[float, ...]


@dataclass
class PointSource:
    cell: Tuple[int, int, int]
    component: str
    function: int
    soft: bool = True
    point: str = "centre"


@dataclass
class OutputPoint:
    cell: Tuple[int, int, int]
    component: str
    point: str = "centre"


@dataclass
class PML:
    """Thickness in cells per face (xmin,xmax,ymin,ymax,zmin,zmax), target
    reflection per face and profile order (read_pml.F90)."""

    thickness: Tuple[int, int, int, int, int, int]
    reflection: Tuple[float, float, float, float, float, float] = (1e-4,) * 6
    order: int = 2


@dataclass
class HuygensBox:
    """Huygens surface on the faces of a cell box (cells inside the box).
    side is the Huygens_surface packet's "side of surface for excitation"
    (read_huygens_surface.F90:97-104): side=+1 places the source faces on the
    cells just outside the box, as the reference's surface mesh of a
    rectangular_block does (pinned against TEST_DATA/HUYGENS_SURFACE);
    side=-1 flips them onto the box's own boundary cells.  theta/phi and the
    E polarisation are as in the packet."""

    box: Box
    function: int
    theta_deg: float
    phi_deg: float
    etheta: float
    ephi: float
    side: int = 1
    wave_impedance: float = Z0
    # explicit surface-mesh faces (cx, cy, cz, face) as the model builder's
    # mesh lists them (side=-1 turns each to the other side); when given,
    # the box only sets the surface extent used for offset_min
    faces: Optional[list] = None


@dataclass
class FrequencyOutput:
    """A frequency-domain output (frequency_output_volumes.F90:309-386,
    frequency_output_surfaces.F90:297-421): every step in the window adds
    value * exp(-j 2 pi f t) per cell (kind 'volume': region, a cell mask
    object) or per face (kind 'surface': faces (cx, cy, cz, face) of the
    surface mesh, side +1 the outward normal, -1 the other side of each
    face).  component: 'ex'..'hz', and 'jx'..'jz' (n x H) for surfaces."""

    kind: str
    frequency: float
    component: str
    region: object = None
    faces: Optional[list] = None
    side: int = 1
    first_timestep: int = 0
    last_timestep: int = 1000000000
    timestep_interval: int = 1


@dataclass
class FieldOutput:
    """A time-domain field output (read_output_surface_list.F90,
    read_output_volume_list.F90; output_surfaces.F90, output_volumes.F90):
    every step in the window records one value per cell (kind 'volume':
    region, component 'ex'..'hz') or per face (kind 'surface': faces of the
    surface mesh; side +1 the outward normal, -1 the other side, 0 both
    sides, which takes the surface current as the jump across the face;
    components 'ex'..'hz', 'jx'..'jz', 'em', 'hm', 'jm' for |E|, |H|,
    |J|)."""

    kind: str
    component: str
    region: object = None
    faces: Optional[list] = None
    side: int = 1
    first_timestep: int = 0
    last_timestep: int = 1000000000
    timestep_interval: int = 1


@dataclass
class FarFieldSurface:
    """A far-field surface (read_far_field_surface.F90, far_field_surfaces.F90):
    faces (cx, cy, cz, face) of the surface mesh and the side (+1 the
    outward normal, -1 the other side of each face), the frequency, and the
    angle ranges (min, max, step) in degrees as the input file gives them.
    dim 2: the 2D transform in the theta = 90 degree plane (theta is not
    read)."""

    faces: list
    frequency: float
    theta: Tuple[float, float, float] = (0.0, 180.0, 1.0)
    phi: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    side: int = 1
    dim: int = 3


@dataclass
class RCSSurface:
    """The RCS surface (read_rcs_surface.F90, RCS_surfaces.F90): faces and
    side as for far-field surfaces, the frequency range (fmin, fmax, fstep)
    of the RCS spectrum, and the monostatic direction in degrees."""

    faces: list
    fmin: float
    fmax: float
    fstep: float
    theta: float = 0.0
    phi: float = 0.0
    side: int = 1


@dataclass
class Problem:
    nx: int
    ny: int
    nz: int
    dl: float
    n_timesteps: Optional[int] = None
    simulation_time: Optional[float] = None
    mesh_min: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    reflection: Tuple[float, float, float, float, float, float] = (-1.0,) * 6
    reduced_c_factor: float = 1.0
    pml: Optional[PML] = None
    materials: List[VolumeMaterial] = field(default_factory=list)
    material_regions: List[Tuple[object, int]] = field(default_factory=list)
    # 'pec'|'pmc'|'free_space' (device) or a host-path type: 'thin_layer',
    # 'anisotropic_thin_layer', 'diode', 'spice', 'switch'
    surface_materials: List[str] = field(default_factory=list)
    surface_faces: List[Tuple[np.ndarray, int]] = field(default_factory=list)
    functions: List[ExcitationFunction] = field(default_factory=list)
    sources: List[PointSource] = field(default_factory=list)
    huygens: Optional[HuygensBox] = None
    outputs: List[OutputPoint] = field(default_factory=list)
    dft_frequencies: Sequence[float] = ()
    # cells whose scatter stays on the host path (cable junction cells,
    # cell_update_code_to_cable_cell_number != 0), 1-based (cx, cy, cz)
    host_cells: List[Tuple[int, int, int]] = field(default_factory=list)
    # faces carrying a cable face junction (face_update_code_to_cable_number
    # = position in this list + 1), (cx, cy, cz, face) as surface_faces
    cable_faces: List[Tuple[int, int, int, int]] = field(default_factory=list)
    frequency_outputs: List[FrequencyOutput] = field(default_factory=list)
    far_field_surfaces: List[FarFieldSurface] = field(default_factory=list)
    rcs_surface: Optional[RCSSurface] = None
    field_outputs: List[FieldOutput] = field(default_factory=list)

    @property
    def dt(self) -> float:
        # set_TLM_solver_parameters.F90:94
        return self.dl / (2.0 * C0 / self.reduced_c_factor)

    @property
    def timesteps(self) -> int:
        if self.n_timesteps is not None:
            return int(self.n_timesteps)
        # NINT(simulation_time/dt), set_TLM_solver_parameters.F90:96
        return int(_nint(self.simulation_time / self.dt))


def _nint(x: float) -> int:
    """Fortran NINT: round half away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def point_to_cell(mesh_min, dl, p):
    """point_to_cell, general_mesh_subroutines.F90:345-383."""
    return tuple(int((p[i] - mesh_min[i]) / dl) + 1 for i in range(3))


def cell_point_coordinate(mesh_min, dl, cell, point):
    """get_cell_point_coordinate, general_mesh_subroutines.F90:84-137."""
    off = [-0.5, -0.5, -0.5]
    if point == FACE_XMIN:
        off[0] = -1.0
    elif point == FACE_XMAX:
        off[0] = 0.0
    elif point == FACE_YMIN:
        off[1] = -1.0
    elif point == FACE_YMAX:
        off[1] = 0.0
    elif point == FACE_ZMIN:
        off[2] = -1.0
    elif point == FACE_ZMAX:
        off[2] = 0.0
    return tuple(mesh_min[i] + float(cell[i]) * dl + dl * off[i] for i in range(3))


# ---------------------------------------------------------------------------
# excitation function tables
# ---------------------------------------------------------------------------

def excitation_tables(functions, n_timesteps, dt):
    """value(1:n) and value_face(1:n) per function, as Fortran
    (n_timesteps, n_functions) arrays (excitation.F90:192-330).  Scalar
    libm calls keep every entry bit-identical to the reference tables."""
    nf = len(functions)
    value = np.zeros((n_timesteps, nf), dtype=np.float64, order="F")
    value_face = np.zeros((n_timesteps, nf), dtype=np.float64, order="F")
    for f, fn in enumerate(functions):
        p = [float(x) for x in fn.params]
        for ts in range(1, n_timesteps + 1):
            time = (ts - 1) * dt
            time2 = time + dt / 2.0
            v = vf = 0.0
            if fn.kind == "impulse":
                v = vf = p[0] if ts == 1 else 0.0
            elif fn.kind == "gaussian":
                amp, width, delay = p[:3]
                if abs(time - delay) < width * 5.0:
                    x = (time - delay) / width
                    v = amp * (math.exp(-(x * x)))
                if abs(time2 - delay) < width * 5:
                    x = (time2 - delay) / width
                    vf = amp * (math.exp(-(x * x)))
            elif fn.kind == "gaussian_step":
                amp, width, delay = p[:3]
                for t, tgt in ((time, "v"), (time2, "vf")):
                    if t < delay:
                        r = 0.0
                        if abs(t - delay) < width * 5:
                            x = (t - delay) / width
                            r = amp * (math.exp(-(x * x)))
                    else:
                        r = amp
                    if tgt == "v":
                        v = r
                    else:
                        vf = r
            elif fn.kind == "step":
                amp, delay = p[:2]
                v = amp if time >= delay else 0.0
                vf = amp if time2 >= delay else 0.0
            elif fn.kind == "sinusoid":
                amp, freq, ph = p[0], p[1], p[2] * PI / 180.0
                v = amp * math.sin(2.0 * PI * freq * time - ph)
                vf = amp * math.sin(2.0 * PI * freq * time2 - ph)
            elif fn.kind == "gaussian_sinusoid":
                amp, width, delay, freq = p[:4]
                ph = p[4] * PI / 180.0
                g = 0.0
                if abs(time - delay) < width * 5.0:
                    x = (time - delay) / width
                    g = math.exp(-(x * x))
                v = amp * g * math.sin(2.0 * PI * freq * (time - delay) - ph)
                g = 0.0
                if abs(time2 - delay) < width * 5:
                    x = (time2 - delay) / width
                    g = math.exp(-(x * x))
                vf = amp * g * math.sin(2.0 * PI * freq * (time2 - delay) - ph)
            else:
                raise ValueError(f"excitation function {fn.kind!r} not supported")
            value[ts - 1, f] = v
            value_face[ts - 1, f] = vf
    return value, value_face


# ---------------------------------------------------------------------------
# table builder
# ---------------------------------------------------------------------------

VMAT_DTYPE = np.dtype([
    ("type", "<i4"), ("eps_filter_exists", "<i4"), ("mu_filter_exists", "<i4"),
    ("eps_order", "<i4"), ("mu_order", "<i4"), ("pad_", "<i4"),
    ("Ys_eps_f", "<f8"), ("Ge", "<f8"), ("Zs_mu_f", "<f8"), ("Rm", "<f8"),
    ("eps_a", "<f8", (MAX_FILTER_ORDER + 1,)), ("eps_b", "<f8", (MAX_FILTER_ORDER + 1,)),
    ("mu_a", "<f8", (MAX_FILTER_ORDER + 1,)), ("mu_b", "<f8", (MAX_FILTER_ORDER + 1,)),
])

PML_PARAM_FIELDS = ("sx", "sy", "sz", "ax", "ay", "az", "exp_x", "exp_y", "exp_z",
                    "Csx", "Csy", "Csz", "Csx2", "Csy2", "Csz2", "Csy_sz", "Csz_sx", "Csx_sy")
PML_DTYPE = np.dtype([(n, "<f8") for n in PML_PARAM_FIELDS])

POINT_SOURCE_DTYPE = np.dtype([("function", "<i4"), ("component", "<i4"),
                               ("source_type", "<i4"), ("at_face", "<i4"), ("row", "<i4")])

HUYGENS_FACE_DTYPE = np.dtype([("row", "<i4"), ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"),
                               ("offset", "<f8")])


@dataclass
class Tables:
    """Reference-layout tables of one solve (names follow the Fortran
    module globals)."""

    nx: int
    ny: int
    nz: int
    dl: float
    dt: float
    n_timesteps: int
    reduced_c_factor: float
    reflection: Tuple[float, ...]
    cell_code: np.ndarray
    n_special_cells: int
    cell_to_material: np.ndarray  # (n,2) Fortran order
    cell_to_pml: np.ndarray
    cell_to_excitation: np.ndarray
    cell_to_output: np.ndarray
    cell_to_cable: np.ndarray
    face_code: np.ndarray
    n_special_faces: int
    face_to_material: np.ndarray  # (n,3) Fortran order
    face_to_excitation: np.ndarray
    face_to_output: np.ndarray
    face_to_cable: np.ndarray
    volume_materials: np.ndarray
    surface_materials: np.ndarray
    pml_params: np.ndarray
    pml_cell_param: np.ndarray
    value: np.ndarray
    value_face: np.ndarray
    point_sources: np.ndarray
    n_cell_excitation_rows: int
    n_face_excitation_rows: int
    huygens_function: int
    huygens_offset_min: float
    huygens_Ei: np.ndarray
    huygens_Hi: np.ndarray
    huygens_faces: np.ndarray
    output_rows: np.ndarray
    output_components: np.ndarray
    n_output_cells: int
    n_output_faces: int
    face_output_points: list
    dft_frequencies: np.ndarray
    # handy extras (not part of the ABI)
    output_cells: np.ndarray = None   # (n_output_cells, 3) 1-based cx,cy,cz
    freq_outputs: list = field(default_factory=list)  # frequency output volumes and surfaces
    far_fields: list = field(default_factory=list)    # far-field surfaces
    field_outputs: list = field(default_factory=list)  # time-domain surface and volume field outputs
    rcs: Optional[dict] = None                        # the RCS surface
    n_pml_cells: int = 0


def _comp(name: str) -> int:
    return COMPONENTS[name.lower()]


def _min_face(nx, ny, nz, cell, face):
    """Map (cell, face) to the interior connect face it names, given as
    (cx, cy, cz, axis) of the face's max-side cell, or None on the mesh
    boundary."""
    cx, cy, cz = cell
    if face == FACE_XMIN:
        t = (cx, cy, cz, 0)
    elif face == FACE_XMAX:
        t = (cx + 1, cy, cz, 0)
    elif face == FACE_YMIN:
        t = (cx, cy, cz, 1)
    elif face == FACE_YMAX:
        t = (cx, cy + 1, cz, 1)
    elif face == FACE_ZMIN:
        t = (cx, cy, cz, 2)
    elif face == FACE_ZMAX:
        t = (cx, cy, cz + 1, 2)
    else:
        raise ValueError("not a face")
    x, y, z, a = t
    if (a == 0 and not 2 <= x <= nx) or (a == 1 and not 2 <= y <= ny) or (a == 2 and not 2 <= z <= nz):
        return None
    if not (1 <= x <= nx and 1 <= y <= ny and 1 <= z <= nz):
        return None
    return t


def huygens_face_list(hb: HuygensBox, nx, ny, nz):
    """Faces of the box's surface as (cx,cy,cz,face) after the side flip of
    set_huygens_surface_data (huygens_surface.F90:398-437), plus normals."""
    b = hb.box
    faces = []
    if hb.faces is not None:
        faces = [tuple(int(v) for v in f) for f in hb.faces]
    # a box's walls as the builder meshes a rectangular_block surface: the
    # face of the cell just outside each wall, facing the box
    for k in (range(b.k0, b.k1 + 1) if hb.faces is None else ()):
        for j in range(b.j0, b.j1 + 1):
            faces.append((b.i0 - 1, j, k, FACE_XMAX))
            faces.append((b.i1 + 1, j, k, FACE_XMIN))
    for k in (range(b.k0, b.k1 + 1) if hb.faces is None else ()):
        for i in range(b.i0, b.i1 + 1):
            faces.append((i, b.j0 - 1, k, FACE_YMAX))
            faces.append((i, b.j1 + 1, k, FACE_YMIN))
    for j in (range(b.j0, b.j1 + 1) if hb.faces is None else ()):
        for i in range(b.i0, b.i1 + 1):
            faces.append((i, j, b.k0 - 1, FACE_ZMAX))
            faces.append((i, j, b.k1 + 1, FACE_ZMIN))
    out = []
    # side -1 turns each face to the other side (huygens_surface.F90:423-453)
    flip = {FACE_XMIN: (-1, 0, 0, FACE_XMAX), FACE_XMAX: (1, 0, 0, FACE_XMIN),
            FACE_YMIN: (0, -1, 0, FACE_YMAX), FACE_YMAX: (0, 1, 0, FACE_YMIN),
            FACE_ZMIN: (0, 0, -1, FACE_ZMAX), FACE_ZMAX: (0, 0, 1, FACE_ZMIN)}
    normal = {FACE_XMIN: (-1, 0, 0), FACE_XMAX: (1, 0, 0), FACE_YMIN: (0, -1, 0),
              FACE_YMAX: (0, 1, 0), FACE_ZMIN: (0, 0, -1), FACE_ZMAX: (0, 0, 1)}
    for (i, j, k, f) in faces:
        if hb.side == -1:
            di, dj, dk, f2 = flip[f]
            i, j, k, f = i + di, j + dj, k + dk, f2
        out.append((i, j, k, f, normal[f]))
    return out


def build_tables(p: Problem) -> Tables:
    nx, ny, nz, dl = p.nx, p.ny, p.nz, p.dl
    dt = p.dt
    nts = p.timesteps
    N = nx * ny * nz
    shape = (nz, ny, nx)

    # ---- cell-level flags (local_cell_* arrays) -------------------------
    local_material = np.zeros(shape, dtype=np.int32)
    for region, mnum in p.material_regions:
        local_material[region.mask(nx, ny, nz)] = mnum

    # PML membership and depth, build_PML_geometry + calc_pml_dist
    pml_flag = np.zeros(shape, dtype=bool)
    dist = [np.zeros(shape, dtype=np.int32) for _ in range(3)]
    if p.pml is not None:
        t = p.pml.thickness
        idx = [np.arange(1, n + 1) for n in (nx, ny, nz)]
        for axis, n in enumerate((nx, ny, nz)):
            tmin, tmax = t[2 * axis], t[2 * axis + 1]
            c = idx[axis]
            dmin = c - (tmin + 1)
            dmin = np.where(dmin > 0, 0, dmin)
            dmax = tmax - (n - c + 1) + 1
            dmax = np.where(dmax < 0, 0, dmax)
            bshape = [1, 1, 1]
            bshape[2 - axis] = n
            if tmin > 0:
                slab = (c <= tmin).reshape(bshape)
                sl = np.broadcast_to(slab, shape)
                pml_flag |= sl
                dist[axis] = np.where(sl, np.broadcast_to(dmin.reshape(bshape), shape), dist[axis])
            if tmax > 0:
                slab = (c >= n - tmax + 1).reshape(bshape)
                sl = np.broadcast_to(slab, shape)
                pml_flag |= sl
                dist[axis] = np.where(sl, np.broadcast_to(dmax.reshape(bshape), shape), dist[axis])
    if (pml_flag & (local_material != 0)).any():
        mats = np.unique(local_material[pml_flag & (local_material != 0)])
        for m in mats:
            vm = p.materials[m - 1]
            if vm.kind in ("pec", "pmc"):
                raise ValueError("PEC/PMC volume material in PML")
            # set_update_codes.F90:300-333: PML cells may hold only order-0
            # (non-dispersive) materials
            for sf in (vm.eps, vm.mu):
                if sf is not None and (len(sf.a) > 1 or len(sf.b) > 1):
                    raise ValueError(f"PML fatal error: dispersive material found, material number {m}")

    local_exc = np.zeros(shape, dtype=bool)
    for s in p.sources:
        if FACES[s.point] == CENTRE:
            local_exc[s.cell[2] - 1, s.cell[1] - 1, s.cell[0] - 1] = True
    local_out = np.zeros(shape, dtype=bool)
    for o in p.outputs:
        if FACES[o.point] == CENTRE:
            local_out[o.cell[2] - 1, o.cell[1] - 1, o.cell[0] - 1] = True
    # frequency output volumes make their cells output cells
    # (set_outputs_in_mesh, output.F90:46-)
    for fq in list(p.frequency_outputs) + list(p.field_outputs):
        if fq.kind == "volume":
            local_out |= fq.region.mask(nx, ny, nz)
    if (pml_flag & local_exc).any():
        raise ValueError("Excitation intersects PML")
    local_host = np.zeros(shape, dtype=bool)
    for c in p.host_cells:
        local_host[c[2] - 1, c[1] - 1, c[0] - 1] = True
    if (local_host & (pml_flag | local_exc | local_out)).any():
        raise ValueError("host-path cells carry no PML, excitation or output on the device path")

    flat_mat = local_material.ravel()
    flat_pml = pml_flag.ravel()
    flat_exc = local_exc.ravel()
    flat_out = local_out.ravel()
    flat_host = local_host.ravel()
    special = (flat_mat != 0) | flat_pml | flat_exc | flat_out | flat_host
    n_special = int(special.sum())
    cell_code = np.zeros(N, dtype=np.int32)
    cell_code[special] = np.arange(1, n_special + 1, dtype=np.int32)
    sp_idx = np.nonzero(special)[0]

    # PML cell numbering (scan order) and parameter table
    pml_cell_number = np.zeros(N, dtype=np.int32)
    n_pml_cells = int(flat_pml.sum())
    pml_cell_number[flat_pml] = np.arange(1, n_pml_cells + 1, dtype=np.int32)
    pml_params = np.zeros(0, dtype=PML_DTYPE)
    pml_cell_param = np.zeros(0, dtype=np.int32)
    if n_pml_cells:
        d = np.stack([dist[a].ravel()[flat_pml] for a in range(3)], axis=1)
        # first appearance in scan order defines the parameter number; the
        # depth triples are packed into one integer key for the uniquing
        dmin = d.min(axis=0).astype(np.int64)
        span = d.max(axis=0).astype(np.int64) - dmin + 1
        key = ((d[:, 0] - dmin[0]) * span[1] + (d[:, 1] - dmin[1])) * span[2] + (d[:, 2] - dmin[2])
        ukey, first, inverse = np.unique(key, return_index=True, return_inverse=True)
        uniq = d[first]
        order = np.argsort(first)
        rank = np.empty(len(order), dtype=np.int32)
        rank[order] = np.arange(1, len(order) + 1, dtype=np.int32)
        pml_cell_param = rank[inverse.ravel()].astype(np.int32)
        lo = d.min(axis=0)
        hi = d.max(axis=0)
        lo = [min(0, int(v)) for v in lo]
        hi = [max(0, int(v)) for v in hi]
        o = p.pml.order
        r = p.pml.reflection

        def sigma0(rr, Nn):
            # calc_sigma_0, TLM_PML_volume_materials.F90:498-513, then /eps0
            den = dl * float(abs(Nn) ** (o + 1))
            if den == 0.0:
                return float("inf")
            return (-(EPS0 * C0 / 2.0) * math.log(rr) / den) / EPS0

        s0 = [(sigma0(r[2 * a], lo[a]), sigma0(r[2 * a + 1], hi[a])) for a in range(3)]

        def cond(px, a):
            # PML_conductivity, TLM_PML_volume_materials.F90:531-581
            if px == 0:
                return 0.0
            L = abs(px) - 1
            s = s0[a][0] if px < 0 else s0[a][1]
            return s * float((L + 1) ** (o + 1) - L ** (o + 1))

        pml_params = np.zeros(len(uniq), dtype=PML_DTYPE)
        for n_idx in range(len(uniq)):
            px, py, pz = (int(v) for v in uniq[order[n_idx]])
            sx, sy, sz = cond(px, 0), cond(py, 1), cond(pz, 2)
            vals = dict(
                sx=sx, sy=sy, sz=sz,
                ax=4.0 / (4.0 + (sy + sz) * dt),
                ay=4.0 / (4.0 + (sz + sx) * dt),
                az=4.0 / (4.0 + (sx + sy) * dt),
                exp_x=math.exp(-dt * sx / 2.0),
                exp_y=math.exp(-dt * sy / 2.0),
                exp_z=math.exp(-dt * sz / 2.0),
                Csx=sx * dt / 4.0, Csy=sy * dt / 4.0, Csz=sz * dt / 4.0,
                Csx2=sx * dt / 2.0, Csy2=sy * dt / 2.0, Csz2=sz * dt / 2.0,
                Csy_sz=1.0 - (sy + sz) * dt / 4.0,
                Csz_sx=1.0 - (sz + sx) * dt / 4.0,
                Csx_sy=1.0 - (sx + sy) * dt / 4.0)
            for k, v in vals.items():
                pml_params[k][n_idx] = v

    # volume materials
    vm = np.zeros(len(p.materials), dtype=VMAT_DTYPE)
    for m, mat in enumerate(p.materials):
        row = volume_material_row(mat.kind, mat.eps, mat.mu, mat.sigma_e, mat.sigma_m,
                                  dl, dt, p.reduced_c_factor)
        for k in ("type", "eps_filter_exists", "mu_filter_exists", "eps_order", "mu_order",
                  "Ys_eps_f", "Ge", "Zs_mu_f", "Rm"):
            vm[k][m] = row[k]
        for k in ("eps_a", "eps_b", "mu_a", "mu_b"):
            c = np.asarray(row[k], dtype=np.float64)
            if len(c) > MAX_FILTER_ORDER + 1:
                raise ValueError("filter order too high")
            vm[k][m, : len(c)] = c

    to_material = np.zeros((n_special, 2), dtype=np.int32, order="F")
    to_material[:, 0] = flat_mat[sp_idx]
    # filter numbers: 3 per dispersive special cell, in code order
    if len(vm):
        disp = np.zeros(len(vm) + 1, dtype=bool)
        disp[1:] = vm["type"] == VMAT_DISPERSIVE
        is_disp = disp[to_material[:, 0]]
        to_material[is_disp, 1] = 1 + 3 * np.arange(int(is_disp.sum()), dtype=np.int32)
    to_pml = pml_cell_number[sp_idx].astype(np.int32)
    to_exc = flat_exc[sp_idx].astype(np.int32)
    to_out = flat_out[sp_idx].astype(np.int32)
    # host-path (cable junction) cells: numbered 1.. in scan order
    host_rank = np.zeros(N, dtype=np.int32)
    host_rank[flat_host] = np.arange(1, int(flat_host.sum()) + 1, dtype=np.int32)
    to_cable = host_rank[sp_idx].astype(np.int32)

    # excitation / output rows (scan-order ranks)
    exc_rank = np.zeros(N, dtype=np.int32)
    exc_rank[flat_exc] = np.arange(1, int(flat_exc.sum()) + 1, dtype=np.int32)
    out_rank = np.zeros(N, dtype=np.int32)
    out_rank[flat_out] = np.arange(1, int(flat_out.sum()) + 1, dtype=np.int32)

    def lin(c):
        return (c[0] - 1) + nx * ((c[1] - 1) + ny * (c[2] - 1))

    # ---- faces ------------------------------------------------------------
    # face flags on the min faces: [cz,cy,cx,axis]
    fshape = (nz, ny, nx, 3)
    f_mat = np.zeros(fshape, dtype=np.int32)
    f_exc = np.zeros(fshape, dtype=bool)
    f_out = np.zeros(fshape, dtype=bool)
    smat_types = []
    for s in p.surface_materials:
        smat_types.append(SURFACE_MATERIALS[s])
    for faces, mnum in p.surface_faces:
        for (cx, cy, cz, face) in np.asarray(faces, dtype=np.int64).reshape(-1, 4):
            t = _min_face(nx, ny, nz, (int(cx), int(cy), int(cz)), int(face))
            if t is not None:
                f_mat[t[2] - 1, t[1] - 1, t[0] - 1, t[3]] = mnum
    for s in p.sources:
        if FACES[s.point] != CENTRE:
            t = _min_face(nx, ny, nz, s.cell, FACES[s.point])
            if t is not None:
                f_exc[t[2] - 1, t[1] - 1, t[0] - 1, t[3]] = True
    hfaces = []
    if p.huygens is not None:
        hfaces = huygens_face_list(p.huygens, nx, ny, nz)
        for (cx, cy, cz, face, _n) in hfaces:
            t = _min_face(nx, ny, nz, (cx, cy, cz), face)
            if t is not None:
                f_exc[t[2] - 1, t[1] - 1, t[0] - 1, t[3]] = True
    f_cab = np.zeros(fshape, dtype=np.int32)
    for k, (cx, cy, cz, face) in enumerate(p.cable_faces):
        t = _min_face(nx, ny, nz, (int(cx), int(cy), int(cz)), int(face))
        if t is not None:
            f_cab[t[2] - 1, t[1] - 1, t[0] - 1, t[3]] = k + 1
    # output surfaces (frequency output, far-field and RCS surfaces): each
    # face, turned to the other side for side -1
    # (frequency_output_surfaces.F90:30-280, far_field_surfaces.F90:36-145,
    # RCS_surfaces.F90:35-143), is an output face
    other = {FACE_XMIN: (-1, 0, 0, FACE_XMAX), FACE_XMAX: (1, 0, 0, FACE_XMIN),
             FACE_YMIN: (0, -1, 0, FACE_YMAX), FACE_YMAX: (0, 1, 0, FACE_YMIN),
             FACE_ZMIN: (0, 0, -1, FACE_ZMAX), FACE_ZMAX: (0, 0, 1, FACE_ZMIN)}

    def output_surface(surface_faces, side, what):
        faces = []
        for (cx, cy, cz, face) in surface_faces:
            cx, cy, cz, face = int(cx), int(cy), int(cz), int(face)
            if side == -1:
                di, dj, dk, face = other[face]
                cx, cy, cz = cx + di, cy + dj, cz + dk
            faces.append((cx, cy, cz, face))
            t = _min_face(nx, ny, nz, (cx, cy, cz), face)
            if t is None:
                raise ValueError(f"a {what} face lies on the mesh boundary")
            f_out[t[2] - 1, t[1] - 1, t[0] - 1, t[3]] = True
        return faces

    fo_faces = [output_surface(fq.faces, fq.side, "frequency output surface") if fq.kind == "surface" else []
                for fq in p.frequency_outputs]
    ff_faces = [output_surface(s.faces, s.side, "far-field surface") for s in p.far_field_surfaces]
    # field output surfaces: side 0 (both sides) turns the faces as -1 does
    # (output_surfaces.F90:42-259)
    fd_faces = [output_surface(fo.faces, 1 if fo.side == 1 else -1, "field output surface")
                if fo.kind == "surface" else [] for fo in p.field_outputs]
    rcs_faces = output_surface(p.rcs_surface.faces, p.rcs_surface.side, "RCS surface") if p.rcs_surface else []
    face_out_pts = []
    for o in p.outputs:
        if FACES[o.point] != CENTRE:
            t = _min_face(nx, ny, nz, o.cell, FACES[o.point])
            if t is not None:
                f_out[t[2] - 1, t[1] - 1, t[0] - 1, t[3]] = True
    valid = np.zeros(fshape, dtype=bool)
    valid[:, :, 1:, 0] = True
    valid[:, 1:, :, 1] = True
    valid[1:, :, :, 2] = True
    vflat = valid.reshape(-1)
    fm = f_mat.reshape(-1)[vflat]
    fe = f_exc.reshape(-1)[vflat]
    fo = f_out.reshape(-1)[vflat]
    fc = f_cab.reshape(-1)[vflat]
    # set_update_codes.F90:383-791: material, cable, excitation or output
    fspecial = (fm != 0) | (fc != 0) | fe | fo
    n_faces = int(vflat.sum())
    n_special_faces = int(fspecial.sum())
    face_code = np.zeros(n_faces, dtype=np.int32)
    face_code[fspecial] = np.arange(1, n_special_faces + 1, dtype=np.int32)
    fsp = np.nonzero(fspecial)[0]
    face_to_material = np.zeros((n_special_faces, 3), dtype=np.int32, order="F")
    face_to_material[:, 0] = fm[fsp]
    axis_of = np.broadcast_to(np.arange(3, dtype=np.int32), fshape).reshape(-1)[vflat]
    face_to_material[:, 2] = np.where(fm[fsp] != 0, axis_of[fsp] + 1, 0)
    face_to_exc = fe[fsp].astype(np.int32)
    face_to_out = fo[fsp].astype(np.int32)
    face_to_cable = fc[fsp].astype(np.int32)
    # per-face ranks over the full [cz,cy,cx,axis] grid
    fexc_rank_full = np.zeros(fshape, dtype=np.int32).reshape(-1)
    fidx_valid = np.nonzero(vflat)[0]
    exc_faces_full = fidx_valid[fe]
    fexc_rank_full[exc_faces_full] = np.arange(1, len(exc_faces_full) + 1, dtype=np.int32)
    fout_rank_full = np.zeros(fshape, dtype=np.int32).reshape(-1)
    out_faces_full = fidx_valid[fo]
    fout_rank_full[out_faces_full] = np.arange(1, len(out_faces_full) + 1, dtype=np.int32)

    def face_flat(t):
        return ((t[2] - 1) * ny + (t[1] - 1)) * nx * 3 + (t[0] - 1) * 3 + t[3]

    # ---- excitation ---------------------------------------------------------
    value, value_face = excitation_tables(p.functions, nts, dt)
    ps = np.zeros(len(p.sources), dtype=POINT_SOURCE_DTYPE)
    for i, s in enumerate(p.sources):
        ps["function"][i] = s.function
        ps["component"][i] = _comp(s.component)
        ps["source_type"][i] = SOURCE_SOFT if s.soft else SOURCE_HARD
        if FACES[s.point] == CENTRE:
            ps["at_face"][i] = 0
            ps["row"][i] = exc_rank[lin(s.cell)]
        else:
            t = _min_face(nx, ny, nz, s.cell, FACES[s.point])
            ps["at_face"][i] = 1
            ps["row"][i] = 0 if t is None else fexc_rank_full[face_flat(t)]
    ps = ps[ps["row"] != 0] if len(ps) else ps

    huy_faces = np.zeros(0, dtype=HUYGENS_FACE_DTYPE)
    huy_Ei = np.zeros(3)
    huy_Hi = np.zeros(3)
    huy_offset_min = 0.0
    huy_function = 0
    if p.huygens is not None:
        hb = p.huygens
        huy_function = hb.function
        theta = (PI / 180.0) * hb.theta_deg
        phi = (PI / 180.0) * hb.phi_deg
        r = 1.0
        kx = r * math.sin(theta) * math.cos(phi)
        ky = r * math.sin(theta) * math.sin(phi)
        kz = r * math.cos(theta)
        dxdt = r * math.cos(theta) * math.cos(phi)
        dxdp = -r * math.sin(phi)
        dydt = r * math.cos(theta) * math.sin(phi)
        dydp = r * math.cos(phi)
        dzdt = -r * math.sin(theta)
        dzdp = 0.0
        vx = hb.etheta * dxdt + hb.ephi * dxdp
        vy = hb.etheta * dydt + hb.ephi * dydp
        vz = hb.etheta * dzdt + hb.ephi * dzdp
        norm = math.sqrt(vx * vx + vy * vy + vz * vz)
        Ei = [vx / norm, vy / norm, vz / norm]
        Hi = [ky * Ei[2] - kz * Ei[1], kz * Ei[0] - kx * Ei[2], kx * Ei[1] - ky * Ei[0]]
        Hi = [h / hb.wave_impedance for h in Hi]
        huy_Ei = np.array(Ei)
        huy_Hi = np.array(Hi)
        huy_faces = np.zeros(len(hfaces), dtype=HUYGENS_FACE_DTYPE)
        for i, (cx, cy, cz, face, nrm) in enumerate(hfaces):
            X, Y, Z = cell_point_coordinate(p.mesh_min, dl, (cx, cy, cz), face)
            t = _min_face(nx, ny, nz, (cx, cy, cz), face)
            huy_faces["row"][i] = 0 if t is None else fexc_rank_full[face_flat(t)]
            huy_faces["nx"][i], huy_faces["ny"][i], huy_faces["nz"][i] = nrm
            huy_faces["offset"][i] = kx * X + ky * Y + kz * Z
        b = hb.box
        xs = (p.mesh_min[0] + (b.i0 - 1) * dl, p.mesh_min[0] + b.i1 * dl)
        ys = (p.mesh_min[1] + (b.j0 - 1) * dl, p.mesh_min[1] + b.j1 * dl)
        zs = (p.mesh_min[2] + (b.k0 - 1) * dl, p.mesh_min[2] + b.k1 * dl)
        offset_min = 1e30
        for z in zs:
            for y in ys:
                for x in xs:
                    offset_min = min(kx * x + ky * y + kz * z, offset_min)
        huy_offset_min = offset_min - dl / 2.0

    # ---- outputs --------------------------------------------------------------
    o_rows, o_comp = [], []
    face_out_pts = []
    for o in p.outputs:
        if FACES[o.point] == CENTRE:
            o_rows.append(int(out_rank[lin(o.cell)]))
            o_comp.append(_comp(o.component))
        else:
            t = _min_face(nx, ny, nz, o.cell, FACES[o.point])
            face_out_pts.append((0 if t is None else int(fout_rank_full[face_flat(t)]),
                                 _comp(o.component)))
    out_cells = np.stack(np.nonzero(local_out)[::-1], axis=1) + 1 if flat_out.any() else np.zeros((0, 3), int)
    # frequency outputs: per item the output row (cell row for volumes, face
    # row for surfaces, 1-based), the side of the face whose fields it takes
    # (1: the face is its cell's min face) and the face normal as a signed
    # axis (J = n x H), in the reference's per-step loop order
    normal_of = {FACE_XMIN: (1, 1), FACE_XMAX: (2, -1), FACE_YMIN: (1, 2), FACE_YMAX: (2, -2),
                 FACE_ZMIN: (1, 3), FACE_ZMAX: (2, -3)}
    freq_outputs = []
    for fq, faces in zip(p.frequency_outputs, fo_faces):
        e = {"kind": 1 if fq.kind == "volume" else 2, "frequency": float(fq.frequency),
             "component": _comp(fq.component), "first_timestep": int(fq.first_timestep),
             "last_timestep": int(fq.last_timestep), "timestep_interval": int(fq.timestep_interval)}
        if fq.kind == "volume":
            if e["component"] > 6:
                raise ValueError("frequency output volumes take field components Ex..Hz")
            m = fq.region.mask(nx, ny, nz).ravel()
            e["rows"] = out_rank[np.nonzero(m)[0]].astype(np.int32)
            e["sides"] = np.zeros(len(e["rows"]), dtype=np.int32)
            e["normals"] = np.zeros(len(e["rows"]), dtype=np.int32)
        else:
            if e["component"] > 9:
                raise ValueError("frequency output surfaces take Ex..Hz and Jx..Jz")
            rows, sides, normals = [], [], []
            for f in faces:
                t = _min_face(nx, ny, nz, f[:3], f[3])
                rows.append(int(fout_rank_full[face_flat(t)]))
                sides.append(normal_of[f[3]][0])
                normals.append(normal_of[f[3]][1])
            e["rows"] = np.asarray(rows, dtype=np.int32)
            e["sides"] = np.asarray(sides, dtype=np.int32)
            e["normals"] = np.asarray(normals, dtype=np.int32)
        freq_outputs.append(e)

    def surface_items(faces):
        """Face output rows, sides, signed normals and centre coordinates
        (get_cell_point_coordinate, general_mesh_subroutines.F90:84-137) of
        output-surface faces."""
        rows, sides, normals, centres = [], [], [], []
        for (cx, cy, cz, face) in faces:
            t = _min_face(nx, ny, nz, (cx, cy, cz), face)
            rows.append(int(fout_rank_full[face_flat(t)]))
            sides.append(normal_of[face][0])
            normals.append(normal_of[face][1])
            off = [-0.5, -0.5, -0.5]
            if face in (FACE_XMIN, FACE_YMIN, FACE_ZMIN):
                off[face - FACE_XMIN] = -1.0
            elif face in (FACE_XMAX, FACE_YMAX, FACE_ZMAX):
                off[face - FACE_XMAX] = 0.0
            centres.append([p.mesh_min[a] + (cx, cy, cz)[a] * dl + dl * off[a] for a in range(3)])
        return (np.asarray(rows, dtype=np.int32), np.asarray(sides, dtype=np.int32),
                np.asarray(normals, dtype=np.int32), np.asarray(centres, dtype=np.float64).reshape(-1, 3))

    field_outputs = []
    extra = {"em": 13, "hm": 14, "jm": 15}
    for fo, faces in zip(p.field_outputs, fd_faces):
        comp = extra[fo.component.lower()] if fo.component.lower() in extra else _comp(fo.component)
        e = {"kind": 1 if fo.kind == "volume" else 2, "component": comp, "both_sides": int(fo.side == 0),
             "first_timestep": int(fo.first_timestep), "last_timestep": int(fo.last_timestep),
             "timestep_interval": int(fo.timestep_interval)}
        if fo.kind == "volume":
            if comp > 6:
                raise ValueError("output volumes take field components Ex..Hz")
            m = fo.region.mask(nx, ny, nz).ravel()
            e["rows"] = out_rank[np.nonzero(m)[0]].astype(np.int32)
            e["sides"] = np.zeros(len(e["rows"]), dtype=np.int32)
            e["normals"] = np.zeros(len(e["rows"]), dtype=np.int32)
            e["cells"] = np.stack(np.unravel_index(np.nonzero(m)[0], (nz, ny, nx))[::-1], axis=1) + 1
        else:
            if fo.side == 0 and comp not in (7, 8, 9, 15):
                raise ValueError("output on both sides of a surface takes Jx, Jy, Jz or Jm")
            rows, sides, normals, _ = surface_items(faces)
            e["rows"], e["sides"], e["normals"] = rows, sides, normals
            e["faces"] = np.asarray(faces, dtype=np.int64).reshape(-1, 4)
        field_outputs.append(e)

    # far-field surfaces: angles in radians and their counts as
    # read_far_field_surface.F90 and write_far_field_surfaces
    # (far_field_surfaces.F90:503-507) form them
    rad = PI / 180.0
    far_fields = []
    for s, faces in zip(p.far_field_surfaces, ff_faces):
        if s.dim == 2:
            th = (PI / 2.0, PI / 2.0, 1.0)
        else:
            th = (rad * s.theta[0], rad * s.theta[1], rad * s.theta[2])
        ph = (rad * s.phi[0], rad * s.phi[1], rad * s.phi[2])
        rows, sides, normals, centres = surface_items(faces)
        far_fields.append({"frequency": float(s.frequency), "dim": int(s.dim),
                           "theta_min": th[0], "theta_step": th[2], "n_theta": int((th[1] - th[0]) / th[2]) + 1,
                           "phi_min": ph[0], "phi_step": ph[2], "n_phi": int((ph[1] - ph[0]) / ph[2]) + 1,
                           "rows": rows, "sides": sides, "normals": normals, "centres": centres})
    # the RCS surface: its far-field points as initialise_rcs_surfaces
    # (RCS_surfaces.F90:159-367) sizes them, n_rcs_points_per_cell = 2
    # (TLM_output_modules.F90:388)
    rcs = None
    if p.rcs_surface is not None:
        s = p.rcs_surface
        theta, phi = rad * s.theta, rad * s.phi
        rows, sides, normals, centres = surface_items(rcs_faces)
        K = (1.0 * math.sin(theta) * math.cos(phi), 1.0 * math.sin(theta) * math.sin(phi), 1.0 * math.cos(theta))
        d = [K[0] * c[0] + K[1] * c[1] + K[2] * c[2] for c in centres]
        dmin, dmax = (min(d), max(d)) if d else (1e30, -1e30)
        sim_time = p.simulation_time if p.simulation_time is not None else nts * dt
        period = (dmax - dmin) / C0 + sim_time
        distance = period * C0
        ppc = 2
        rcs = {"theta": theta, "phi": phi, "points_per_cell": ppc,
               "n_points": int(ppc * distance / dl + 2), "point_offset": int(ppc * dmin / dl - 1),
               "fmin": float(s.fmin), "fmax": float(s.fmax), "fstep": float(s.fstep),
               "rows": rows, "sides": sides, "normals": normals, "centres": centres}

    return Tables(
        nx=nx, ny=ny, nz=nz, dl=dl, dt=dt, n_timesteps=nts,
        reduced_c_factor=p.reduced_c_factor, reflection=tuple(p.reflection),
        cell_code=cell_code, n_special_cells=n_special,
        cell_to_material=to_material, cell_to_pml=to_pml, cell_to_excitation=to_exc,
        cell_to_output=to_out, cell_to_cable=to_cable,
        face_code=face_code, n_special_faces=n_special_faces,
        face_to_material=face_to_material, face_to_excitation=face_to_exc,
        face_to_output=face_to_out, face_to_cable=face_to_cable,
        volume_materials=vm, surface_materials=np.asarray(smat_types, dtype=np.int32),
        pml_params=pml_params, pml_cell_param=np.asarray(pml_cell_param, dtype=np.int32),
        value=value, value_face=value_face, point_sources=ps,
        n_cell_excitation_rows=int(flat_exc.sum()), n_face_excitation_rows=len(exc_faces_full),
        huygens_function=huy_function, huygens_offset_min=huy_offset_min,
        huygens_Ei=huy_Ei, huygens_Hi=huy_Hi, huygens_faces=huy_faces,
        output_rows=np.asarray(o_rows, dtype=np.int32),
        output_components=np.asarray(o_comp, dtype=np.int32),
        n_output_cells=int(flat_out.sum()), n_output_faces=len(out_faces_full),
        face_output_points=face_out_pts,
        dft_frequencies=np.asarray(p.dft_frequencies, dtype=np.float64),
        output_cells=out_cells, n_pml_cells=n_pml_cells, freq_outputs=freq_outputs,
        far_fields=far_fields, rcs=rcs, field_outputs=field_outputs,
    )
