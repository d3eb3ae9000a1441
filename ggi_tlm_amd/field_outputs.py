"""Time-domain surface and volume field output files in the reference's
formats (OUTPUT/output_surfaces.F90, OUTPUT/output_volumes.F90).

The device records one value per face or cell for every step in an
output's window (tlm_set_field_outputs, tlm_read_field_output).  These
writers lay out the reference's `.surface_field.tout` and
`.volume_field.tout` files:
* a header with the number of outputs, then per output a mesh template:
  the corner points of each face (4) or of each cell's six faces (24),
  format 3E14.5, and per face the polygon record (I3,4I12);
* then, timestep by timestep and output by output, a frame block with the
  output number, frame number and value count (list-directed integers) and
  one value per record (E14.5); a volume writes each cell's value six
  times, once per face of its hexahedron.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from .constants import FACE_XMAX, FACE_XMIN, FACE_YMAX, FACE_YMIN, FACE_ZMAX, FACE_ZMIN
from .tout import fortran_e

# corner offsets (in cells, from the cell's high corner) of each face's four
# corners, in the order get_cell_face_corner_coordinates
# (MESH/general_mesh_subroutines.F90:154-269) lists them
_FACE_CORNERS = {
    FACE_XMIN: ((-1, -1, -1), (-1, -1, 0), (-1, 0, 0), (-1, 0, -1)),
    FACE_XMAX: ((0, -1, -1), (0, 0, -1), (0, 0, 0), (0, -1, 0)),
    FACE_YMIN: ((-1, -1, -1), (0, -1, -1), (0, -1, 0), (-1, -1, 0)),
    FACE_YMAX: ((-1, 0, -1), (-1, 0, 0), (0, 0, 0), (0, 0, -1)),
    FACE_ZMIN: ((-1, -1, -1), (-1, 0, -1), (0, 0, -1), (0, -1, -1)),
    FACE_ZMAX: ((-1, -1, 0), (0, -1, 0), (0, 0, 0), (-1, 0, 0)),
}
# a cell's eight corners (get_cell_corner_coordinates) and the six faces the
# volume template lists, as indices into them
_CELL_CORNERS = ((-1, -1, -1), (-1, -1, 0), (-1, 0, 0), (-1, 0, -1),
                 (0, -1, -1), (0, -1, 0), (0, 0, 0), (0, 0, -1))
_CELL_FACES = ((0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (3, 2, 6, 7), (0, 3, 7, 4), (1, 5, 6, 2))


def list_int(v: int) -> str:
    """A list-directed write of one default integer (leading blank, I12)."""
    return " " + f"{int(v):12d}"


def due_steps(e: dict, n_timesteps: int) -> int:
    """Frames of an output over the run (get_output_flag over 1..n)."""
    return sum(1 for t in range(1, n_timesteps + 1)
               if e["first_timestep"] <= t <= e["last_timestep"]
               and (t - e["first_timestep"]) % e["timestep_interval"] == 0)


def _point(mesh_min, dl, cell, off):
    return [mesh_min[a] + cell[a] * dl + off[a] * dl for a in range(3)]


def _e14_5(v: float) -> str:
    return fortran_e(float(v), 14, 5)


def write_field_outputs(path: str, kind: str, outputs: Sequence[dict], frames: Sequence[Tuple[np.ndarray, int]],
                        n_timesteps: int, mesh_min, dl: float):
    """Write the surface ('surface') or volume ('volume') field outputs of a
    run: outputs are the tables' field-output entries of that kind, frames
    per output (values (n_frames, n_items), timestep of the first frame)."""
    surf = kind == "surface"
    word = "SURFACE" if surf else "VOLUME"
    L = [f"# NUMBER OF OUTPUT {word}S", list_int(len(outputs))]
    for k, e in enumerate(outputs, 1):
        n = len(e["rows"])
        if surf:
            L.append(f"# START SURFACE FIELD FILE TEMPLATE, OUTPUT SURFACE NUMBER:{k:5d}")
            L += ["# Number of points:", f"{4 * n:10d}", "# Number of faces:", f"{n:10d}"]
        else:
            L.append(f"# START volume FIELD FILE TEMPLATE, OUTPUT volume NUMBER:{k:5d}")
            L += ["# Number of points:", f"{24 * n:10d}", "# Number of faces:", f"{6 * n:10d}"]
        L += ["# Number of frames:", f"{due_steps(e, n_timesteps):10d}"]
        if surf:
            for (cx, cy, cz, face) in e["faces"]:
                for off in _FACE_CORNERS[int(face)]:
                    L.append("".join(_e14_5(v) for v in _point(mesh_min, dl, (cx, cy, cz), off)))
            polys = n
        else:
            for c in e["cells"]:
                corners = [_point(mesh_min, dl, c, off) for off in _CELL_CORNERS]
                for fc in _CELL_FACES:
                    for q in fc:
                        L.append("".join(_e14_5(v) for v in corners[q]))
            polys = 6 * n
        for q in range(polys):
            L.append(f"{4:3d}" + "".join(f"{4 * q + j:12d}" for j in range(4)))
        if surf:
            L.append(f"# END SURFACE FIELD FILE TEMPLATE, OUTPUT SURFACE NUMBER:{k:5d}")
        else:
            L.append(f"# END VOLUME FIELD FILE TEMPLATE, OUTPUT VOLUME NUMBER:{k:5d}")
    # frames in timestep order, outputs in order within a timestep
    events: List[Tuple[int, int, int]] = []
    for k, (vals, first) in enumerate(frames):
        for f in range(vals.shape[0]):
            events.append((first + f * outputs[k]["timestep_interval"], k, f))
    events.sort()
    for (_, k, f) in events:
        vals = frames[k][0][f]
        if surf:
            L += ["# START OF SURFACE FIELD OUTPUT DATA", "# OUTPUT SURFACE NUMBER", list_int(k + 1),
                  "# FRAME NUMBER", list_int(f + 1), "# NUMBER OF FACES", list_int(len(vals)),
                  "# SURFACE FIELD DATA"]
            L += [_e14_5(v) for v in vals]
            L.append(f"# END OF SURFACE FIELD OUTPUT DATA, OUTPUT SURFACE NUMBER:{k + 1:5d}")
        else:
            L += ["# START OF VOLUME FIELD OUTPUT DATA", "# OUTPUT VOLUME NUMBER", list_int(k + 1),
                  "# FRAME NUMBER", list_int(f + 1), "# NUMBER OF cellS", list_int(6 * len(vals)),
                  "# volume FIELD DATA"]
            for v in vals:
                L += [_e14_5(v)] * 6
            L.append(f"# END OF VOLUME FIELD OUTPUT DATA, OUTPUT volume NUMBER:{k + 1:5d}")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")
