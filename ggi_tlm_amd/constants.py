"""Physical constants, port and component numbering of GGI_TLM.

Values follow SRC/CONSTANTS/constants.F90:27-39; port numbering
SRC/TLM_MODULES/Mesh_Modules.F90:83-94; field components
SRC/TLM_MODULES/TLM_Modules.F90:164-176; cell faces
SRC/TLM_MODULES/Mesh_Modules.F90:40-46.  These numeric values define parity
with the reference, so they are kept exactly.
"""

PI = 3.141592653589793
Z0 = 376.73031346177
Y0 = 1.0 / Z0
EPS0 = 8.8541878176e-12
MU0 = 3.141592653589793 * 4e-7
C0 = 2.99792458e8
SMALL = 1e-12

# SCN link ports (1-based, as in the reference)
VY_XMIN, VZ_XMIN, VY_XMAX, VZ_XMAX = 1, 2, 3, 4
VX_YMIN, VZ_YMIN, VX_YMAX, VZ_YMAX = 5, 6, 7, 8
VX_ZMIN, VY_ZMIN, VX_ZMAX, VY_ZMAX = 9, 10, 11, 12

# field components
COMPONENTS = {
    "ex": 1, "ey": 2, "ez": 3, "hx": 4, "hy": 5, "hz": 6,
    "jx": 7, "jy": 8, "jz": 9, "mx": 10, "my": 11, "mz": 12,
}

# cell faces / points
FACE_XMIN, FACE_YMIN, FACE_ZMIN, FACE_XMAX, FACE_YMAX, FACE_ZMAX, CENTRE = 1, 2, 3, 4, 5, 6, 7
FACES = {
    "xmin": FACE_XMIN, "ymin": FACE_YMIN, "zmin": FACE_ZMIN,
    "xmax": FACE_XMAX, "ymax": FACE_YMAX, "zmax": FACE_ZMAX, "centre": CENTRE,
}

# volume / surface material types (SRC/TLM_MODULES/TLM_material_modules.F90:72-169)
VMAT_PEC, VMAT_PMC, VMAT_DISPERSIVE = 1, 2, 3
SMAT_PEC, SMAT_PMC, SMAT_FREE_SPACE = 1, 2, 4
# host-path surface materials (their face models stay on the host)
SMAT_DISPERSIVE, SMAT_ANISOTROPIC_DISPERSIVE, SMAT_DIODE, SMAT_SPICE, SMAT_SWITCH = 3, 5, 6, 7, 9
SURFACE_MATERIALS = {"pec": SMAT_PEC, "pmc": SMAT_PMC, "free_space": SMAT_FREE_SPACE,
                     "thin_layer": SMAT_DISPERSIVE, "anisotropic_thin_layer": SMAT_ANISOTROPIC_DISPERSIVE,
                     "diode": SMAT_DIODE, "spice": SMAT_SPICE, "switch": SMAT_SWITCH}

SOURCE_HARD, SOURCE_SOFT = 0, 1

MAX_FILTER_ORDER = 8
