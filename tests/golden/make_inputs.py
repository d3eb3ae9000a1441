"""Copy the reference's own test inputs used by tests/test_ingest.py into
tests/golden/inputs/ (data files of TEST_DATA: problem specification
files, material files, one model-builder mesh, and the RCS output files
of one case, the larger ones gzip-compressed).

    python tests/golden/make_inputs.py     # in the container holding the reference
"""
import gzip
import os
import shutil

SRC = "/root/reference/GGI_TLM/TEST_DATA"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "inputs")
FILES = {
    "CAVITY_FREE_SPACE/PROBLEM_SPECIFICATION_FILES/cavity_1_2_3.inp": "cavity_free_space.inp",
    "CAVITY_EPSR=2/PROBLEM_SPECIFICATION_FILES/cavity_1_2_3.inp": "cavity_epsr2.inp",
    "CAVITY_MUR=2/PROBLEM_SPECIFICATION_FILES/cavity_1_2_3.inp": "cavity_mur2.inp",
    "MATERIAL_TEST2/PROBLEM_SPECIFICATION_FILES/mat_test.inp": "material_test2.inp",
    "HUYGENS_SURFACE/PROBLEM_SPECIFICATION_FILES/huygens_test.inp": "huygens_surface.inp",
    "TEM_PML_TEST/PROBLEM_SPECIFICATION_FILES/TEM_cell.inp": "tem_pml.inp",
    "HELIX_ANTENNA/PROBLEM_SPECIFICATION_FILES/helix.inp": "helix.inp",
    "RCS_PEC_PLATE_ROUGH_SURFACE/PROBLEM_SPECIFICATION_FILES/plate.inp": "rcs_plate.inp",
    "RCS_PEC_PLATE_ROUGH_SURFACE/PROBLEM_SPECIFICATION_FILES/plate.rcs.fout_ref": "rcs_plate.rcs.fout_ref",
    "HELIX_ANTENNA/helix.info": "helix.info",
}
MATERIALS = ("epsr=2.vmat", "mur=2.vmat", "mur_lorentz.vmat", "epsr=1.5.vmat")


def main():
    os.makedirs(os.path.join(OUT, "materials"), exist_ok=True)
    for src, dst in FILES.items():
        shutil.copyfile(os.path.join(SRC, src), os.path.join(OUT, dst))
    for m in MATERIALS:
        shutil.copyfile(os.path.join(SRC, "MATERIAL_DATA", m), os.path.join(OUT, "materials", m))
    with open(os.path.join(SRC, "RCS_PEC_PLATE_ROUGH_SURFACE", "PROBLEM_SPECIFICATION_FILES", "plate.rcs.tout_ref"),
              "rb") as f, gzip.GzipFile(os.path.join(OUT, "rcs_plate.rcs.tout_ref.gz"), "wb", mtime=0) as g:
        g.write(f.read())
    with open(os.path.join(SRC, "HELIX_ANTENNA", "helix.mesh"), "rb") as f, \
            gzip.GzipFile(os.path.join(OUT, "helix.mesh.gz"), "wb", mtime=0) as g:
        g.write(f.read())


if __name__ == "__main__":
    main()
