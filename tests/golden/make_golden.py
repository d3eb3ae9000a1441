"""Extract the reference's own golden result files for the hot path into
small numeric fixtures (tests/golden/reference_goldens.npz).

Run in a container where the reference checkout is mounted:
    python tests/golden/make_golden.py /root/reference/GGI_TLM/TEST_DATA

Only data is kept (the numbers printed in the reference's *_ref files); no
reference source is copied.  Each array is named after its case and file.
"""
import os
import sys

import numpy as np

FILES = {
    "cavity_free_space_fout": "CAVITY_FREE_SPACE/PROBLEM_SPECIFICATION_FILES/cavity_1_2_3.fout_ref",
    "cavity_epsr2_fout": "CAVITY_EPSR=2/PROBLEM_SPECIFICATION_FILES/cavity_1_2_3.fout_ref",
    "cavity_mur2_fout": "CAVITY_MUR=2/PROBLEM_SPECIFICATION_FILES/cavity_1_2_3.fout_ref",
    "material_test2_tout": "MATERIAL_TEST2/PROBLEM_SPECIFICATION_FILES/mat_test.field.tout_ref",
    "tem_pml_ey_tout": "TEM_PML_TEST/PROBLEM_SPECIFICATION_FILES/Ey.tout_ref",
}
for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
    FILES[f"huygens_{c.lower()}_centre_tout"] = f"HUYGENS_SURFACE/PROBLEM_SPECIFICATION_FILES/{c}_centre.tout_ref"


# The first lines of one probe-output data file, kept as text to pin the
# .tout format byte for byte (header and (E16.7,I5,E16.7) records).
TOUT_TEXT = ("TEM_PML_TEST/PROBLEM_SPECIFICATION_FILES/Ey.tout_ref", "tem_pml_ey_head.tout", 400)


def main(root):
    out = {}
    here = os.path.dirname(os.path.abspath(__file__))
    for key, rel in FILES.items():
        out[key] = np.loadtxt(os.path.join(root, rel), comments="#")
    path = os.path.join(here, "reference_goldens.npz")
    np.savez_compressed(path, **out)
    print(path, {k: v.shape for k, v in out.items()})
    rel, name, n = TOUT_TEXT
    with open(os.path.join(root, rel)) as f:
        lines = [next(f) for _ in range(n)]
    with open(os.path.join(here, name), "w") as f:
        f.writelines(lines)
    print(os.path.join(here, name), len(lines), "lines")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/GGI_TLM/TEST_DATA")
