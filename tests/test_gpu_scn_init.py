"""gpu_scn_init (ggi_tlm_amd/fortran/gpu_scn_init.F90) against the
reference's own module layouts.

The reference's module files (TLM_general, mesh, PML_module,
TLM_volume_materials, TLM_surface_materials, TLM_excitation, TLM_output and
the types they use) are compiled from the reference checkout by
oracle/ref_modules.mk; gpu_scn_init and a driver are compiled against them.
The driver fills those modules' globals with a deck's tables (the state the
reference's set-up leaves), calls gpu_scn_init and steps the solve through
the engine ABI, here bound to the CPU oracle (the objects' tlm_* references
renamed to the oracle's oracle_* exports).  Its probe rows must equal the
oracle's run of the same tables bit for bit, so every module field reaches
the ABI argument it belongs to.  Skipped where the reference checkout or
amdflang is absent (the reference never leaves this container).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import problems
from ggi_tlm_amd.problem import build_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/GGI_TLM/SRC"
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")

pytestmark = pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(FC) and shutil.which("objcopy")),
                                reason="reference checkout, amdflang or objcopy not available")


def abi_names():
    txt = open(os.path.join(ROOT, "include", "tlmscn.h")).read()
    return sorted(set(re.findall(r"\b(tlm_[a-z_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory, oracle_lib):
    d = tmp_path_factory.mktemp("scn_init")
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref_modules.mk")], check=True)
    ref = os.path.join(ROOT, "oracle", "_ref")
    fort = os.path.join(ROOT, "ggi_tlm_amd", "fortran")
    objs = []
    for src in (os.path.join(fort, "tlm_gpu_iface.F90"), os.path.join(fort, "gpu_scn_init.F90"),
                os.path.join(ROOT, "tests", "fortran", "scn_init_driver.F90")):
        obj = str(d / (os.path.basename(src)[:-4] + ".o"))
        subprocess.run([FC, "-O0", "-cpp", "-J", str(d), "-I", ref, "-c", src, "-o", obj], check=True, cwd=d)
        objs.append(obj)
    redef = []
    for n in abi_names():
        redef += ["--redefine-sym", f"{n}=oracle_{n[len('tlm_'):]}"]
    for o in objs:
        subprocess.run(["objcopy", *redef, o], check=True)
    ref_objs = sorted(os.path.join(ref, f) for f in os.listdir(ref) if f.endswith(".o"))
    exe = str(d / "scn_init_driver")
    oracle_dir = os.path.join(ROOT, "oracle")
    subprocess.run([FC, *objs, *ref_objs, "-L", oracle_dir, "-loracle_scn", f"-Wl,-rpath,{oracle_dir}", "-o", exe],
                   check=True, cwd=d)
    return exe, d


def ints(a):
    return " ".join(str(int(v)) for v in np.asarray(a).ravel(order="F"))


def reals(a):
    return " ".join(repr(float(v)) for v in np.asarray(a).ravel(order="F"))


def write_tables(t, path):
    L = [f"{t.nx} {t.ny} {t.nz} {t.n_timesteps}", f"{t.dl!r} {t.dt!r} {t.reduced_c_factor!r}",
         reals(t.reflection), ints(t.cell_code), str(t.n_special_cells)]
    if t.n_special_cells:
        cm = np.asarray(t.cell_to_material).reshape(t.n_special_cells, 2, order="F")
        L += [ints(cm[:, 0]), ints(cm[:, 1]), ints(t.cell_to_pml), ints(t.cell_to_excitation),
              ints(t.cell_to_output), ints(t.cell_to_cable)]
    L += [str(len(t.face_code))]
    if len(t.face_code):
        L.append(ints(t.face_code))
    L.append(str(t.n_special_faces))
    if t.n_special_faces:
        fm = np.asarray(t.face_to_material).reshape(t.n_special_faces, 3, order="F")
        L += [ints(fm[:, 0]), ints(fm[:, 1]), ints(fm[:, 2]), ints(t.face_to_excitation), ints(t.face_to_output),
              ints(t.face_to_cable)]
    vm = t.volume_materials
    L.append(str(len(vm)))
    for r in vm:
        eo, mo = int(r["eps_order"]), int(r["mu_order"])
        L += [f"{int(r['type'])} {int(r['eps_filter_exists'])} {int(r['mu_filter_exists'])} {eo} {mo}",
              reals([r["Ys_eps_f"], r["Ge"], r["Zs_mu_f"], r["Rm"]]),
              reals(r["eps_a"][:eo + 1]), reals(r["eps_b"][:eo + 1]), reals(r["mu_a"][:mo + 1]),
              reals(r["mu_b"][:mo + 1])]
    L.append(str(len(t.surface_materials)))
    L += [str(int(v)) for v in t.surface_materials]
    pp = t.pml_params
    L.append(f"{len(pp)} {len(t.pml_cell_param)}")
    for r in pp:
        L.append(reals([r[n] for n in pp.dtype.names]))
    L += [str(int(v)) for v in t.pml_cell_param]
    nf = t.value.shape[1] if t.value.ndim == 2 else 0
    L.append(str(nf))
    for k in range(nf):
        L += [reals(t.value[:, k]), reals(t.value_face[:, k])]
    ps = t.point_sources
    L.append(f"{len(ps)} {t.n_cell_excitation_rows} {t.n_face_excitation_rows}")
    for r in ps:
        L.append(ints([r["function"], r["component"], r["source_type"], r["at_face"], r["row"]]))
    hf = t.huygens_faces
    L.append(str(len(hf)))
    if len(hf):
        L += [f"{int(t.huygens_function)} {t.huygens_offset_min!r}", reals(list(t.huygens_Ei) + list(t.huygens_Hi))]
        for r in hf:
            L.append(f"{int(r['row'])} {int(r['nx'])} {int(r['ny'])} {int(r['nz'])} {float(r['offset'])!r}")
    L.append(f"{len(t.output_rows)} {t.n_output_cells}")
    for r, c in zip(t.output_rows, t.output_components):
        L.append(f"{int(r)} {int(c)}")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


DECKS = {
    "pml_sphere": lambda: problems.pml_sphere(n=20, steps=40, npml=3),
    "special_mix": lambda: problems.special_mix(14, 40),
    "dispersive_blocks": problems.dispersive_blocks,
    "pml_material": lambda: problems.pml_material(steps=40),
}


@pytest.mark.parametrize("name", sorted(DECKS))
def test_gpu_scn_init_tables(driver, name):
    from oracle import oracle

    exe, d = driver
    t = build_tables(DECKS[name]())
    steps = min(t.n_timesteps, 40)
    fin, fout = str(d / f"{name}.tables"), str(d / f"{name}.out")
    write_tables(t, fin)
    r = subprocess.run([exe, fin, fout, str(steps)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(fout) as f:
        n_read, first = (int(v) for v in f.readline().split())
        vals = np.array([float(v) for v in f.read().split()])
    assert n_read == steps and first == 1
    got = vals.reshape(steps, t.n_output_cells, 6)
    o = oracle.handle(t, output_capacity=steps)
    o.step(1, steps)
    want, _ = o.read_cell_outputs()
    assert np.abs(want).max() > 0
    assert np.array_equal(got, want)
