"""Standalone ingest (ggi_tlm_amd/ingest.py) against the reference's own
test inputs (tests/golden/inputs/, copied from TEST_DATA by
tests/golden/make_inputs.py).

* The six golden decks' `.inp` files give the same tables, bit for bit, as
  the decks built by hand in tests/problems.py, whose oracle runs match the
  reference's golden outputs (test_oracle_golden.py).  MATERIAL_TEST2's two
  points lie exactly on cell boundaries: the ingest flags them, and with
  the cells the reference's model builder gave them (the `.mesh` point
  records; pinned by the golden series) the tables match.
* HUYGENS_SURFACE ingested with all its output points (centre and face)
  runs on the oracle to the reference's golden centre-field series.
* The model builder's own `.mesh` (TEST_DATA/HELIX_ANTENNA/helix.mesh)
  reads to the counts the reference's run log (helix.info) states, and
  survives a write/read round trip.
* `.vmat` files read to the filters the hand-built decks use.
"""
import os

import numpy as np
import pytest

import problems
from ggi_tlm_amd import ingest
from ggi_tlm_amd.capi import probe_series
from ggi_tlm_amd.filters import Sfilter
from ggi_tlm_amd.problem import build_tables

D = os.path.join(os.path.dirname(__file__), "golden", "inputs")
MAT = os.path.join(D, "materials")
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.npz"))


def same_tables(a, b):
    diff = []
    for k in vars(a):
        x, y = getattr(a, k), getattr(b, k)
        if isinstance(x, np.ndarray) or isinstance(y, np.ndarray):
            ok = np.array_equal(np.asarray(x), np.asarray(y))
        else:
            ok = x == y
        if not ok:
            diff.append(k)
    return diff


@pytest.mark.parametrize("tok,want", [("3.0-3", 3.0e-3), ("-3.0e-3", -3.0e-3), ("1.0d2", 100.0), (".5", 0.5),
                                      ("5.", 5.0), ("-2", -2.0), ("1E+3", 1000.0), ("2.5+1", 25.0)])
def test_fortran_reals(tok, want):
    assert ingest.fortran_real(tok) == want


def test_fortran_real_rejects_words():
    with pytest.raises(ingest.IngestError):
        ingest.fortran_real("Ex")


DECKS = {
    "cavity_free_space": problems.cavity,
    "cavity_epsr2": lambda: problems.cavity(Sfilter(1e8, [2.0], [1.0]), None),
    "cavity_mur2": lambda: problems.cavity(None, Sfilter(1e8, [2.0], [1.0])),
    "tem_pml": problems.tem_pml,
}


@pytest.mark.parametrize("name", sorted(DECKS))
def test_reference_inputs_give_the_hand_built_tables(name):
    ing = ingest.load(os.path.join(D, name + ".inp"), material_dir=MAT)
    assert ing.unsupported == [] and ing.boundary_points == []
    assert same_tables(build_tables(DECKS[name]()), build_tables(ing.problem)) == []


def test_material_test2_boundary_points():
    path = os.path.join(D, "material_test2.inp")
    ing = ingest.load(path, material_dir=MAT)
    assert ing.boundary_points == [1, 2]        # x = -1.2e-3 and +1.2e-3 on cell boundaries
    ing = ingest.load(path, material_dir=MAT, point_cells={1: (18, 31, 31), 2: (42, 31, 31)})
    assert same_tables(build_tables(problems.material_test2()), build_tables(ing.problem)) == []


def test_huygens_surface_inp_runs_to_the_golden_series(oracle_lib):
    from oracle import oracle

    ing = ingest.load(os.path.join(D, "huygens_surface.inp"))
    assert ing.unsupported == []
    p = ing.problem
    # its three output surfaces (|E| on surfaces 2, 3, 4, side -1) and three
    # output volumes (Hx, Hy, Hz of volume 1), every 4th step to 120
    assert [(o.kind, o.component, o.side, o.first_timestep, o.last_timestep, o.timestep_interval)
            for o in p.field_outputs] == [("surface", "em", -1, 0, 120, 4)] * 3 + \
        [("volume", c, 1, 0, 120, 4) for c in ("hx", "hy", "hz")]
    hand = problems.huygens_surface()
    centre = [o for o in p.outputs if o.point == "centre"]
    assert [(o.cell, o.component) for o in centre] == [(o.cell, o.component) for o in hand.outputs]
    assert len(p.outputs) == 30 and p.huygens == hand.huygens
    # the same tables as the hand-built deck once the face outputs are left out
    q = ingest.load(os.path.join(D, "huygens_surface.inp")).problem
    q.outputs = centre
    q.field_outputs = []
    assert same_tables(build_tables(hand), build_tables(q)) == []
    # all output points (face outputs make their faces special) on the oracle
    t = build_tables(p)
    h = oracle.handle(t, output_capacity=t.n_timesteps)
    h.step(1, t.n_timesteps)
    s = probe_series(h.read_cell_outputs()[0], t)     # one column per centre output point, in order
    assert s.shape[1] == len(centre) == 6
    for i, o in enumerate(centre):
        r = GOLD[f"huygens_{o.component}_centre_tout"][:, 2]
        err = np.max(np.abs(s[:, i] - r)) / np.max(np.abs(r))
        assert err < 1e-6, (o.component, err)
    fo = h.read_face_outputs()[0]
    assert fo.shape[0] == t.n_timesteps and np.abs(fo).max() > 0


def info_counts():
    counts = {}
    with open(os.path.join(D, "helix.info")) as f:
        for line in f:
            if "Total number of material cells=" in line:
                counts["material_cells"] = int(line.split("=")[1])
            elif "Total number of material faces=" in line:
                counts["material_faces"] = int(line.split("=")[1])
            elif "Number of cell_faces=" in line:
                counts.setdefault("frequency_surface_faces", []).append(int(line.split("=")[-1]))
            elif "Total number of cells=" in line:
                counts["cells"] = int(line.split("=")[1])
    return counts


def test_helix_mesh_matches_the_reference_run_log(tmp_path):
    m = ingest.read_mesh(os.path.join(D, "helix.mesh.gz"))
    c = info_counts()
    assert m.dl == 0.5 and m.n == (42, 42, 42) and m.mesh_min == (-10.5, -10.5, -10.5)
    assert int(np.prod(m.n)) == c["cells"]
    assert len(m.volumes) == 1 and len(m.volumes[0]) == c["material_cells"]
    # helix.inp: PEC on surfaces 3-6, frequency output surfaces on surface 2
    assert sum(len(m.surfaces[k - 1]["faces"]) for k in (3, 4, 5, 6)) == c["material_faces"]
    assert c["frequency_surface_faces"] == [len(m.surfaces[1]["faces"])] * 2
    assert [len(ln) for ln in m.lines] == [10, 10] and len(m.points) == 4
    assert all(np.all((v[:, :3] >= 1) & (v[:, :3] <= 42)) for v in m.volumes)
    # round trip through the writer
    out = str(tmp_path / "helix.mesh")
    ingest.write_mesh(m, out)
    m2 = ingest.read_mesh(out)
    assert (m2.dl, m2.n, m2.mesh_min, m2.mesh_max) == (m.dl, m.n, m.mesh_min, m.mesh_max)
    for a, b in zip(m.volumes + m.lines, m2.volumes + m2.lines):
        assert np.array_equal(a, b)
    for a, b in zip(m.surfaces, m2.surfaces):
        assert np.array_equal(a["faces"], b["faces"]) and a["bounds"] == b["bounds"]
    assert m.points == m2.points


def test_surfaces_meshed_as_the_model_builder_meshes_them():
    """helix.inp's rectangular block, y plane and two z planes, meshed here
    without the `.mesh` file, give exactly the face lists of the model
    builder's helix.mesh: the walls' grid lines (ties of the block's
    +-8.25 m walls rounded away from zero), the cells on each wall, and the
    side each face is kept on (the outside cell of a block wall; the side
    its plane's triangle normal points away from)."""
    deck = ingest.read_inp(os.path.join(D, "helix.inp"))
    m = ingest.read_mesh(os.path.join(D, "helix.mesh.gz"))
    for k in (1, 2, 3, 4):
        mine = sorted(ingest.mesh_surface(deck.surfaces[k], m.mesh_min, m.dl, m.n))
        ref = sorted(tuple(int(v) for v in f) for f in m.surfaces[k - 1]["faces"])
        assert mine == ref, k


def test_mesh_volumes_become_cell_regions(tmp_path):
    """A deck meshed here and the same deck read back from a written `.mesh`
    (volume cell lists, surface face lists, point cells) give the same
    tables."""
    path = os.path.join(D, "cavity_epsr2.inp")
    a = ingest.load(path, material_dir=MAT).problem
    lo = np.array(a.mesh_min)
    cells = np.array([(i, j, k, 7) for k in range(1, 31) for j in range(1, 21) for i in range(1, 11)])
    pts = []
    for s, o in ((a.sources[0], None), (None, a.outputs[0])):
        cell = (s or o).cell
        pts.append({"xyz": (0.0, 0.0, 0.0), "cell": cell, "face": (*cell, 7)})
    # the packets number the excitation point 1 and the output point 2
    m = ingest.MeshFile(a.dl, (a.nx, a.ny, a.nz), tuple(lo), tuple(lo + a.dl * np.array([a.nx, a.ny, a.nz])),
                        [cells], [], [], pts)
    out = str(tmp_path / "cavity.mesh")
    ingest.write_mesh(m, out)
    b = ingest.load(path, mesh_path=out, material_dir=MAT).problem
    assert same_tables(build_tables(a), build_tables(b)) == []


def test_frequency_output_packets(tmp_path):
    """Frequency output surfaces and volumes (read_frequency_output_surface.F90:70-101,
    read_frequency_output_volume.F90:120-140): helix.inp's two surfaces on the
    model builder's mesh hold the face counts of the reference's run log, and
    a cavity deck with appended packets gives the windows and regions they
    name."""
    ing = ingest.load(os.path.join(D, "helix.inp"), os.path.join(D, "helix.mesh.gz"), material_dir=MAT)
    fo = ing.problem.frequency_outputs
    assert [(o.kind, o.frequency, o.component, o.side) for o in fo] == [("surface", 1.3e7, "ex", -1),
                                                                           ("surface", 1.3e7, "ez", -1)]
    assert "frequency_output_surface_list" not in ing.unsupported
    t = build_tables(ing.problem)
    assert [len(e["rows"]) for e in t.freq_outputs] == info_counts()["frequency_surface_faces"]
    assert [e["component"] for e in t.freq_outputs] == [1, 3]
    text = open(os.path.join(D, "cavity_free_space.inp")).read()
    text += ("\nsurface_list\n1\n1\nrectangular_block\n0.4 0.8 1.2\n0 0 0\n0 0 0\n\n"
             "Frequency_output_surface_list\n1\n1\n1\n1\n2.0e8\nJy\n\n"
             "Frequency_output_volume_list\n2\n1\n1\n3.0e8 frequency\nHz\n"
             "output_timestep_information\n5\n40\n3\n2\n1\n2.5e8\nEx\n\n")
    text += "volume_list\n1\n1\nrectangular_block\n0.2 0.4 0.6\n0 0 0\n0 0 0\n"
    path = tmp_path / "cavity_fo.inp"
    path.write_text(text)
    ing = ingest.load(str(path))
    s, v1, v2 = ing.problem.frequency_outputs
    assert (s.kind, s.component, s.side, s.frequency, len(s.faces)) == ("surface", "jy", 1, 2.0e8,
                                                                         2 * (8 * 12 + 4 * 12 + 4 * 8))
    assert (v1.kind, v1.component, v1.first_timestep, v1.last_timestep, v1.timestep_interval) == \
        ("volume", "hz", 5, 40, 3)
    assert (v2.component, v2.frequency, v2.timestep_interval) == ("ex", 2.5e8, 1)
    t = build_tables(ing.problem)
    assert [len(e["rows"]) for e in t.freq_outputs] == [len(s.faces), 2 * 4 * 6, 2 * 4 * 6]
    e = t.freq_outputs[1]
    assert (e["first_timestep"], e["last_timestep"], e["timestep_interval"]) == (5, 40, 3)


def test_far_field_and_rcs_packets(tmp_path):
    """Far_field_surface_list and RCS_surface (read_far_field_surface.F90,
    read_rcs_surface.F90) of RCS_PEC_PLATE_ROUGH_SURFACE's deck.  Its plate
    is a rough rectangle (random surface points): meshing it needs the
    model builder's `.mesh`; with the roughness record taken out, the deck
    builds to tables with the RCS surface and a 361-angle far-field cut."""
    path = os.path.join(D, "rcs_plate.inp")
    deck = ingest.read_inp(path)
    assert deck.surfaces[1].roughness == (0.02, 0.025)
    assert deck.rcs_surface == {"number": 3, "side": 1, "fmin": 0.3e9, "fmax": 3e9, "fstep": 0.1e8,
                                "theta": 180.0, "phi": 0.0}
    assert deck.far_field_surfaces == [{"number": 3, "side": 1, "frequency": 3e9, "theta": (-180.0, 180.0, 1.0),
                                        "phi": (0.0, 0.0, 1.0), "dim": 3}]
    with pytest.raises(ingest.IngestError, match="rough"):
        ingest.problem_from(deck)
    flat = tmp_path / "plate.inp"
    flat.write_bytes(open(path, "rb").read().replace(b"surface_roughness \n0.02 0.025\n", b""))
    t = build_tables(ingest.load(str(flat)).problem)
    ff, rcs = t.far_fields[0], t.rcs
    assert (ff["n_theta"], ff["n_phi"], ff["dim"], ff["frequency"]) == (361, 1, 3, 3e9)
    assert len(ff["rows"]) == len(rcs["rows"]) == 6 * 29 * 29        # the 0.29 m block, dl = 0.01
    assert np.all(np.abs(ff["centres"]) <= 0.145 + 1e-12) and np.abs(ff["centres"]).max() > 0.145 - 1e-12
    # the far-field series has as many points as the reference's own run of
    # this deck wrote records to plate.rcs.tout (initialise_rcs_surfaces'
    # sizing from the block's K.r range and the 10 ns simulation time)
    import gzip
    with gzip.open(os.path.join(D, "rcs_plate.rcs.tout_ref.gz"), "rt") as f:
        records = [line.split() for line in f if line.strip()]
    assert [int(r[0]) for r in records] == list(range(1, len(records) + 1))
    assert rcs["points_per_cell"] == 2 and rcs["n_points"] == len(records) == 659


def test_vmat_filters():
    m = ingest.read_vmat(os.path.join(MAT, "epsr=2.vmat"))
    assert (m.eps.wnorm, list(m.eps.a), list(m.eps.b)) == (1e8, [2.0], [1.0])
    assert (list(m.mu.a), list(m.mu.b), m.sigma_e, m.sigma_m) == ([1.0], [1.0], 0.0, 0.0)
    lor = ingest.read_vmat(os.path.join(MAT, "mur_lorentz.vmat"))
    want = problems.material_test2().materials[0]
    for f, g in ((lor.eps, want.eps), (lor.mu, want.mu)):
        assert (f.wnorm, list(f.a), list(f.b)) == (g.wnorm, list(g.a), list(g.b))


def test_unknown_geometry_needs_the_mesh_file(tmp_path):
    p = tmp_path / "sphere.inp"
    p.write_text("Mesh_outer_boundary_dimension\n0 1 0 1 0 1\n\nMesh_dimensions_in_cells\n10 10 10\n\n"
                 "volume_list\n1\n1\nsphere\n0.3\n0 0 0\n0.5 0.5 0.5\n\n"
                 "volume_material_list\n1\n1\nPEC\n1\n1\n\nSimulation_time\n1e-9\n")
    deck = ingest.read_inp(str(p))
    assert deck.volumes[1].kind == "sphere"
    with pytest.raises(ingest.IngestError, match="mesh"):
        ingest.problem_from(deck)


@pytest.mark.gpu
def test_serving_takes_reference_input_files(oracle_lib):
    """Requests given as the reference's `.inp` files, served with continuous
    batching, in the plain form: the same bits as the oracle."""
    from oracle import oracle

    from ggi_tlm_amd import capi, serving

    reqs = [serving.request_from_files(0, os.path.join(D, "tem_pml.inp"), 0.0, steps=300),
            serving.request_from_files(1, os.path.join(D, "cavity_free_space.inp"), 0.0, steps=200),
            serving.request_from_files(2, os.path.join(D, "cavity_epsr2.inp"), 0.01, material_dir=MAT,
                                       steps=400)]
    st = serving.run_batched(reqs, capi.FORM_PLAIN_FP64, budget_bytes=4 << 30, chunk=32)
    print(f"[ingest serve] {st}", flush=True)
    for r in reqs:
        h = oracle.handle(r.tables, output_capacity=r.steps)
        h.step(1, r.steps)
        want = probe_series(h.read_cell_outputs()[0], r.tables)
        assert np.array_equal(r.outputs, want), r.rid
