"""Time-domain surface and volume field outputs (output_surfaces.F90:601-948,
output_volumes.F90:496-761): the oracle's frames against the series it
records independently, the files, and the engine against the oracle.

Pins (oracle):
* every frame is the due step's value of the recorded cell or face output
  series (window and interval, get_output_flag): the field, J = n x H with
  the face's outward normal, |E|, |H| and |J| as root sums of squares of the
  recorded fields;
* side 0 (both sides) takes J as the jump n x (H1 - H2) across the face:
  zero on ordinary faces (the two sides agree to the connect's rounding),
  and across a PEC plate the sum of the two one-sided surface currents
  (sides +1 and -1, each with its own outward normal).
"""
import numpy as np
import pytest

import problems
from ggi_tlm_amd import capi, field_outputs
from ggi_tlm_amd.capi import FORM_EXTENDED_FP32, FORM_PLAIN_FP64, FORM_PRECISION_FP32
from ggi_tlm_amd.constants import FACE_XMIN, FACE_YMAX, FACE_ZMAX, FACE_ZMIN
from ggi_tlm_amd.problem import Box, FieldOutput, build_tables

STEPS = 90


def deck():
    """pml_sphere with volume and surface field outputs (every component
    kind, windows with intervals, both-sides outputs on free-space faces and
    on a PEC plate)."""
    p = problems.pml_sphere(24, STEPS)
    faces = [(12, j, k, FACE_XMIN) for j in range(9, 13) for k in range(10, 13)] + \
            [(i, 14, 11, FACE_YMAX) for i in range(8, 12)] + [(10, 10, 15, FACE_ZMIN), (11, 10, 15, FACE_ZMIN)]
    plate = [(i, j, 8, FACE_ZMAX) for i in range(10, 15) for j in range(10, 15)]
    p.surface_materials = ["pec"]
    p.surface_faces = [(np.array(plate), 1)]
    p.field_outputs = [
        FieldOutput("volume", "ez", region=Box(11, 13, 11, 13, 11, 13)),
        FieldOutput("volume", "hy", region=Box(8, 9, 10, 11, 12, 12), first_timestep=7, last_timestep=70,
                    timestep_interval=3),
        FieldOutput("surface", "jx", faces=faces),
        FieldOutput("surface", "ey", faces=faces, side=-1, first_timestep=0, timestep_interval=5),
        FieldOutput("surface", "em", faces=faces),
        FieldOutput("surface", "hm", faces=faces, first_timestep=10, last_timestep=60, timestep_interval=2),
        FieldOutput("surface", "jm", faces=faces),
        FieldOutput("surface", "jy", faces=faces, side=0),
        FieldOutput("surface", "jx", faces=plate, side=0),
        FieldOutput("surface", "jx", faces=plate, side=-1),
        FieldOutput("surface", "jz", faces=faces, side=1),
        FieldOutput("surface", "jx", faces=plate, side=1),
    ]
    return p


def run_oracle(t, cap=STEPS):
    from oracle import oracle

    h = oracle.handle(t, output_capacity=cap)
    h.step(1, STEPS)
    return h


def due(e, n):
    return e["first_timestep"] <= n <= e["last_timestep"] and (n - e["first_timestep"]) % e["timestep_interval"] == 0


def expected(e, cells, faces):
    """The output's frames from the recorded series."""
    out = []
    for n in range(1, STEPS + 1):
        if not due(e, n):
            continue
        if e["kind"] == 1:
            out.append(cells[n - 1, e["rows"] - 1, e["component"] - 1])
            continue
        f1 = faces[n - 1, e["rows"] - 1, 0, :]
        f2 = faces[n - 1, e["rows"] - 1, 1, :]
        own = faces[n - 1, e["rows"] - 1, e["sides"] - 1, :]
        fld = f1 if e["both_sides"] else own
        nrm = np.zeros((len(e["rows"]), 3))
        nrm[np.arange(len(e["rows"])), np.abs(e["normals"]) - 1] = np.sign(e["normals"])
        J = np.cross(nrm, fld[:, 3:6])
        if e["both_sides"]:
            J = J - np.cross(nrm, f2[:, 3:6])
        c = e["component"]
        if c <= 6:
            v = fld[:, c - 1]
        elif c <= 9:
            v = J[:, c - 7]
        elif c == 13:
            v = np.sqrt(np.sum(fld[:, 0:3] ** 2, axis=1))
        elif c == 14:
            v = np.sqrt(np.sum(fld[:, 3:6] ** 2, axis=1))
        else:
            v = np.sqrt(np.sum(J ** 2, axis=1))
        out.append(v)
    return np.array(out)


def first_due(e):
    return next(n for n in range(1, STEPS + 1) if due(e, n))


def test_frames_are_the_recorded_series(oracle_lib):
    t = build_tables(deck())
    h = run_oracle(t)
    cells, _ = h.read_cell_outputs()
    faces, _ = h.read_face_outputs()
    for k, e in enumerate(t.field_outputs):
        got, first = h.read_field_output(k)
        want = expected(e, cells, faces)
        assert got.shape == want.shape and first == first_due(e), k
        assert np.abs(want).max() > 0, k
        assert np.max(np.abs(got - want)) <= 1e-13 * np.abs(want).max(), k
        # consumed
        again, _ = h.read_field_output(k)
        assert again.shape[0] == 0


def test_both_sides_is_the_jump_across_the_face(oracle_lib):
    t = build_tables(deck())
    h = run_oracle(t)
    one_side = h.read_field_output(2)[0]        # jx on free-space faces
    jump = h.read_field_output(7)[0]            # jy, both sides, same faces
    assert np.abs(jump).max() <= 1e-12 * np.abs(one_side).max()
    # across the PEC plate: the jump is the sum of the two one-sided
    # currents, each with its own side's outward normal
    plate_jump = h.read_field_output(8)[0]      # jx, both sides
    below = h.read_field_output(9)[0]           # jx, side -1: the cells above the faces' own, normal +z
    above = h.read_field_output(11)[0]          # jx, side +1: the faces' own cells, normal -z
    assert np.abs(plate_jump).max() > 0 and np.abs(below).max() > 0 and np.abs(above).max() > 0
    assert np.max(np.abs(plate_jump - (below + above))) <= 1e-13 * np.abs(plate_jump).max()


def test_capacity_drains_keep_every_frame(oracle_lib):
    """A ring of a few frames on the engine's side drains as it fills; on the
    oracle the same frames come back whatever the capacity."""
    t = build_tables(deck())
    a, b = run_oracle(t), run_oracle(t, cap=4)
    for k in range(len(t.field_outputs)):
        assert np.array_equal(a.read_field_output(k)[0], b.read_field_output(k)[0])


def test_bad_outputs_are_refused(oracle_lib):
    from oracle import oracle

    t = build_tables(deck())
    t.field_outputs[0]["rows"] = t.field_outputs[0]["rows"] + 10000
    with pytest.raises(capi.TlmError):
        oracle.handle(t, output_capacity=8)
    p = deck()
    p.field_outputs = [FieldOutput("surface", "ex", faces=p.field_outputs[2].faces, side=0)]
    with pytest.raises(ValueError):
        build_tables(p)


def test_files(oracle_lib, tmp_path):
    p = deck()
    t = build_tables(p)
    h = run_oracle(t)
    surf = [e for e in t.field_outputs if e["kind"] == 2]
    frames = [h.read_field_output(k) for k, e in enumerate(t.field_outputs) if e["kind"] == 2]
    path = str(tmp_path / "deck.surface_field.tout")
    field_outputs.write_field_outputs(path, "surface", surf, frames, t.n_timesteps, p.mesh_min, p.dl)
    lines = open(path).read().splitlines()
    assert lines[0] == "# NUMBER OF OUTPUT SURFACES" and lines[1] == " " + f"{len(surf):12d}"
    # the first template: points, faces, frames, then 4 corners per face
    n0 = len(surf[0]["rows"])
    assert lines[2].endswith("OUTPUT SURFACE NUMBER:    1")
    assert int(lines[4]) == 4 * n0 and int(lines[6]) == n0 and int(lines[8]) == STEPS
    # the first face (12, 9, 10) xmin: its first corner is the cell's low corner
    x, y, z = (float(v) for v in lines[9].split())
    dl = p.dl
    assert np.allclose((x, y, z), (11 * dl, 8 * dl, 9 * dl), rtol=1e-5)
    assert lines[9 + 4 * n0].split() == ["4", "0", "1", "2", "3"]
    # frames: as many blocks as due steps over all outputs, in timestep order
    starts = [i for i, s in enumerate(lines) if s == "# START OF SURFACE FIELD OUTPUT DATA"]
    assert len(starts) == sum(f[0].shape[0] for f in frames)
    first = starts[0]
    assert int(lines[first + 2]) == 1 and int(lines[first + 4]) == 1 and int(lines[first + 6]) == n0
    assert abs(float(lines[first + 8]) - frames[0][0][0, 0]) <= 1e-5 * abs(frames[0][0]).max()
    vol = [e for e in t.field_outputs if e["kind"] == 1]
    vframes = [h.read_field_output(k) for k, e in enumerate(t.field_outputs) if e["kind"] == 1]
    vpath = str(tmp_path / "deck.volume_field.tout")
    field_outputs.write_field_outputs(vpath, "volume", vol, vframes, t.n_timesteps, p.mesh_min, p.dl)
    v = open(vpath).read().splitlines()
    assert int(v[4]) == 24 * len(vol[0]["rows"]) and int(v[6]) == 6 * len(vol[0]["rows"])


@pytest.mark.gpu
@pytest.mark.parametrize("form", [FORM_PLAIN_FP64, FORM_PRECISION_FP32, FORM_EXTENDED_FP32])
def test_engine_field_outputs_vs_oracle(oracle_lib, form):
    """The engine's frames against the oracle's: the same bits in the plain
    form (a copy or a root sum of squares of bitwise face fields), with an
    output capacity of 8 frames so that the device rings drain mid-run."""
    t = build_tables(deck())
    h = run_oracle(t)
    g = capi.Handle(capi.engine_library(), t, form, output_capacity=8)
    print(f"[field outputs] form={form}", flush=True)
    g.step(1, STEPS)
    tol = {FORM_PLAIN_FP64: 0.0, FORM_PRECISION_FP32: 1e-5, FORM_EXTENDED_FP32: 1e-7}[form]
    want = [h.read_field_output(k) for k in range(len(t.field_outputs))]
    for k, e in enumerate(t.field_outputs):
        got, gf = g.read_field_output(k)
        assert got.shape == want[k][0].shape and gf == want[k][1], k
        if tol == 0.0:
            assert np.array_equal(got, want[k][0]), k
            continue
        # the jump across free-space faces is rounding: measure it against
        # the one-sided current of the same faces
        scale = np.abs(want[2 if k == 7 else k][0]).max()
        assert np.max(np.abs(got - want[k][0])) <= tol * scale, k
