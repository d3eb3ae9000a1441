"""Far-field and RCS surfaces (OUTPUT/far_field_surfaces.F90,
OUTPUT/RCS_surfaces.F90): the oracle against closed forms, and the engine
against the oracle.

Pins (oracle; each fails for a dropped or mis-signed J or M term, a swapped
theta/phi projection, a missing dt, time-derivative or normalisation
factor):
* 3D transform: a z-directed current element J dl^3 radiates
  |E_theta| = Z0 k |J(f)| dl^3 sin(theta) / (4 pi) at r = 1 m and no E_phi,
  for every phi (the Hertzian dipole); matched to 0.3% at 30 cells per
  wavelength.
* 2D transform: a z-directed line current I = J dl^2 (one cell between PEC
  walls in z) radiates |E| = Z0 |I(f)| sqrt(k / (8 pi)) at r = 1 m in every
  direction (the large-argument Hankel function); matched to 0.3%.
* RCS: the monostatic RCS of a square PEC plate at normal incidence,
  computed from the time-domain far-field series with the reference's end
  transform, is physical optics' 4 pi A^2 / lambda^2 within the plate
  edges' ripple (0.6 dB at 1.3 wavelengths, 0.15 dB at 3); the
  cross-polarised field vanishes; with the plate removed, the currents of
  the total field on the closed surface radiate nothing back (extinction).
"""
import os

import numpy as np
import pytest

from ggi_tlm_amd import capi, far_field
from ggi_tlm_amd.capi import FORM_EXTENDED_FP32, FORM_PLAIN_FP64, FORM_PRECISION_FP32
from ggi_tlm_amd.constants import C0, FACE_XMAX, FACE_XMIN, FACE_YMAX, FACE_YMIN, FACE_ZMAX, FACE_ZMIN, Z0
from ggi_tlm_amd.problem import (PML, Box, ExcitationFunction, FarFieldSurface, HuygensBox, PointSource, Problem,
                                 RCSSurface, build_tables)


def box_faces(i0, i1, j0, j1, k0, k1):
    """A closed box surface around cells i0..i1 x j0..j1 x k0..k1 as the
    model builder meshes a rectangular block: faces of the outside cells,
    facing the box (normal into the listed cell: outward)."""
    f = []
    for k in range(k0, k1 + 1):
        for j in range(j0, j1 + 1):
            f += [(i0 - 1, j, k, FACE_XMAX), (i1 + 1, j, k, FACE_XMIN)]
    for k in range(k0, k1 + 1):
        for i in range(i0, i1 + 1):
            f += [(i, j0 - 1, k, FACE_YMAX), (i, j1 + 1, k, FACE_YMIN)]
    for j in range(j0, j1 + 1):
        for i in range(i0, i1 + 1):
            f += [(i, j, k0 - 1, FACE_ZMAX), (i, j, k1 + 1, FACE_ZMIN)]
    return f


def spectrum(t, steps, f):
    """Fourier integral of excitation function 1 at the cell-source times
    (t-1) dt."""
    v = np.asarray(t.value)[:steps, 0]
    return np.sum(v * t.dt * np.exp(-2j * np.pi * f * np.arange(steps) * t.dt))


def dipole(n=31, steps=1000, f=1.0e9):
    """A soft Jz current element at the centre of a PML-bounded free-space
    cube, a zero-mean pulse (a Gaussian-modulated sine: no static charge
    left behind), and a closed far-field box 8 cells in from the walls."""
    c = (n + 1) // 2
    return Problem(nx=n, ny=n, nz=n, dl=0.01, n_timesteps=steps, reflection=(0.0,) * 6,
                   pml=PML((6,) * 6, (1e-4,) * 6, 2),
                   functions=[ExcitationFunction("gaussian_sinusoid", (1.0, 2.0 / f, 8.0 / f, f, 0.0))],
                   sources=[PointSource((c, c, c), "jz", 1)],
                   far_field_surfaces=[FarFieldSurface(box_faces(9, n - 8, 9, n - 8, 9, n - 8), f,
                                                       theta=(0.0, 180.0, 15.0), phi=(0.0, 90.0, 45.0))])


def line_source(n=61, steps=1200, f=1.0e9):
    """A Jz source in a one-cell-thick mesh between PEC z walls: an
    infinite line current; the far-field surface is the x and y walls of a
    box (the 2D transform)."""
    c, b = (n + 1) // 2, 10
    faces = []
    for j in range(b + 1, n - b + 1):
        faces += [(b, j, 1, FACE_XMAX), (n - b + 1, j, 1, FACE_XMIN)]
    for i in range(b + 1, n - b + 1):
        faces += [(i, b, 1, FACE_YMAX), (i, n - b + 1, 1, FACE_YMIN)]
    return Problem(nx=n, ny=n, nz=1, dl=0.01, n_timesteps=steps, reflection=(0.0, 0.0, 0.0, 0.0, -1.0, -1.0),
                   pml=PML((8, 8, 8, 8, 0, 0), (1e-4,) * 6, 2),
                   functions=[ExcitationFunction("gaussian_sinusoid", (1.0, 2.0 / f, 8.0 / f, f, 0.0))],
                   sources=[PointSource((c, c, 1), "jz", 1)],
                   far_field_surfaces=[FarFieldSurface(faces, f, phi=(0.0, 180.0, 30.0), dim=2)])


def plate(dl=0.005, steps=600, with_plate=True, length=0.2):
    """A square PEC plate (length x length, in a z plane) lit at normal
    incidence by an x-polarised plane wave travelling +z from a Huygens box;
    the RCS surface is a closed box between the plate and the Huygens box,
    looking back along -z (theta = 180 degrees).  The mesh is centred on the
    origin: the reference sizes the far-field series from the surface's
    K.r range as if it were (initialise_rcs_surfaces), and drops what falls
    outside."""
    npl = int(round(length / dl))
    m, g = 8, 3
    nxy, nz = npl + 2 * (m + 2 * g + 2), 2 * (m + 2 * g + 6)
    p0, kp = m + 2 * g + 3, nz // 2
    faces = np.array([[i, j, kp, FACE_ZMAX] for i in range(p0, p0 + npl) for j in range(p0, p0 + npl)])
    rb = (m + 3 + g, nxy - m - 2 - g, m + 3 + g, nxy - m - 2 - g, m + 3 + g, nz - m - 2 - g)
    return Problem(nx=nxy, ny=nxy, nz=nz, dl=dl, n_timesteps=steps, reflection=(0.0,) * 6,
                   mesh_min=(-nxy * dl / 2, -nxy * dl / 2, -nz * dl / 2), pml=PML((m,) * 6, (1e-4,) * 6, 2),
                   surface_materials=["pec"], surface_faces=[(faces, 1)] if with_plate else [],
                   functions=[ExcitationFunction("gaussian", (1.0, 1.0e-10, 4.0e-10))],
                   huygens=HuygensBox(Box(m + 3, nxy - m - 2, m + 3, nxy - m - 2, m + 3, nz - m - 2), 1,
                                      0.0, 0.0, 1.0, 0.0, side=1),
                   rcs_surface=RCSSurface(box_faces(*rb), 1.5e9, 4.5e9, 0.5e9, theta=180.0, phi=0.0))


def run_oracle(t, steps):
    from oracle import oracle

    h = oracle.handle(t, output_capacity=8)
    h.step(1, steps)
    return h


def test_dipole_far_field_is_the_hertzian_dipole(oracle_lib):
    f, steps = 1.0e9, 1000
    t = build_tables(dipole(steps=steps, f=f))
    e = t.far_fields[0]
    assert (e["n_theta"], e["n_phi"]) == (13, 3)
    Et, Ep = run_oracle(t, steps).read_far_field(0)
    k = 2 * np.pi * f / C0
    th, _ = far_field.far_field_angles(e)
    want = Z0 * k * abs(spectrum(t, steps, f)) * 0.01 ** 3 * np.sin(th) / (4 * np.pi)
    inner = slice(1, -1)      # theta = 0 and 180 degrees: no radiation
    ratio = np.abs(Et[inner]) / want[inner, None]
    assert np.all(np.abs(ratio - 1.0) < 3e-3), ratio
    assert np.abs(Et[[0, -1]]).max() < 1e-6 * np.abs(Et).max()
    assert np.abs(Ep).max() < 1e-6 * np.abs(Et).max()


def test_line_current_far_field_is_the_2d_solution(oracle_lib):
    f, steps = 1.0e9, 1200
    t = build_tables(line_source(steps=steps, f=f))
    e = t.far_fields[0]
    assert (e["n_theta"], e["n_phi"], e["dim"]) == (1, 7, 2)
    Et, Ep = run_oracle(t, steps).read_far_field(0)
    k = 2 * np.pi * f / C0
    want = Z0 * abs(spectrum(t, steps, f)) * 0.01 ** 2 * np.sqrt(k / (8 * np.pi))
    assert np.all(np.abs(np.abs(Et) / want - 1.0) < 3e-3), np.abs(Et) / want
    assert np.abs(Ep).max() < 1e-6 * np.abs(Et).max()


@pytest.fixture(scope="module")
def plate_runs(oracle_lib):
    out = {}
    for with_plate in (True, False):
        t = build_tables(plate(with_plate=with_plate))
        out[with_plate] = (t, run_oracle(t, t.n_timesteps).read_rcs())
    return out


def test_plate_rcs_is_physical_optics(plate_runs, tmp_path):
    t, series = plate_runs[True]
    fr, et, ep, einc, db = far_field.rcs_spectrum(series, t.rcs, np.asarray(t.value)[:, 0], t.dt, t.dl)
    po = 10 * np.log10(4 * np.pi * 0.2 ** 4 / (C0 / fr) ** 2)
    assert np.all(np.abs(db - po) < 0.7), db - po
    assert np.all(np.abs(db[fr >= 3.5e9] - po[fr >= 3.5e9]) < 0.3), db - po
    assert np.abs(ep).max() < 1e-10 * np.abs(et).max()
    # the reference's two files
    far_field.write_rcs(str(tmp_path / "plate"), series, t.rcs, np.asarray(t.value)[:, 0], t.dt, t.dl)
    tout = open(tmp_path / "plate.rcs.tout").read().splitlines()
    fout = open(tmp_path / "plate.rcs.fout").read().splitlines()
    assert len(tout) == t.rcs["n_points"] and all(len(x) == 10 + 5 * 14 for x in tout)
    assert len(fout) == len(fr) and all(len(x) == 5 * 14 for x in fout)
    assert abs(float(fout[3].split()[4]) - db[3]) < 1e-5 * max(1.0, abs(db[3]))


def test_empty_surface_radiates_nothing_back(plate_runs):
    """Extinction: the incident plane wave's own currents on the closed
    surface cancel in the monostatic direction (J and M of each face, and
    front against back face)."""
    s_plate, s_empty = plate_runs[True][1], plate_runs[False][1]
    assert np.abs(s_plate[:, 0]).max() > 0
    assert np.abs(s_empty).max() < 1e-12 * np.abs(s_plate[:, 0]).max()


def test_plate_deck_against_the_reference_rcs_files(oracle_lib):
    """TEST_DATA/RCS_PEC_PLATE_ROUGH_SURFACE through the ingest, with its
    plate made flat (the reference's plate is rough: random surface points
    of +-2 cm, which only its model builder's run fixes), against that
    run's plate.rcs.fout: the incident spectrum |E_inc(f)| of its Huygens
    pulse equals the reference's to the file's 6 digits, and the flat
    plate's monostatic RCS is within 1.3 dB of the rough plate's up to
    2.5 GHz (wavelengths of 6 roughness amplitudes and more; above, the
    roughness scatters and the two part by up to 1.9 dB)."""
    from ggi_tlm_amd import ingest

    d = os.path.join(os.path.dirname(__file__), "golden", "inputs")
    raw = open(os.path.join(d, "rcs_plate.inp"), "rb").read().replace(b"surface_roughness \n0.02 0.025\n", b"")
    deck = ingest.read_inp_text(raw.decode("latin-1"), "plate.inp")
    t = build_tables(ingest.problem_from(deck).problem)
    s = run_oracle(t, t.n_timesteps).read_rcs()
    fr, et, ep, einc, db = far_field.rcs_spectrum(s, t.rcs, np.asarray(t.value)[:, 0], t.dt, t.dl)
    gold = np.loadtxt(os.path.join(d, "rcs_plate.rcs.fout_ref"))
    assert np.allclose(fr, gold[:, 0], rtol=1e-6, atol=0)
    assert np.abs(np.abs(einc) - gold[:, 3]).max() <= 5e-6 * gold[:, 3].max()
    low = fr <= 2.5e9
    assert np.abs(db[low] - gold[low, 4]).max() < 1.3, db - gold[:, 4]


def test_rcs_points_hold_every_contribution():
    """initialise_rcs_surfaces sizes the far-field series (n_points,
    point_offset) so that every face's retarded-time point of every step
    falls inside it: checked over the whole run, with both ends reached to
    within the sizing's two-point margins."""
    t = build_tables(plate(with_plate=False))
    r = t.rcs
    th, ph = r["theta"], r["phi"]
    K = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    d = r["centres"] @ K
    times = (np.arange(1, t.n_timesteps + 1) - 1) * t.dt + t.dt / 2.0
    pts = np.trunc(r["points_per_cell"] * (-d[None, :] + times[:, None] * C0) / t.dl).astype(int) - r["point_offset"]
    assert pts.min() >= 1 and pts.max() <= r["n_points"]
    assert pts.min() <= 3 and pts.max() >= r["n_points"] - 4


def test_far_field_file_format(oracle_lib, tmp_path):
    t = build_tables(dipole(n=21, steps=40))
    e = t.far_fields[0]
    h = run_oracle(t, 40)
    Et, Ep = h.read_far_field(0)
    far_field.write_far_field(str(tmp_path / "ff.fout"), e, Et, Ep, t.dt)
    lines = open(tmp_path / "ff.fout").read().splitlines()
    assert lines[0][:2] == "# " and lines[0][2:14].split() == ["13", "3"]
    assert abs(float(lines[0][14:]) - 1.0 / t.dt) <= 1e-6 / t.dt
    assert len(lines) == 1 + 13 * 3 and all(len(x) == 4 * 14 for x in lines[1:])
    assert [float(v) for v in lines[2].split()[:2]] == [0.0, 45.0]


def test_bad_far_field_rows_are_refused(oracle_lib):
    from oracle import oracle

    t = build_tables(dipole(n=21, steps=10))
    t.far_fields[0]["rows"] = t.far_fields[0]["rows"] + 100000
    with pytest.raises(capi.TlmError):
        oracle.handle(t, output_capacity=8)


@pytest.mark.gpu
@pytest.mark.parametrize("form", [FORM_PLAIN_FP64, FORM_PRECISION_FP32, FORM_EXTENDED_FP32])
def test_engine_far_field_and_rcs_vs_oracle(oracle_lib, form):
    """The device's accumulation and transform against the oracle's
    sequential restatement: the face sums run in another order and sin/cos
    differ in the last bits, so the plain form agrees to 1e-12 of the
    pattern's or series' largest value."""
    steps = 300
    p = plate(dl=0.01, steps=steps)
    n = p.nx
    p.far_field_surfaces = [FarFieldSurface(box_faces(12, n - 11, 12, n - 11, 12, p.nz - 11), 2.0e9,
                                            theta=(0.0, 180.0, 10.0), phi=(0.0, 350.0, 10.0)),
                            FarFieldSurface(box_faces(12, n - 11, 12, n - 11, 12, p.nz - 11), 3.0e9,
                                            theta=(30.0, 60.0, 15.0), phi=(0.0, 90.0, 30.0), side=-1)]
    t = build_tables(p)
    h = run_oracle(t, steps)
    g = capi.Handle(capi.engine_library(), t, form, output_capacity=8)
    print(f"[far field] form={form} faces={len(t.far_fields[0]['rows'])} points={t.rcs['n_points']}", flush=True)
    g.step(1, steps)
    # measured: 1.4e-14 (plain), 6.5e-8 (precision), 1.9e-10 (extended)
    tol = {FORM_PLAIN_FP64: 1e-12, FORM_PRECISION_FP32: 1e-6, FORM_EXTENDED_FP32: 1e-8}[form]
    for k in range(2):
        for want, got in zip(h.read_far_field(k), g.read_far_field(k)):
            err = np.abs(got - want).max() / np.abs(want).max()
            print(f"[far field] surface {k}: {err:.2e}", flush=True)
            assert err <= tol, k
    want, got = h.read_rcs(), g.read_rcs()
    # E_theta, E_phi against the largest E; H_theta, H_phi against the
    # largest H (the cross-polarised columns are rounding noise)
    for cols in ((0, 1), (2, 3)):
        err = np.abs(got[:, cols] - want[:, cols]).max() / np.abs(want[:, cols]).max()
        print(f"[far field] rcs columns {cols}: {err:.2e}", flush=True)
        assert err <= tol, cols
