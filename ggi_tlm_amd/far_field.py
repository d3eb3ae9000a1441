"""Host post-processing of the device's far-field and RCS accumulations, and
their output files in the reference's formats.

* Far-field surfaces: the device transforms the accumulated surface
  currents itself (tlm_read_far_field); write_far_field writes the result as
  write_far_field_surfaces does (OUTPUT/far_field_surfaces.F90:641-700).
* RCS surface: the device accumulates the time-domain far field in the
  monostatic direction (tlm_read_rcs); rcs_spectrum restates the end-of-run
  transform of write_rcs_surfaces (OUTPUT/RCS_surfaces.F90:575-762): the
  incident field's spectrum from the Huygens excitation function, the
  Fourier transform of the far-field series with its time derivative, and
  RCS = 10 log10(4 pi Re(E_theta H_phi* - E_phi H_theta*) / (|E_inc|^2 / Z0)).
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np

from .constants import C0, PI, Z0
from .tout import fortran_e


def fortran_es(x: float, w: int = 16, d: int = 6) -> str:
    """Fortran ESw.d edit descriptor: d.dddddd E+ee, right-justified."""
    if x == 0.0:
        return f"0.{'0' * d}E+00".rjust(w)
    s = f"{x:.{d}E}"
    mant, exp = s.split("E")
    e = int(exp)
    return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}".rjust(w)


def far_field_angles(entry: dict):
    """Theta and phi (radians) of the rows of tlm_read_far_field, as the
    transform steps them."""
    th = entry["theta_min"] + np.arange(entry["n_theta"]) * entry["theta_step"]
    ph = entry["phi_min"] + np.arange(entry["n_phi"]) * entry["phi_step"]
    return th, ph


def write_far_field(path: str, entry: dict, e_theta: np.ndarray, e_phi: np.ndarray, dt: float):
    """One far-field surface's file: for the 3D transform a header
    '# n_theta n_phi 1/dt' (format (A2,2I6,ES16.6)), then per angle
    theta, phi (degrees), |E_theta|, |E_phi| (format 4E14.6); the 2D
    transform writes the rows only."""
    th, ph = far_field_angles(entry)
    lines = []
    if entry["dim"] == 3:
        lines.append("# " + f"{entry['n_theta']:6d}{entry['n_phi']:6d}" + fortran_es(1.0 / dt))
    for i, t in enumerate(th):
        for j, p in enumerate(ph):
            lines.append("".join(fortran_e(v, 14, 6) for v in (t * 180.0 / PI, p * 180.0 / PI,
                                                               abs(e_theta[i, j]), abs(e_phi[i, j]))))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def incident_spectrum(values: Sequence[float], dt: float, frequency: float) -> complex:
    """E_inc(f) = sum_t value(t) dt exp(-j 2 pi f (t-1) dt) over the
    excitation function's timesteps (RCS_surfaces.F90:690-700)."""
    v = np.asarray(values, dtype=np.float64)
    t = np.arange(len(v)) * dt
    return complex(np.sum(v * (dt * np.exp(-1j * 2.0 * PI * frequency * t))))


def rcs_spectrum(series: np.ndarray, rcs: dict, values: Sequence[float], dt: float, dl: float):
    """RCS spectrum of the far-field series (n_points, 4: E_theta, E_phi,
    H_theta, H_phi; point p at time p dl / (points_per_cell c0)).
    Returns frequencies, E_theta(f), E_phi(f), E_inc(f) and the RCS in
    dB m^2 (-1000 where the scattered power is not positive)."""
    n_f = int((rcs["fmax"] - rcs["fmin"]) / rcs["fstep"]) + 1
    freqs = rcs["fmin"] + np.arange(n_f) * rcs["fstep"]
    ppc = rcs["points_per_cell"]
    dt_local = dl / (ppc * C0)
    d = np.arange(1, series.shape[0] + 1)
    t_local = d * dl / (ppc * C0)
    out_et, out_ep, out_inc, out_rcs = [], [], [], []
    for f in freqs:
        einc = incident_spectrum(values, dt, f)
        k = dt_local * np.exp(-1j * 2.0 * PI * f * t_local)
        et, ep, ht, hp = (np.sum(series[:, c] * k) * (1j * 2.0 * PI * f / C0) for c in range(4))
        rp = (et * np.conj(hp) - ep * np.conj(ht)).real
        ip = abs(einc) ** 2 / Z0
        out_rcs.append(10.0 * math.log10(4.0 * PI * rp / ip) if (rp > 0.0 and ip != 0.0) else -1000.0)
        out_et.append(et)
        out_ep.append(ep)
        out_inc.append(einc)
    return freqs, np.array(out_et), np.array(out_ep), np.array(out_inc), np.array(out_rcs)


def write_rcs(prefix: str, series: np.ndarray, rcs: dict, values: Sequence[float], dt: float, dl: float):
    """The reference's two RCS files: `prefix.rcs.tout` (per far-field
    point: point, time, E_theta, E_phi, H_theta, H_phi; format (I10,5E14.6),
    |v| < 1e-30 printed as 0) and `prefix.rcs.fout` (per frequency:
    frequency, |E_theta|, |E_phi|, |E_inc|, RCS; format 5E14.6)."""
    ppc = rcs["points_per_cell"]
    lines = []
    for p in range(series.shape[0]):
        vals = [0.0 if abs(v) < 1e-30 else float(v) for v in series[p]]
        t = (p + 1) * dl / (ppc * C0)
        lines.append(f"{p + 1:10d}" + "".join(fortran_e(v, 14, 6) for v in [t] + vals))
    with open(prefix + ".rcs.tout", "w") as f:
        f.write("\n".join(lines) + "\n")
    freqs, et, ep, einc, db = rcs_spectrum(series, rcs, values, dt, dl)
    with open(prefix + ".rcs.fout", "w") as f:
        for i, fr in enumerate(freqs):
            f.write("".join(fortran_e(v, 14, 6) for v in (fr, abs(et[i]), abs(ep[i]), abs(einc[i]), db[i])) + "\n")
