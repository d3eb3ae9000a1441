"""Probe time-series files in the reference's .tout text format.

write_time_domain_header_data (FILE_HANDLING/read_write_header_data.F90:
37-57) writes three header lines; cell_output_point (OUTPUT/
output_points.F90:545-597) then writes one record per output point per
timestep in format (E16.7,I5,E16.7) (File_information.F90:235), printing
values with abs(v) < 1e-30 as 0.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np


def fortran_e(x: float, w: int = 16, d: int = 7) -> str:
    """Fortran Ew.d edit descriptor: 0.ddddddd E+ee, right-justified."""
    if x == 0.0:
        return f"0.{'0' * d}E+00".rjust(w)
    if not math.isfinite(x):
        return ("NaN" if math.isnan(x) else ("Infinity" if x > 0 else "-Infinity")).rjust(w)
    s = f"{abs(x):.{d - 1}e}"          # d significant digits: m.mmmmmmE+xx
    mant, exp = s.split("e")
    digits = mant.replace(".", "")
    e = int(exp) + 1
    if abs(e) <= 99:
        es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:                                # three exponent digits drop the E
        es = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    body = ("-" if x < 0 else "") + "0." + digits + es
    return body.rjust(w)


def header(n_points: int, n_timesteps: int) -> str:
    return ("# Time Domain Data\n"
            f"# Number of data points: {n_points:10d}\n"
            f"# Number of Timesteps  : {n_timesteps:10d}\n")


def write_time_domain(path: str, times: Sequence[float], series: np.ndarray, n_timesteps: int = None):
    """series: (steps, n_points) probe values at the given times."""
    series = np.asarray(series, dtype=np.float64)
    if series.ndim == 1:
        series = series[:, None]
    steps, npts = series.shape
    with open(path, "w") as f:
        f.write(header(npts, steps if n_timesteps is None else n_timesteps))
        for n in range(steps):
            t = fortran_e(float(times[n]))
            for p in range(npts):
                v = float(series[n, p])
                if abs(v) < 1e-30:
                    v = 0.0
                f.write(f"{t}{p + 1:5d}{fortran_e(v)}\n")


def read_time_domain(path: str):
    """(times, series (steps, n_points)) of a .tout file."""
    data = np.loadtxt(path, comments="#")
    if data.ndim == 1:
        data = data[None, :]
    npts = int(data[:, 1].max())
    steps = data.shape[0] // npts
    d = data[: steps * npts].reshape(steps, npts, 3)
    return d[:, 0, 0], d[:, :, 2]
