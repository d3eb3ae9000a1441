"""Host-side material filter set-up: s-domain material models to the
z-domain stub filters the scatter consumes.

Follows SRC/MATERIALS/TLM_volume_materials.F90:190-385
(calculate_volume_material_filter_coefficients):

  1. stub impedance S-filter
     eps:  Z(s) = 1 / (s K (eps_r(s) - 1)),  K = eps0 * reduced_c * dl
           (calculate_electric_susceptibility_impedance_Sfilter,
            SRC/FILTER/filter_subroutines.F90:1379-1431)
     mu:   Z(s) = s K (mu_r(s) - 1),          K = mu0 * reduced_c * dl
           (calculate_magnetic_susceptibility_impedance_Sfilter, :1310-1370)
  2. bilinear s -> z transform with T = dt / frequency_scale
     (s_to_z_warp, SRC/FILTER/filter_functions.F90:343-389;
      bilinear_s_to_z, filter_subroutines.F90:601-886), normalised so b0 = 1
  3. fast/slow split: f_fast = a0/b0, slow a0 = 0, a_i -= b_i f_fast
     (Z_fast_slow_decomposition, filter_subroutines.F90:1145-1199)

The reference builds step 2 from polynomial roots (EISPACK); this module
expands the substitution s_n = (2/(T wnorm)) (1 - z^-1)/(1 + z^-1) directly.
Both are the same rational function; the coefficients agree to rounding
(checked against the frequency-response identity in the tests).  The
coefficients are inputs of the hot path, shared by the oracle and the
device, so this choice does not affect hot-path parity.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .constants import EPS0, MU0


@dataclass
class Sfilter:
    """Rational s-domain function a(s_n)/b(s_n) in s_n = s / wnorm
    (SRC/FILTER/filter_types.F90:37-43; read_Sfilter,
    filter_subroutines.F90:54-78)."""

    wnorm: float = 1.0
    a: list = field(default_factory=lambda: [1.0])
    b: list = field(default_factory=lambda: [1.0])

    @staticmethod
    def constant(value: float) -> "Sfilter":
        return Sfilter(1.0, [float(value)], [1.0])


@dataclass
class Zfilter:
    """Polynomials in z^-1 (filter_types.F90:45-52)."""

    a: np.ndarray
    b: np.ndarray

    @property
    def order(self) -> int:
        return len(self.a) - 1


def _trim(c):
    c = [float(x) for x in c]
    order = 0
    for i, x in enumerate(c):
        if x != 0.0:
            order = i
    return c[: order + 1]


def electric_stub_impedance(eps: Sfilter, K: float) -> Sfilter:
    """filter_subroutines.F90:1379-1431: b has order a.order+1 with b0 = 0,
    b_{i+1} = K wnorm (a_i - b_i); numerator = eps.b."""
    na, nb = len(eps.a) - 1, len(eps.b) - 1
    bden = [0.0] * (na + 2)
    for i in range(na + 1):
        bden[i + 1] = K * eps.wnorm * eps.a[i]
    for i in range(nb + 1):
        bden[i + 1] = bden[i + 1] - K * eps.wnorm * eps.b[i]
    return Sfilter(eps.wnorm, list(eps.b), _trim(bden))


def magnetic_stub_impedance(mu: Sfilter, K: float) -> Sfilter:
    """filter_subroutines.F90:1310-1370: numerator order a.order+1, a0 = 0,
    a_{i+1} = K wnorm (a_i - b_i); denominator = mu.b."""
    na, nb = len(mu.a) - 1, len(mu.b) - 1
    anum = [0.0] * (na + 2)
    for i in range(na + 1):
        anum[i + 1] = K * mu.wnorm * mu.a[i]
    for i in range(nb + 1):
        anum[i + 1] = anum[i + 1] - K * mu.wnorm * mu.b[i]
    return Sfilter(mu.wnorm, _trim(anum), list(mu.b))


def _poly_mul(p, q):
    return np.convolve(np.asarray(p, dtype=np.float64), np.asarray(q, dtype=np.float64))


def bilinear_s_to_z(s: Sfilter, T: float) -> Zfilter:
    """Substitute s_n = dt2 (1 - z^-1)/(1 + z^-1), dt2 = 2/(T wnorm), and
    multiply through by (1 + z^-1)^M, M = max order; normalise b0 = 1."""
    a = _trim(s.a)
    b = _trim(s.b)
    M = max(len(a), len(b)) - 1
    dt2 = 2.0 / (T * s.wnorm)
    one_m = np.array([1.0, -1.0])
    one_p = np.array([1.0, 1.0])

    def expand(c):
        out = np.zeros(M + 1)
        for i, ci in enumerate(c):
            if ci == 0.0:
                continue
            term = np.array([ci * dt2 ** i])
            for _ in range(i):
                term = _poly_mul(term, one_m)
            for _ in range(M - i):
                term = _poly_mul(term, one_p)
            out[: len(term)] += term
        return out

    za = expand(a)
    zb = expand(b)
    b0 = zb[0]
    if b0 != 0.0:
        za = za / b0
        zb = zb / b0
    return Zfilter(za, zb)


def fast_slow(z: Zfilter):
    """Z_fast_slow_decomposition: returns (f_fast, slow filter)."""
    if len(z.a) != len(z.b):
        raise ValueError("fast/slow split needs equal numerator/denominator order")
    f_fast = z.a[0] / z.b[0]
    a = z.a.copy()
    a[0] = 0.0
    for i in range(1, len(z.b)):
        a[i] = a[i] - z.b[i] * f_fast
    return f_fast, Zfilter(a, z.b.copy())


def is_free_space(s: Sfilter) -> bool:
    """TLM_volume_materials.F90:246-248: order-0 numerator equal to 1."""
    a = list(s.a)
    return len(a) == 1 and a[0] == 1.0


def volume_material_row(kind: str, eps: Sfilter, mu: Sfilter, sigma_e: float,
                        sigma_m: float, dl: float, dt: float,
                        reduced_c_factor: float = 1.0, frequency_scale: float = 1.0):
    """One volume_material_list row, as a dict of the fields the scatter
    reads (TLM_volume_materials.F90:213-380)."""
    from .constants import VMAT_DISPERSIVE, VMAT_PEC, VMAT_PMC

    row = dict(type=0, eps_filter_exists=0, mu_filter_exists=0, eps_order=0,
               mu_order=0, Ys_eps_f=0.0, Ge=0.0, Zs_mu_f=0.0, Rm=0.0,
               eps_a=[0.0], eps_b=[1.0], mu_a=[0.0], mu_b=[1.0])
    if kind == "pec":
        row["type"] = VMAT_PEC
        return row
    if kind == "pmc":
        row["type"] = VMAT_PMC
        return row
    if kind != "dispersive":
        raise ValueError(f"unknown volume material kind {kind!r}")
    row["type"] = VMAT_DISPERSIVE
    Ge = sigma_e * dl / reduced_c_factor
    if Ge < 0.0:
        if Ge < -1e-12:
            raise ValueError("negative electric loss")
        Ge = 0.0
    row["Ge"] = Ge
    T = dt / frequency_scale
    if not is_free_space(eps):
        K = (EPS0 * reduced_c_factor) * dl
        zs = bilinear_s_to_z(electric_stub_impedance(eps, K), T)
        f, slow = fast_slow(zs)
        row.update(eps_filter_exists=1, eps_order=slow.order, Ys_eps_f=1.0 / f,
                   eps_a=list(slow.a), eps_b=list(slow.b))
    Rm = sigma_m * dl
    if Rm < 0.0:
        if Rm < -1e-12:
            raise ValueError("negative magnetic loss")
        Rm = 0.0
    row["Rm"] = Rm
    if not is_free_space(mu):
        K = (MU0 * reduced_c_factor) * dl
        zs = bilinear_s_to_z(magnetic_stub_impedance(mu, K), T)
        f, slow = fast_slow(zs)
        row.update(mu_filter_exists=1, mu_order=slow.order, Zs_mu_f=f,
                   mu_a=list(slow.a), mu_b=list(slow.b))
    return row


def debye_lorentz_eps(eps_inf: float, poles) -> Sfilter:
    """Rational eps_r(s) = eps_inf + sum_k r_k / (1 + s/w_k) (Debye poles,
    given as ('debye', r_k, w_k)) or r_k w_k^2 / (s^2 + 2 d_k w_k s + w_k^2)
    (Lorentz pairs, ('lorentz', r_k, w_k, d_k)), combined over a common
    denominator in s_n = s / wnorm with wnorm = the largest pole frequency.
    This mirrors what a pole-residue to rational conversion
    (Convert_filter_S_PR_to_S, SRC/FILTER/filter_functions.F90:837-) produces
    for real poles and conjugate pairs."""
    wn = max(p[2] for p in poles) if poles else 1.0
    num = np.array([float(eps_inf)])
    den = np.array([1.0])
    for p in poles:
        if p[0] == "debye":
            _, r, w = p
            pd = np.array([1.0, wn / w])           # 1 + s_n wn / w
            pn = np.array([float(r)])
        elif p[0] == "lorentz":
            _, r, w, d = p
            pd = np.array([1.0, 2.0 * d * wn / w, (wn / w) ** 2])
            pn = np.array([float(r)])
        else:
            raise ValueError(p[0])
        num = np.polynomial.polynomial.polyadd(
            np.polynomial.polynomial.polymul(num, pd),
            np.polynomial.polynomial.polymul(pn, den))
        den = np.polynomial.polynomial.polymul(den, pd)
    return Sfilter(wn, list(num), list(den))


def sfilter_response(s: Sfilter, f):
    """Evaluate a(jw/wnorm)/b(jw/wnorm) at frequencies f (Hz)."""
    sn = 1j * 2.0 * np.pi * np.asarray(f, dtype=np.float64) / s.wnorm
    return np.polynomial.polynomial.polyval(sn, s.a) / np.polynomial.polynomial.polyval(sn, s.b)


def zfilter_response(z: Zfilter, f, dt: float):
    """Evaluate a(z^-1)/b(z^-1) at z = exp(j w dt)."""
    zi = np.exp(-1j * 2.0 * np.pi * np.asarray(f, dtype=np.float64) * dt)
    return np.polynomial.polynomial.polyval(zi, z.a) / np.polynomial.polynomial.polyval(zi, z.b)
