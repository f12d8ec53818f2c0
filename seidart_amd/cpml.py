"""C-PML damping profiles (FP64 host setup).

The reference builds its C-PML coefficients in
src/seidart/routines/definitions.py:560-830 (``cpmlcompute``) and
:862-998 (``cpml_parameters``) and writes them as 2-D ``(nx, nz)`` fields.
Outside the corners each of those fields is a pure function of one
coordinate: the x slabs hold the x profile (``cpml_parameters`` :967-975),
the z slabs the z profile (:977-984).  The legacy solver consumed exactly
those 1-D profiles, one stream file per axis and staggering
(legacy/fortran/seismicfdtd.f95:199-217, readwrite_routines.f95:26-37), and
applies the x profile to x derivatives and the z profile to z derivatives.
This module builds the 1-D profiles with the reference formulas and can
recover them from the reference's 2-D files (their mid lines).

Profiles per axis: kappa, alpha, sigma, a, b on the integer grid and on the
half grid, indexed 0..n-1 over the padded axis.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

EPS0 = 8.85418782e-12       # definitions.py:68
CLIGHT = 2.99792458e8       # definitions.py:67
MU0 = 4.0 * np.pi * 1.0e-7  # definitions.py:69

__all__ = ["AxisProfile", "axis_profile", "sigma_max_seismic", "sigma_max_em",
           "alpha_max_seismic", "alpha_max_em", "profiles_from_2d", "slab_width",
           "EPS0", "CLIGHT", "MU0"]


@dataclass
class AxisProfile:
    """1-D C-PML coefficients of one axis (integer and half grid)."""
    kappa: np.ndarray
    a: np.ndarray
    b: np.ndarray
    kappa_h: np.ndarray
    a_h: np.ndarray
    b_h: np.ndarray
    sigma: np.ndarray | None = None
    alpha: np.ndarray | None = None
    sigma_h: np.ndarray | None = None
    alpha_h: np.ndarray | None = None

    @property
    def n(self) -> int:
        return int(self.kappa.size)

    @classmethod
    def trivial(cls, n: int) -> "AxisProfile":
        """No absorption: kappa = 1, a = 0, b = 1 (interior values)."""
        one, zero = np.ones(n), np.zeros(n)
        return cls(one.copy(), zero.copy(), one.copy(), one.copy(), zero.copy(), one.copy())


def sigma_max_seismic(vmax: float, cpml: int, NP: float, Rcoef: float,
                      velocity_scaling_factor: float = 1.0) -> float:
    """definitions.py:649-652 / :675-681 ("global" smoothing): the thickness
    enters in cells, not metres, exactly as the reference writes it."""
    return float(-np.log(Rcoef) * (NP + 1) * velocity_scaling_factor * 0.7 * vmax / (2.0 * cpml))


def sigma_max_em(cpml: int, deltamin: float, NP: float, Rcoef: float,
                 cmax: float = CLIGHT) -> float:
    """definitions.py:682-689.  The reference's EM "velocity map" holds
    n/c (``compute_max_velocities`` inverts the speed, classes.py:1112-1114),
    so ``clight / sqrt(max(vmap, 1))`` is ``clight`` for every physical
    medium; cmax defaults to that value."""
    return float(-(NP + 1) * np.log(Rcoef) * EPS0 * cmax / (2.0 * cpml * deltamin))


def alpha_max_seismic(alpha_max_scalar: float, f0: float) -> float:
    return float(alpha_max_scalar * np.pi * f0)        # definitions.py:648


def alpha_max_em(alpha_max_scalar: float, f0: float) -> float:
    return float(alpha_max_scalar * np.pi * EPS0 * f0)  # definitions.py:658


def _one(n: int, cpml: int, d: float, half: bool, sig_max: float, alpha_max: float,
         kappa_max: float, NP: float, NPA: float, dt: float, seismic: bool):
    """One staggering of one axis, restating cpml_parameters (:940-996)
    on the x-slab rows (the 2-D field away from the corners)."""
    dist = d * np.arange(0, cpml, dtype=np.float64)
    if half:
        dist = dist + d / 2
    dist = (d * cpml - dist) / (d * cpml)          # cpmlcompute :600-608
    dist = dist / dist.max()                        # cpml_parameters :940-941
    r = dist ** NP
    sigma = np.zeros(n)
    kappa = np.ones(n)
    alpha = np.zeros(n)
    for ind in range(cpml):                         # :967-975
        sigma[ind] = sig_max * r[ind]
        kappa[ind] = 1.0 + (kappa_max - 1.0) * r[ind]
        alpha[ind] = alpha_max * (1 - dist[ind]) ** NPA
        sigma[-(ind + 1)] = sig_max * r[ind]
        kappa[-(ind + 1)] = 1 + (kappa_max - 1) * r[ind]
        alpha[-(ind + 1)] = alpha_max * (1 - dist[ind]) ** NPA
    if seismic:                                     # :986-989
        b = np.exp(-(sigma / kappa + alpha) * dt)
    else:
        b = np.exp(-(sigma / kappa + alpha) * (dt / EPS0))
    alpha[alpha < 0.0] = 0.0
    a = np.zeros(n)
    idx = np.abs(sigma) > 1.0e-6                    # :992-996
    a[idx] = sigma[idx] * (b[idx] - 1) / (kappa[idx] * sigma[idx] + kappa[idx] * alpha[idx])
    return sigma, kappa, alpha, a, b


def axis_profile(n: int, cpml: int, d: float, sig_max: float, alpha_max: float,
                 kappa_max: float, NP: float, NPA: float, dt: float,
                 seismic: bool = True) -> AxisProfile:
    """Integer- and half-grid profiles of a padded axis of n cells."""
    if cpml <= 0:
        return AxisProfile.trivial(n)
    if 2 * cpml > n:
        raise ValueError(f"C-PML of {cpml} cells does not fit an axis of {n} cells")
    s, k, al, a, b = _one(n, cpml, d, False, sig_max, alpha_max, kappa_max, NP, NPA, dt, seismic)
    sh, kh, alh, ah, bh = _one(n, cpml, d, True, sig_max, alpha_max, kappa_max, NP, NPA, dt, seismic)
    return AxisProfile(k, a, b, kh, ah, bh, s, al, sh, alh)


def profiles_from_2d(kappa, acoef, bcoef, kappa_h, acoef_h, bcoef_h, cpml: int):
    """Recover the x and z profiles from the reference's 2-D C-PML fields
    (arrays (nz, nx) as read from the ``*_cpml.dat`` records): the x profile
    is any row away from the z slabs, the z profile any column away from the
    x slabs (cpml_parameters :967-984)."""
    nz, nx = np.asarray(kappa).shape
    jm, im = nz // 2, nx // 2
    if not (cpml <= jm < nz - cpml and cpml <= im < nx - cpml):
        raise ValueError("grid too small to separate the C-PML profiles")
    px = AxisProfile(*(np.asarray(f)[jm, :].astype(np.float64).copy()
                       for f in (kappa, acoef, bcoef, kappa_h, acoef_h, bcoef_h)))
    pz = AxisProfile(*(np.asarray(f)[:, im].astype(np.float64).copy()
                       for f in (kappa, acoef, bcoef, kappa_h, acoef_h, bcoef_h)))
    return px, pz


def slab_width(p: AxisProfile) -> int:
    """Smallest l such that every cell with l <= i < n-l has a = 0 and
    kappa = 1 on both staggerings: those cells need no memory variable."""
    n = p.n
    nontriv = ((p.a != 0) | (p.a_h != 0) | (p.kappa != 1) | (p.kappa_h != 1))
    idx = np.nonzero(nontriv)[0]
    if idx.size == 0:
        return 0
    lo = int(idx[idx < (n + 1) // 2].max() + 1) if np.any(idx < (n + 1) // 2) else 0
    hi = int(n - idx[idx >= (n + 1) // 2].min()) if np.any(idx >= (n + 1) // 2) else 0
    return min(max(lo, hi), (n + 1) // 2)
