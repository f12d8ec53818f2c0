"""Source time functions and orientation (FP64 host setup).

Restates src/seidart/routines/sourcefunction.py:25-57 (``wavelet``) and the
force branch of ``pointsource`` (:239-383), and
src/seidart/routines/definitions.py:2101-2183 (``rotate_sdr``).
"""
from __future__ import annotations

import numpy as np

__all__ = ["wavelet", "rotate_sdr", "point_force"]


def wavelet(timevec, f: float, stype: str = "gaus1") -> np.ndarray:
    """Source wavelet normalised to max |x| = 1 (sourcefunction.py:25-57).

    'gaus1' is the Ricker wavelet (1 - 2a^2) exp(-a^2), a = pi f (t - 1/f).
    """
    t = np.asarray(timevec, dtype=np.float64)
    to = 1.0 / f
    if stype == "gaus0":
        bw = 0.2 / (np.pi * f)
        a = (t - to) / bw
        x = np.exp(-a * a)
    elif stype == "gaus1":
        a = np.pi * f * (t - to)
        e = np.exp(-a * a)
        x = (1.0 - 2.0 * a * a) * e
    elif stype == "gaus2":
        dt = t - to
        k = np.pi * np.pi * f * f
        e = np.exp(-k * dt * dt)
        x = 2.0 * k * dt * (2.0 * k * dt * dt - 3.0) * e
    else:
        raise ValueError(f"unsupported source wavelet {stype!r} (gaus0, gaus1, gaus2)")
    return x / np.max(np.abs(x))


def rotate_sdr(strike: float, dip: float, rake: float, is_degree: bool = True) -> np.ndarray:
    """R = Rz(strike) @ Ry(dip) @ Rx(rake) (definitions.py:2101-2183)."""
    if is_degree:
        s, d, r = np.deg2rad([strike, dip, rake])
    else:
        s, d, r = strike, dip, rake
    Rz = np.array([[np.cos(s), -np.sin(s), 0.0], [np.sin(s), np.cos(s), 0.0], [0.0, 0.0, 1.0]])
    Ry = np.array([[np.cos(d), 0.0, np.sin(d)], [0.0, 1.0, 0.0], [-np.sin(d), 0.0, np.cos(d)]])
    Rx = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(r), -np.sin(r)], [0.0, np.sin(r), np.cos(r)]])
    return Rz @ Ry @ Rx


def point_force(nt: int, dt: float, f0: float, amplitude: float = 1.0, stype: str = "gaus1",
                xy_rotation: float = 0.0, xz_rotation: float = 0.0, yz_rotation: float = 0.0):
    """(fx, fy, fz) series of an 'ac' point force (sourcefunction.py:283-295,
    :365-372): R = rotate_sdr(phi, theta, psi) with phi = x-y rotation,
    theta = x-z rotation, psi = y-z rotation; fz = R[0,2] s, fy = R[1,2] s,
    fx = R[2,2] s.  All angles zero therefore give an x force and
    ``x-z_rotation = 90`` a vertical (z) force."""
    timevec = np.arange(int(nt)) * float(dt)
    s = float(amplitude) * wavelet(timevec, f0, stype)
    R = rotate_sdr(xy_rotation, xz_rotation, yz_rotation)
    return R[2, 2] * s, R[1, 2] * s, R[0, 2] * s
