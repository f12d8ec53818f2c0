"""Source time functions, orientation and moment tensors (FP64 host setup).

Restates src/seidart/routines/sourcefunction.py: ``wavelet`` (:25-57),
``wavelet_center0`` (:60-78), ``multimodesrc`` (:80-101),
``bandlimited_impulse_fir`` / ``broadbandsrc`` (:104-142),
``double_couple_tensor`` (:166-197), ``isotropic_tensor`` (:199-204) and
``pointsource`` (:239-383, without its file writes), and
src/seidart/routines/definitions.py:2101-2183 (``rotate_sdr``).
"""
from __future__ import annotations

import numpy as np

__all__ = ["wavelet", "wavelet_center0", "multimodesrc", "broadbandsrc", "rotate_sdr",
           "point_force", "double_couple_tensor", "isotropic_tensor", "pointsource",
           "SOURCE_TYPES", "WAVELETS"]

WAVELETS = ("gaus0", "gaus1", "gaus2", "chirp", "chirplet")
SOURCE_TYPES = ("ac", "dc", "tnt", "omni_ps", "omni_s")


def wavelet(timevec, f: float, stype="gaus1") -> np.ndarray:
    """Source wavelet normalised to max |x| = 1 (sourcefunction.py:25-57).

    'gaus0' a Gaussian, 'gaus1' the Ricker wavelet (1 - 2a^2) exp(-a^2),
    a = pi f (t - 1/f), 'gaus2' its derivative; 'chirp' a cosine sweep from
    10 f down to f over 1/f (scipy.signal.chirp, phase -90 deg), 'chirplet'
    a sweep from f up to 20 f under a Gaussian window exp(-pi^2 f^2 (t-1/f)^2).
    Any other value is taken as the series itself (normalised).
    """
    t = np.asarray(timevec, dtype=np.float64)
    to = 1.0 / f
    a2 = (2.0 * np.pi * f) ** 2
    if isinstance(stype, str) and stype == "gaus0":
        bw = 0.2 / (np.pi * f)
        a = (t - to) / bw
        x = np.exp(-a * a)
    elif isinstance(stype, str) and stype == "gaus1":
        a = np.pi * f * (t - to)
        e = np.exp(-a * a)
        x = (1.0 - 2.0 * a * a) * e
    elif isinstance(stype, str) and stype == "gaus2":
        dt = t - to
        k = np.pi * np.pi * f * f
        e = np.exp(-k * dt * dt)
        x = 2.0 * k * dt * (2.0 * k * dt * dt - 3.0) * e
    elif isinstance(stype, str) and stype == "chirp":
        from scipy.signal import chirp
        x = chirp(t, 10 * f, to, f, phi=-90)
    elif isinstance(stype, str) and stype == "chirplet":
        from scipy.signal import chirp
        x = chirp(t, f, to, 20 * f, phi=-90) * np.exp(-(a2 / 4) * (t - to) ** 2)
    else:
        if isinstance(stype, str):
            raise ValueError(f"unknown source wavelet {stype!r} ({', '.join(WAVELETS)} or a series)")
        x = np.asarray(stype, dtype=np.float64)
    return x / np.max(np.abs(x))


def wavelet_center0(timevec, f: float, stype: str) -> np.ndarray:
    """The Gaussian family centred on t = 0 (simultaneous onset), normalised
    (sourcefunction.py:60-78): exp(-a t^2), -2 a t exp(-a t^2),
    2 a exp(-a t^2)(2 a t^2 - 1) with a = (2 pi f)^2; other wavelets are
    shifted by 1/f."""
    t = np.asarray(timevec, dtype=np.float64)
    a = (2.0 * np.pi * f) ** 2
    if stype == "gaus0":
        x = np.exp(-a * t ** 2)
    elif stype == "gaus1":
        x = -2.0 * a * t * np.exp(-a * t ** 2)
    elif stype == "gaus2":
        x = 2.0 * a * np.exp(-a * t ** 2) * (2.0 * a * t ** 2 - 1)
    else:
        x = wavelet(t + 1.0 / f, f, stype)
    return x / np.max(np.abs(x))


def multimodesrc(timevec, f: float, stype: str, center: bool = False, num_octaves: int = 3):
    """Sum of 4 * num_octaves - 1 wavelets at f * 2^linspace(-num_octaves, 0),
    normalised (sourcefunction.py:80-101)."""
    t = np.asarray(timevec, dtype=np.float64)
    freqs = f * 2 ** np.linspace(-num_octaves, 0, num_octaves * 4 - 1)
    stf = np.zeros_like(t)
    for fr in freqs:
        stf += wavelet_center0(t, fr, stype) if center else wavelet(t, fr, stype)
    return stf / np.max(np.abs(stf))


def broadbandsrc(timevec, fmin: float, fmax: float, numtaps: int = 201, window: str = "hann"):
    """Minimum-phase band-pass impulse (sourcefunction.py:104-142): a
    linear-phase FIR band-pass [fmin, fmax] (scipy.signal.firwin) turned
    minimum-phase (scipy.signal.minimum_phase), placed at t = 0, normalised."""
    from scipy.signal import firwin, minimum_phase
    t = np.asarray(timevec, dtype=np.float64)
    fs = 1.0 / np.mean(np.diff(t))
    mp = minimum_phase(firwin(numtaps, [fmin, fmax], pass_zero=False, fs=fs, window=window))
    h = np.zeros(t.size)
    h[: mp.size] = mp
    return h / np.max(np.abs(h))


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


def double_couple_tensor(strike_deg: float, dip_deg: float, rake_deg: float, M0: float = 1.0):
    """Unit double couple scaled by M0 (sourcefunction.py:166-197): fault
    strike direction s, down-dip direction d, normal n = s x d, slip
    cos(rake) s + sin(rake) d; M = M0 (slip n^T + n slip^T), symmetrised and
    made traceless."""
    st, dp, rk = np.deg2rad([strike_deg, dip_deg, rake_deg])
    s_hat = np.array([-np.sin(st), np.cos(st), 0.0])
    d_hat = np.array([np.cos(st) * np.cos(dp), np.sin(st) * np.cos(dp), -np.sin(dp)])
    s_hat /= np.linalg.norm(s_hat)
    d_hat /= np.linalg.norm(d_hat)
    n_hat = np.cross(s_hat, d_hat)
    n_hat /= np.linalg.norm(n_hat)
    slip = np.cos(rk) * s_hat + np.sin(rk) * d_hat
    slip /= np.linalg.norm(slip)
    M = M0 * (np.outer(slip, n_hat) + np.outer(n_hat, slip))
    M = 0.5 * (M + M.T)
    return M - np.trace(M) / 3.0 * np.eye(3)


def isotropic_tensor(M0: float = 1.0) -> np.ndarray:
    """M0 I: explosion for M0 > 0 (sourcefunction.py:199-204)."""
    return M0 * np.eye(3)


def pointsource(nt: int, dt: float, f0: float, amplitude: float = 1.0, source_wavelet="gaus1",
                source_type: str = "ac", phi: float = 0.0, theta: float = 0.0, psi: float = 0.0,
                is_seismic: bool = True, multimodal: bool = False, broadband: bool = False,
                fmin: float | None = None, fmax: float | None = None, center: bool = False,
                num_octaves: int = 2, a_iso: float = 1.0, b_dc: float = 0.7):
    """Point-source series of sourcefunction.py:239-383 (the arrays, not the
    files it writes).  Returns (force, moment): force = (fx, fy, fz), moment =
    (sxx, syy, szz, sxy, sxz, syz) stress-rate series, each of nt samples.

    Seismic source types: 'ac' (a force, rotated as in point_force), 'dc'
    (double couple of strike phi, dip theta, rake psi), 'tnt' (isotropic),
    'omni_ps' (isotropic part a_iso plus b_dc times the mean of three double
    couples on the axes of rotate_sdr(phi, theta, psi)), 'omni_s' (the
    double couples alone).  EM sources are always a rotated force."""
    N = int(nt)
    t = np.arange(N) * float(dt)
    amp = float(amplitude)
    if broadband:
        srcfn = amp * broadbandsrc(t, fmin, fmax)
    elif multimodal:
        srcfn = amp * multimodesrc(t, f0, source_wavelet, center=center, num_octaves=num_octaves)
    elif center:
        srcfn = amp * wavelet_center0(t, f0, source_wavelet)
    else:
        srcfn = amp * wavelet(t, f0, source_wavelet)
    z = np.zeros(N)
    fx, fy, fz = z.copy(), z.copy(), z.copy()
    sxx, syy, szz, sxy, sxz, syz = (z.copy() for _ in range(6))
    if not is_seismic or source_type == "ac":
        R = rotate_sdr(phi, theta, psi)
        fz, fy, fx = R[0, 2] * srcfn, R[1, 2] * srcfn, R[2, 2] * srcfn
    elif source_type == "dc":
        M = double_couple_tensor(phi, theta, psi, M0=1.0)
        sxx, syy, szz = srcfn * M[0, 0], srcfn * M[1, 1], srcfn * M[2, 2]
        sxy, sxz, syz = srcfn * M[0, 1], srcfn * M[0, 2], srcfn * M[1, 2]
    elif source_type == "tnt":
        sxx, syy, szz = srcfn.copy(), srcfn.copy(), srcfn.copy()
    elif source_type in ("omni_ps", "omni_s"):
        R = rotate_sdr(phi, theta, psi)
        if source_type == "omni_ps":
            s_iso = a_iso * srcfn
            sxx, syy, szz = sxx + s_iso, syy + s_iso, szz + s_iso
        s_dc = b_dc * srcfn
        e1, e2, e3 = R[:, 0], R[:, 1], R[:, 2]
        dcs = [np.outer(e1, e2) + np.outer(e2, e1), np.outer(e1, e3) + np.outer(e3, e1),
               np.outer(e2, e3) + np.outer(e3, e2)]
        Mavg = sum(dcs) / 3.0
        sxx, syy, szz = sxx + s_dc * Mavg[0, 0], syy + s_dc * Mavg[1, 1], szz + s_dc * Mavg[2, 2]
        sxy, sxz, syz = sxy + s_dc * Mavg[0, 1], sxz + s_dc * Mavg[0, 2], syz + s_dc * Mavg[1, 2]
    else:
        raise ValueError(f"unknown source type {source_type!r} ({', '.join(SOURCE_TYPES)})")
    return (fx, fy, fz), (sxx, syy, szz, sxy, sxz, syz)
