"""Synthetic workloads: the snow / firn / ice configurations the FDTD core is
measured on, built in FP64 on the host with fixed seeds.

* ``plain_psv``      2-D P-SV elastic, 256 x 256 cells, dx = dz = 0.5 m,
                     snow / firn / ice, C-PML 20, 200 Hz, 2000 steps, FP64.
* ``gpr``            2-D GPR (EM), 1024 x 768 cells, dx = 0.01 m, banded snow,
                     rotated-permittivity ice, 500 MHz, 4000 steps, CFL 0.95.
* ``aniso25``        2.5-D elastic, 800 x 64 x 400 cells, dx = 1 m, 8 regions
                     of rotated ice single-crystal stiffness (21 Cij),
                     attenuation, 100 Hz, 3000 steps.
* ``batched_layered`` 2-D elastic, 3 layered models (512 x 256, 1024 x 512,
                     2048 x 768; dx = 0.5 m), 32 shots each, 150 Hz, 3000 steps.
* ``fwi_vti``        2-D VTI elastic FWI gradient, 1024 x 512, dx = 1 m,
                     16 shots, 80 Hz, 6000 steps, 256 receivers (true and
                     smoothed start model).

Cell indices given for sources and receivers are those of the model without
the C-PML; the padded (device) index adds the C-PML width on each axis.
Where a configuration leaves a quantity open the choice is recorded here:
firn Vs = Vp / 2; C-PML 20 for the batched models; GPR snow occupies the top
256 rows; batched receivers every 4 cells at the source depth.
"""
from __future__ import annotations

import numpy as np

from .cpml import (CLIGHT, alpha_max_em, alpha_max_seismic, axis_profile, sigma_max_em,
                   sigma_max_seismic)
from .problem import STIFFNESS, Problem, Shot
from .sources import point_force

__all__ = ["plain_psv", "gpr", "aniso25", "batched_layered", "fwi_vti", "snow_firn_ice",
           "rotate_stiffness", "voigt_to_tensor", "tensor_to_voigt", "CONFIGS"]

SNOW = (400.0, 1800.0, 900.0)     # rho, Vp, Vs
ICE = (917.0, 3840.0, 1920.0)
FIRN_TOP = (400.0, 1800.0)         # rho, Vp at the top of the firn
FIRN_BOT = (830.0, 3600.0)

# ice single crystal (GPa), c-axis along z
ICE_C = dict(c11=13.93, c12=7.08, c13=5.77, c33=15.01, c44=3.01)

_VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def voigt_to_tensor(C6: np.ndarray) -> np.ndarray:
    """6x6 Voigt stiffness -> 3x3x3x3 tensor."""
    c = np.zeros((3, 3, 3, 3))
    for I, (i, j) in enumerate(_VOIGT):
        for J, (k, l) in enumerate(_VOIGT):
            for a, b in ((i, j), (j, i)):
                for cc, d in ((k, l), (l, k)):
                    c[a, b, cc, d] = C6[I, J]
    return c


def tensor_to_voigt(c: np.ndarray) -> np.ndarray:
    C6 = np.zeros((6, 6))
    for I, (i, j) in enumerate(_VOIGT):
        for J, (k, l) in enumerate(_VOIGT):
            C6[I, J] = c[i, j, k, l]
    return C6


def rotate_stiffness(C6: np.ndarray, R: np.ndarray) -> np.ndarray:
    """C'_ijkl = R_ip R_jq R_kr R_ls C_pqrs, in Voigt form.  Equal to the
    Bond-matrix product M C M^T the reference uses (materials.py:717-758)."""
    c = voigt_to_tensor(C6)
    return tensor_to_voigt(np.einsum("ip,jq,kr,ls,pqrs->ijkl", R, R, R, R, c, optimize=True))


def _ice_crystal_voigt() -> np.ndarray:
    c = ICE_C
    c66 = (c["c11"] - c["c12"]) / 2
    C6 = np.array([
        [c["c11"], c["c12"], c["c13"], 0, 0, 0],
        [c["c12"], c["c11"], c["c13"], 0, 0, 0],
        [c["c13"], c["c13"], c["c33"], 0, 0, 0],
        [0, 0, 0, c["c44"], 0, 0],
        [0, 0, 0, 0, c["c44"], 0],
        [0, 0, 0, 0, 0, c66],
    ]) * 1e9
    return C6


def _caxis_rotation(dip_deg: float, azimuth_deg: float) -> np.ndarray:
    """Rotation taking the z axis onto a c-axis of the given dip (from the
    vertical) and azimuth (from x in the x-y plane)."""
    t, f = np.deg2rad(dip_deg), np.deg2rad(azimuth_deg)
    Ry = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    Rz = np.array([[np.cos(f), -np.sin(f), 0], [np.sin(f), np.cos(f), 0], [0, 0, 1]])
    return Rz @ Ry


def _iso(rho, vp, vs):
    return dict(c11=rho * vp * vp, c33=rho * vp * vp, c55=rho * vs * vs,
                c13=rho * (vp * vp - 2 * vs * vs), rho=rho)


def _pad_edge(a: np.ndarray, c: int) -> np.ndarray:
    """Edge extension into the C-PML, as tensor2dat does (classes.py:664-675)."""
    return np.pad(a, c, mode="edge")


def _full_stiffness(nz, nx):
    return {n: np.zeros((nz, nx)) for n in STIFFNESS}


def snow_firn_ice(nx: int, nz: int, dz: float, rng, snow_bot: float, firn_bot: float,
                  amp_cells: float = 3.0):
    """Isotropic layered snow / firn / ice stiffness on an (nz, nx) grid.

    Interfaces are horizontal lines at rows ``snow_bot`` and ``firn_bot``
    perturbed by a seeded sinusoid of ``amp_cells`` amplitude; firn density
    and Vp rise linearly with depth through the firn."""
    x = np.arange(nx)
    lam = rng.uniform(0.25, 1.0) * nx
    ph1, ph2 = rng.uniform(0, 2 * np.pi, size=2)
    top = snow_bot + amp_cells * np.sin(2 * np.pi * x / lam + ph1)
    bot = firn_bot + amp_cells * np.sin(2 * np.pi * x / lam + ph2)
    z = np.arange(nz)[:, None].astype(np.float64)
    rho = np.empty((nz, nx))
    vp = np.empty((nz, nx))
    vs = np.empty((nz, nx))
    snow = z < top[None, :]
    ice = z >= bot[None, :]
    firn = ~snow & ~ice
    frac = np.clip((z - top[None, :]) / np.maximum(bot - top, 1.0)[None, :], 0.0, 1.0)
    rho[snow], vp[snow], vs[snow] = SNOW
    rho[ice], vp[ice], vs[ice] = ICE
    fr = FIRN_TOP[0] + (FIRN_BOT[0] - FIRN_TOP[0]) * frac
    fv = FIRN_TOP[1] + (FIRN_BOT[1] - FIRN_TOP[1]) * frac
    rho[firn], vp[firn] = fr[firn], fv[firn]
    vs[firn] = vp[firn] / 2
    labels = np.where(snow, 0, np.where(firn, 1, 2))
    return rho, vp, vs, labels


def _elastic2_problem(rho, vp, vs, dx, dz, cpml, dt, nt, f0, srcs, recs, labels=None):
    nz0, nx0 = rho.shape
    C = _full_stiffness(nz0, nx0)
    iso = _iso(rho, vp, vs)
    for k, v in iso.items():
        C[k] = v
    C["c22"], C["c12"], C["c23"], C["c44"], C["c66"] = iso["c11"], iso["c13"], iso["c13"], \
        iso["c55"], iso["c55"]
    coef = {k: _pad_edge(v, cpml) for k, v in C.items()}
    nz, nx = nz0 + 2 * cpml, nx0 + 2 * cpml
    vmax = float(np.max(vp))
    sm = sigma_max_seismic(vmax, cpml, 2, 0.001)
    am = alpha_max_seismic(1.0, f0)
    prof = {"x": axis_profile(nx, cpml, dx, sm, am, 5.0, 2, 2, dt, True),
            "z": axis_profile(nz, cpml, dz, sm, am, 5.0, 2, 2, dt, True)}
    fx, _, fz = point_force(nt, dt, f0, 1.0, "gaus1", xz_rotation=90.0)
    rec = np.array([[r[0] + cpml, 0, r[1] + cpml] for r in recs], dtype=np.int32)
    shots = [Shot(src=(s[0] + cpml, 0, s[1] + cpml), force=(fx, None, fz), rec=rec) for s in srcs]
    return Problem("elastic2", nx, 1, nz, dx, 1.0, dz, dt, nt, coef, prof, shots)


def _elastic2_problem_c(C0, dx, dz, cpml, dt, nt, f0, srcs, recs, vmax):
    """Like _elastic2_problem, from stiffness fields (Pa) and density given
    per cell on the unpadded grid."""
    nz0, nx0 = C0["rho"].shape
    C = _full_stiffness(nz0, nx0)
    C.update(C0)
    coef = {k: _pad_edge(v, cpml) for k, v in C.items()}
    nz, nx = nz0 + 2 * cpml, nx0 + 2 * cpml
    sm = sigma_max_seismic(vmax, cpml, 2, 0.001)
    am = alpha_max_seismic(1.0, f0)
    prof = {"x": axis_profile(nx, cpml, dx, sm, am, 5.0, 2, 2, dt, True),
            "z": axis_profile(nz, cpml, dz, sm, am, 5.0, 2, 2, dt, True)}
    fx, _, fz = point_force(nt, dt, f0, 1.0, "gaus1", xz_rotation=90.0)
    rec = np.array([[r[0] + cpml, 0, r[1] + cpml] for r in recs], dtype=np.int32)
    shots = [Shot(src=(s[0] + cpml, 0, s[1] + cpml), force=(fx, None, fz), rec=rec) for s in srcs]
    return Problem("elastic2", nx, 1, nz, dx, 1.0, dz, dt, nt, coef, prof, shots)


FWI_PARAMS = ("c11", "c13", "c33", "c55", "rho")


def fwi_vti(nt: int = 6000, seed: int = 1005, nshots: int = 16, nx0: int = 1024, nz0: int = 512,
            cpml: int = 20, nrec: int = 256, window=None):
    """FWI gradient workload, 2-D VTI elastic: 1024 x 512 cells, dx = dz = 1 m,
    C-PML 20, per-cell (C11, C13, C33, C55, rho).

    True model: snow / firn / ice layers (snow to row 0.06 nz, firn to row
    0.25 nz, isotropic from Vp / Vs), with the ice VTI (the ice single crystal
    with a vertical c-axis: C11 13.93, C13 5.77, C33 15.01, C55 3.01 GPa),
    plus 3 seeded Gaussian anomalies (sigma 20 cells) scaling all five
    parameters by +-5%.  Start model: the true model Gaussian-smoothed with
    sigma 30 cells.  16 vertical-force Ricker shots of 80 Hz evenly spaced
    along x at z = 4, 256 receivers every 4 cells at z = 2, dt with CFL
    1/sqrt(2) (classes.py:469-485), split C-PML variant.  Returns (true,
    start) Problems that share sources, receivers, dt and C-PML.

    ``window`` = (x0, x1, z0, z1): the same models cut to cells x0..x1-1,
    z0..z1-1 (with their own C-PML), the dt of the full model, ``nshots``
    sources spread along x at z = 4 and receivers every 4 cells at z = 2
    inside the window."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    dx = 1.0
    rho, vp, vs, labels = snow_firn_ice(nx0, nz0, dx, rng, 0.06 * nz0, 0.25 * nz0)
    C = {k: v.copy() for k, v in _iso(rho, vp, vs).items()}
    C["rho"] = rho.copy()
    ice = labels == 2
    for k, v in (("c11", ICE_C["c11"]), ("c13", ICE_C["c13"]), ("c33", ICE_C["c33"]),
                 ("c55", ICE_C["c44"])):
        C[k][ice] = v * 1e9
    z, x = np.mgrid[0:nz0, 0:nx0].astype(np.float64)
    pert = np.ones((nz0, nx0))
    for _ in range(3):
        cz = rng.uniform(0.3 * nz0, 0.9 * nz0)
        cx = rng.uniform(0.1 * nx0, 0.9 * nx0)
        amp = rng.choice([-0.05, 0.05])
        pert += amp * np.exp(-((z - cz) ** 2 + (x - cx) ** 2) / (2 * 20.0 ** 2))
    true = {k: C[k] * pert for k in FWI_PARAMS}
    start = {k: gaussian_filter(v, 30.0, mode="nearest") for k, v in true.items()}
    vmax = max(float(np.sqrt(np.max(np.maximum(M["c11"], M["c33"]) / M["rho"]))) for M in (true, start))
    dt = (1 / np.sqrt(2.0)) / (vmax * np.sqrt(2.0 / dx ** 2))
    if window is not None:
        x0, x1, z0, z1 = window
        true = {k: v[z0:z1, x0:x1].copy() for k, v in true.items()}
        start = {k: v[z0:z1, x0:x1].copy() for k, v in start.items()}
        nx0 = x1 - x0
        nrec = nx0 // 4
    srcs = [(int((k + 0.5) * nx0 / nshots), 4) for k in range(nshots)]
    step = nx0 // nrec
    recs = [(step // 2 + step * k, 2) for k in range(nrec)]
    out = tuple(_elastic2_problem_c(M, dx, dx, cpml, dt, nt, 80.0, srcs, recs, vmax)
                for M in (true, start))
    for p in out:
        # the legacy C-PML (shared memory arrays, b_x_half(j)) blows up on
        # this model after ~4400 steps; gradients use the split variant
        p.cpml_split = True
    return out


def plain_psv(nt: int = 2000, seed: int = 1001, cpml_split: bool = True) -> Problem:
    """2-D P-SV, 256 x 256, dx = dz = 0.5 m, C-PML 20, 200 Hz vertical force at
    (128, 30), dt = 0.9 dx / (sqrt(2) Vmax), 64 receivers at z = 30 every
    4 cells.  ``cpml_split``: the split C-PML (stable; the default) or the
    legacy sharing of the memory arrays, whose C-PML feeds a growing mode
    (its energy rises by ~1e13 over the 2000 steps)."""
    rng = np.random.default_rng(seed)
    rho, vp, vs, _ = snow_firn_ice(256, 256, 0.5, rng, 40.0, 100.0)
    dt = 0.9 * 0.5 / (np.sqrt(2.0) * ICE[1])
    recs = [(2 + 4 * k, 30) for k in range(64)]
    p = _elastic2_problem(rho, vp, vs, 0.5, 0.5, 20, dt, nt, 200.0, [(128, 30)], recs)
    p.cpml_split = bool(cpml_split)
    return p


def batched_layered(nt: int = 3000, seed: int = 1004, shots_per_model: int = 32,
                    sizes=((512, 256), (1024, 512), (2048, 768)), cpml_split: bool = True):
    """Three layered models of uneven size, dx = dz = 0.5 m, C-PML 20, 32 shots
    each at z = 10, x_k = (k + 1/2) nx / 32, 150 Hz vertical force, dt from the
    reference formula with CFL 1/sqrt(2) (classes.py:469-485), receivers every
    4 cells at z = 10.  ``cpml_split``: the split C-PML (stable; the default)
    or the legacy sharing of the memory arrays, which grows without bound
    from step ~1200 on these models."""
    rng = np.random.default_rng(seed)
    dx = 0.5
    vmax = ICE[1]
    dt = (1 / np.sqrt(2.0)) / (vmax * np.sqrt(2.0 / dx ** 2))
    out = []
    for nx0, nz0 in sizes:
        rho, vp, vs, _ = snow_firn_ice(nx0, nz0, dx, rng, 0.15 * nz0, 0.4 * nz0)
        srcs = [(int((k + 0.5) * nx0 / shots_per_model), 10) for k in range(shots_per_model)]
        recs = [(x, 10) for x in range(2, nx0 - 1, 4)]
        out.append(_elastic2_problem(rho, vp, vs, dx, dx, 20, dt, nt, 150.0, srcs, recs))
        out[-1].cpml_split = bool(cpml_split)
    return out


def gpr(nt: int = 4000, seed: int = 1002) -> Problem:
    """2-D GPR: 1024 x 768, dx = dz = 0.01 m, C-PML 20.  Snow (top 256 rows)
    in 10-row bands of isotropic eps_r ~ U[1.5, 2.2]; ice below with
    eps_r = R diag(3.15, 3.15, 3.18) R^T, R a rotation about y by a seeded
    per-column angle in [0, 30] deg; sigma log-uniform in [1e-5, 1e-4] S/m
    per 10-row band.  500 MHz Ez source at (512, 10); 128 receivers at z = 10
    every 8 cells; dt from CFL 0.95 with the reference formula
    (classes.py:476-485)."""
    rng = np.random.default_rng(seed)
    nx0, nz0, cpml, d = 1024, 768, 20, 0.01
    nb = nz0 // 10 + 1
    band_eps = rng.uniform(1.5, 2.2, size=nb)
    band_sig = 10 ** rng.uniform(-5, -4, size=nb)
    theta = np.deg2rad(rng.uniform(0, 30, size=nx0))
    z = np.arange(nz0)
    band = z // 10
    C = {n: np.zeros((nz0, nx0)) for n in ("e11", "e12", "e13", "e22", "e23", "e33",
                                            "s11", "s12", "s13", "s22", "s23", "s33")}
    snow = z < 256
    for n in ("e11", "e22", "e33"):
        C[n][snow, :] = band_eps[band[snow]][:, None]
    c, s = np.cos(theta), np.sin(theta)
    e1, e3 = 3.15, 3.18
    ice = ~snow
    C["e11"][ice, :] = (e1 * c * c + e3 * s * s)[None, :]
    C["e33"][ice, :] = (e1 * s * s + e3 * c * c)[None, :]
    C["e13"][ice, :] = ((e3 - e1) * s * c)[None, :]
    C["e22"][ice, :] = e1
    for n in ("s11", "s22", "s33"):
        C[n][:, :] = band_sig[band][:, None]
    coef = {k: _pad_edge(v, cpml) for k, v in C.items()}
    nx, nz = nx0 + 2 * cpml, nz0 + 2 * cpml
    emin = min(C["e11"].min(), C["e22"].min(), C["e33"].min())
    vmax = CLIGHT / np.sqrt(emin)
    dt = 0.95 / (vmax * np.sqrt(2.0 / d ** 2))
    f0 = 500e6
    sm = sigma_max_em(cpml, d, 2, 0.001)
    am = alpha_max_em(1.0, f0)
    prof = {"x": axis_profile(nx, cpml, d, sm, am, 5.0, 2, 2, dt, False),
            "z": axis_profile(nz, cpml, d, sm, am, 5.0, 2, 2, dt, False)}
    fx, _, fz = point_force(nt, dt, f0, 1.0, "gaus1", xz_rotation=90.0)
    rec = np.array([[4 + 8 * k + cpml, 0, 10 + cpml] for k in range(128)], dtype=np.int32)
    shot = Shot(src=(512 + cpml, 0, 10 + cpml), force=(fx, None, fz), rec=rec)
    return Problem("em2", nx, 1, nz, d, 1.0, d, dt, nt, coef, prof, [shot])


def gpr25(nt: int = 1000, seed: int = 1006, nx0=512, ny0=64, nz0=384, cpml=16) -> Problem:
    """2.5-D GPR, the y-invariant extension of the 2-D GPR workload at half its
    x and z extent: 512 x 64 x 384, dx = dy = dz = 0.01 m, C-PML 16; the same
    snow bands over c-axis-rotated ice (rotation about y per column, so the
    tensor is y-invariant), a 500 MHz Ez source at (256, 32, 10), 64
    receivers at z = 10, y = 32 every 8 cells, dt from CFL 0.95 in 3-D.  Not
    a configuration of the reference benchmark: a measured line for the 2.5-D
    EM kernels."""
    rng = np.random.default_rng(seed)
    d = 0.01
    nb = nz0 // 10 + 1
    band_eps = rng.uniform(1.5, 2.2, size=nb)
    band_sig = 10 ** rng.uniform(-5, -4, size=nb)
    theta = np.deg2rad(rng.uniform(0, 30, size=nx0))
    z = np.arange(nz0)
    band = z // 10
    C = {n: np.zeros((nz0, nx0)) for n in ("e11", "e12", "e13", "e22", "e23", "e33",
                                            "s11", "s12", "s13", "s22", "s23", "s33")}
    snow = z < nz0 // 3
    for n in ("e11", "e22", "e33"):
        C[n][snow, :] = band_eps[band[snow]][:, None]
    c, s = np.cos(theta), np.sin(theta)
    e1, e3 = 3.15, 3.18
    ice = ~snow
    C["e11"][ice, :] = (e1 * c * c + e3 * s * s)[None, :]
    C["e33"][ice, :] = (e1 * s * s + e3 * c * c)[None, :]
    C["e13"][ice, :] = ((e3 - e1) * s * c)[None, :]
    C["e22"][ice, :] = e1
    for n in ("s11", "s22", "s33"):
        C[n][:, :] = band_sig[band][:, None]
    coef = {k: _pad_edge(v, cpml) for k, v in C.items()}
    nx, ny, nz = nx0 + 2 * cpml, ny0 + 2 * cpml, nz0 + 2 * cpml
    emin = min(C["e11"].min(), C["e22"].min(), C["e33"].min())
    vmax = CLIGHT / np.sqrt(emin)
    dt = 0.95 / (vmax * np.sqrt(3.0 / d ** 2))
    f0 = 500e6
    sm = sigma_max_em(cpml, d, 2, 0.001)
    am = alpha_max_em(1.0, f0)
    prof = {a: axis_profile(n, cpml, d, sm, am, 5.0, 2, 2, dt, False)
            for a, n in (("x", nx), ("y", ny), ("z", nz))}
    fx, fy, fz = point_force(nt, dt, f0, 1.0, "gaus1", xz_rotation=90.0)
    rec = np.array([[4 + 8 * k + cpml, ny0 // 2 + cpml, 10 + cpml] for k in range(64)], dtype=np.int32)
    shot = Shot(src=(nx0 // 2 + cpml, ny0 // 2 + cpml, 10 + cpml), force=(fx, fy, fz), rec=rec)
    return Problem("em25", nx, ny, nz, d, d, d, dt, nt, coef, prof, [shot])


def aniso25(nt: int = 3000, seed: int = 1003, nx0=800, ny0=64, nz0=400, cpml=16,
            window=None) -> Problem:
    """2.5-D elastic: 800 x 64 x 400, dx = 1 m, C-PML 16; 8 seeded Voronoi
    regions of ice single-crystal stiffness rotated to a c-axis of dip
    U[0, 60] deg and azimuth U[0, 360] deg (21 Cij), rho 917, attenuation
    gamma ~ U[0, 5e-4] per region on every velocity component; 100 Hz
    vertical force at (400, 32, 10); 200 receivers at z = 2, y = 32, x = 4k;
    dt from the reference formula with CFL 1/sqrt(3).

    ``window`` = (x0, x1, z0, z1): the same model cut to cells x0..x1-1,
    z0..z1-1 (with its own C-PML), the dt of the full model, the source and
    the receivers that fall inside the window (source coordinates shift with
    the window)."""
    rng = np.random.default_rng(seed)
    d = 1.0
    nreg = 8
    cx = rng.uniform(0, nx0, nreg)
    cz = rng.uniform(0, nz0, nreg)
    X, Z = np.meshgrid(np.arange(nx0), np.arange(nz0))
    lab = np.argmin((X[None] - cx[:, None, None]) ** 2 + (Z[None] - cz[:, None, None]) ** 2, axis=0)
    base = _ice_crystal_voigt()
    C = {n: np.zeros((nz0, nx0)) for n in STIFFNESS}
    gam = np.zeros((nz0, nx0))
    vmax = 0.0
    for r in range(nreg):
        R = _caxis_rotation(rng.uniform(0, 60), rng.uniform(0, 360))
        C6 = rotate_stiffness(base, R)
        mask = lab == r
        for I in range(6):
            for J in range(I, 6):
                C[f"c{I + 1}{J + 1}"][mask] = C6[I, J]
        C["rho"][mask] = 917.0
        gam[mask] = rng.uniform(0, 5e-4)
        for n in np.eye(3):
            Gam = np.einsum("ijkl,j,l->ik", voigt_to_tensor(C6), n, n) / 917.0
            vmax = max(vmax, float(np.sqrt(np.max(np.linalg.eigvalsh(Gam)))))
    sx, sz = 400, 10
    rx = 4 * np.arange(200)
    if window is not None:
        x0, x1, z0, z1 = window
        C = {k: v[z0:z1, x0:x1] for k, v in C.items()}
        gam = gam[z0:z1, x0:x1]
        nx0, nz0 = x1 - x0, z1 - z0
        sx, sz = sx - x0, sz - z0
        rx = rx[(rx >= x0) & (rx < x1)] - x0
        if not (0 <= sx < nx0 and 0 <= sz < nz0 and 0 <= 2 - z0 < nz0):
            raise ValueError("the window must hold the source and the receiver line")
    coef = {k: _pad_edge(v, cpml) for k, v in C.items()}
    gpad = _pad_edge(gam, cpml)
    nx, ny, nz = nx0 + 2 * cpml, ny0 + 2 * cpml, nz0 + 2 * cpml
    dt = (1 / np.sqrt(3.0)) / (vmax * np.sqrt(3.0 / d ** 2))
    f0 = 100.0
    sm = sigma_max_seismic(vmax, cpml, 2, 0.001)
    am = alpha_max_seismic(1.0, f0)
    prof = {a: axis_profile(n, cpml, d, sm, am, 5.0, 2, 2, dt, True)
            for a, n in (("x", nx), ("y", ny), ("z", nz))}
    fx, fy, fz = point_force(nt, dt, f0, 1.0, "gaus1", xz_rotation=90.0)
    rz = 2 - (window[2] if window is not None else 0)
    rec = np.array([[x + cpml, 32 + cpml, rz + cpml] for x in rx], dtype=np.int32)
    shot = Shot(src=(sx + cpml, 32 + cpml, sz + cpml), force=(fx, fy, fz), rec=rec)
    return Problem("elastic25", nx, ny, nz, d, d, d, dt, nt, coef, prof, [shot],
                   gamma={"x": gpad, "y": gpad, "z": gpad})



def gpr_survey(nt: int = 1000, nshots: int = 32, seed: int = 1002) -> Problem:
    """The 2-D GPR model as a common-offset survey: `nshots` source positions
    along z = 10 (every 30 cells from x = 20), the same 128 surface
    receivers for each.  Not a configuration of the reference benchmark: the
    multi-shot EM line (what `CommonOffset.co_run` runs as one plan)."""
    p = gpr(nt=nt, seed=seed)
    sh = p.shots[0]
    p.shots = [Shot(src=(20 + 30 * k + 20, 0, sh.src[2]), force=sh.force, rec=sh.rec)
               for k in range(nshots)]
    return p


CONFIGS = {
    "gpr_survey": gpr_survey,
    "gpr25": gpr25,
    "plain_psv": plain_psv,
    "gpr": gpr,
    "aniso25": aniso25,
    "batched_layered": batched_layered,
    "fwi_vti": fwi_vti,
}
