"""Small seeded problems for the parity tests (FP64 host data)."""
from __future__ import annotations

import numpy as np

from seidart_amd.cpml import (alpha_max_em, alpha_max_seismic, axis_profile, sigma_max_em,
                              sigma_max_seismic)
from seidart_amd.problem import Problem, Shot
from seidart_amd.sources import point_force


def _smooth_field(rng, shape, lo, hi, blobs=4):
    """Heterogeneous positive field: a few Gaussian blobs between lo and hi."""
    nz, nx = shape
    z, x = np.mgrid[0:nz, 0:nx].astype(np.float64)
    f = np.zeros(shape)
    for _ in range(blobs):
        cz, cx = rng.uniform(0, nz), rng.uniform(0, nx)
        s = rng.uniform(2.0, max(3.0, min(nz, nx) / 3))
        f += rng.uniform(-1, 1) * np.exp(-((z - cz) ** 2 + (x - cx) ** 2) / (2 * s * s))
    f = (f - f.min()) / (np.ptp(f) + 1e-30)
    return lo + (hi - lo) * f


def elastic_coefs(rng, shape, full21=False, aniso=0.08):
    """Stiffness fields (Pa) of a heterogeneous, mildly anisotropic solid."""
    rho = _smooth_field(rng, shape, 900.0, 2400.0)
    vp = _smooth_field(rng, shape, 1800.0, 3800.0)
    vs = vp / _smooth_field(rng, shape, 1.8, 2.2)
    c11 = rho * vp ** 2
    c55 = rho * vs ** 2
    lam = c11 - 2 * c55
    C = {n: np.zeros(shape) for n in ("c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23",
                                      "c24", "c25", "c26", "c33", "c34", "c35", "c36", "c44",
                                      "c45", "c46", "c55", "c56", "c66")}
    C["c11"] = c11 * (1 + aniso * rng.uniform(-1, 1))
    C["c22"] = c11.copy()
    C["c33"] = c11 * (1 + aniso * rng.uniform(-1, 1))
    C["c12"] = lam.copy()
    C["c13"] = lam * (1 + aniso * rng.uniform(-1, 1))
    C["c23"] = lam.copy()
    C["c44"] = c55.copy()
    C["c55"] = c55.copy()
    C["c66"] = c55.copy()
    C["c15"] = aniso * 0.3 * c55 * rng.uniform(-1, 1)
    C["c35"] = aniso * 0.3 * c55 * rng.uniform(-1, 1)
    if full21:
        for n in ("c14", "c16", "c24", "c25", "c26", "c34", "c36", "c45", "c46", "c56"):
            C[n] = aniso * 0.2 * c55 * rng.uniform(-1, 1)
    C["rho"] = rho
    return C


def em_coefs(rng, shape, full=False):
    C = {}
    C["e11"] = _smooth_field(rng, shape, 1.5, 4.0)
    C["e22"] = _smooth_field(rng, shape, 1.5, 4.0)
    C["e33"] = _smooth_field(rng, shape, 1.5, 4.0)
    C["e12"] = 0.05 * _smooth_field(rng, shape, -1, 1) if full else np.zeros(shape)
    C["e13"] = 0.1 * _smooth_field(rng, shape, -1, 1)
    C["e23"] = 0.05 * _smooth_field(rng, shape, -1, 1) if full else np.zeros(shape)
    for n in ("s11", "s22", "s33"):
        C[n] = _smooth_field(rng, shape, 1e-4, 5e-3)
    C["s12"] = np.zeros(shape)
    C["s13"] = 1e-4 * _smooth_field(rng, shape, 0, 1)
    C["s23"] = np.zeros(shape)
    return C


def tiny_problem(kind, seed=0, nx=40, ny=1, nz=36, npml=6, nt=80, dt=None, gamma=0.0,
                 nshots=1, nrec=6, src=None, full21=True, dx=1.0, dy=1.0, dz=1.0):
    """A seeded heterogeneous problem on a small padded grid (nx, ny, nz)."""
    rng = np.random.default_rng(seed)
    is3d = kind in ("elastic25", "em25")
    if not is3d:
        ny, dy = 1, 1.0
    shape = (nz, nx)
    if kind.startswith("elastic"):
        C = elastic_coefs(rng, shape, full21=full21)
        vmax = float(np.sqrt(np.max(np.maximum(C["c11"], C["c33"]) / C["rho"])))
        f0 = 100.0 / max(dx, dz)
        if dt is None:
            dt = 0.5 * min(dx, dy, dz) / (vmax * np.sqrt(3.0))
        sm = lambda d: sigma_max_seismic(vmax, npml, 2, 0.001)
        am = alpha_max_seismic(1.0, f0)
        seismic = True
    else:
        C = em_coefs(rng, shape, full=is3d)
        f0 = 3.0e8 / (12 * max(dx, dz) * np.sqrt(4.0))
        if dt is None:
            dt = 0.5 * min(dx, dy, dz) / (3.0e8 * np.sqrt(3.0))
        sm = lambda d: sigma_max_em(npml, min(dx, dy, dz) if is3d else min(dx, dz), 2, 0.001)
        am = alpha_max_em(1.0, f0)
        seismic = False
    prof = {}
    for a, n, d in (("x", nx, dx), ("y", ny, dy), ("z", nz, dz)):
        if a == "y" and not is3d:
            continue
        prof[a] = axis_profile(n, npml, d, sm(d), am, 5.0, 2, 2, dt, seismic=seismic)
    shots = []
    for s in range(nshots):
        if src is None:
            si = int(rng.integers(npml + 2, nx - npml - 2))
            sk = int(rng.integers(npml + 2, nz - npml - 2))
            sj = int(rng.integers(1, ny - 1)) if is3d else 0
        else:
            si, sj, sk = src
        fx, fy, fz = point_force(nt, dt, f0, amplitude=1.0 if seismic else 1.0,
                                 xz_rotation=float(rng.uniform(0, 90)),
                                 xy_rotation=float(rng.uniform(0, 90)) if is3d else 0.0)
        if not is3d:
            fy = None
        ri = rng.integers(1, nx - 1, size=nrec)
        rj = rng.integers(1, ny - 1, size=nrec) if is3d else np.zeros(nrec, dtype=int)
        rk = rng.integers(1, nz - 1, size=nrec)
        rec = np.stack([ri, rj, rk], axis=1).astype(np.int32)
        shots.append(Shot(src=(si, sj, sk), force=(fx, fy, fz), rec=rec))
    gam = {}
    if gamma and kind.startswith("elastic"):
        for a in ("x", "y", "z"):
            gam[a] = gamma * _smooth_field(rng, shape, 0.0, 1.0)
    p = Problem(kind=kind, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, dt=float(dt), nt=nt,
                coef=C, prof=prof, shots=shots, gamma=gam)
    p.validate()
    return p


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def write_reference_project(p, tmp, cpml):
    """A 2-D Problem as the reference's solver inputs (the files Model.build
    leaves, classes.py:393-547): the project JSON, coefficient records,
    legacy 1-D C-PML stream files, source records, and a receiver file
    (x, y, z in metres).  The source sits where the file route places it
    (0-based padded index xind + cpml - 1).  Returns (jsonfile, receiver
    file)."""
    import json
    import os
    from seidart_amd.fortranio import write_record, write_stream
    from seidart_amd.problem import CONDUCTIVITY, PERMITTIVITY, STIFFNESS
    seismic = p.kind.startswith("elastic")
    i, _, k = (int(v) for v in p.shots[0].src)
    xind, zind = i - cpml + 1, k - cpml + 1
    src = {"dt": p.dt, "time_steps": p.nt, "x": xind * p.dx, "y": 0.0, "z": zind * p.dz,
           "xind": xind, "yind": 0, "zind": zind, "source_frequency": 100.0, "x-z_rotation": 0,
           "x-y_rotation": 0, "y-z_rotation": 0, "amplitude": 1.0, "source_wavelet": "gaus1",
           "source_type": "ac"}
    prj = {"Domain": {"dim": 2, "nx": p.nx - 2 * cpml, "ny": 1, "nz": p.nz - 2 * cpml, "dx": p.dx,
                      "dy": 1, "dz": p.dz, "cpml": cpml, "nmats": 1, "alpha_max_scalar": 1.0,
                      "kappa_max": 5, "NP": 2, "NPA": 2, "Rcoef": 0.001, "image_file": None},
           "Seismic" if seismic else "Electromagnetic": {"Source": src}}
    jf = os.path.join(tmp, "golden.json")
    with open(jf, "w") as fh:
        json.dump(prj, fh)
    for n in (STIFFNESS if seismic else PERMITTIVITY + CONDUCTIVITY):
        if n in p.coef:
            write_record(os.path.join(tmp, f"{n}.dat"), p.coef[n])
    for a, g in p.gamma.items():
        write_record(os.path.join(tmp, f"gamma_{a}.dat"), g)
    for a in ("x", "z"):
        pr = p.prof[a]
        for q, h, v in (("kappa", "", pr.kappa), ("acoef", "", pr.a), ("bcoef", "", pr.b),
                        ("kappa", "_half", pr.kappa_h), ("acoef", "_half", pr.a_h),
                        ("bcoef", "_half", pr.b_h)):
            write_stream(os.path.join(tmp, f"{q}{a}{h}_cpml.dat"), v)
    stem = "seismicsource" if seismic else "electromagneticsource"
    for a, f in zip("xyz", p.shots[0].force):
        write_record(os.path.join(tmp, f"{stem}{a}.dat"), np.zeros(p.nt) if f is None else f)
    rec = np.asarray(p.shots[0].rec)
    rf = os.path.join(tmp, "receivers.xyz")
    np.savetxt(rf, np.stack([(rec[:, 0] - cpml) * p.dx, np.zeros(len(rec)), (rec[:, 2] - cpml) * p.dz], 1),
               delimiter=",", header="x,y,z", comments="")
    return jf, rf
