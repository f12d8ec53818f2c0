"""Drop-in replacement for SeidarT's solver call.

The reference runs its solver as a separate process on a file contract
(src/seidart/routines/classes.py:878-924)::

    seidartfdtd <project.json> seismic=<true|false> density_method=<...>

reading the coefficient, C-PML and source files that ``Model.build`` wrote
into the working directory, and writing one snapshot file per step, from
which ``Array.getrcx`` gathers receiver traces (arraybuild.py:199-295).

This module offers the same call in three forms:

* :func:`run_files` / ``python -m seidart_amd.run <json> seismic=... ...``:
  the same arguments and input files as the reference binary;
* :func:`run`: the ``Model.run`` signature, for a Model / Domain pair (the
  reference's objects or any object with the same attributes);
* :func:`problem_from_files` / :func:`problem_from_objects`: the host setup
  alone, giving a :class:`~seidart_amd.problem.Problem`.

Outputs: receiver traces (returned, and written as ``Array.save``-style CSV
files when receivers are given), and snapshot files in the reference's
format at a chosen step interval.  Everything between the inputs and the
outputs runs in the HIP kernels of libseidart_hip.so.
"""
from __future__ import annotations

import json
import os
import sys
from dataclasses import dataclass

import numpy as np

from .cpml import (AxisProfile, alpha_max_em, alpha_max_seismic, axis_profile,
                   profiles_from_2d, sigma_max_em, sigma_max_seismic)
from .fortranio import em25_block_name, read_record, read_stream, snapshot_name, write_em25_block, \
    write_snapshot
from .problem import (CONDUCTIVITY, DENSITY_METHODS, FIELDS, MOMENT_COMPONENTS, PERMITTIVITY,
                      STIFFNESS, TRACE_COMPONENTS, Problem, Shot)

__all__ = ["run", "run_files", "problem_from_files", "problem_from_objects", "RunResult",
           "load_receivers", "receiver_indices", "source_index", "christoffel_vmax",
           "write_trace_csv", "main"]

@dataclass
class RunResult:
    problem: Problem
    traces: np.ndarray          # [nt, ncomp, nrec] (first shot)
    components: list
    unstable_step: int
    csv_files: list
    snapshot_files: int


# --------------------------------------------------------------------------
# geometry helpers
# --------------------------------------------------------------------------
def source_index(xind: int, yind: int, zind: int, cpml: int, dim: float,
                 convention: str = "legacy"):
    """Padded 0-based source cell.  The legacy solver places the source at
    Fortran index src + npoints_pml with src = int(x/dx) (seismicfdtd.f95:
    162-163, legacy/v0.0.2/materials/definitions.py:532-546), i.e. 0-based
    xind + cpml - 1; ``convention='receiver'`` uses xind + cpml, the index the
    trace reader uses for receivers (arraybuild.py:185-192)."""
    off = cpml - 1 if convention == "legacy" else cpml
    j = yind + off if dim == 2.5 else 0
    return (int(xind + off), int(j), int(zind + off))


def receiver_indices(xyz, dx, dy, dz, cpml, dim):
    """int(xyz / d) + cpml per receiver (Array.loadxyz, arraybuild.py:185-192),
    as (i, j, k) padded indices; j = 0 in 2-D."""
    xyz = np.atleast_2d(np.asarray(xyz, dtype=np.float64))
    idx = (xyz / np.array([dx, dy, dz])).astype(int) + cpml
    if dim != 2.5:
        idx[:, 1] = 0
    return idx.astype(np.int32)


def load_receivers(path: str) -> np.ndarray:
    """Receiver file with an x, y, z header row (Array.loadxyz reads it with
    pandas.read_csv, arraybuild.py:162-176)."""
    a = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
    if a.shape[1] != 3 and a.shape[0] == 3:
        a = a.T
    return a


# --------------------------------------------------------------------------
# material physics needed by the C-PML setup
# --------------------------------------------------------------------------
def _voigt6(row):
    names = STIFFNESS[:-1]
    v = dict(zip(names, [float(row[n]) for n in names]))
    C = np.zeros((6, 6))
    k = 0
    for i in range(6):
        for j in range(i, 6):
            C[i, j] = C[j, i] = v[names[k]]
            k += 1
    return C


def christoffel_vmax(row, direction) -> float:
    """Largest phase velocity along a direction, as the reference's C-PML
    setup computes it (``get_christoffel_matrix``, materials.py:1014-1172):
    the Voigt matrix is expanded to a 3x3x3x3 tensor with a factor 1/sqrt(2)
    on each shear index pair (``voigt_to_full_tensor``, materials.py:903-966,
    so that C_1212 = c66 / 2 and C_1112 = c16 / sqrt(2)), then
    Gamma_ik = C_ijkl n_j n_l / rho and v = sqrt(max(eig(Gamma), 0)).  The
    factor is kept so that sigma_max, and with it every C-PML coefficient,
    equals the reference's; the P speed of a medium without shear couplings
    is unaffected."""
    from .configs import voigt_to_tensor
    n = np.asarray(direction, dtype=np.float64)
    n = n / np.linalg.norm(n)
    rho = float(row["rho"])
    f = np.array([1.0, 1.0, 1.0, 2 ** -0.5, 2 ** -0.5, 2 ** -0.5])
    G = np.einsum("ijkl,j,l->ik", voigt_to_tensor(_voigt6(row) * np.outer(f, f)), n, n) / rho
    return float(np.sqrt(max(np.max(np.linalg.eigvalsh(G)), 0.0)))


# --------------------------------------------------------------------------
# problem from the reference's files
# --------------------------------------------------------------------------
def _jget(d, *keys, default=None):
    for k in keys:
        if k in d and d[k] is not None:
            return d[k]
    return default


def _y_profile_like(px: AxisProfile, ny: int, cpml: int) -> AxisProfile:
    """2.5-D y profile: the reference's 2-D C-PML files carry x and z only,
    so the y axis takes the x slab values (same thickness in cells)."""
    out = AxisProfile.trivial(ny)
    for name in ("kappa", "a", "b", "kappa_h", "a_h", "b_h"):
        src = getattr(px, name)
        dst = getattr(out, name)
        c = min(cpml, ny // 2)
        dst[:c] = src[:c]
        dst[ny - c:] = src[px.n - c:]
    return out


def problem_from_files(jsonfile: str, seismic: bool = True, workdir: str | None = None,
                       density_method: str = "legacy", source_convention: str = "legacy",
                       receivers=None) -> Problem:
    """Read the project JSON and the files Model.build wrote (classes.py:
    393-547) -- the same inputs the reference binary reads."""
    wd = workdir or os.path.dirname(os.path.abspath(jsonfile))
    P = lambda f: os.path.join(wd, f)
    with open(jsonfile) as fh:
        prj = json.load(fh)
    D = prj["Domain"]
    dim = float(D["dim"])
    cpml = int(D["cpml"])
    nx, nz = int(D["nx"]) + 2 * cpml, int(D["nz"]) + 2 * cpml
    ny = int(D["ny"]) + 2 * cpml if dim == 2.5 else 1
    dx, dz = float(D["dx"]), float(D["dz"])
    dy = float(D["dy"]) if dim == 2.5 else 1.0
    sec = prj["Seismic" if seismic else "Electromagnetic"]
    S = sec["Source"]
    nt = int(S["time_steps"])
    dt = float(S["dt"])
    shape2 = (nz, nx)
    coef = {}
    names = STIFFNESS if seismic else PERMITTIVITY + CONDUCTIVITY
    for n in names:
        f = P(f"{n}.dat")
        if os.path.exists(f):
            coef[n] = read_record(f, shape=shape2)
        else:
            coef[n] = np.zeros(shape2)
    if seismic and dim == 2.0 and os.path.exists(P("density_gradient.dat")) and "rho" in coef:
        # (nx, nz) interior factor written by airsurf (definitions.py:553-555)
        grad = read_record(P("density_gradient.dat")).reshape(nx - 2 * cpml, nz - 2 * cpml)
        coef["rho"] = _apply_density_gradient(coef["rho"], grad, cpml)
    gamma = {}
    if seismic:
        for a in ("x", "y", "z"):
            f = P(f"gamma_{a}.dat")
            if os.path.exists(f):
                gamma[a] = read_record(f, shape=shape2)
    # C-PML: 2-D fields of the current writer, else legacy 1-D stream files
    if os.path.exists(P("kappa_cpml.dat")):
        flds = [read_record(P(f"{q}{h}_cpml.dat"), shape=shape2)
                for h in ("", "_half") for q in ("kappa", "acoef", "bcoef")]
        px, pz = profiles_from_2d(flds[0], flds[1], flds[2], flds[3], flds[4], flds[5], cpml)
        prof = {"x": px, "z": pz}
        if dim == 2.5:
            prof["y"] = _y_profile_like(px, ny, cpml)
    else:
        prof = {}
        for a, n in (("x", nx), ("y", ny), ("z", nz)):
            if a == "y" and dim != 2.5:
                continue
            vals = [read_stream(P(f"{q}{a}{h}_cpml.dat"), n)
                    for h in ("", "_half") for q in ("kappa", "acoef", "bcoef")]
            prof[a] = AxisProfile(*vals)
    # sources
    stem = "seismicsource" if seismic else "electromagneticsource"
    force = []
    for a in ("x", "y", "z"):
        f = P(f"{stem}{a}.dat")
        force.append(read_record(f)[:nt] if os.path.exists(f) else None)
    if dim != 2.5:
        force[1] = None
    # stress-rate (moment-tensor) series, sourcefunction.py:358-363
    moment = [None] * 6
    if seismic:
        for q, comp in enumerate(MOMENT_COMPONENTS):
            f = P(f"seismicsource{comp}.dat")
            if os.path.exists(f):
                m = read_record(f)[:nt]
                if np.any(m):
                    moment[q] = m
    xind = int(_jget(S, "xind", default=int(float(S["x"]) / dx)))
    zind = int(_jget(S, "zind", default=int(float(S["z"]) / dz)))
    yind = int(_jget(S, "yind", default=0)) if dim == 2.5 else 0
    src = source_index(xind, yind, zind, cpml, dim, source_convention)
    rec = np.zeros((0, 3), dtype=np.int32)
    if receivers is not None:
        xyz = load_receivers(receivers) if isinstance(receivers, str) else receivers
        rec = receiver_indices(xyz, dx, dy, dz, cpml, dim)
    # initial conditions (Model.run writes them as .T records, classes.py:882-898)
    kind = ("elastic" if seismic else "em") + ("25" if dim == 2.5 else "2")
    init = {}
    for a in ("x", "y", "z"):
        f = P(f"initialcondition{'V' if seismic else 'E'}{a}.dat")
        name = ("v" if seismic else "E") + a
        if os.path.exists(f) and name in FIELDS[kind]:
            arr = read_record(f, shape=(nz, ny, nx) if dim == 2.5 else shape2)
            if np.any(arr):
                init[name] = arr
    p = Problem(kind, nx, ny, nz, dx, dy, dz, dt, nt, coef, prof,
                [Shot(src=src, force=tuple(force), rec=rec, moment=tuple(moment))], gamma,
                [init] if init else None)
    p.density_method = density_method
    p.src_index = (xind, yind, zind)
    p.cpml = cpml
    p.dim = dim
    p.validate()
    return p


# --------------------------------------------------------------------------
# problem from in-memory objects (Model / Domain attributes of the reference)
# --------------------------------------------------------------------------
def _paint(vals, geometry, cpml):
    """coef[x, z] = table[label[x, z]], edge-extended into the C-PML
    (tensor2dat / atten2dat, classes.py:608-750); returned as (nz, nx)."""
    a = np.asarray(vals, dtype=np.float64)[np.asarray(geometry)]   # (nx, nz)
    return np.pad(a, cpml, mode="edge").T.copy()


def air_density_gradient(geometry, air_label, N: int = 2) -> np.ndarray:
    """Density factor near air (``airsurf``, definitions.py:475-557) on a
    (nx, nz) label image: a non-air cell next to air (across either axis)
    gets 0.15, the cell one further in 0.15 + 0.85/2, and so on for N
    levels, each level halving what is left to 1; every other cell 1.  The
    nearest level wins.  As in the reference, an air label of None or 0
    (air listed first) means no gradient.  Returned as (nx, nz)."""
    g = np.asarray(geometry)
    if not air_label:
        return np.ones(g.shape)
    solid = (g != air_label).astype(np.int64)
    pct = [0.15]
    for _ in range(1, N):
        pct.append(pct[-1] + (1.0 - pct[-1]) / 2.0)
    out = np.zeros(g.shape)
    for level in range(N - 1, -1, -1):
        hit = np.zeros(g.shape)
        for ax in (0, 1):
            d = np.diff(solid, axis=ax)          # +1: air -> solid, -1: solid -> air
            zero = np.zeros_like(np.take(d, [0], axis=ax))
            up = np.concatenate([zero, d], axis=ax)     # +1 on the solid cell after air
            dn = np.concatenate([d, zero], axis=ax)     # -1 on the solid cell before air
            for _ in range(level):                       # walk `level` cells into the solid
                up = np.concatenate([zero, up], axis=ax)[tuple(
                    slice(0, -1) if a == ax else slice(None) for a in range(2))]
                dn = np.concatenate([dn, zero], axis=ax)[tuple(
                    slice(1, None) if a == ax else slice(None) for a in range(2))]
            hit = hit + up - dn
        out[hit > 0] = pct[level]
    out[out == 0] = 1.0
    return out


def _apply_density_gradient(rho_padded: np.ndarray, grad: np.ndarray, cpml: int) -> np.ndarray:
    """rho (nz, nx padded) scaled by the (nx, nz) interior gradient, then the
    C-PML rows and columns refilled from the scaled edge (the legacy order,
    legacy/v0.0.2/fdtd/seismicFDTD2D.f95:53-80)."""
    r = rho_padded.copy()
    nz, nx = r.shape
    inner = r[cpml:nz - cpml, cpml:nx - cpml] * np.asarray(grad, dtype=np.float64).T
    return np.pad(inner, cpml, mode="edge")


def _sponge(eta, cpml, NPS, eta_max):
    """Additive sponge of definitions.py:1001-1047 on an (nx, nz) array:
    corner products plus full-length edge rows and columns."""
    d = np.arange(cpml) / (cpml - 1)
    w = eta_max * (1.0 - d) ** NPS
    corner = eta_max * np.outer((1.0 - d) ** NPS, (1.0 - d) ** NPS)
    eta[:cpml, :cpml] += corner
    eta[-cpml:, :cpml] += corner[::-1, :]
    eta[:cpml, -cpml:] += corner[:, ::-1]
    eta[-cpml:, -cpml:] += corner[::-1, ::-1]
    eta[:cpml, :] += w[:, None]
    eta[-cpml:, :] += w[::-1, None]
    eta[:, :cpml] += w[None, :]
    eta[:, -cpml:] += w[None, ::-1]
    return eta


def problem_from_objects(model, domain, density_method: str | None = None,
                         source_convention: str = "legacy", receivers=None,
                         material=None) -> Problem:
    """Host setup from a Model / Domain pair after ``Model.build``
    (classes.py:393-547): tables painted on domain.geometry, C-PML profiles
    from the reference formulas, the model's source series and dt."""
    dim = float(domain.dim)
    cpml = int(domain.cpml)
    nx0, nz0 = int(domain.nx), int(domain.nz)
    nx, nz = nx0 + 2 * cpml, nz0 + 2 * cpml
    ny = int(domain.ny) + 2 * cpml if dim == 2.5 else 1
    dx, dz = float(domain.dx), float(domain.dz)
    dy = float(domain.dy) if dim == 2.5 else 1.0
    geom = np.asarray(domain.geometry)
    seismic = bool(model.is_seismic)
    nt = int(model.time_steps)
    f0 = float(model.f0)
    coef, gamma = {}, {}
    if seismic:
        tab = model.stiffness_coefficients
        for n in STIFFNESS:
            coef[n] = _paint(tab[n].values, geom, cpml)
        # density factor near air (classes.py:529-534: 2-D, more than one material)
        if dim == 2.0 and int(getattr(domain, "nmats", 1)) > 1:
            # the material list of the Material object (definitions.py:502-505)
            mlist = getattr(material, "material", None)
            names = [str(v).strip().lower() for v in
                     (np.asarray(mlist).ravel() if mlist is not None else [])]
            air = names.index("air") if "air" in names else None
            grad = air_density_gradient(geom, air, int(getattr(model, "air_gradient_integer", 2) or 2))
            coef["rho"] = _apply_density_gradient(coef["rho"], grad, cpml)
        att = getattr(model, "attenuation_coefficients", None)
        if att is not None:
            for a in ("x", "y", "z"):
                col = f"gamma_{a}"
                if col in att:
                    g = np.asarray(att[col].values, dtype=np.float64) * f0
                    arr = np.pad(g[geom], cpml, mode="edge")
                    gmax = getattr(model, "gamma_max", None)
                    if gmax:
                        arr = _sponge(arr, cpml, float(getattr(domain, "NPS", 1.0)), float(gmax))
                    gamma[a] = arr.T.copy()
    else:
        for n in PERMITTIVITY:
            coef[n] = _paint(np.real(model.permittivity_coefficients[n].values), geom, cpml)
        for n in CONDUCTIVITY:
            coef[n] = _paint(np.real(model.conductivity_coefficients[n].values), geom, cpml)
    dt = float(model.dt)
    NP, NPA = float(domain.NP), float(domain.NPA)
    Rcoef, kmax = float(domain.Rcoef), float(domain.kappa_max)
    ams = float(domain.alpha_max_scalar)
    vsf = float(getattr(model, "velocity_scaling_factor", 1.0))
    prof = {}
    if seismic:
        tab = model.stiffness_coefficients
        mats = np.unique(geom)
        vx = max(christoffel_vmax(tab.loc[m], (1, 0, 0)) for m in mats)
        vzv = max(christoffel_vmax(tab.loc[m], (0, 0, 1)) for m in mats)
        am = alpha_max_seismic(ams, f0)
        prof["x"] = axis_profile(nx, cpml, dx, sigma_max_seismic(vx, cpml, NP, Rcoef, vsf), am, kmax,
                                 NP, NPA, dt, True)
        prof["z"] = axis_profile(nz, cpml, dz, sigma_max_seismic(vzv, cpml, NP, Rcoef, vsf), am, kmax,
                                 NP, NPA, dt, True)
        if dim == 2.5:
            prof["y"] = _y_profile_like(prof["x"], ny, cpml)
    else:
        dmin = min(dx, dy, dz) if dim == 2.5 else min(dx, dz)
        am = alpha_max_em(ams, f0)
        sm = sigma_max_em(cpml, dmin, NP, Rcoef)
        prof["x"] = axis_profile(nx, cpml, dx, sm, am, kmax, NP, NPA, dt, False)
        prof["z"] = axis_profile(nz, cpml, dz, sm, am, kmax, NP, NPA, dt, False)
        if dim == 2.5:
            prof["y"] = _y_profile_like(prof["x"], ny, cpml)
    force = [getattr(model, f"sourcefunction_{a}", None) for a in ("x", "y", "z")]
    if all(f is None for f in force):
        from .sources import point_force
        force = list(point_force(nt, dt, f0, float(model.source_amplitude),
                                 str(model.source_wavelet), float(model.phi or 0),
                                 float(model.theta or 0), float(getattr(model, "psi", 0) or 0)))
    if dim != 2.5:
        force[1] = None
    moment = [None] * 6
    if seismic:  # pointsource sets source_sigma_* (sourcefunction.py:365-370)
        for q, comp in enumerate(MOMENT_COMPONENTS):
            m = getattr(model, f"source_sigma_{comp}", None)
            if m is not None and np.any(m):
                moment[q] = np.asarray(m, dtype=np.float64)
    xind, zind = int(float(model.x) / dx), int(float(model.z) / dz)
    yind = int(float(model.y) / dy) if dim == 2.5 else 0
    src = source_index(xind, yind, zind, cpml, dim, source_convention)
    rec = np.zeros((0, 3), dtype=np.int32)
    if receivers is not None:
        xyz = load_receivers(receivers) if isinstance(receivers, str) else receivers
        rec = receiver_indices(xyz, dx, dy, dz, cpml, dim)
    init = {}
    fields = FIELDS[("elastic" if seismic else "em") + ("25" if dim == 2.5 else "2")]
    for a, n in zip(("x", "y", "z"), ("x", "y", "z")):
        ic = getattr(model, f"initialcondition_{a}", None)
        name = ("v" if seismic else "E") + n
        if ic is not None and name in fields and np.any(ic):
            arr = np.asarray(ic, dtype=np.float64)
            init[name] = arr.T.copy() if dim != 2.5 else np.transpose(arr, (2, 1, 0)).copy()
    kind = ("elastic" if seismic else "em") + ("25" if dim == 2.5 else "2")
    p = Problem(kind, nx, ny, nz, dx, dy, dz, dt, nt, coef, prof,
                [Shot(src=src, force=tuple(force), rec=rec, moment=tuple(moment))], gamma,
                [init] if init else None)
    p.density_method = density_method or getattr(model, "density_method", "legacy") or "legacy"
    p.src_index = (xind, yind, zind)
    p.cpml = cpml
    p.dim = dim
    p.validate()
    return p


# --------------------------------------------------------------------------
# run
# --------------------------------------------------------------------------
def write_trace_csv(trace2d: np.ndarray, basefile: str, channel: str, source_xyz) -> str:
    """nt x nrec CSV without header, named like Array.save
    ('{base}-{channel}-{sx}-{sy}-{sz}.csv', arraybuild.py:1031-1056)."""
    import pandas as pd
    src = "-".join(str(v) for v in np.asarray(source_xyz, dtype=np.float64))
    fn = f"{basefile}-{channel}-{src}.csv"
    # the reference's writer (DataFrame.to_csv, shortest round-trip floats),
    # so the files are byte-identical to Array.save's
    pd.DataFrame(np.asarray(trace2d)).to_csv(fn, header=False, index=False)
    return fn


# Extents (z, [y,] x slices) of each snapshot channel as the legacy solver
# writes it (write_image2 / write_image3 calls, electromagfdtd.f95:323-324,
# :1086-1088; the elastic velocities are written on the full padded grid).
def _legacy_extent(kind: str, ch: str):
    if kind == "em2":
        return {"Ex": (slice(None), slice(0, -1)), "Ez": (slice(0, -1), slice(None))}[ch]
    if kind == "em25":
        return {"Ex": (slice(None), slice(None), slice(0, -1)),
                "Ey": (slice(None), slice(0, -1), slice(None)),
                "Ez": (slice(0, -1), slice(None), slice(None))}[ch]
    return (slice(None),) * (3 if kind.endswith("25") else 2)


class SnapshotWriter:
    """Snapshots of a running Simulation without stalling it.

    At a snapshot step the compute stream records an event; a copy stream
    waits for it, converts each channel's field into a device staging buffer
    of the file's precision (FP32 ``sngl`` records by default,
    readwrite_routines.f95:489-494) and copies that into a pinned host
    buffer, while the compute stream waits only for the device-side staging
    copy before its next steps.  A writer thread writes each record once its
    copy has landed.  Two pinned buffers per channel alternate, so the host
    write of one snapshot overlaps the device work of the next.

    ``layout``: 'padded' writes every channel on the full padded grid (what
    the reference's reader ``read_dat`` reshapes, definitions.py:77-136);
    'legacy' writes each channel at the extent the legacy solver writes
    (EM: Ex (nx-1) x nz, Ez nx x (nz-1), ...).  ``block``: collect the frames
    for one EM25 block file instead of per-step records."""

    def __init__(self, sim, p: Problem, outdir: str, single: bool, legacy_names: bool,
                 layout: str = "padded", block: bool = False):
        import queue
        import threading
        import torch
        if layout not in ("padded", "legacy"):
            raise ValueError("snapshot layout must be 'padded' or 'legacy'")
        self.sim, self.p, self.outdir = sim, p, outdir
        self.single, self.legacy_names, self.block = single, legacy_names, block
        self.comps = TRACE_COMPONENTS[p.kind]
        self.names = FIELDS[p.kind][: len(self.comps)]
        si = getattr(p, "src_index", None) or p.shots[0].src
        self.idx = (si[0], si[2]) if not p.is3d else tuple(si)
        self.ext = {ch: (_legacy_extent(p.kind, ch) if layout == "legacy"
                         else (slice(None),) * len(p.field_shape)) for ch in self.comps}
        dt = torch.float32 if (single or block) else torch.float64
        dev = sim.device
        views = {ch: sim.shots[0][2][n][self.ext[ch]] for ch, n in zip(self.comps, self.names)}
        self.views = views
        self.staging = {ch: torch.empty(v.shape, dtype=dt, device=dev) for ch, v in views.items()}
        self.host = [{ch: torch.empty(v.shape, dtype=dt, pin_memory=True) for ch, v in views.items()}
                     for _ in range(2)]
        self.copy_stream = torch.cuda.Stream(device=dev)
        self.free = [threading.Semaphore(1), threading.Semaphore(1)]
        self.q = queue.Queue()
        self.frames = {ch: [] for ch in self.comps}
        self.count = 0
        self.error = None
        self.slot = 0
        self.thread = threading.Thread(target=self._loop, daemon=True)
        self.thread.start()

    def capture(self, step: int):
        """Snapshot of the state after ``step`` steps (1-based file index)."""
        import torch
        if self.error is not None:
            raise self.error
        slot = self.slot
        self.slot ^= 1
        self.free[slot].acquire()               # the writer is done with this buffer
        compute = torch.cuda.current_stream(self.sim.device)
        done = torch.cuda.Event()
        done.record(compute)
        staged = torch.cuda.Event()
        landed = torch.cuda.Event()
        with torch.cuda.stream(self.copy_stream):
            self.copy_stream.wait_event(done)
            for ch in self.comps:
                self.staging[ch].copy_(self.views[ch])
            staged.record(self.copy_stream)
            for ch in self.comps:
                self.host[slot][ch].copy_(self.staging[ch], non_blocking=True)
            landed.record(self.copy_stream)
        compute.wait_event(staged)              # later steps overwrite the fields
        self.q.put((slot, step, landed))

    def _loop(self):
        while True:
            item = self.q.get()
            if item is None:
                return
            slot, step, landed = item
            try:
                landed.synchronize()
                for ch in self.comps:
                    a = self.host[slot][ch].numpy()
                    if self.block:
                        self.frames[ch].append(a.copy())
                        continue
                    fn = os.path.join(self.outdir, snapshot_name(ch, step, self.idx, self.legacy_names))
                    write_snapshot(fn, a, single=self.single)
                    self.count += 1
            except Exception as e:  # surfaced by capture / close
                self.error = e
            finally:
                self.free[slot].release()

    def close(self):
        """Wait for every pending record; write the EM25 block file if any."""
        self.q.put(None)
        self.thread.join()
        if self.error is not None:
            raise self.error
        if self.block and self.frames[self.comps[0]]:
            # steps snapshot_every, 2*snapshot_every, ... stored as block steps 1, 2, ...
            write_em25_block(os.path.join(self.outdir, em25_block_name(1, self.idx)),
                             {ch: np.stack(v) for ch, v in self.frames.items()}, first_step=1)
            self.count = 1
        return self.count


def _execute(p: Problem, precision: str, snapshot_every: int, snapshot_single: bool,
             snapshot_dir: str, graph_block: int, legacy_names: bool, snapshot_format: str = "dat",
             snapshot_layout: str = "padded"):
    """Run one problem; write snapshots every ``snapshot_every`` steps, as one
    record per field and step (``"dat"``), or for 2.5-D EM optionally as one
    int16 block file of the E fields (``"em25block"``).  Snapshot copies and
    file writes overlap the following steps (SnapshotWriter)."""
    from .solver import Simulation
    from ._hip import UnstableError
    if snapshot_format not in ("dat", "em25block"):
        raise ValueError("snapshot_format must be 'dat' or 'em25block'")
    if snapshot_format == "em25block" and p.kind != "em25":
        raise ValueError("em25block snapshots apply to 2.5-D EM runs")
    sim = Simulation(p, precision=precision)
    nsnap = 0
    if snapshot_every and snapshot_every > 0:
        w = SnapshotWriter(sim, p, snapshot_dir, snapshot_single, legacy_names, snapshot_layout,
                           block=snapshot_format == "em25block")
        try:
            done = 0
            while done < p.nt:
                n = min(snapshot_every, p.nt - done)
                sim.run(n, graph_block=graph_block if n >= graph_block else 0)
                done += n
                w.capture(done)
        finally:
            nsnap = w.close()
    else:
        sim.run(graph_block=graph_block)
    unstable = -1
    try:
        sim.check()
    except UnstableError:
        unstable = sim.unstable[0]
    tr = sim.traces()[0]
    sim.close()
    return tr, unstable, nsnap


def run_files(jsonfile: str, seismic: bool = True, density_method: str = "legacy",
              precision: str = "fp64", snapshot_every: int = 0, snapshot_single: bool = True,
              receivers=None, graph_block: int = 50, raise_unstable: bool = True,
              legacy_names: bool = False, snapshot_format: str = "dat",
              snapshot_layout: str = "padded") -> RunResult:
    """The reference binary's call (classes.py:910-924) on the same files."""
    p = problem_from_files(jsonfile, seismic, density_method=density_method, receivers=receivers)
    wd = os.path.dirname(os.path.abspath(jsonfile))
    tr, unstable, nsnap = _execute(p, precision, snapshot_every, snapshot_single, wd, graph_block,
                                   legacy_names, snapshot_format, snapshot_layout)
    csvs = []
    if receivers is not None and len(p.shots[0].rec):
        base = os.path.splitext(jsonfile)[0]
        with open(jsonfile) as fh:
            S = json.load(fh)["Seismic" if seismic else "Electromagnetic"]["Source"]
        sxyz = (S.get("x", 0), S.get("y", 0), S.get("z", 0))
        for c, ch in enumerate(TRACE_COMPONENTS[p.kind]):
            csvs.append(write_trace_csv(tr[:, c, :], base, ch, sxyz))
    if unstable >= 0 and raise_unstable:
        from ._hip import UnstableError
        raise UnstableError(3, f"code became unstable and blew up at step {unstable}")
    return RunResult(p, tr, TRACE_COMPONENTS[p.kind], unstable, csvs, nsnap)


def run(model, domain=None, material=None, *, num_threads: int = 1, jsonfile: str | None = None,
        precision: str = "fp64", snapshot_every: int = 0, snapshot_single: bool = True,
        receivers=None, graph_block: int = 50, density_method: str | None = None,
        raise_unstable: bool = True) -> RunResult:
    """``Model.run(num_threads=1, jsonfile=None)`` (classes.py:878-924) with the
    solver on the GPU.  ``model`` is a Model built by ``Model.build`` (or a
    project JSON path, which takes the file route).  ``num_threads`` is
    accepted for signature compatibility; the host does only setup."""
    del num_threads
    if isinstance(model, str):
        return run_files(model, True, density_method or "legacy", precision, snapshot_every,
                         snapshot_single, receivers, graph_block, raise_unstable)
    if domain is None:
        raise ValueError("run(model, domain): the domain is required with a Model object")
    p = problem_from_objects(model, domain, density_method, receivers=receivers, material=material)
    wd = os.path.dirname(os.path.abspath(jsonfile or getattr(model, "project_file", None) or "."))
    tr, unstable, nsnap = _execute(p, precision, snapshot_every, snapshot_single, wd, graph_block,
                                   False)
    if unstable >= 0 and raise_unstable:
        from ._hip import UnstableError
        raise UnstableError(3, f"code became unstable and blew up at step {unstable}")
    return RunResult(p, tr, TRACE_COMPONENTS[p.kind], unstable, [], nsnap)


def main(argv=None) -> int:
    """``python -m seidart_amd.run project.json seismic=true density_method=geometric
    [precision=fp32] [snapshot_every=N] [single=true] [receivers=file.xyz]
    [snapshot_format=dat|em25block] [snapshot_layout=padded|legacy] [legacy_names=false]``"""
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv:
        print(main.__doc__, file=sys.stderr)
        return 2
    jsonfile = argv[0]
    kw = dict(a.split("=", 1) for a in argv[1:])
    seismic = kw.get("seismic", "true").lower() == "true"
    dm = kw.get("density_method", "legacy").lower()
    if dm not in DENSITY_METHODS:
        print(f"density_method must be one of {DENSITY_METHODS}", file=sys.stderr)
        return 2
    try:
        res = run_files(jsonfile, seismic, dm, kw.get("precision", "fp64"),
                        int(kw.get("snapshot_every", "1")), kw.get("single", "true") == "true",
                        kw.get("receivers"), int(kw.get("graph_block", "50")),
                        legacy_names=kw.get("legacy_names", "false").lower() == "true",
                        snapshot_format=kw.get("snapshot_format", "dat"),
                        snapshot_layout=kw.get("snapshot_layout", "padded"))
    except Exception as e:  # report like the reference solver, but with a status
        print(f"seidart_amd.run: {e}", file=sys.stderr)
        return 1
    print(f"seidart_amd.run: {res.problem.kind} {res.problem.nx}x{res.problem.ny}x{res.problem.nz}"
          f" nt={res.problem.nt} snapshots={res.snapshot_files} csv={len(res.csv_files)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
