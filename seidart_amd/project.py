"""Project-file host route: the solver inputs from the project JSON and its
geometry alone, with no SeidarT install.

The reference prepares a run in Python before it calls its solver: it loads
the project (``loadproject``, src/seidart/routines/definitions.py:299-472),
reads the geometry image (``image2int``, src/seidart/routines/prjbuild.py:
303-339), and ``Model.build`` (src/seidart/routines/classes.py:393-547)
computes dt from the largest phase velocity (:469-492,
``compute_max_velocities`` :1070-1120), the source series (``pointsource``,
sourcefunction.py:239-383) and the C-PML coefficients (``cpmlcompute`` /
``cpml_parameters``, definitions.py:560-998), and writes the painted
coefficient files (``tensor2dat`` / ``atten2dat``, classes.py:608-750), the
density gradient near air (``airsurf``, definitions.py:475-557) and the
source files.  This module does all of that from the JSON (whose material
tables ``save_to_json`` filled, classes.py:787-836) and ``geometry.dat`` or
the project image:

* :func:`load_project` -> (domain, material, seismic, electromag) objects with
  the attributes loadproject sets;
* :func:`build` -> the attributes Model.build adds (xind.., CFL, dt, source
  series), and with ``write=True`` the files it writes, in its formats;
* :func:`problem_from_project` -> a :class:`~seidart_amd.problem.Problem`
  (through :func:`seidart_amd.run.problem_from_objects`);
* :func:`run_project` -> the run, like :func:`seidart_amd.run.run_files`.

Material-property models (stiffness from temperature, porosity, fabric;
``get_seismic`` / ``get_perm``) are not part of this route: the coefficient
tables come from the project file.
"""
from __future__ import annotations

import json
import os
from types import SimpleNamespace

import numpy as np

from .cpml import CLIGHT, EPS0
from .fortranio import read_record, write_record
from .problem import CONDUCTIVITY, PERMITTIVITY, STIFFNESS

__all__ = ["load_project", "image2int", "read_geometry", "max_velocity", "model_dt", "build",
           "write_build_files", "save_to_json", "cpml_fields_2d", "problem_from_project",
           "run_project", "ATTENUATION"]

ATTENUATION = ("gamma_x", "gamma_y", "gamma_z", "gamma_yz", "gamma_xz", "gamma_xy",
               "reference_frequency")
MATERIAL_KEYS = ("id", "name", "rgb", "temperature", "density", "porosity", "water_content",
                 "is_anisotropic", "euler_angles")


class _Table:
    """A per-material coefficient table: ``tab[col].values`` (one value per
    material id) and ``tab.loc[m][col]``, the access the painting and the
    Christoffel speeds use."""

    def __init__(self, cols, rows):
        self.columns = list(cols)
        self._a = np.asarray(rows, dtype=np.float64).reshape(len(rows), len(self.columns))

    def __getitem__(self, col):
        return SimpleNamespace(values=self._a[:, self.columns.index(col)].copy())

    def __contains__(self, col):
        return col in self.columns

    @property
    def loc(self):
        tab = self

        class _Loc:
            def __getitem__(self, m):
                return dict(zip(tab.columns, tab._a[int(m)]))
        return _Loc()

    @property
    def index(self):
        return range(self._a.shape[0])

    def values_of(self, col):
        return self._a[:, self.columns.index(col)]


def _rows(entries, cols):
    out = []
    for e in entries:
        if all(c in e for c in cols):
            out.append([float(e[c]) for c in cols])
        else:  # positional, as loadproject reads them (the first value is the id)
            out.append([float(v) for v in list(e.values())[1:1 + len(cols)]])
    return out


def image2int(imfile: str):
    """Label image of the project PNG (prjbuild.py:303-339): each pixel's
    colour keyed as 65536 R + 255 G + B, labels 0.. in ascending key order.
    Returns (labels (rows, cols) int, rgb (nlabels, 3))."""
    import matplotlib.image as mpimg
    img = mpimg.imread(imfile)
    key = np.array(65536 * img[:, :, 0] + 255 * img[:, :, 1] + img[:, :, 2])
    uni, inv = np.unique(key, return_inverse=True)
    labels = inv.reshape(key.shape).astype(int)
    flat = img[:, :, :3].reshape(-1, 3)
    first = np.array([np.flatnonzero(inv.ravel() == u)[0] for u in range(uni.size)])
    rgb = flat[first].astype(np.float64)
    if rgb.max() <= 1.0:
        rgb = (rgb * 255).astype(int)
    return labels, rgb


def read_geometry(domain, workdir: str, prefer: str = "dat") -> np.ndarray:
    """The (nx, nz) label array: ``geometry.dat`` (the image2int record of the
    (nz, nx) label image, int64) when present and ``prefer == 'dat'``, else
    the project image (loadproject: ``image2int(...)``, transposed)."""
    nx, nz = int(domain.nx), int(domain.nz)
    gdat = os.path.join(workdir, "geometry.dat")
    if prefer == "dat" and os.path.exists(gdat):
        raw = read_record(gdat, dtype=np.int64)
        if raw.size != nx * nz:
            raw = read_record(gdat, dtype=np.int32)
        if raw.size != nx * nz:
            raise ValueError(f"geometry.dat holds {raw.size} labels, the domain {nx} x {nz}")
        return raw.reshape(nz, nx).T.astype(int)
    im = domain.imfile
    if not im:
        raise ValueError("no geometry.dat and no image_file in the project")
    path = im if os.path.isabs(im) else os.path.join(workdir, im)
    labels, _ = image2int(path)
    if labels.shape != (nz, nx):
        raise ValueError(f"image {labels.shape} does not match the domain ({nz}, {nx})")
    return labels.T.astype(int)


def load_project(jsonfile: str, geometry=None, prefer_geometry: str = "dat"):
    """Domain, Material and the seismic and electromagnetic Model attributes
    of a project file (definitions.py:299-472).  ``geometry``: an (nx, nz)
    label array to use instead of the file's."""
    with open(jsonfile) as fh:
        data = json.load(fh)
    wd = os.path.dirname(os.path.abspath(jsonfile))
    D = data["Domain"]
    domain = SimpleNamespace(
        dim=float(D["dim"]), nx=int(D["nx"]), ny=int(D.get("ny", 1) or 1), nz=int(D["nz"]),
        dx=float(D["dx"]), dy=float(D.get("dy", 1) or 1), dz=float(D["dz"]), cpml=int(D["cpml"]),
        nmats=int(D["nmats"]), alpha_max_scalar=float(D["alpha_max_scalar"]),
        kappa_max=float(D["kappa_max"]), NP=float(D["NP"]), NPA=float(D["NPA"]),
        Rcoef=float(D["Rcoef"]), imfile=D.get("image_file"), NPS=float(D.get("NPS", 1.0) or 1.0))
    domain.geometry = np.asarray(geometry).astype(int) if geometry is not None else \
        read_geometry(domain, wd, prefer_geometry)
    names = []
    for m in data.get("Materials", []):
        names.append(str(m.get("name", list(m.values())[1] if len(m) > 1 else "")))
    material = SimpleNamespace(material=np.array(names, dtype=object), material_list=data.get("Materials", []))
    models = []
    for sec, seismic in (("Seismic", True), ("Electromagnetic", False)):
        S = data.get(sec, {}).get("Source", {})
        m = SimpleNamespace(
            dt=float(S.get("dt", 0) or 0), time_steps=int(S.get("time_steps", 0) or 0),
            x=float(S.get("x", 0) or 0), y=float(S.get("y", 0) or 0), z=float(S.get("z", 0) or 0),
            f0=float(S.get("source_frequency", 1.0) or 1.0),
            theta=float(S.get("x-z_rotation", 0) or 0), phi=float(S.get("x-y_rotation", 0) or 0),
            psi=float(S.get("y-z_rotation", 0) or 0), source_amplitude=float(S.get("amplitude", 1.0)),
            source_wavelet=S.get("source_wavelet", "gaus1"),
            source_type=S.get("source_type", "ac" if seismic else "ps"),
            is_seismic=seismic, project_file=jsonfile, CFL=1 / np.sqrt(2),
            density_method="geometric", air_gradient_integer=2, velocity_scaling_factor=1.0,
            gamma_max=None, use_multimodal=False, center=False, use_broadband=False,
            broadband_fmin=None, broadband_fmax=None, source_n_octaves=2,
            initialcondition_x=None, initialcondition_y=None, initialcondition_z=None)
        sec_d = data.get(sec, {})
        if seismic:
            m.stiffness_coefficients = _Table(STIFFNESS, _rows(sec_d.get("Stiffness_Coefficients", []), STIFFNESS))
            m.attenuation_coefficients = _Table(ATTENUATION, _rows(sec_d.get("Attenuation", []), ATTENUATION))
        else:
            m.permittivity_coefficients = _Table(PERMITTIVITY,
                                                 _rows(sec_d.get("Permittivity_Coefficients", []), PERMITTIVITY))
            m.conductivity_coefficients = _Table(CONDUCTIVITY,
                                                 _rows(sec_d.get("Conductivity_Coefficients", []), CONDUCTIVITY))
        models.append(m)
    return domain, material, models[0], models[1]


_DIRS = {"x": (1, 0, 0), "y": (0, 1, 0), "z": (0, 0, 1), "xy": (1, 1, 0), "xz": (1, 0, 1),
         "yz": (0, 1, 1)}


def max_velocity(model, dim: float) -> float:
    """The speed dt is set from (classes.py:469-481): seismic, the largest
    Christoffel phase velocity of any material along x, z, xz (2-D) or x, y,
    z, xy, xz, yz (2.5-D) (compute_max_velocities, :1070-1120, with the
    reference's Christoffel matrix); EM, c / sqrt(min over materials of e11,
    e22, e33)."""
    if not model.is_seismic:
        tab = model.permittivity_coefficients
        emin = min(float(np.min(tab.values_of(n))) for n in ("e11", "e22", "e33"))
        return float(CLIGHT / np.sqrt(emin))
    from .run import christoffel_vmax
    dirs = ("x", "z", "xz") if dim == 2 else ("x", "y", "z", "xy", "xz", "yz")
    tab = model.stiffness_coefficients
    return max(christoffel_vmax(tab.loc[m], _DIRS[d]) for m in tab.index for d in dirs)


def model_dt(model, domain) -> float:
    """dt = CFL / (vmax sqrt(sum 1/d^2)) (classes.py:469-485); the CFL is
    capped at 1/sqrt(3) in 2.5-D (:419-420)."""
    cfl = float(model.CFL)
    if domain.dim == 2.5:
        cfl = min(cfl, 1 / np.sqrt(3))
    if domain.dim in (2.5, 3):
        denom = np.sqrt((1 / domain.dx) ** 2 + (1 / domain.dy) ** 2 + (1 / domain.dz) ** 2)
    else:
        denom = np.sqrt((1 / domain.dx) ** 2 + (1 / domain.dz) ** 2)
    return float(cfl / (max_velocity(model, domain.dim) * denom))


def build(model, domain, material=None, recompute_dt: bool = True, write: bool = False,
          workdir: str | None = None):
    """What Model.build adds before the solver runs (classes.py:393-547):
    source cell indices, the 2.5-D CFL cap, dt (unless ``recompute_dt`` is
    False: then the project's dt is kept), the source series (force and
    stress-rate components).  ``write``: also write the files Model.build and
    pointsource write, into ``workdir`` (write_build_files)."""
    from .sources import pointsource
    model.xind = int(model.x / domain.dx)
    model.yind = int(model.y / domain.dy)
    model.zind = int(model.z / domain.dz)
    if domain.dim == 2.5:
        model.CFL = min(float(model.CFL), 1 / np.sqrt(3))
    if recompute_dt or not model.dt:
        model.dt = model_dt(model, domain)
    force, moment = pointsource(
        model.time_steps, model.dt, model.f0, model.source_amplitude, model.source_wavelet,
        model.source_type, model.phi, model.theta, model.psi, model.is_seismic,
        multimodal=model.use_multimodal, broadband=model.use_broadband, fmin=model.broadband_fmin,
        fmax=model.broadband_fmax, center=model.center, num_octaves=model.source_n_octaves)
    model.sourcefunction_x, model.sourcefunction_y, model.sourcefunction_z = force
    (model.source_sigma_xx, model.source_sigma_yy, model.source_sigma_zz,
     model.source_sigma_xy, model.source_sigma_xz, model.source_sigma_yz) = moment
    if write:
        write_build_files(model, domain, material, workdir or os.path.dirname(
            os.path.abspath(model.project_file or ".")))
        if model.project_file:
            save_to_json(model, model.project_file)
    return model


def save_to_json(model, jsonfile: str) -> None:
    """Write the source settings build computed back into the project file's
    Source section (save_to_json, classes.py:787-836): dt, x, y, z, xind,
    yind, zind, source_frequency, the rotations, amplitude, wavelet and
    source type.  The coefficient tables are left as they are."""
    with open(jsonfile) as fh:
        data = json.load(fh)
    S = data["Seismic" if model.is_seismic else "Electromagnetic"]["Source"]
    S.update({"dt": model.dt, "x": model.x, "y": model.y, "z": model.z, "xind": model.xind,
              "yind": model.yind, "zind": model.zind, "source_frequency": model.f0,
              "x-z_rotation": model.theta, "x-y_rotation": model.phi,
              "amplitude": model.source_amplitude})
    if isinstance(model.source_wavelet, str):
        S["source_wavelet"] = model.source_wavelet
    if model.source_type:
        S["source_type"] = model.source_type
    with open(jsonfile, "w") as fh:
        json.dump(data, fh, indent=4)


def cpml_fields_2d(model, domain, half: bool):
    """The 2-D C-PML fields cpmlcompute writes for 2 / 2.5-D domains in its
    default 'global' mode (definitions.py:560-830 with cpml_parameters
    :862-998): sigma_max per axis from the largest Christoffel speed along
    that axis (seismic) or from c (EM), graded profiles on the slabs, sums
    (sigma) and products (kappa, alpha) in the corners before the slab loops
    overwrite kappa and alpha there; b = exp(-(sigma/kappa + alpha) dt) (EM:
    dt / eps0), a = sigma (b - 1) / (kappa sigma + kappa alpha) where
    |sigma| > 1e-6.  Returns (sigma, kappa, alpha, a, b), each (nx, nz) padded."""
    from .cpml import alpha_max_em, alpha_max_seismic, sigma_max_em, sigma_max_seismic
    from .run import christoffel_vmax
    c = int(domain.cpml)
    nxf, nzf = domain.nx + 2 * c, domain.nz + 2 * c
    distx = domain.dx * np.arange(0, c)
    distz = domain.dz * np.arange(0, c)
    if half:
        distx = distx + domain.dx / 2
        distz = distz + domain.dz / 2
    distx = (domain.dx * c - distx) / (domain.dx * c)
    distz = (domain.dz * c - distz) / (domain.dz * c)
    mats = np.unique(domain.geometry)
    vsf = float(getattr(model, "velocity_scaling_factor", 1.0))
    if model.is_seismic:
        tab = model.stiffness_coefficients
        vgx = max(christoffel_vmax(tab.loc[m], (1, 0, 0)) for m in mats)
        vgz = max(christoffel_vmax(tab.loc[m], (0, 0, 1)) for m in mats)
        sx = sigma_max_seismic(vgx, c, domain.NP, domain.Rcoef, vsf)
        sz = sigma_max_seismic(vgz, c, domain.NP, domain.Rcoef, vsf)
        amax = alpha_max_seismic(domain.alpha_max_scalar, model.f0)
    else:
        dmin = min(domain.dx, domain.dy, domain.dz) if domain.dim == 2.5 else min(domain.dx, domain.dz)
        sx = sz = sigma_max_em(c, dmin, domain.NP, domain.Rcoef)
        amax = alpha_max_em(domain.alpha_max_scalar, model.f0)
    kmax, NP, NPA, dt = float(domain.kappa_max), float(domain.NP), float(domain.NPA), float(model.dt)
    kappa = np.ones([nxf, nzf])
    alpha = np.zeros([nxf, nzf])
    sigma = np.zeros([nxf, nzf])
    a = np.zeros([nxf, nzf])
    dx_ = distx / distx.max()
    dz_ = distz / distz.max()
    rx = dx_ ** NP
    rz = dz_ ** NP
    m, n = len(dx_), len(dz_)
    for ii in range(m):
        for jj in range(n):
            sg = sx * rx[ii] + sz * rz[jj]
            kp = (1.0 + (kmax - 1.0) * rx[ii]) * (1.0 + (kmax - 1.0) * rz[jj])
            al = amax * ((1.0 - dx_[ii]) ** NPA) * ((1.0 - dz_[jj]) ** NPA)
            for I, J in ((ii, jj), (-(ii + 1), jj), (ii, -(jj + 1)), (-(ii + 1), -(jj + 1))):
                sigma[I, J], kappa[I, J], alpha[I, J] = sg, kp, al
    for ind in range(m):
        sigma[ind, m:-m] = sx * rx[ind]
        kappa[ind, :] = 1.0 + (kmax - 1.0) * rx[ind]
        alpha[ind, :] = amax * (1 - dx_[ind]) ** NPA
        sigma[-(ind + 1), m:-m] = sx * rx[ind]
        kappa[-(ind + 1), :] = 1 + (kmax - 1) * rx[ind]
        alpha[-(ind + 1), :] = amax * (1 - dx_[ind]) ** NPA
    for ind in range(n):
        sigma[n:-n, ind] = sz * rz[ind]
        kappa[:, ind] = 1.0 + (kmax - 1.0) * rz[ind]
        alpha[:, ind] = amax * (1 - dz_[ind]) ** NPA
        sigma[n:-n, -(ind + 1)] = sz * rz[ind]
        kappa[:, -(ind + 1)] = 1 + (kmax - 1) * rz[ind]
        alpha[:, -(ind + 1)] = amax * (1 - dz_[ind]) ** NPA
    if model.is_seismic:
        b = np.exp(-(sigma / kappa + alpha) * dt)
    else:
        b = np.exp(-(sigma / kappa + alpha) * (dt / EPS0))
    alpha[alpha < 0.0] = 0.0
    idx = np.abs(sigma) > 1.0e-6
    a[idx] = sigma[idx] * (b[idx] - 1) / (kappa[idx] * sigma[idx] + kappa[idx] * alpha[idx])
    return sigma, kappa, alpha, a, b


def _extend2d(vals, geometry, cpml):
    """tensor2dat's painting (classes.py:664-675): table values on the
    (nx, nz) labels, edge rows and columns repeated into the C-PML."""
    return np.pad(np.asarray(vals, dtype=np.float64)[np.asarray(geometry)], cpml, mode="edge")


# materials that Model.writemask marks as fluid (classes.py:556)
FLUIDS = ("air", "water", "oil")


def geometry_mask(domain, material) -> np.ndarray:
    """Model.writemask (classes.py:549-606) for a 2-D domain: 1 on the cells
    whose material is named air, water or oil, 0 elsewhere, the edge rows and
    columns repeated into the C-PML.  Returns the padded (nx + 2c, nz + 2c)
    array; ``geometry_mask.dat`` holds its transpose as one int32 record.
    (Nothing in the reference calls writemask or reads the file; the kernels
    do not use it.)"""
    ids = [int(m.get("id", k)) for k, m in enumerate(material.material_list)
           if str(m.get("name", "")) in FLUIDS]
    mask = np.isin(np.asarray(domain.geometry), ids).astype(int)
    return np.pad(mask, int(domain.cpml), mode="edge")


def write_build_files(model, domain, material, workdir: str) -> list:
    """The files Model.build and pointsource write for a 2-D or 2.5-D run, in
    their formats (one Fortran record each, 2-D arrays as ``arr.T``):
    coefficient tables painted on the geometry (``c11.dat`` ... ``rho.dat`` or
    ``e11.dat`` ... ``s33.dat``), ``gamma_*.dat`` (attenuation x f0, with the
    sponge when gamma_max is set), ``density_gradient.dat`` (airsurf, written
    for nmats > 1), ``geometry_mask.dat`` (writemask, 2-D), the C-PML fields
    ``{sigma,kappa,alpha,acoef,bcoef}[_half]_cpml.dat`` and the source
    series.  Returns the paths written."""
    from .run import _sponge, air_density_gradient
    os.makedirs(workdir, exist_ok=True)
    P = lambda f: os.path.join(workdir, f)
    c = int(domain.cpml)
    geom = domain.geometry
    out = []

    def rec(name, arr):
        write_record(P(name), arr)
        out.append(P(name))

    if model.is_seismic:
        tab = model.stiffness_coefficients
        for col in STIFFNESS:
            rec(f"{col}.dat", _extend2d(tab[col].values, geom, c).T)
        att = model.attenuation_coefficients
        for col in ATTENUATION[:-1]:
            ext = _extend2d(att[col].values * model.f0, geom, c)
            if model.gamma_max:
                ext = _sponge(ext, c, float(domain.NPS), float(model.gamma_max))
            rec(f"{col}.dat", ext.T)
    else:
        for col in PERMITTIVITY:
            rec(f"{col}.dat", _extend2d(model.permittivity_coefficients[col].values, geom, c).T)
        for col in CONDUCTIVITY:
            rec(f"{col}.dat", _extend2d(model.conductivity_coefficients[col].values, geom, c).T)
    if int(domain.nmats) > 1 and domain.dim != 3:
        names = [str(v).strip().lower() for v in getattr(material, "material", [])]
        air = names.index("air") if "air" in names else None
        grad = air_density_gradient(geom, air, int(getattr(model, "air_gradient_integer", 2) or 2))
        rec("density_gradient.dat", np.asarray(grad, dtype=np.float64))
    if domain.dim != 3:
        write_record(P("geometry_mask.dat"), geometry_mask(domain, material).T, dtype=np.int32)
        out.append(P("geometry_mask.dat"))
    for half, suf in ((False, "_cpml.dat"), (True, "_half_cpml.dat")):
        for name, arr in zip(("sigma", "kappa", "alpha", "acoef", "bcoef"),
                             cpml_fields_2d(model, domain, half)):
            rec(name + suf, arr.T)
    if model.is_seismic:
        for a in ("x", "y", "z"):
            rec(f"seismicsource{a}.dat", np.asarray(getattr(model, f"sourcefunction_{a}")))
        for comp in ("xx", "xy", "xz", "yy", "yz", "zz"):
            rec(f"seismicsource{comp}.dat", np.asarray(getattr(model, f"source_sigma_{comp}")))
    else:
        for a in ("x", "y", "z"):
            rec(f"electromagneticsource{a}.dat", np.asarray(getattr(model, f"sourcefunction_{a}")))
    return out


def problem_from_project(jsonfile: str, seismic: bool = True, *, geometry=None, receivers=None,
                         density_method: str | None = None, source_convention: str = "legacy",
                         recompute_dt: bool = True, **model_options):
    """A Problem from the project file and its geometry alone: load_project,
    build (dt from the materials unless ``recompute_dt`` is False), then the
    host setup of :func:`seidart_amd.run.problem_from_objects`.
    ``model_options`` set Model attributes first (CFL, source_type,
    air_gradient_integer, gamma_max, use_multimodal, ...)."""
    from .run import problem_from_objects
    domain, material, sm, em = load_project(jsonfile, geometry=geometry)
    model = sm if seismic else em
    for k, v in model_options.items():
        setattr(model, k, v)
    build(model, domain, material, recompute_dt=recompute_dt)
    return problem_from_objects(model, domain, density_method or "legacy", source_convention,
                                receivers=receivers, material=material)


def run_project(jsonfile: str, seismic: bool = True, precision: str = "fp64", snapshot_every: int = 0,
                snapshot_single: bool = True, receivers=None, graph_block: int = 50,
                raise_unstable: bool = True, density_method: str | None = None, **model_options):
    """Run a project from its JSON and geometry alone (no SeidarT install):
    problem_from_project, then the same device run, snapshots and trace CSVs
    as :func:`seidart_amd.run.run_files`."""
    from .run import RunResult, _execute, write_trace_csv
    from .problem import TRACE_COMPONENTS
    p = problem_from_project(jsonfile, seismic, receivers=receivers, density_method=density_method,
                             **model_options)
    wd = os.path.dirname(os.path.abspath(jsonfile))
    tr, unstable, nsnap = _execute(p, precision, snapshot_every, snapshot_single, wd, graph_block, False)
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
