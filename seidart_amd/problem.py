"""In-memory form of one run of the FDTD core, and its FP64 coefficient setup.

A :class:`Problem` carries what the reference hands its solver through
files (src/seidart/routines/classes.py:393-547, 878-924): the padded grid,
the per-cell coefficient fields (``c11.dat`` ... ``rho.dat``,
``gamma_*.dat``, ``e11.dat`` ... ``s33.dat``), the C-PML profiles, the
time step, and per shot the source cell, source series and receivers.

All arrays are float64 and C-ordered (nz, nx) [2-D, and the y-invariant
coefficients of 2.5-D] or (nz, ny, nx) [2.5-D fields], i.e. x fastest.
Indices are 0-based on the padded grid.

``kernel_coefs`` does the FP64 setup the device kernels consume: the
staggered-node coefficient averages, dt/rho, 1 - gamma and the EM update
coefficients, each in the evaluation order of the legacy loop it replaces,
so the FP64 device path differs from the legacy arithmetic only by the
rounding of these precomputed products.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .cpml import EPS0, slab_width

__all__ = ["KINDS", "FIELDS", "STIFFNESS", "PERMITTIVITY", "CONDUCTIVITY", "Shot", "Problem",
           "kernel_coefs", "source_increments", "aux_b_x_half_by_row", "ncomp",
           "DENSITY_METHODS", "node_density", "elastic2_coefs", "em2_coefs"]

# Density averaging at staggered velocity nodes.  The legacy solver has one
# rule (weighted arithmetic mean); the reference binary takes a
# density_method option whose arithmetic is not published: 'geometric' and
# 'harmonic' are the weighted geometric and harmonic means with the same
# weights, 'none' the cell's own density.
DENSITY_METHODS = ("legacy", "arithmetic", "geometric", "harmonic", "none")


def node_density(parts, weights, method: str = "legacy"):
    """Weighted mean of the density samples ``parts`` with integer weights,
    in the legacy expression order for 'legacy' / 'arithmetic'
    (e.g. (2 rho + rho_e + rho_n) / 4, seismicfdtd.f95:342)."""
    tot = float(sum(weights))
    if method in ("legacy", "arithmetic"):
        acc = weights[0] * parts[0] if weights[0] != 1 else parts[0]
        for w, x in zip(weights[1:], parts[1:]):
            acc = acc + (w * x if w != 1 else x)
        return acc / tot
    if method == "geometric":
        if _is_torch(parts[0]):
            import torch
            return torch.exp(sum(w * torch.log(x) for w, x in zip(weights, parts)) / tot)
        return np.exp(sum(w * np.log(x) for w, x in zip(weights, parts)) / tot)
    if method == "harmonic":
        return tot / sum(w / x for w, x in zip(weights, parts))
    if method == "none":
        return parts[0].clone() if _is_torch(parts[0]) else np.array(parts[0], copy=True)
    raise ValueError(f"density_method must be one of {DENSITY_METHODS}")

KINDS = {"elastic2": 0, "elastic25": 1, "em2": 2, "em25": 3}
FIELDS = {
    "elastic2": ["vx", "vz", "sxx", "szz", "sxz"],
    "elastic25": ["vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz"],
    "em2": ["Ex", "Ez", "Hy"],
    "em25": ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"],
}
TRACE_COMPONENTS = {
    "elastic2": ["Vx", "Vz"], "elastic25": ["Vx", "Vy", "Vz"],
    "em2": ["Ex", "Ez"], "em25": ["Ex", "Ey", "Ez"],
}
# Column order of the project JSON (src/seidart/routines/definitions.py:322-330)
STIFFNESS = ["c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26",
             "c33", "c34", "c35", "c36", "c44", "c45", "c46", "c55", "c56", "c66", "rho"]
PERMITTIVITY = ["e11", "e12", "e13", "e22", "e23", "e33"]
CONDUCTIVITY = ["s11", "s12", "s13", "s22", "s23", "s33"]


def field_extent(kind: str, name: str):
    """Index (slices over the C-ordered (z, [y,] x) array) of the cells a
    field owns in the legacy arrays: the EM fields are staggered with one
    cell fewer along their own axis (Ex (nx-1) x nz, Ez nx x (nz-1), Hy
    (nx-1) x (nz-1); electromagfdtd.f95:100-102, :668-673); the elastic
    fields span the whole padded grid.  Cells outside stay zero."""
    full = slice(None)
    short = slice(0, -1)
    if kind == "em2":
        return {"Ex": (full, short), "Ez": (short, full), "Hy": (short, short)}[name]
    if kind == "em25":
        return {"Ex": (full, full, short), "Ey": (full, short, full), "Ez": (short, full, full),
                "Hx": (short, short, full), "Hy": (short, full, short),
                "Hz": (full, short, short)}[name]
    return (full,) * (3 if kind == "elastic25" else 2)


def ncomp(kind: str) -> int:
    return 2 if kind in ("elastic2", "em2") else 3


MOMENT_COMPONENTS = ("xx", "yy", "zz", "xy", "xz", "yz")


@dataclass
class Shot:
    """Source cell (i, j, k), the source series per component (x, y, z; None
    = absent; length nt), receiver cells (nrec, 3) and, for elastic runs,
    optional stress-rate (moment-tensor) series per component
    MOMENT_COMPONENTS (sourcefunction.py:296-372)."""
    src: tuple
    force: tuple = (None, None, None)
    rec: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), dtype=np.int32))
    moment: tuple = (None,) * 6


@dataclass
class Problem:
    kind: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dt: float
    nt: int
    coef: dict
    prof: dict
    shots: list
    gamma: dict = field(default_factory=dict)
    init: list | None = None
    # elastic2 only: run the split C-PML variant (own velocity-derivative
    # memory arrays for the shear loop, b_x_half(i) in the normal loop)
    # instead of the legacy shared arrays and b_x_half(j).  Not in the
    # reference; stable in long runs where the legacy C-PML grows.
    cpml_split: bool = False

    @property
    def is3d(self) -> bool:
        return self.kind in ("elastic25", "em25")

    @property
    def field_shape(self):
        return (self.nz, self.ny, self.nx) if self.is3d else (self.nz, self.nx)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def with_shots(self, shots) -> "Problem":
        """The same model (coefficients, profiles, attenuation, C-PML variant,
        density method and other set-up attributes) with other shots; the
        initial conditions are dropped unless there is one per new shot."""
        import copy
        q = copy.copy(self)
        q.shots = list(shots)
        if self.init is not None and len(self.init) != len(q.shots):
            q.init = None
        return q

    def slabs(self):
        lx = slab_width(self.prof["x"])
        lz = slab_width(self.prof["z"])
        ly = slab_width(self.prof["y"]) if self.is3d else 0
        return lx, ly, lz

    def validate(self) -> None:
        if self.kind not in KINDS:
            raise ValueError(f"unknown kind {self.kind!r}")
        if self.nx < 3 or self.nz < 3 or (self.is3d and self.ny < 3) or (not self.is3d and self.ny != 1):
            raise ValueError("grid too small")
        for a in ("x", "z") + (("y",) if self.is3d else ()):
            n = {"x": self.nx, "y": self.ny, "z": self.nz}[a]
            p = self.prof.get(a)
            if p is None or p.n != n:
                raise ValueError(f"C-PML profile for axis {a} missing or of wrong length")
        names = (STIFFNESS if self.kind.startswith("elastic") else PERMITTIVITY + CONDUCTIVITY)
        need = names if self.kind in ("elastic25", "em25") else (
            ["c11", "c13", "c15", "c33", "c35", "c55", "rho"] if self.kind == "elastic2"
            else ["e11", "e13", "e33", "s11", "s13", "s33"])
        for n in need:
            a = self.coef.get(n)
            if a is None or a.shape != (self.nz, self.nx):
                raise ValueError(f"coefficient {n} missing or not of shape {(self.nz, self.nx)}")
            if not np.all(np.isfinite(a)):
                raise ValueError(f"coefficient {n} holds non-finite values")
        if self.kind.startswith("elastic") and np.any(self.coef["rho"] <= 0):
            raise ValueError("density must be positive")
        if not (self.dt > 0 and self.dx > 0 and self.dz > 0 and (not self.is3d or self.dy > 0)):
            raise ValueError("dx, dy, dz, dt must be positive")
        for s, sh in enumerate(self.shots):
            i, j, k = sh.src
            if not (0 <= i < self.nx and 0 <= k < self.nz and 0 <= j < self.ny):
                raise ValueError(f"shot {s}: source {sh.src} outside the padded grid")
            r = np.asarray(sh.rec, dtype=np.int64).reshape(-1, 3)
            if r.size and ((r < 0).any() or (r[:, 0] >= self.nx).any() or
                           (r[:, 1] >= self.ny).any() or (r[:, 2] >= self.nz).any()):
                raise ValueError(f"shot {s}: receiver outside the padded grid")
            for c in sh.force:
                if c is not None and len(c) < self.nt:
                    raise ValueError(f"shot {s}: source series shorter than nt")


# The elastic 2-D setup below accepts numpy arrays or torch tensors (the
# adjoint path differentiates it with torch autograd); both go through the
# same expressions in the same order.
def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _zeros_like(a):
    if _is_torch(a):
        import torch
        return torch.zeros_like(a)
    return np.zeros_like(a)


def _avg4_normal(c):
    """(c(i+1,j) + 2 c(i,j) + c(i,j-1)) / 4 at the sigma_xx node,
    seismicfdtd.f95:269-278; zero where the loop does not reach."""
    out = _zeros_like(c)
    out[1:, :-1] = (c[1:, 1:] + 2 * c[1:, :-1] + c[:-1, :-1]) / 4
    return out


def _avg4_shear(c):
    """(c(i,j+1) + 2 c(i,j) + c(i-1,j)) / 4 at the sigma_xz node,
    seismicfdtd.f95:307-310."""
    out = _zeros_like(c)
    out[:-1, 1:] = (c[1:, 1:] + 2 * c[:-1, 1:] + c[:-1, :-1]) / 4
    return out


def _sum4(c, di1, dk1, di2, dk2, di3, dk3, di4, dk4, irange, krange):
    """((c[p1] + c[p2]) + c[p3]) + c[p4] over the given (i, k) index ranges,
    in the order the legacy 2.5-D loops write the four-point sums."""
    nz, nx = c.shape
    out = np.zeros_like(c)
    i0, i1 = irange
    k0, k1 = krange
    I = slice(i0, nx + i1)
    K = slice(k0, nz + k1)

    def sh(di, dk):
        return c[k0 + dk: nz + k1 + dk, i0 + di: nx + i1 + di]

    out[K, I] = ((sh(di1, dk1) + sh(di2, dk2)) + sh(di3, dk3)) + sh(di4, dk4)
    return out


def kernel_coefs(p: Problem):
    """Device coefficient arrays of a problem, in the slot order of
    include/seidart_hip.h (FP64, shape (nz, nx) each)."""
    C = p.coef
    dt = p.dt
    if p.kind == "elastic2":
        return elastic2_coefs(C, p.gamma, dt, getattr(p, "density_method", "legacy"))
    if p.kind == "elastic25":
        rho = C["rho"]
        # loop 1 stencil (i+1,k) (i,k) (i,k-1) (i+1,k-1): seismicfdtd.f95:651-673
        s1 = lambda n: _sum4(C[n], 1, 0, 0, 0, 0, -1, 1, -1, (0, -1), (1, 0))
        # loop 2 stencil (i,k) (i-1,k) (i,k-1) (i-1,k-1): :716-721
        s2 = lambda n: _sum4(C[n], 0, 0, -1, 0, 0, -1, -1, -1, (1, 0), (1, 0))
        # loop 3 stencil (i,k+1) (i,k) (i-1,k) (i-1,k+1): :761-767
        s3 = lambda n: _sum4(C[n], 0, 1, 0, 0, -1, 0, -1, 1, (1, 0), (0, -1))
        # loop 4 stencil (i+1,k) (i,k) (i,k+1) (i+1,k+1): :806-811
        s4 = lambda n: _sum4(C[n], 1, 0, 0, 0, 0, 1, 1, 1, (0, -1), (0, -1))
        L1 = [s1("c11") / 4, s1("c12") / 4, s1("c13") / 4, s1("c14") / 8, s1("c15") / 8, s1("c16") / 8,
              s1("c22") / 4, s1("c23") / 4, s1("c24") / 8, s1("c25") / 8, s1("c26") / 8,
              s1("c33") / 4, s1("c34") / 8, s1("c35") / 8, s1("c36") / 8]
        L2 = [s2("c16") / 4, s2("c26") / 4, s2("c36") / 4, s2("c46") / 8, s2("c56") / 8, s2("c66") / 8]
        L3 = [s3("c15") / 4, s3("c25") / 4, s3("c35") / 4, s3("c45") / 8, s3("c55") / 8, s3("c56") / 8]
        L4 = [s4("c14") / 4, s4("c24") / 4, s4("c34") / 4, s4("c44") / 8, s4("c45") / 8, s4("c46") / 8]
        nz, nx = rho.shape
        dx_ = np.zeros_like(rho)
        dy_ = np.zeros_like(rho)
        dz_ = np.zeros_like(rho)
        dm = getattr(p, "density_method", "legacy")
        # (4 rho(i,k) + rho(i-1,k) + rho(i,k-1)) / 6, seismicfdtd.f95:826
        dx_[1:, 1:] = dt / node_density([rho[1:, 1:], rho[1:, :-1], rho[:-1, 1:]], [4, 1, 1], dm)
        # (4 rho(i,k) + rho(i+1,k) + rho(i,k-1)) / 6, :853
        dy_[1:, :-1] = dt / node_density([rho[1:, :-1], rho[1:, 1:], rho[:-1, :-1]], [4, 1, 1], dm)
        # (rho(i+1,k) + rho(i,k+1) + 4 rho(i,k)) / 6, :878
        dz_[:-1, :-1] = dt / node_density([rho[:-1, 1:], rho[1:, :-1], rho[:-1, :-1]], [1, 1, 4], dm)
        z = np.zeros_like(rho)
        om = [1 - p.gamma.get(a, z) for a in ("x", "y", "z")]
        return L1 + L2 + L3 + L4 + [dx_, dy_, dz_] + om
    if p.kind == "em2":
        return em2_coefs(C, dt)
    if p.kind == "em25":
        # electromagfdtd.f95:760-781
        ca_x, cb_x = _cacb(C["s11"] + C["s12"] + C["s13"], C["e11"] + C["e12"] + C["e13"], dt)
        ca_y, cb_y = _cacb(C["s12"] + C["s22"] + C["s23"], C["e12"] + C["e22"] + C["e23"], dt)
        ca_z, cb_z = _cacb(C["s13"] + C["s23"] + C["s33"], C["e13"] + C["e23"] + C["e33"], dt)
        ay, by_, az, bz = (np.zeros_like(ca_x) for _ in range(4))
        # (4 ca(i,k) + ca(i-1,k) + ca(i,k-1)) / 6 for Ey, :1001-1003
        ay[1:, 1:] = (4 * ca_y[1:, 1:] + ca_y[1:, :-1] + ca_y[:-1, 1:]) / 6
        by_[1:, 1:] = (4 * cb_y[1:, 1:] + cb_y[1:, :-1] + cb_y[:-1, 1:]) / 6
        # (4 ca(i,k) + ca(i-1,k) + ca(i,k+1)) / 6 for Ez, :1024-1026
        az[:-1, 1:] = (4 * ca_z[:-1, 1:] + ca_z[:-1, :-1] + ca_z[1:, 1:]) / 6
        bz[:-1, 1:] = (4 * cb_z[:-1, 1:] + cb_z[:-1, :-1] + cb_z[1:, 1:]) / 6
        return [ca_x, cb_x, ay, by_, az, bz]
    raise ValueError(p.kind)


def elastic2_coefs(C, gamma, dt, density_method="legacy"):
    """ELASTIC2 coefficient slots from the per-cell fields C (c11 c13 c15
    c33 c35 c55 rho) and gamma (x, z; absent = 0), as numpy arrays or torch
    tensors (nz, nx)."""
    rho = C["rho"]
    gx = gamma.get("x")
    gz = gamma.get("z")
    gx = _zeros_like(rho) if gx is None else gx
    gz = _zeros_like(rho) if gz is None else gz
    dtr_z = _zeros_like(rho)
    # deltarho = (2 rho(i,j) + rho(i+1,j) + rho(i,j+1)) / 4, seismicfdtd.f95:342
    dtr_z[:-1, :-1] = dt / node_density([rho[:-1, :-1], rho[:-1, 1:], rho[1:, :-1]],
                                        [2, 1, 1], density_method)
    return [
        _avg4_normal(C["c11"]), _avg4_normal(C["c13"]), _avg4_normal(C["c15"]),
        _avg4_normal(C["c33"]), _avg4_normal(C["c35"]),
        _avg4_shear(C["c15"]), _avg4_shear(C["c35"]), _avg4_shear(C["c55"]),
        dt / rho, dtr_z, 1 - gx, 1 - gz,
    ]


def em2_coefs(C, dt):
    """EM2 coefficient slots (caEx cbEx caEz cbEz, 2-point averaged) from the
    per-cell relative permittivity e11 e13 e33 and conductivity s11 s13 s33,
    as numpy arrays or torch tensors (nz, nx)."""
    # electromagfdtd.f95:164-177
    ca_x, cb_x = _cacb(C["s11"] + C["s13"], C["e11"] + C["e13"], dt)
    ca_z, cb_z = _cacb(C["s13"] + C["s33"], C["e13"] + C["e33"], dt)
    o = [_zeros_like(ca_x) for _ in range(4)]
    # (ca(i,j) + ca(i,j-1)) / 2 for Ex, :281-282; (ca(i,j) + ca(i-1,j)) / 2 for Ez, :293-294
    o[0][1:, :] = (ca_x[1:, :] + ca_x[:-1, :]) / 2
    o[1][1:, :] = (cb_x[1:, :] + cb_x[:-1, :]) / 2
    o[2][:, 1:] = (ca_z[:, 1:] + ca_z[:, :-1]) / 2
    o[3][:, 1:] = (cb_z[:, 1:] + cb_z[:, :-1]) / 2
    return o


def _cacb(sig, epsr_sum, dt):
    """caE = (1 - s dt / 2e) / (1 + s dt / 2e), cbE = (dt / e) / (1 + s dt / 2e)
    with e = eps_r * eps0 (electromagfdtd.f95:171-177)."""
    eps = epsr_sum * EPS0
    ca = (1.0 - sig * dt / (2.0 * eps)) / (1.0 + sig * dt / (2.0 * eps))
    cb = (dt / eps) / (1.0 + sig * dt / (2.0 * eps))
    return ca, cb


def em_dbh(dt: float) -> float:
    """dbH = dt / mu0 with mu0 = 4 pi 1e-7 (electromagfdtd.f95:81-84, :179-181)."""
    return dt / (4.0 * 3.141592653589793238462643 * 1.0e-7)


def source_increments(p: Problem, sh: Shot):
    """Per-step increment added at the source cell, per component:
    s(it) dt / rho(src) (seismicfdtd.f95:359-361, :898-900) or
    s(it) dt / eps_r(src) with eps11 / eps22 / eps33 (electromagfdtd.f95:300-301,
    :1033-1035)."""
    i, j, k = sh.src
    out = []
    for c, f in enumerate(sh.force):
        if f is None:
            out.append(None)
            continue
        f = np.asarray(f, dtype=np.float64)[: p.nt]
        if p.kind.startswith("elastic"):
            div = p.coef["rho"][k, i]
        else:
            div = p.coef[("e11", "e22", "e33")[c]][k, i]
        out.append(f * p.dt / div)
    return out


def stress_source_increments(p: Problem, sh: Shot):
    """Per-step stress increments at the source cell, per MOMENT_COMPONENTS
    entry: rate * dt (None = absent).  Only elastic runs carry them; a 2-D run
    uses xx, zz and xz.  The reference writes these series
    (seismicsource{xx,...}.dat, sourcefunction.py:358-363) for a solver whose
    source is not in its sources, so the injection rule is this package's."""
    if not p.kind.startswith("elastic"):
        return [None] * 6
    out = []
    for q, m in enumerate(getattr(sh, "moment", (None,) * 6)):
        if m is None or (p.kind == "elastic2" and MOMENT_COMPONENTS[q] not in ("xx", "zz", "xz")):
            out.append(None)
            continue
        out.append(np.asarray(m, dtype=np.float64)[: p.nt] * p.dt)
    return out


def aux_b_x_half_by_row(p: Problem) -> np.ndarray:
    """b_x_half(j) for each z row j: the decay the legacy normal-stress loop
    applies to memory_dvx_dx (seismicfdtd.f95:255).  Rows past the x extent
    take the last entry (the legacy reads out of bounds there)."""
    bxh = p.prof["x"].b_h
    j = np.minimum(np.arange(p.nz), p.nx - 1)
    return bxh[j].astype(np.float64)
