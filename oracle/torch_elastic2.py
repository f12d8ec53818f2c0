"""Pure-PyTorch FP64 restatement of the elastic 2-D scheme, differentiable by
torch autograd.  TEST INFRASTRUCTURE ONLY: the reference gradient that the
hand-written adjoint kernels are checked against (used by tests/, never by
the product path).

It follows the same legacy loops as oracle/fdtd_oracle.c elastic2()
(legacy/fortran/seismicfdtd.f95:12-385), vectorised over each loop's index
range, with the coefficient averages, 1/kappa divisions and dt/rho written
inside the step exactly as the legacy loops write them (so it shares no code
with the device coefficient setup):
  normal stresses  :247-281  (A2: memory_dvx_dx decays with b_x_half(j))
  shear stress     :283-312  (A1: same memory arrays as the normal loop)
(or, with split=True, the split C-PML variant of oracle/fdtd_oracle.h)
  vx               :317-337  (A6: dt / rho(i,j))
  vz               :339-357  (deltarho = (2 rho(i,j) + rho(i+1,j) + rho(i,j+1)) / 4)
  source           :359-361
  Dirichlet edges  :365-375
Its forward traces are pinned against the C oracle by tests/test_adjoint.py.
Every loop updates cells of a rectangular index range; `_put` writes such a
range out of place (keep * old + zero-padded new), which is exact and keeps
the autograd graph free of in-place writes.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

__all__ = ["run_elastic2"]

# (rows, cols) of each loop's range as slices, and the F.pad spec that puts
# the range back on the (nz, nx) grid
_N = ((slice(1, None), slice(None, -1)), (0, 1, 1, 0))   # j >= 1,     i <= nx-2
_S = ((slice(None, -1), slice(1, None)), (1, 0, 0, 1))   # j <= nz-2,  i >= 1
_X = ((slice(1, None), slice(1, None)), (1, 0, 1, 0))    # j >= 1,     i >= 1
_Z = ((slice(None, -1), slice(None, -1)), (0, 1, 0, 1))  # j <= nz-2,  i <= nx-2


def _keep(shape, reg, dtype, device):
    m = torch.ones(shape, dtype=dtype, device=device)
    m[reg[0]] = 0.0
    return m


def _put(old, reg, keep, new):
    return old * keep + F.pad(new, reg[1])


def run_elastic2(prm, prof, dx, dz, dt, nt, src, force, rec, brow, state0=None, split=False):
    """Run nt steps; return traces [nt, 2, nrec] (vx, vz after each step).

    prm:   dict of (nz, nx) FP64 tensors c11 c13 c15 c33 c35 c55 rho, and
           optionally gx gz (attenuation, default 0).
    prof:  dict axis 'x' / 'z' -> (kappa, a, b, kappa_h, a_h, b_h) 1-D tensors.
    src:   (i, k) source cell; force: (fx, fz) per-step series or None.
    rec:   (nrec, 2) integer tensor of (i, k).
    brow:  b_x_half(min(j, nx-1)) per row j (seismicfdtd.f95:255).
    state0: optional dict of initial vx vz sxx szz sxz and m_* memory arrays.
    split: the split C-PML variant of oracle/fdtd_oracle.h (the shear loop has
           its own memory arrays s_*; memory_dvx_dx of the normal loop decays
           with b_x_half(i)).
    """
    c11, c13, c15 = prm["c11"], prm["c13"], prm["c15"]
    c33, c35, c55, rho = prm["c33"], prm["c35"], prm["c55"], prm["rho"]
    nz, nx = rho.shape
    dev, dtp = rho.device, rho.dtype
    gx = prm.get("gx")
    gz = prm.get("gz")
    kx, ax, bx, kxh, axh, bxh = (v.view(1, nx) for v in prof["x"])
    kz, az, bz, kzh, azh, bzh = (v.view(nz, 1) for v in prof["z"])
    brow = brow.view(nz, 1)
    DT = dt
    keep = {k: _keep((nz, nx), r, dtp, dev) for k, r in (("N", _N), ("S", _S), ("X", _X), ("Z", _Z))}
    names = ["vx", "vz", "sxx", "szz", "sxz", "m_dvx_dx", "m_dvx_dz", "m_dvz_dx", "m_dvz_dz",
             "m_dsxx_dx", "m_dszz_dz", "m_dsxz_dx", "m_dsxz_dz",
             "s_dvx_dx", "s_dvx_dz", "s_dvz_dx", "s_dvz_dz"]
    st = {n: torch.zeros((nz, nx), dtype=dtp, device=dev) for n in names}
    if state0:
        for n, v in state0.items():
            st[n] = v
    vx, vz, sxx, szz, sxz = st["vx"], st["vz"], st["sxx"], st["szz"], st["sxz"]
    mvxx, mvxz, mvzx, mvzz = st["m_dvx_dx"], st["m_dvx_dz"], st["m_dvz_dx"], st["m_dvz_dz"]
    msxx, mszz, msxzx, msxzz = st["m_dsxx_dx"], st["m_dszz_dz"], st["m_dsxz_dx"], st["m_dsxz_dz"]
    svxx, svxz, svzx, svzz = st["s_dvx_dx"], st["s_dvx_dz"], st["s_dvz_dx"], st["s_dvz_dz"]
    bn = bxh[:, :-1] if split else brow[1:]  # decay of memory_dvx_dx in the normal loop
    interior = torch.zeros((nz, nx), dtype=dtp, device=dev)
    interior[1:-1, 1:-1] = 1.0
    isrc, ksrc = int(src[0]), int(src[1])
    onehot = torch.zeros((nz, nx), dtype=dtp, device=dev)
    onehot[ksrc, isrc] = 1.0
    rho_src = rho[ksrc, isrc]
    ri, rk = rec[:, 0].long(), rec[:, 1].long()
    out = []
    # coefficient averages (the legacy computes them inside the loops)
    R = _N[0]
    a11 = (c11[1:, 1:] + 2 * c11[R] + c11[:-1, :-1]) / 4
    a13 = (c13[1:, 1:] + 2 * c13[R] + c13[:-1, :-1]) / 4
    a15 = (c15[1:, 1:] + 2 * c15[R] + c15[:-1, :-1]) / 4
    a33 = (c33[1:, 1:] + 2 * c33[R] + c33[:-1, :-1]) / 4
    a35 = (c35[1:, 1:] + 2 * c35[R] + c35[:-1, :-1]) / 4
    R = _S[0]
    s15 = (c15[1:, 1:] + 2 * c15[R] + c15[:-1, :-1]) / 4
    s35 = (c35[1:, 1:] + 2 * c35[R] + c35[:-1, :-1]) / 4
    s55 = (c55[1:, 1:] + 2 * c55[R] + c55[:-1, :-1]) / 4
    deltarho = (2 * rho[:-1, :-1] + rho[:-1, 1:] + rho[1:, :-1]) / 4
    omx = None if gx is None else (1 - gx)[_X[0]]
    omz = None if gz is None else (1 - gz)[_Z[0]]
    for it in range(nt):
        # ---- normal stresses, j >= 1, i <= nx-2 ----
        R = _N[0]
        e1 = (vx[1:, 1:] - vx[R]) / dx
        e2 = (vz[R] - vz[:-1, :-1]) / dz
        e3 = (vz[1:, 1:] - vz[R]) / dx
        e4 = (vx[R] - vx[:-1, :-1]) / dz
        p1 = bn * mvxx[R] + axh[:, :-1] * e1
        p2 = bz[1:] * mvzz[R] + az[1:] * e2
        p4 = bzh[1:] * mvxz[R] + azh[1:] * e4
        p3 = bx[:, :-1] * mvzx[R] + ax[:, :-1] * e3
        mvxx, mvzz = _put(mvxx, _N, keep["N"], p1), _put(mvzz, _N, keep["N"], p2)
        mvxz, mvzx = _put(mvxz, _N, keep["N"], p4), _put(mvzx, _N, keep["N"], p3)
        d1 = e1 / kxh[:, :-1] + p1
        d2 = e2 / kz[1:] + p2
        d3 = e3 / kx[:, :-1] + p3
        d4 = e4 / kzh[1:] + p4
        sxx = _put(sxx, _N, keep["N"], sxx[R] + (a11 * d1 + a13 * d2 + a15 * (d3 + d4)) * DT)
        szz = _put(szz, _N, keep["N"], szz[R] + (a13 * d1 + a33 * d2 + a35 * (d3 + d4)) * DT)
        # ---- shear stress, j <= nz-2, i >= 1 ----
        R = _S[0]
        f1 = (vx[R] - vx[:-1, :-1]) / dx
        f2 = (vz[1:, 1:] - vz[R]) / dz
        f3 = (vz[R] - vz[:-1, :-1]) / dx
        f4 = (vx[1:, 1:] - vx[R]) / dz
        if split:
            q1 = bxh[:, 1:] * svxx[R] + axh[:, 1:] * f1
            q2 = bz[:-1] * svzz[R] + az[:-1] * f2
            q4 = bzh[:-1] * svxz[R] + azh[:-1] * f4
            q3 = bx[:, 1:] * svzx[R] + ax[:, 1:] * f3
            svxx, svzz = _put(svxx, _S, keep["S"], q1), _put(svzz, _S, keep["S"], q2)
            svxz, svzx = _put(svxz, _S, keep["S"], q4), _put(svzx, _S, keep["S"], q3)
        else:
            q1 = bxh[:, 1:] * mvxx[R] + axh[:, 1:] * f1
            q2 = bz[:-1] * mvzz[R] + az[:-1] * f2
            q4 = bzh[:-1] * mvxz[R] + azh[:-1] * f4
            q3 = bx[:, 1:] * mvzx[R] + ax[:, 1:] * f3
            mvxx, mvzz = _put(mvxx, _S, keep["S"], q1), _put(mvzz, _S, keep["S"], q2)
            mvxz, mvzx = _put(mvxz, _S, keep["S"], q4), _put(mvzx, _S, keep["S"], q3)
        g1 = f1 / kxh[:, 1:] + q1
        g2 = f2 / kz[:-1] + q2
        g3 = f3 / kx[:, 1:] + q3
        g4 = f4 / kzh[:-1] + q4
        sxz = _put(sxz, _S, keep["S"], sxz[R] + (s15 * g1 + s35 * g2 + s55 * (g3 + g4)) * DT)
        # ---- vx, j >= 1, i >= 1 ----
        R = _X[0]
        h1 = (sxx[R] - sxx[1:, :-1]) / dx
        h2 = (sxz[R] - sxz[:-1, 1:]) / dz
        r1 = bx[:, 1:] * msxx[R] + ax[:, 1:] * h1
        r2 = bz[1:] * msxzz[R] + az[1:] * h2
        msxx, msxzz = _put(msxx, _X, keep["X"], r1), _put(msxzz, _X, keep["X"], r2)
        vxo = vx[R] if omx is None else vx[R] * omx
        vx_new = vxo + ((h1 / kx[:, 1:] + r1) + (h2 / kz[1:] + r2)) * DT / rho[R]
        # ---- vz, j <= nz-2, i <= nx-2 ----
        R = _Z[0]
        k1 = (sxz[:-1, 1:] - sxz[R]) / dx
        k2 = (szz[1:, :-1] - szz[R]) / dz
        r3 = bxh[:, :-1] * msxzx[R] + axh[:, :-1] * k1
        r4 = bzh[:-1] * mszz[R] + azh[:-1] * k2
        msxzx, mszz = _put(msxzx, _Z, keep["Z"], r3), _put(mszz, _Z, keep["Z"], r4)
        vzo = vz[R] if omz is None else vz[R] * omz
        vz_new = vzo + ((k1 / kxh[:, :-1] + r3) + (k2 / kzh[:-1] + r4)) * DT / deltarho
        vx = _put(vx, _X, keep["X"], vx_new)
        vz = _put(vz, _Z, keep["Z"], vz_new)
        # ---- source, then Dirichlet edges ----
        if force[0] is not None:
            vx = vx + onehot * (force[0][it] * DT / rho_src)
        if force[1] is not None:
            vz = vz + onehot * (force[1][it] * DT / rho_src)
        vx = vx * interior
        vz = vz * interior
        out.append(torch.stack([vx[rk, ri], vz[rk, ri]]))
    return torch.stack(out)
