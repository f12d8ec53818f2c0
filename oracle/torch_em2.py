"""Pure-PyTorch FP64 restatement of the EM 2-D scheme, differentiable by torch
autograd.  TEST INFRASTRUCTURE ONLY: the reference gradient that the
hand-written EM adjoint kernels are checked against (used by tests/, never by
the product path).

It follows oracle/fdtd_oracle.c em2() (legacy/fortran/electromagfdtd.f95
:10-326, electromag2), vectorised over each loop's index range, with ca / cb
(:164-177) and their 2-point averages written as the legacy writes them:
  Hy               :249-266
  Ex               :273-283  ((ca(i,j) + ca(i,j-1)) / 2)
  Ez               :285-295  ((ca(i,j) + ca(i-1,j)) / 2)
  source           :300-301  (divided by the relative permittivity e11 / e33)
  Dirichlet edges  :304-317
Its forward traces are pinned against the C oracle by tests/test_adjoint.py.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

__all__ = ["run_em2"]

EPS0 = 8.85418782e-12
MU0 = 4.0 * 3.141592653589793238462643 * 1.0e-7


def _cacb(sig, epsr_sum, dt):
    eps = epsr_sum * EPS0
    ca = (1.0 - sig * dt / (2.0 * eps)) / (1.0 + sig * dt / (2.0 * eps))
    cb = (dt / eps) / (1.0 + sig * dt / (2.0 * eps))
    return ca, cb


def _range(shape, rows, cols, dtype, device):
    """keep-mask (0 on the range) and F.pad spec of a (rows, cols) range."""
    nz, nx = shape
    m = torch.ones(shape, dtype=dtype, device=device)
    m[rows, cols] = 0.0
    r0, r1, _ = rows.indices(nz)
    c0, c1, _ = cols.indices(nx)
    return m, (c0, nx - c1, r0, nz - r1)


def run_em2(prm, prof, dx, dz, dt, nt, src, force, rec, state0=None):
    """Run nt steps; return traces [nt, 2, nrec] (Ex, Ez after each step).

    prm:   dict of (nz, nx) FP64 tensors e11 e13 e33 s11 s13 s33.
    prof:  dict axis 'x' / 'z' -> (kappa, a, b, kappa_h, a_h, b_h) 1-D tensors.
    src:   (i, k) source cell; force: (fx, fz) per-step series or None.
    rec:   (nrec, 2) integer tensor of (i, k).
    state0: optional dict of initial Ex Ez Hy and memory arrays.
    """
    e11, e13, e33 = prm["e11"], prm["e13"], prm["e33"]
    s11, s13, s33 = prm["s11"], prm["s13"], prm["s33"]
    nz, nx = e11.shape
    dev, dtp = e11.device, e11.dtype
    kx, ax, bx, kxh, axh, bxh = (v.view(1, nx) for v in prof["x"])
    kz, az, bz, kzh, azh, bzh = (v.view(nz, 1) for v in prof["z"])
    caEx, cbEx = _cacb(s11 + s13, e11 + e13, dt)
    caEz, cbEz = _cacb(s13 + s33, e13 + e33, dt)
    dbH = dt / MU0
    names = ["Ex", "Ez", "Hy", "m_dEx_dz", "m_dEz_dx", "m_dHy_dx", "m_dHy_dz"]
    st = {n: torch.zeros((nz, nx), dtype=dtp, device=dev) for n in names}
    if state0:
        for n, v in state0.items():
            st[n] = v
    Ex, Ez, Hy = st["Ex"], st["Ez"], st["Hy"]
    mExz, mEzx, mHyx, mHyz = st["m_dEx_dz"], st["m_dEz_dx"], st["m_dHy_dx"], st["m_dHy_dz"]
    RH = (slice(None, -1), slice(None, -1))     # Hy: j <= nz-2, i <= nx-2
    RX = (slice(1, -1), slice(None, -1))        # Ex: 1 <= j <= nz-2, i <= nx-2
    RZ = (slice(None, -1), slice(1, -1))        # Ez: j <= nz-2, 1 <= i <= nx-2
    kH, pH = _range((nz, nx), *RH, dtp, dev)
    kX, pX = _range((nz, nx), *RX, dtp, dev)
    kZ, pZ = _range((nz, nx), *RZ, dtp, dev)
    put = lambda old, keep, pad, new: old * keep + F.pad(new, pad)
    caX = (caEx[1:-1, :-1] + caEx[:-2, :-1]) / 2
    cbX = (cbEx[1:-1, :-1] + cbEx[:-2, :-1]) / 2
    caZ = (caEz[:-1, 1:-1] + caEz[:-1, :-2]) / 2
    cbZ = (cbEz[:-1, 1:-1] + cbEz[:-1, :-2]) / 2
    dEx = torch.ones((nz, nx), dtype=dtp, device=dev)   # Dirichlet masks, :304-317
    dEx[:, 0] = 0; dEx[:, nx - 2] = 0; dEx[0, :] = 0; dEx[nz - 1, :] = 0
    dEz = torch.ones((nz, nx), dtype=dtp, device=dev)
    dEz[:, 0] = 0; dEz[:, nx - 1] = 0; dEz[0, :] = 0; dEz[nz - 2, :] = 0
    dHy = torch.ones((nz, nx), dtype=dtp, device=dev)
    dHy[:, 0] = 0; dHy[:, nx - 2] = 0; dHy[0, :] = 0; dHy[nz - 2, :] = 0
    isrc, ksrc = int(src[0]), int(src[1])
    onehot = torch.zeros((nz, nx), dtype=dtp, device=dev)
    onehot[ksrc, isrc] = 1.0
    ri, rk = rec[:, 0].long(), rec[:, 1].long()
    out = []
    for it in range(nt):
        # Hy
        a = (Ex[1:, :-1] - Ex[:-1, :-1]) / dz
        ma = bz[:-1] * mExz[RH] + az[:-1] * a
        b = (Ez[:-1, 1:] - Ez[:-1, :-1]) / dx
        mb = bx[:, :-1] * mEzx[RH] + ax[:, :-1] * b
        mExz, mEzx = put(mExz, kH, pH, ma), put(mEzx, kH, pH, mb)
        Hy = put(Hy, kH, pH, 1.0 * Hy[RH] + dbH * ((b / kx[:, :-1] + mb) + (a / kz[:-1] + ma)))
        # Ex
        d = (Hy[1:-1, :-1] - Hy[:-2, :-1]) / dz
        md = bzh[1:-1] * mHyz[RX] + azh[1:-1] * d
        mHyz = put(mHyz, kX, pX, md)
        Ex = put(Ex, kX, pX, caX * Ex[RX] + cbX * (d / kzh[1:-1] + md))
        # Ez
        e = (Hy[:-1, 1:-1] - Hy[:-1, :-2]) / dx
        me = bxh[:, 1:-1] * mHyx[RZ] + axh[:, 1:-1] * e
        mHyx = put(mHyx, kZ, pZ, me)
        Ez = put(Ez, kZ, pZ, caZ * Ez[RZ] + cbZ * (e / kxh[:, 1:-1] + me))
        # source, then Dirichlet
        if force[0] is not None:
            Ex = Ex + onehot * (force[0][it] * dt / e11[ksrc, isrc])
        if force[1] is not None:
            Ez = Ez + onehot * (force[1][it] * dt / e33[ksrc, isrc])
        Ex, Ez, Hy = Ex * dEx, Ez * dEz, Hy * dHy
        out.append(torch.stack([Ex[rk, ri], Ez[rk, ri]]))
    return torch.stack(out)
