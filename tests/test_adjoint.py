"""Adjoint (FWI gradient) of the elastic 2-D scheme.

CPU: the pure-PyTorch FP64 restatement (oracle/torch_elastic2.py) is pinned
against the C oracle's traces, and the torch form of the coefficient setup
against the numpy form the forward uses.

GPU (through the C ABI): the hand-written adjoint kernels are checked
  * by the dot-product test <A x, y> = <x, A^T y> in FP64, where A maps an
    initial state (fields and C-PML memory) to the receiver traces;
  * against torch autograd of the FP64 restatement (gradients of an L2
    trace misfit with respect to every per-cell parameter and the source
    series);
  * against central finite differences of the FP64 device forward at 20
    seeded cells per parameter (relative step 1e-4);
  * FP32 gradients against FP64 ones.
"""
import numpy as np
import pytest
import torch

from helpers import rel_l2, tiny_problem
from seidart_amd.problem import elastic2_coefs, kernel_coefs

PHYS = ("c11", "c13", "c15", "c33", "c35", "c55", "rho")


def _torch_ref_args(p, shot, params=None):
    """Arguments of oracle.torch_elastic2.run_elastic2 for one shot (and the
    C-PML variant of the problem)."""
    T = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    prm = params if params is not None else {k: T(p.coef[k]) for k in PHYS}
    if p.gamma and params is None:
        prm["gx"] = T(p.gamma["x"])
        prm["gz"] = T(p.gamma["z"])
    prof = {a: tuple(T(v) for v in (pr.kappa, pr.a, pr.b, pr.kappa_h, pr.a_h, pr.b_h))
            for a, pr in p.prof.items()}
    # b_x_half(j) for row j, rows past nx take the last entry
    brow = T(p.prof["x"].b_h[np.minimum(np.arange(p.nz), p.nx - 1)])
    sh = p.shots[shot]
    force = tuple(None if f is None else T(f) for f in (sh.force[0], sh.force[2]))
    rec = torch.as_tensor(np.asarray(sh.rec)[:, [0, 2]])
    return dict(prm=prm, prof=prof, dx=p.dx, dz=p.dz, dt=p.dt, nt=p.nt, src=(sh.src[0], sh.src[2]),
                force=force, rec=rec, brow=brow, split=bool(getattr(p, "cpml_split", False)))


def test_torch_reference_matches_oracle(oracle_lib):
    from torch_elastic2 import run_elastic2
    p = tiny_problem("elastic2", seed=41, nx=40, nz=34, npml=6, nt=70, gamma=1e-3, nrec=7)
    with torch.no_grad():
        tr = run_elastic2(**_torch_ref_args(p, 0)).numpy()
    o = oracle_lib.OracleRun(p, shot=0).run()
    assert np.abs(o.trace).max() > 0
    assert rel_l2(tr, o.trace) < 1e-13


def test_torch_reference_matches_oracle_no_attenuation(oracle_lib):
    from torch_elastic2 import run_elastic2
    p = tiny_problem("elastic2", seed=42, nx=36, nz=40, npml=5, nt=60, nrec=5, full21=False)
    with torch.no_grad():
        tr = run_elastic2(**_torch_ref_args(p, 0)).numpy()
    o = oracle_lib.OracleRun(p, shot=0).run()
    assert rel_l2(tr, o.trace) < 1e-13


def test_torch_reference_matches_oracle_split(oracle_lib):
    from torch_elastic2 import run_elastic2
    p = tiny_problem("elastic2", seed=44, nx=42, nz=38, npml=6, nt=70, gamma=1e-3, nrec=6)
    p.cpml_split = True
    with torch.no_grad():
        tr = run_elastic2(**_torch_ref_args(p, 0)).numpy()
    o = oracle_lib.OracleRun(p, shot=0).run()
    assert rel_l2(tr, o.trace) < 1e-13
    # the variant really differs from the legacy scheme in the C-PML
    p.cpml_split = False
    assert rel_l2(oracle_lib.OracleRun(p, shot=0).run().trace, o.trace) > 1e-8


@pytest.mark.parametrize("dm", ["legacy", "geometric", "harmonic", "none"])
def test_coefs_torch_matches_numpy(dm):
    p = tiny_problem("elastic2", seed=43, nx=30, nz=26, npml=4, nt=10, gamma=1e-3)
    p.density_method = dm
    ref = kernel_coefs(p)
    Ct = {k: torch.as_tensor(v) for k, v in p.coef.items()}
    gt = {a: torch.as_tensor(v) for a, v in p.gamma.items()}
    got = elastic2_coefs(Ct, gt, p.dt, dm)
    # torch evaluates scalar / tensor as scalar * (1 / tensor): 1-ulp differences
    for q, (a, b) in enumerate(zip(got, ref)):
        assert np.max(np.abs(a.numpy() - b) / (np.abs(b) + 1e-300)) <= 1e-15, (dm, q)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _engine(p, precision="fp64", every=16, **kw):
    from seidart_amd.adjoint import Elastic2Engine
    return Elastic2Engine(p, precision=precision, every=every, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("every,split,overlap", [(16, False, True), (25, False, True),
                                                 (16, True, True), (16, False, False)])
def test_dot_product(every, split, overlap):
    """<A x, y> = <x, A^T y> for A: initial state -> traces (no source)."""
    p = tiny_problem("elastic2", seed=51, nx=52, nz=44, npml=7, nt=110, gamma=1e-3, nshots=2, nrec=9)
    p.cpml_split = split
    for sh in p.shots:
        sh.force = (np.zeros(p.nt), None, np.zeros(p.nt))
    eng = _engine(p, "fp64", every=every, overlap=overlap)
    g = torch.Generator(device="cpu").manual_seed(5)
    xs = []
    for s in range(eng.nshots):
        x = torch.randn(eng.init[s].numel(), generator=g, dtype=torch.float64)
        eng.init[s].copy_(x.to(eng.device))
        xs.append(x)
    tr = eng.forward()
    ys = [torch.randn(t.shape, generator=g, dtype=torch.float64) for t in tr]
    lhs = sum(float((t.cpu() * y).sum()) for t, y in zip(tr, ys))
    eng.backward(ys)
    rhs = sum(float((x * eng.adjoint_state(s).cpu()).sum()) for s, x in enumerate(xs))
    assert abs(lhs) > 0
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,split", [("fp64", False), ("fp64", True), ("fp32", False)])
def test_fused_adjoint_matches_reference_kernels(precision, split):
    """The fused single-pass adjoint kernel against the three-kernel one:
    same lambda_0 and gradients up to rounding (sums are reassociated),
    on several models of uneven size with several shots each."""
    probs = [tiny_problem("elastic2", seed=80 + m, nt=70, gamma=1e-3, nshots=2, nrec=7, **d)
             for m, d in enumerate([dict(nx=300, nz=75, npml=6), dict(nx=57, nz=45, npml=5)])]
    for p in probs:
        p.cpml_split = split
    out = []
    for ref in (False, True):
        eng = _engine(probs, precision, every=24, reference=ref)
        g = torch.Generator().manual_seed(9)
        for s in range(eng.nshots):
            eng.init[s].copy_(torch.randn(eng.init[s].numel(), generator=g, dtype=torch.float64))
        tr = eng.forward()
        ys = [torch.randn(t.shape, generator=g, dtype=torch.float64) for t in tr]
        grad, gsrc = eng.backward(ys)
        out.append(([x.cpu().numpy() for x in grad], [x.cpu().numpy() for x in gsrc],
                    [eng.adjoint_state(s).double().cpu().numpy() for s in range(eng.nshots)]))
    tol = 1e-12 if precision == "fp64" else 1e-5
    (g0, s0, l0), (g1, s1, l1) = out
    for a, b in zip(g0, g1):
        for q in range(12):
            assert rel_l2(a[q], b[q]) < tol, ("grad slot", q, rel_l2(a[q], b[q]))
    for a, b in zip(s0, s1):
        assert rel_l2(a, b) < tol
    for a, b in zip(l0, l1):
        assert rel_l2(a, b) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_fused_adjoint_many_shots_per_gradient(precision):
    """Eleven shots adding into one gradient array at once (in FP32 through
    one staging array, with atomics): gradients, source gradients and
    lambda_0 against the three-kernel adjoint."""
    p = tiny_problem("elastic2", seed=87, nx=140, nz=70, npml=6, nt=60, gamma=1e-3, nshots=11, nrec=7)
    out = []
    for ref in (False, True):
        eng = _engine(p, precision, every=20, reference=ref)
        g = torch.Generator().manual_seed(12)
        for s in range(eng.nshots):
            eng.init[s].copy_(torch.randn(eng.init[s].numel(), generator=g, dtype=torch.float64))
        tr = eng.forward()
        ys = [torch.randn(t.shape, generator=g, dtype=torch.float64) for t in tr]
        grad, gsrc = eng.backward(ys)
        out.append(([x.cpu().numpy() for x in grad], [x.cpu().numpy() for x in gsrc],
                    [eng.adjoint_state(s).double().cpu().numpy() for s in range(eng.nshots)]))
    tol = 1e-12 if precision == "fp64" else 1e-5
    (g0, s0, l0), (g1, s1, l1) = out
    assert len(g0) == 1
    for q in range(12):
        if np.abs(g1[0][q]).max() > 0:
            assert rel_l2(g0[0][q], g1[0][q]) < tol, ("grad slot", q, rel_l2(g0[0][q], g1[0][q]))
    for a, b in zip(s0, s1):
        assert rel_l2(a, b) < tol
    for a, b in zip(l0, l1):
        assert rel_l2(a, b) < tol


@pytest.mark.gpu
def test_skipped_trivial_slots_give_same_gradients():
    """VTI without attenuation: the engine leaves the c15 / c35 and gamma
    slots off the device, yet its gradients (including those slots') equal
    the ones of an engine that keeps every slot."""
    p = tiny_problem("elastic2", seed=86, nx=50, nz=42, npml=6, nt=80, nrec=6, full21=False)
    for k in ("c15", "c35"):
        p.coef[k] = np.zeros_like(p.coef[k])
    res = []
    for keep in (False, True):
        eng = _engine(p, "fp64", every=32, all_slots=keep)
        assert (eng.sim.coef_t[0][2] is None) == (not keep)
        tr = eng.forward()
        g, _ = eng.backward([torch.ones_like(t) for t in tr])
        res.append(g[0].cpu().numpy())
    for q in range(12):
        assert rel_l2(res[0][q], res[1][q]) < 1e-13, q
    with pytest.raises(ValueError, match="all_slots"):
        eng0 = _engine(p, "fp64", every=32)
        eng0.set_coefs(0, [np.ones((p.nz, p.nx))] * 12)


@pytest.mark.gpu
def test_grad_mask_limits_accumulation():
    p = tiny_problem("elastic2", seed=85, nx=48, nz=40, npml=6, nt=40, nrec=5)
    eng = _engine(p, "fp64", every=20, grad_mask=(1 << 0) | (1 << 8))
    tr = eng.forward()
    grad, _ = eng.backward([torch.ones_like(t) for t in tr])
    g = grad[0].cpu().numpy()
    assert np.abs(g[0]).max() > 0 and np.abs(g[8]).max() > 0
    assert all(np.abs(g[q]).max() == 0 for q in range(12) if q not in (0, 8))


@pytest.mark.gpu
def test_trail_recompute_is_bit_exact():
    """The recomputed states equal the checkpoints of the forward run."""
    p = tiny_problem("elastic2", seed=52, nx=48, nz=40, npml=6, nt=64, gamma=1e-3, nrec=4)
    eng = _engine(p, "fp32", every=16)
    eng.forward()
    eng.backward([torch.zeros_like(eng.resid[0])])
    # after the backward pass trail buffer 0 holds segment 0's recomputed states
    assert torch.equal(eng.trail[0][0][16], eng.ckpt[0][1])
    assert torch.count_nonzero(eng.ckpt[0][1]) > 0


def _gpu_grads(p, engine, d_obs):
    """dJ/dparam of J = 1/2 sum (traces - d_obs)^2 through the autograd Function."""
    from seidart_amd.adjoint import elastic2_traces
    dev = engine.device
    prm = {k: torch.tensor(p.coef[k], dtype=torch.float64, device=dev, requires_grad=True) for k in PHYS}
    if p.gamma:
        prm["gamma_x"] = torch.tensor(p.gamma["x"], device=dev, requires_grad=True)
        prm["gamma_z"] = torch.tensor(p.gamma["z"], device=dev, requires_grad=True)
    tr = elastic2_traces(engine, [prm])
    J = 0.5 * ((tr - d_obs.to(dev)) ** 2).sum()
    J.backward()
    return float(J), {k: v.grad.cpu().numpy() for k, v in prm.items()}


def _ref_grads(p, d_obs):
    from torch_elastic2 import run_elastic2
    prm = {k: torch.tensor(p.coef[k], dtype=torch.float64, requires_grad=True) for k in PHYS}
    if p.gamma:
        prm["gx"] = torch.tensor(p.gamma["x"], requires_grad=True)
        prm["gz"] = torch.tensor(p.gamma["z"], requires_grad=True)
    J = 0
    for s in range(len(p.shots)):
        tr = run_elastic2(**_torch_ref_args(p, s, params=prm))
        nrec = tr.shape[2]
        J = J + 0.5 * ((tr - d_obs[s, :, :, :nrec]) ** 2).sum()
    J.backward()
    out = {k: v.grad.numpy() for k, v in prm.items()}
    if p.gamma:
        out["gamma_x"] = out.pop("gx")
        out["gamma_z"] = out.pop("gz")
    return float(J), out


def _observed(p, seed=7):
    """Synthetic 'observed' data: the oracle traces of a perturbed model."""
    from pyoracle import OracleRun
    q = tiny_problem("elastic2", seed=p._seed, **p._kw)
    q.cpml_split = p.cpml_split
    rng = np.random.default_rng(seed)
    for k in ("c11", "c33", "c55", "rho"):
        q.coef[k] = q.coef[k] * (1 + 0.03 * rng.standard_normal(q.coef[k].shape).clip(-2, 2))
    nrec = max(np.asarray(sh.rec).shape[0] for sh in q.shots)
    d = np.zeros((len(q.shots), q.nt, 2, nrec))
    for s in range(len(q.shots)):
        t = OracleRun(q, shot=s).run().trace
        d[s, :, :, : t.shape[2]] = t
    return torch.as_tensor(d)


def _grad_problem(split=False):
    kw = dict(nx=64, nz=50, npml=8, nt=160, gamma=1e-3, nshots=2, nrec=10)
    p = tiny_problem("elastic2", seed=61, **kw)
    p._seed, p._kw = 61, kw
    p.cpml_split = split
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("precision,split,every", [("fp32", False, 32), ("fp32", True, 25),
                                                   ("fp64", True, 32)])
def test_two_step_forward_and_recompute_keep_gradients(precision, split, every):
    """Two steps per launch in the checkpointed forward (odd segments end
    with one step) give the same misfit and checkpoints bit for bit as one
    step per launch, the recomputed trail equals the checkpoints, and the
    gradients agree up to the order of the atomic accumulation."""
    p = _grad_problem(split)
    d = _observed(p)
    out = []
    for spl in (1, 2):
        eng = _engine(p, precision, every=every)
        eng.sim.steps_per_launch = spl
        J, g = _gpu_grads(p, eng, d)
        out.append((J, g, eng.ckpt[0][1].clone()))
        if spl == 2:
            # after the backward pass trail buffer 0 holds segment 0's states
            assert torch.equal(eng.trail[0][0][every], eng.ckpt[0][1])
    assert out[0][0] == out[1][0]
    tol = 1e-6 if precision == "fp32" else 1e-13
    for k in out[0][1]:
        assert np.abs(out[0][1][k]).max() > 0, k
        assert rel_l2(out[1][1][k], out[0][1][k]) < tol, (k, rel_l2(out[1][1][k], out[0][1][k]))
    assert torch.equal(out[0][2], out[1][2])


@pytest.mark.gpu
@pytest.mark.parametrize("split", [False, True])
def test_gradient_matches_torch_autograd_fp64(split):
    p = _grad_problem(split)
    d = _observed(p)
    eng = _engine(p, "fp64", every=32)
    Jg, gg = _gpu_grads(p, eng, d)
    Jr, gr = _ref_grads(p, d)
    assert abs(Jg - Jr) <= 1e-12 * abs(Jr)
    for k in gr:
        assert np.abs(gr[k]).max() > 0, k
        assert rel_l2(gg[k], gr[k]) < 1e-9, (k, rel_l2(gg[k], gr[k]))


@pytest.mark.gpu
def test_gradient_fp32_against_fp64():
    p = _grad_problem()
    d = _observed(p)
    _, g64 = _gpu_grads(p, _engine(p, "fp64", every=32), d)
    _, g32 = _gpu_grads(p, _engine(p, "fp32", every=32), d)
    for k in ("c11", "c13", "c33", "c55", "rho"):
        assert rel_l2(g32[k], g64[k]) < 1e-3, (k, rel_l2(g32[k], g64[k]))


@pytest.mark.gpu
def test_source_series_gradient():
    """dJ/d(force series) from the adjoint equals the torch autograd one."""
    from seidart_amd.adjoint import Elastic2Traces
    from seidart_amd.problem import elastic2_coefs
    from torch_elastic2 import run_elastic2
    p = tiny_problem("elastic2", seed=62, nx=48, nz=42, npml=6, nt=120, nrec=6)
    eng = _engine(p, "fp64", every=40)
    sh = p.shots[0]
    dev = eng.device
    fx = torch.tensor(sh.force[0], device=dev, requires_grad=True)
    fz = torch.tensor(sh.force[2], device=dev, requires_grad=True)
    rho_s = float(p.coef["rho"][sh.src[2], sh.src[0]])
    inc = torch.stack([fx * p.dt / rho_s, torch.zeros_like(fx), fz * p.dt / rho_s])
    slots = torch.stack([torch.as_tensor(v, device=dev) for v in kernel_coefs(p)])
    w = torch.randn((p.nt, 2, 6), generator=torch.Generator().manual_seed(3), dtype=torch.float64)
    tr = Elastic2Traces.apply(eng, slots, inc)
    (tr[0] * w.to(dev)).sum().backward()
    a = _torch_ref_args(p, 0)
    rfx = a["force"][0].clone().requires_grad_(True)
    rfz = a["force"][1].clone().requires_grad_(True)
    a["force"] = (rfx, rfz)
    (run_elastic2(**a) * w).sum().backward()
    assert rel_l2(fx.grad.cpu().numpy(), rfx.grad.numpy()) < 1e-10
    assert rel_l2(fz.grad.cpu().numpy(), rfz.grad.numpy()) < 1e-10


@pytest.mark.gpu
def test_gradient_finite_differences():
    """Central differences of the FP64 device forward, relative step 1e-4,
    at 20 seeded cells per parameter among the cells whose gradient is at
    least 1% of the parameter's largest (smaller ones are round-off bound)."""
    from seidart_amd.adjoint import elastic2_traces
    p = _grad_problem()
    d = _observed(p)
    eng = _engine(p, "fp64", every=32)
    _, gg = _gpu_grads(p, eng, d)
    dev = eng.device
    base = {k: torch.tensor(p.coef[k], dtype=torch.float64, device=dev) for k in PHYS}
    base["gamma_x"] = torch.tensor(p.gamma["x"], device=dev)
    base["gamma_z"] = torch.tensor(p.gamma["z"], device=dev)
    dd = d.to(dev)

    def J(prm):
        with torch.no_grad():
            return float(0.5 * ((elastic2_traces(eng, [prm]) - dd) ** 2).sum())

    rng = np.random.default_rng(2024)
    worst = {}
    for k in ("c11", "c13", "c33", "c55", "rho"):
        g = gg[k]
        cand = np.argwhere(np.abs(g) >= 0.01 * np.abs(g).max())
        pick = cand[rng.choice(len(cand), size=min(20, len(cand)), replace=False)]
        errs = []
        for j, i in pick:
            h = 1e-4 * abs(float(p.coef[k][j, i]))
            up = {q: v.clone() for q, v in base.items()}
            dn = {q: v.clone() for q, v in base.items()}
            up[k][j, i] += h
            dn[k][j, i] -= h
            fd = (J(up) - J(dn)) / (2 * h)
            errs.append(abs(fd - g[j, i]) / abs(g[j, i]))
        worst[k] = max(errs)
    assert all(v < 1e-3 for v in worst.values()), worst


# ---------------------------------------------------------------------------
# EM 2-D (permittivity and conductivity gradients)
# ---------------------------------------------------------------------------
EMP = ("e11", "e13", "e33", "s11", "s13", "s33")


def _em_ref_args(p, shot, params=None):
    T = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    prm = params if params is not None else {k: T(p.coef[k]) for k in EMP}
    prof = {a: tuple(T(v) for v in (pr.kappa, pr.a, pr.b, pr.kappa_h, pr.a_h, pr.b_h))
            for a, pr in p.prof.items()}
    sh = p.shots[shot]
    force = tuple(None if f is None else T(f) for f in (sh.force[0], sh.force[2]))
    rec = torch.as_tensor(np.asarray(sh.rec)[:, [0, 2]])
    return dict(prm=prm, prof=prof, dx=p.dx, dz=p.dz, dt=p.dt, nt=p.nt, src=(sh.src[0], sh.src[2]),
                force=force, rec=rec)


def test_torch_em_reference_matches_oracle(oracle_lib):
    from torch_em2 import run_em2
    p = tiny_problem("em2", seed=45, nx=40, nz=36, npml=6, nt=80, nrec=7)
    with torch.no_grad():
        tr = run_em2(**_em_ref_args(p, 0)).numpy()
    o = oracle_lib.OracleRun(p, shot=0).run()
    assert np.abs(o.trace).max() > 0
    assert rel_l2(tr, o.trace) < 1e-13


def test_em_coefs_torch_matches_numpy():
    from seidart_amd.problem import em2_coefs
    p = tiny_problem("em2", seed=46, nx=30, nz=26, npml=4, nt=10)
    ref = kernel_coefs(p)
    got = em2_coefs({k: torch.as_tensor(v) for k, v in p.coef.items()}, p.dt)
    for a, b in zip(got, ref):
        assert np.max(np.abs(a.numpy() - b) / (np.abs(b) + 1e-300)) <= 1e-15


@pytest.mark.gpu
def test_em_dot_product():
    p = tiny_problem("em2", seed=53, nx=52, nz=44, npml=7, nt=110, nshots=2, nrec=9)
    for sh in p.shots:
        sh.force = (np.zeros(p.nt), None, np.zeros(p.nt))
    eng = _engine(p, "fp64", every=16)
    g = torch.Generator(device="cpu").manual_seed(6)
    xs = []
    for s in range(eng.nshots):
        x = torch.randn(eng.init[s].numel(), generator=g, dtype=torch.float64)
        eng.init[s].copy_(x.to(eng.device))
        xs.append(x)
    tr = eng.forward()
    ys = [torch.randn(t.shape, generator=g, dtype=torch.float64) for t in tr]
    lhs = sum(float((t.cpu() * y).sum()) for t, y in zip(tr, ys))
    eng.backward(ys)
    rhs = sum(float((x * eng.adjoint_state(s).cpu()).sum()) for s, x in enumerate(xs))
    assert abs(lhs) > 0
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_em_adjoint_matches_reference_kernels(precision):
    """The fused single-pass EM adjoint kernel against the three-kernel one
    (E^T, H^T, gather): same lambda_0, gradients and source gradients up to
    rounding, on two models of uneven size (one several tiles wide and tall,
    one narrower than a tile) with two shots each, C-PML on every side."""
    probs = [tiny_problem("em2", seed=90 + m, nt=70, nshots=2, nrec=7, **d)
             for m, d in enumerate([dict(nx=300, nz=75, npml=6), dict(nx=57, nz=45, npml=5)])]
    out = []
    for ref in (False, True):
        eng = _engine(probs, precision, every=24, reference=ref)
        g = torch.Generator().manual_seed(11)
        for s in range(eng.nshots):
            eng.init[s].copy_(torch.randn(eng.init[s].numel(), generator=g, dtype=torch.float64))
        tr = eng.forward()
        ys = [torch.randn(t.shape, generator=g, dtype=torch.float64) for t in tr]
        grad, gsrc = eng.backward(ys)
        out.append(([x.cpu().numpy() for x in grad], [x.cpu().numpy() for x in gsrc],
                    [eng.adjoint_state(s).double().cpu().numpy() for s in range(eng.nshots)]))
    tol = 1e-12 if precision == "fp64" else 1e-5
    (g0, s0, l0), (g1, s1, l1) = out
    for a, b in zip(g0, g1):
        for q in range(4):
            assert np.abs(b[q]).max() > 0, ("grad slot", q)
            assert rel_l2(a[q], b[q]) < tol, ("grad slot", q, rel_l2(a[q], b[q]))
    for a, b in zip(s0, s1):
        assert rel_l2(a, b) < tol
    for a, b in zip(l0, l1):
        assert rel_l2(a, b) < tol


def _em_problem():
    kw = dict(nx=60, nz=50, npml=8, nt=160, nshots=2, nrec=10)
    p = tiny_problem("em2", seed=63, **kw)
    p._seed, p._kw = 63, kw
    return p


def _em_observed(p, seed=8):
    from pyoracle import OracleRun
    q = tiny_problem("em2", seed=p._seed, **p._kw)
    rng = np.random.default_rng(seed)
    for k in ("e11", "e33", "s11", "s33"):
        q.coef[k] = q.coef[k] * (1 + 0.03 * rng.standard_normal(q.coef[k].shape).clip(-2, 2))
    nrec = max(np.asarray(sh.rec).shape[0] for sh in q.shots)
    d = np.zeros((len(q.shots), q.nt, 2, nrec))
    for s in range(len(q.shots)):
        t = OracleRun(q, shot=s).run().trace
        d[s, :, :, : t.shape[2]] = t
    return torch.as_tensor(d)


def _em_gpu_grads(p, engine, d_obs):
    from seidart_amd.adjoint import em2_traces
    dev = engine.device
    prm = {k: torch.tensor(p.coef[k], dtype=torch.float64, device=dev, requires_grad=True) for k in EMP}
    tr = em2_traces(engine, [prm])
    J = 0.5 * ((tr - d_obs.to(dev)) ** 2).sum()
    J.backward()
    return float(J.detach()), {k: v.grad.cpu().numpy() for k, v in prm.items()}


@pytest.mark.gpu
def test_em_gradient_matches_torch_autograd():
    from torch_em2 import run_em2
    p = _em_problem()
    d = _em_observed(p)
    Jg, gg = _em_gpu_grads(p, _engine(p, "fp64", every=32), d)
    prm = {k: torch.tensor(p.coef[k], dtype=torch.float64, requires_grad=True) for k in EMP}
    J = 0
    for s in range(len(p.shots)):
        tr = run_em2(**_em_ref_args(p, s, params=prm))
        J = J + 0.5 * ((tr - d[s, :, :, : tr.shape[2]]) ** 2).sum()
    J.backward()
    assert abs(Jg - float(J)) <= 1e-12 * abs(float(J))
    for k in EMP:
        ref = prm[k].grad.numpy()
        assert np.abs(ref).max() > 0, k
        assert rel_l2(gg[k], ref) < 1e-9, (k, rel_l2(gg[k], ref))
    _, g32 = _em_gpu_grads(p, _engine(p, "fp32", every=32), d)
    for k in ("e11", "e33", "s11", "s33"):
        assert rel_l2(g32[k], gg[k]) < 1e-3, (k, rel_l2(g32[k], gg[k]))


@pytest.mark.gpu
def test_em_gradient_finite_differences():
    from seidart_amd.adjoint import em2_traces
    p = _em_problem()
    d = _em_observed(p)
    eng = _engine(p, "fp64", every=32)
    _, gg = _em_gpu_grads(p, eng, d)
    dev = eng.device
    base = {k: torch.tensor(p.coef[k], dtype=torch.float64, device=dev) for k in EMP}
    dd = d.to(dev)

    def J(prm):
        with torch.no_grad():
            return float(0.5 * ((em2_traces(eng, [prm]) - dd) ** 2).sum())

    rng = np.random.default_rng(2026)
    worst = {}
    for k in ("e11", "e33", "s11", "s33"):
        g = gg[k]
        cand = np.argwhere(np.abs(g) >= 0.01 * np.abs(g).max())
        pick = cand[rng.choice(len(cand), size=min(20, len(cand)), replace=False)]
        errs = []
        for j, i in pick:
            h = 1e-4 * abs(float(p.coef[k][j, i]))
            up = {q: v.clone() for q, v in base.items()}
            dn = {q: v.clone() for q, v in base.items()}
            up[k][j, i] += h
            dn[k][j, i] -= h
            errs.append(abs((J(up) - J(dn)) / (2 * h) - g[j, i]) / abs(g[j, i]))
        worst[k] = max(errs)
    assert all(v < 1e-3 for v in worst.values()), worst
