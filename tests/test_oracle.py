"""The oracle (liboracle.so, a C restatement of the legacy loops) against the
reference's own arithmetic: the legacy Fortran compiled unmodified
(oracle/_ref) when it is built here, and the golden fixtures it wrote
(tests/golden) everywhere."""
import os

import numpy as np
import pytest

import legacy_ref
from golden.make_golden import CASES, case_problem
from helpers import rel_l2, tiny_problem

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CHAN = {"Vx": "vx", "Vy": "vy", "Vz": "vz", "Ex": "Ex", "Ey": "Ey", "Ez": "Ez"}


def _trace_index(kind, ch):
    return {"elastic2": ["Vx", "Vz"], "em2": ["Ex", "Ez"], "elastic25": ["Vx", "Vy", "Vz"],
            "em25": ["Ex", "Ey", "Ez"]}[kind].index(ch)


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_golden(oracle_lib, name):
    """Oracle traces and final fields equal the legacy solver's to round-off."""
    g = np.load(os.path.join(GOLDEN, f"legacy_{name}.npz"))
    p = case_problem(name)
    assert p.dt == pytest.approx(float(g["dt"]), rel=0, abs=0)
    np.testing.assert_array_equal(np.asarray(p.shots[0].rec), g["rec"])
    o = oracle_lib.OracleRun(p).run()
    kind = p.kind
    for key in g.files:
        if key.startswith("trace_"):
            ch = key[6:]
            got = o.trace[:, _trace_index(kind, ch), :]
            assert rel_l2(got, g[key]) < 1e-13, (name, ch)
        elif key.startswith("final_"):
            ch = key[6:]
            f = o.field(CHAN[ch])
            exp = g[key]
            sl = tuple(slice(0, s) for s in exp.shape)
            assert rel_l2(f[sl], exp) < 1e-13, (name, ch)
            assert np.abs(exp).max() > 0


@pytest.mark.parametrize("kind,kw", [
    ("elastic2", dict(seed=21, nx=34, nz=40, npml=5, nt=60, gamma=1e-3)),   # nz > nx (A2 clamp off-range rows unused by legacy? no: skip)
    ("em2", dict(seed=22, nx=30, nz=26, npml=4, nt=70)),
    ("elastic25", dict(seed=23, nx=18, ny=10, nz=16, npml=3, nt=30)),
    ("em25", dict(seed=24, nx=18, ny=10, nz=16, npml=3, nt=30)),
])
def test_oracle_matches_legacy_build(oracle_lib, kind, kw):
    """Fresh seeded cases straight through the compiled legacy solver."""
    if not legacy_ref.available(kind):
        pytest.skip("oracle/_ref not built (needs the reference tree and amdflang)")
    if kind == "elastic2" and kw["nz"] > kw["nx"]:
        kw = dict(kw, nx=kw["nz"] + 2)  # the legacy reads b_x_half(j) past its end when nz > nx
    p0 = tiny_problem(kind, **kw)
    dt = legacy_ref.legacy_dt(kind, p0.coef, p0.dx, p0.dy, p0.dz)
    p = tiny_problem(kind, dt=dt, **kw)
    if "z" in p.gamma:
        p.gamma = {"x": p.gamma["x"]}
    leg = legacy_ref.run_legacy(p)
    o = oracle_lib.OracleRun(p).run()
    for ch, arr in leg.items():
        exp = arr[-1]
        sl = tuple(slice(0, s) for s in exp.shape)
        assert rel_l2(o.field(CHAN[ch])[sl], exp) < 1e-13, ch


def test_oracle_segments_equal_full_run(oracle_lib):
    """A run split into segments (state and psi kept) equals one call."""
    p = tiny_problem("elastic2", seed=5, nt=50)
    full = oracle_lib.OracleRun(p).run()
    seg = oracle_lib.OracleRun(p)
    for it0, n in ((0, 7), (7, 20), (27, 23)):
        seg.run(it0, n)
    np.testing.assert_array_equal(full.trace, seg.trace)
    for n in full.names:
        np.testing.assert_array_equal(full.field(n), seg.field(n))


def test_oracle_threads_deterministic(oracle_lib):
    p = tiny_problem("elastic25", seed=6, nx=20, ny=10, nz=16, npml=3, nt=20)
    a = oracle_lib.OracleRun(p, threads=1).run()
    b = oracle_lib.OracleRun(p, threads=4).run()
    np.testing.assert_array_equal(a.trace, b.trace)


def test_oracle_blowup_flag(oracle_lib):
    """dt far above the CFL limit trips the 1e25 check (seismicfdtd.f95:378-379)."""
    p = tiny_problem("elastic2", seed=7, nt=200)
    p.dt *= 8
    o = oracle_lib.OracleRun(p).run()
    assert o.blowup >= 0
    q = tiny_problem("elastic2", seed=7, nt=200)
    assert oracle_lib.OracleRun(q).run().blowup == -1


def test_oracle_energy_conserved_without_pml(oracle_lib):
    """No C-PML, no attenuation, no source after it is switched off: the
    discrete energy of the staggered scheme stays bounded and the leapfrog
    solution does not grow (closed rigid box)."""
    from seidart_amd.cpml import AxisProfile
    p = tiny_problem("elastic2", seed=8, nt=400, full21=False)
    p.prof = {"x": AxisProfile.trivial(p.nx), "z": AxisProfile.trivial(p.nz)}
    f = np.asarray(p.shots[0].force[2]).copy()
    f[40:] = 0.0
    p.shots[0].force = (None, None, f)
    o = oracle_lib.OracleRun(p)
    e = []
    for it in range(0, p.nt, 40):
        o.run(it, 40)
        e.append(sum(float(np.sum(o.field(n) ** 2)) for n in ("vx", "vz")))
    assert max(e[2:]) < 50 * min(e[2:])


def _homogeneous_iso(kind, n=41, ny=1, nt=24):
    """Homogeneous isotropic elastic grid without C-PML, a source in the
    middle and no force: the mirror-symmetry setting of the moment tests."""
    from seidart_amd.cpml import AxisProfile
    from seidart_amd.problem import STIFFNESS, Problem, Shot
    lam, mu, rho = 2.0e9, 1.0e9, 1000.0
    is3d = kind == "elastic25"
    shape = (n, n)
    C = {k: np.zeros(shape) for k in STIFFNESS}
    for k in ("c11", "c22", "c33"):
        C[k][:] = lam + 2 * mu
    for k in ("c12", "c13", "c23"):
        C[k][:] = lam
    for k in ("c44", "c55", "c66"):
        C[k][:] = mu
    C["rho"][:] = rho
    dx = 1.0
    dt = 0.3 * dx / np.sqrt((lam + 2 * mu) / rho)
    prof = {"x": AxisProfile.trivial(n), "z": AxisProfile.trivial(n)}
    if is3d:
        prof["y"] = AxisProfile.trivial(ny)
    c = n // 2
    t = np.arange(nt) * dt
    rate = np.exp(-((t - 8 * dt) / (3 * dt)) ** 2) * 1e6
    moment = (rate, rate if is3d else None, rate, None, None, None)   # explosive (tnt)
    sh = Shot(src=(c, ny // 2 if is3d else 0, c), rec=np.zeros((0, 3), np.int32), moment=moment)
    p = Problem(kind, n, ny if is3d else 1, n, dx, dx, dx, float(dt), nt, C, prof, [sh])
    p.validate()
    return p, c


def test_oracle_explosive_moment_source_mirror_symmetry(oracle_lib):
    """An isotropic stress-rate source (xx = zz) at a stress node of a
    homogeneous isotropic medium: the field is mirror symmetric about the
    source, vx odd about the node in x (vx[i] at i - 1/2: vx[2c + 1 - i] =
    -vx[i]) and vz odd in z (vz[j] at j + 1/2: vz[2c - 1 - j] = -vz[j]),
    sxx, szz even; the wave has not reached the edges."""
    p, c = _homogeneous_iso("elastic2")
    o = oracle_lib.OracleRun(p).run()
    vx, vz, sxx = o.field("vx"), o.field("vz"), o.field("sxx")
    n = p.nx
    assert np.abs(vx).max() > 0 and np.abs(vz).max() > 0
    i = np.arange(2, n - 2)
    np.testing.assert_allclose(vx[:, 2 * c + 1 - i], -vx[:, i], rtol=0, atol=1e-12 * np.abs(vx).max())
    np.testing.assert_allclose(vz[2 * c - 1 - i, :], -vz[i, :], rtol=0, atol=1e-12 * np.abs(vz).max())
    np.testing.assert_allclose(sxx[:, 2 * c - i], sxx[:, i], rtol=0, atol=1e-12 * np.abs(sxx).max())
    # the increment is rate * dt at the source cell: after one step (zero
    # velocities before it) sigma_xx there is exactly that
    o1 = oracle_lib.OracleRun(p).run(0, 1)
    rate0 = p.shots[0].moment[0][0]
    assert o1.field("sxx")[c, c] == rate0 * p.dt
    assert o1.field("szz")[c, c] == rate0 * p.dt and o1.field("sxz")[c, c] == 0.0
