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


# ---------------------------------------------------------------------------
# Pins of the oracle where the legacy Fortran leaves the behaviour open or
# has none (attenuation in 2-D z and in 2.5-D, the 2.5-D stress-rate source,
# the split C-PML variant): closed forms and invariants.
# ---------------------------------------------------------------------------
def _homogeneous_decay(kind, n=48, ny=1, nt=8, v0=(0.7, -0.4, 1.3), gam=(2e-3, 5e-3, 9e-3)):
    """Homogeneous isotropic grid without C-PML or source, every velocity
    component uniform (v0) and a distinct uniform attenuation per component
    (gam = x, y, z): away from the Dirichlet edges the stresses stay zero and
    each component decays as v0 * (1 - gamma)^n."""
    p, _ = _homogeneous_iso(kind, n=n, ny=ny, nt=nt)
    p.shots[0].moment = (None,) * 6
    is3d = kind == "elastic25"
    comps = ("vx", "vy", "vz") if is3d else ("vx", "vz")
    axes = ("x", "y", "z") if is3d else ("x", "z")
    g = dict(zip(("x", "y", "z"), gam))
    v = dict(zip(("vx", "vy", "vz"), v0))
    p.gamma = {a: np.full((p.nz, p.nx), g[a]) for a in axes}
    p.init = [{c: np.full(p.field_shape, v[c]) for c in comps}]
    expect = {}
    for c, a in zip(comps, axes):
        e = v[c]
        for _ in range(nt):
            e = e * (1.0 - g[a])
        expect[c] = e
    return p, expect


def _interior(a, m):
    if a.ndim == 2:
        return a[m:-m, m:-m]
    return a[m:-m, m:-m, m:-m]


@pytest.mark.parametrize("kind", ["elastic2", "elastic25"])
def test_oracle_attenuation_closed_form(oracle_lib, kind):
    """1 - gamma_x scales vx, 1 - gamma_y vy, 1 - gamma_z vz once per step
    (seismicfdtd.f95:334, :354 in 2-D; the same rule in 2.5-D, where the
    legacy has no attenuation): uniform fields far from the edges decay as
    v0 (1 - gamma)^n with each component's own gamma."""
    nt = 8
    p, expect = _homogeneous_decay(kind, ny=41 if kind == "elastic25" else 1, nt=nt)
    o = oracle_lib.OracleRun(p).run()
    m = 2 * nt + 4
    for c, e in expect.items():
        got = _interior(o.field(c), m)
        assert got.size > 0
        np.testing.assert_allclose(got, e, rtol=1e-14, atol=0, err_msg=c)
    for s in (("sxx", "szz", "sxz") if kind == "elastic2" else ("sxx", "syy", "szz", "sxy", "sxz", "syz")):
        assert np.abs(_interior(o.field(s), m)).max() == 0.0, s


def test_oracle_explosive_moment_source_mirror_symmetry_25d(oracle_lib):
    """2.5-D counterpart of the 2-D mirror test: an isotropic stress-rate
    source (xx = yy = zz) at a normal-stress node.  Staggering of seismic25
    (seismicfdtd.f95:821-896): vx at i - 1/2, vy at j + 1/2, vz at k + 1/2,
    normal stresses on the nodes.  So vx is odd about the source in x and
    even in y and z, vy odd in y, vz odd in z, the normal stresses even; and
    one step from rest leaves exactly rate * dt in sxx, syy, szz at the
    source."""
    p, c = _homogeneous_iso("elastic25", n=41, ny=41, nt=24)
    cy = p.ny // 2
    o = oracle_lib.OracleRun(p).run()
    vx, vy, vz, sxx, syy = (o.field(n) for n in ("vx", "vy", "vz", "sxx", "syy"))
    for a in (vx, vy, vz, sxx):
        assert np.abs(a).max() > 0
    i = np.arange(2, p.nx - 2)
    j = np.arange(2, p.ny - 2)
    tol = lambda a: 1e-12 * np.abs(a).max()
    np.testing.assert_allclose(vx[:, :, 2 * c + 1 - i], -vx[:, :, i], rtol=0, atol=tol(vx))
    np.testing.assert_allclose(vx[:, 2 * cy - j, :], vx[:, j, :], rtol=0, atol=tol(vx))
    np.testing.assert_allclose(vx[2 * c - i, :, :], vx[i, :, :], rtol=0, atol=tol(vx))
    np.testing.assert_allclose(vy[:, 2 * cy - 1 - j, :], -vy[:, j, :], rtol=0, atol=tol(vy))
    np.testing.assert_allclose(vy[:, :, 2 * c - i], vy[:, :, i], rtol=0, atol=tol(vy))
    np.testing.assert_allclose(vz[2 * c - 1 - i, :, :], -vz[i, :, :], rtol=0, atol=tol(vz))
    np.testing.assert_allclose(sxx[:, :, 2 * c - i], sxx[:, :, i], rtol=0, atol=tol(sxx))
    np.testing.assert_allclose(syy[:, 2 * cy - j, :], syy[:, j, :], rtol=0, atol=tol(syy))
    o1 = oracle_lib.OracleRun(p).run(0, 1)
    r0 = p.shots[0].moment[0][0] * p.dt
    for n in ("sxx", "syy", "szz"):
        assert o1.field(n)[c, cy, c] == r0, n
    for n in ("sxy", "sxz", "syz"):
        assert o1.field(n)[c, cy, c] == 0.0, n


def _homogeneous_ice_with_cpml(nx0, nz0, cpml, dt, nt, src, recs, f0=80.0):
    """Homogeneous isotropic ice (Vp 3840, Vs 1920, rho 917), dx = 1 m, a
    45-degree Ricker force; C-PML of `cpml` cells (0: a rigid box)."""
    from seidart_amd.cpml import AxisProfile, alpha_max_seismic, axis_profile, sigma_max_seismic
    from seidart_amd.problem import STIFFNESS, Problem, Shot
    from seidart_amd.sources import point_force
    vp, vs, rho = 3840.0, 1920.0, 917.0
    nx, nz = nx0 + 2 * cpml, nz0 + 2 * cpml
    C = {k: np.zeros((nz, nx)) for k in STIFFNESS}
    for k in ("c11", "c22", "c33"):
        C[k][:] = rho * vp * vp
    for k in ("c12", "c13", "c23"):
        C[k][:] = rho * (vp * vp - 2 * vs * vs)
    for k in ("c44", "c55", "c66"):
        C[k][:] = rho * vs * vs
    C["rho"][:] = rho
    if cpml:
        sm, am = sigma_max_seismic(vp, cpml, 2, 0.001), alpha_max_seismic(1.0, f0)
        prof = {a: axis_profile(n, cpml, 1.0, sm, am, 5.0, 2, 2, dt, True) for a, n in (("x", nx), ("z", nz))}
    else:
        prof = {"x": AxisProfile.trivial(nx), "z": AxisProfile.trivial(nz)}
    fx, _, fz = point_force(nt, dt, f0, 1.0, "gaus1", xz_rotation=45.0)
    rec = np.array([[x + cpml, 0, z + cpml] for x, z in recs], np.int32)
    sh = Shot(src=(src[0] + cpml, 0, src[1] + cpml), force=(fx, None, fz), rec=rec)
    p = Problem("elastic2", nx, 1, nz, 1.0, 1.0, 1.0, dt, nt, C, prof, [sh])
    p.validate()
    return p


REFL_DT = 1.0 / (2 * 3840.0)       # CFL 1/sqrt(2) in 2-D at dx = 1 m
REFL_SRC = (60, 100)
REFL_RECS = [(10, z) for z in range(20, 181, 20)] + [(x, 10) for x in range(20, 181, 20)]
REFL_BOUND = 0.02                  # -34 dB of the direct wave


def reflection_case(nt=900):
    """A 200 x 200 ice domain with C-PML 20 and the same source and
    receivers (10 cells from the left and top C-PML) placed in the middle of
    a 600 x 600 rigid box, whose walls no wave reaches the receivers from
    within nt = 900 steps (P speed 0.5 cells per step; the nearest wall path
    is 940 steps).  The difference of the traces is what the C-PML
    reflects."""
    small = _homogeneous_ice_with_cpml(200, 200, 20, REFL_DT, nt, REFL_SRC, REFL_RECS)
    off = 200
    big = _homogeneous_ice_with_cpml(600, 600, 0, REFL_DT, nt, (REFL_SRC[0] + off, REFL_SRC[1] + off),
                                     [(x + off, z + off) for x, z in REFL_RECS])
    return small, big


def test_oracle_split_cpml_without_memory_equals_legacy(oracle_lib):
    """Profiles with kappa != 1 but a = 0 keep every memory variable at zero,
    where the split variant and the legacy sharing (quirks A1, A2) coincide:
    bit-identical runs."""
    from seidart_amd.cpml import AxisProfile
    p = tiny_problem("elastic2", seed=71, nx=40, nz=36, npml=6, nt=60, gamma=1e-3, nrec=6)
    for a in ("x", "z"):
        pr = p.prof[a]
        p.prof[a] = AxisProfile(pr.kappa, np.zeros_like(pr.a), pr.b, pr.kappa_h, np.zeros_like(pr.a_h), pr.b_h)
    assert max(np.abs(p.prof[a].kappa - 1).max() for a in ("x", "z")) > 0.5
    a = oracle_lib.OracleRun(p).run()
    p.cpml_split = True
    b = oracle_lib.OracleRun(p).run()
    np.testing.assert_array_equal(a.trace, b.trace)
    for n in a.names:
        np.testing.assert_array_equal(a.field(n), b.field(n))
    assert np.abs(a.trace).max() > 0


def test_oracle_split_cpml_reflection_below_bound(oracle_lib):
    """The split C-PML reflects less than -34 dB of the direct wave back to
    receivers 10 cells from it (measured -40 dB).  The legacy sharing of the
    memory arrays (quirks A1, A2) does not absorb at all on this case: its
    C-PML feeds a growing mode whose 'reflection' exceeds the direct wave."""
    small, big = reflection_case()
    ref = oracle_lib.OracleRun(big).run().trace
    small.cpml_split = True
    r = rel_l2(oracle_lib.OracleRun(small).run().trace, ref)
    assert r < REFL_BOUND, r
    small.cpml_split = False
    assert rel_l2(oracle_lib.OracleRun(small).run().trace, ref) > 10 * REFL_BOUND


def test_oracle_split_cpml_energy_decays(oracle_lib):
    """Once the wave has left the 200 x 200 C-PML domain the split variant's
    energy falls by 1e7 and keeps falling over 4000 steps; the legacy
    sharing grows without bound on the same case."""
    small, _ = reflection_case(nt=4000)
    e = {}
    for split in (True, False):
        small.cpml_split = split
        o = oracle_lib.OracleRun(small)
        e[split] = []
        for it in range(0, small.nt, 500):
            o.run(it, 500)
            e[split].append(sum(float(np.sum(o.field(n) ** 2)) for n in ("vx", "vz")))
    s = e[True]
    assert max(s[3:]) < 1e-7 * max(s[:2])
    assert all(b <= a * 1.0001 for a, b in zip(s[3:], s[4:])), s
    assert e[False][-1] > 1e6 * max(e[False][:2])


@pytest.mark.parametrize("name", ["elastic2", "em2"])
def test_legacy_snapshot_fixtures_match_oracle(oracle_lib, name):
    """The legacy writer's snapshot records (tests/golden/legacy_snapshots_*):
    one Fortran record of FP64 values at the channel's legacy extent (EM:
    Ex (nx-1) x nz, Ez nx x (nz-1); readwrite_routines.f95:473-497), equal to
    the oracle's fields after those steps."""
    import struct
    g = np.load(os.path.join(GOLDEN, f"legacy_snapshots_{name}.npz"))
    p = case_problem(name)
    n = len([k for k in g.files if k.startswith("bytes_")])
    steps = sorted({int(g[f"step_{q}"]) for q in range(n)})
    o = oracle_lib.OracleRun(p)
    done = 0
    fields = {}
    for it in steps:
        o.run(done, it - done)
        done = it
        for ch in ("Vx", "Vz", "Ex", "Ez"):
            if CHAN[ch] in o.names:
                fields[(ch, it)] = o.field(CHAN[ch]).copy()
    ext = {"Vx": (slice(None), slice(None)), "Vz": (slice(None), slice(None)),
           "Ex": (slice(None), slice(0, -1)), "Ez": (slice(0, -1), slice(None))}
    for q in range(n):
        ch, it, raw = str(g[f"channel_{q}"]), int(g[f"step_{q}"]), g[f"bytes_{q}"].tobytes()
        exp = fields[(ch, it)][ext[ch]]
        (m,) = struct.unpack("<i", raw[:4])
        assert m == exp.size * 8 == len(raw) - 8 and struct.unpack("<i", raw[-4:])[0] == m
        got = np.frombuffer(raw[4:-4], dtype=np.float64).reshape(exp.shape)
        assert np.abs(exp).max() > 0
        assert rel_l2(got, exp) < 1e-13, (ch, it)


@pytest.mark.parametrize("name", ["elastic2", "em2"])
def test_reference_project_files_give_the_golden_problem(tmp_path, name):
    """The golden case written in the reference's file contract reads back
    (seidart_amd.run.problem_from_files) as the same Problem."""
    from helpers import write_reference_project
    from seidart_amd import run as dropin
    p = case_problem(name)
    jf, rf = write_reference_project(p, str(tmp_path), cpml=CASES[name]["npml"])
    q = dropin.problem_from_files(jf, seismic=p.kind.startswith("elastic"), receivers=rf)
    assert (q.nx, q.nz, q.nt, q.dt) == (p.nx, p.nz, p.nt, p.dt)
    assert tuple(q.shots[0].src) == tuple(p.shots[0].src)
    np.testing.assert_array_equal(q.shots[0].rec, p.shots[0].rec)
    for k, v in p.coef.items():
        np.testing.assert_array_equal(q.coef[k], v)
    for a in ("x", "z"):
        for f in ("kappa", "a", "b", "kappa_h", "a_h", "b_h"):
            np.testing.assert_array_equal(getattr(q.prof[a], f), getattr(p.prof[a], f))
