"""Host-side setup (FP64): C-PML profiles, sources, rotations, file formats,
the drop-in file loader.  Pinned by closed forms and invariants."""
import json
import os

import numpy as np
import pytest

from seidart_amd import configs, cpml, fortranio, problem, run as dropin, sources, survey


# --------------------------------------------------------------------- C-PML
@pytest.mark.parametrize("seismic", [True, False])
def test_cpml_interior_and_bounds(seismic):
    n, c = 80, 12
    dt = 1e-4 if seismic else 1e-11
    sm = cpml.sigma_max_seismic(3000.0, c, 2, 1e-3) if seismic else cpml.sigma_max_em(c, 0.01, 2, 1e-3)
    am = cpml.alpha_max_seismic(1.0, 100.0) if seismic else cpml.alpha_max_em(1.0, 5e8)
    p = cpml.axis_profile(n, c, 0.5, sm, am, 5.0, 2, 2, dt, seismic)
    inner = slice(c, n - c)
    for arr, val in ((p.kappa, 1.0), (p.a, 0.0), (p.b, 1.0), (p.kappa_h, 1.0), (p.a_h, 0.0), (p.b_h, 1.0)):
        np.testing.assert_array_equal(arr[inner], val)
    for b in (p.b, p.b_h):
        assert np.all((b > 0) & (b <= 1))
    for a in (p.a, p.a_h):
        assert np.all(a <= 0)
    # mirror symmetry of the two slabs
    np.testing.assert_array_equal(p.kappa[:c], p.kappa[::-1][:c])
    np.testing.assert_array_equal(p.a_h[:c], p.a_h[::-1][:c])
    assert cpml.slab_width(p) == c


def test_cpml_closed_form_at_the_outer_cell():
    """At the outermost cell the normalised distance is 1: sigma = sigma_max,
    kappa = kappa_max, alpha = 0, b = exp(-sigma_max/kappa_max dt),
    a = (b - 1)/kappa_max (definitions.py:940-996)."""
    n, c, dt, kmax = 60, 10, 2e-4, 5.0
    sm, am = 1234.0, 300.0
    p = cpml.axis_profile(n, c, 1.0, sm, am, kmax, 2, 2, dt, True)
    b = np.exp(-(sm / kmax) * dt)
    for idx in (0, n - 1):
        assert p.sigma[idx] == pytest.approx(sm, rel=1e-15)
        assert p.kappa[idx] == pytest.approx(kmax, rel=1e-15)
        assert p.b[idx] == pytest.approx(b, rel=1e-14)
        assert p.a[idx] == pytest.approx((b - 1) / kmax, rel=1e-13)
    # innermost slab cell: d = 1/c, r = d^2
    r = (1.0 / c) ** 2
    assert p.kappa[c - 1] == pytest.approx(1 + (kmax - 1) * r, rel=1e-14)
    assert p.sigma[c - 1] == pytest.approx(sm * r, rel=1e-14)
    assert p.alpha[c - 1] == pytest.approx(am * (1 - 1.0 / c) ** 2, rel=1e-14)


def test_sigma_max_formulas():
    # -ln(R) (NP+1) 0.7 v / (2 cpml) and -(NP+1) ln(R) eps0 c / (2 cpml dmin)
    assert cpml.sigma_max_seismic(2000.0, 10, 2, 1e-3) == pytest.approx(
        np.log(1e3) * 3 * 0.7 * 2000.0 / 20, rel=1e-15)
    assert cpml.sigma_max_em(10, 0.02, 2, 1e-3) == pytest.approx(
        3 * np.log(1e3) * cpml.EPS0 * cpml.CLIGHT / (20 * 0.02), rel=1e-15)


def test_profiles_from_2d_fields():
    n_x, n_z, c = 50, 40, 8
    px = cpml.axis_profile(n_x, c, 1.0, 500.0, 100.0, 4.0, 2, 1, 1e-4, True)
    pz = cpml.axis_profile(n_z, c, 1.0, 700.0, 100.0, 4.0, 2, 1, 1e-4, True)

    def field(ax, az):
        f = np.ones((n_z, n_x))
        f[:, :c] = ax[None, :c]
        f[:, n_x - c:] = ax[None, n_x - c:]
        f[:c, c:n_x - c] = az[:c, None]
        f[n_z - c:, c:n_x - c] = az[n_z - c:, None]
        return f
    flds = [field(getattr(px, q), getattr(pz, q)) for q in ("kappa", "a", "b", "kappa_h", "a_h", "b_h")]
    # the interior of a and the exterior of the other axis hold the trivial value
    for q, f in zip(("a", "a_h"), (flds[1], flds[4])):
        f[c:n_z - c, c:n_x - c] = 0.0
    gx, gz = cpml.profiles_from_2d(*flds, cpml=c)
    np.testing.assert_array_equal(gx.kappa, px.kappa)
    np.testing.assert_array_equal(gz.b_h, pz.b_h)


# ------------------------------------------------------------------- sources
def test_ricker_closed_form():
    f = 100.0
    dt = 1e-5
    t = np.arange(4000) * dt
    w = sources.wavelet(t, f, "gaus1")
    k = int(round(1.0 / f / dt))
    assert w[k] == pytest.approx(1.0, abs=1e-12)
    assert np.argmax(w) == k
    tz = 1 / f + 1 / (np.sqrt(2) * np.pi * f)   # zero crossing (1 - 2a^2 = 0)
    a = np.pi * f * (tz - 1 / f)
    assert (1 - 2 * a * a) == pytest.approx(0.0, abs=1e-12)
    assert abs(np.interp(tz, t, w)) < 1e-3


def test_rotate_sdr_and_force_mapping():
    R = sources.rotate_sdr(30, 40, 50)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)
    fx, fy, fz = sources.point_force(50, 1e-4, 100.0)
    assert np.abs(fy).max() == 0 and np.abs(fz).max() < 1e-16 and np.abs(fx).max() == pytest.approx(1.0)
    fx, fy, fz = sources.point_force(50, 1e-4, 100.0, xz_rotation=90.0)
    assert np.abs(fx).max() < 1e-15 and np.abs(fz).max() == pytest.approx(1.0)


# -------------------------------------------------------------- anisotropy
def _rand_rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_rotate_stiffness_invariants():
    rng = np.random.default_rng(1)
    lam, mu = 3e9, 2e9
    iso = np.zeros((6, 6))
    iso[:3, :3] = lam
    iso[np.arange(3), np.arange(3)] = lam + 2 * mu
    iso[np.arange(3, 6), np.arange(3, 6)] = mu
    R1, R2 = _rand_rot(rng), _rand_rot(rng)
    np.testing.assert_allclose(configs.rotate_stiffness(iso, R1), iso, rtol=1e-12, atol=1e-3)
    C = configs._ice_crystal_voigt()
    a = configs.rotate_stiffness(configs.rotate_stiffness(C, R1), R2)
    b = configs.rotate_stiffness(C, R2 @ R1)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-2)
    np.testing.assert_allclose(configs.rotate_stiffness(C, np.eye(3)), C, rtol=0, atol=0)
    # phase velocities along R n in the rotated medium equal those along n
    n = np.array([0.3, -0.5, 0.8])

    def speeds(C6, d):
        G = np.einsum("ijkl,j,l->ik", configs.voigt_to_tensor(C6), d, d) / 917.0
        return np.linalg.eigvalsh(G)
    np.testing.assert_allclose(speeds(configs.rotate_stiffness(C, R1), R1 @ n), speeds(C, n), rtol=1e-12)


def _row(C6, rho=917.0):
    return dict({f"c{i + 1}{j + 1}": C6[i, j] for i in range(6) for j in range(i, 6)}, rho=rho)


def test_christoffel_vmax_follows_the_reference_convention():
    """sigma_max of the object route uses the reference's Christoffel
    velocities (materials.py:903-966, 1014-1172: shear index pairs scaled
    by 1/sqrt(2)); closed forms of its matrix along x and z."""
    C = configs._ice_crystal_voigt()
    # ice along the c-axis: Gamma = diag(c55/2, c44/2, c33) / rho
    assert dropin.christoffel_vmax(_row(C), (0, 0, 1)) == pytest.approx(np.sqrt(15.01e9 / 917), rel=1e-12)
    # a shear-dominated medium along x: Gamma = diag(c11, c66/2, c55/2) / rho
    S = np.diag([1e9, 1e9, 1e9, 3e9, 5e9, 8e9])
    assert dropin.christoffel_vmax(_row(S, 1000.0), (1, 0, 0)) == pytest.approx(np.sqrt(4e9 / 1000), rel=1e-12)
    assert dropin.christoffel_vmax(_row(S, 1000.0), (0, 0, 1)) == pytest.approx(np.sqrt(2.5e9 / 1000), rel=1e-12)
    # a c16 coupling along x: the block [[c11, c16/sqrt2], [c16/sqrt2, c66/2]]
    K = np.diag([2e9, 1e9, 1e9, 1e9, 1e9, 6e9])
    K[0, 5] = K[5, 0] = 1.5e9
    a, b, c = 2e9, 3e9, 1.5e9 / np.sqrt(2)
    lam = 0.5 * (a + b + np.sqrt((a - b) ** 2 + 4 * c * c))
    assert dropin.christoffel_vmax(_row(K, 1000.0), (1, 0, 0)) == pytest.approx(np.sqrt(lam / 1000), rel=1e-12)
    # an isotropic medium: the P speed along the axes; along (1, 0, 1) the
    # scaled shear terms give the xz block [[c11 + c55/2, c13 + c55/2],
    # [c13 + c55/2, c33 + c55/2]] / 2, largest eigenvalue (c11 + c13 + c55) / 2
    iso = np.zeros((6, 6))
    iso[:3, :3] = 3e9
    iso[np.arange(3), np.arange(3)] = 7e9
    iso[np.arange(3, 6), np.arange(3, 6)] = 2e9
    for d in ((1, 0, 0), (0, 0, 1)):
        assert dropin.christoffel_vmax(_row(iso, 2000.0), d) == pytest.approx(np.sqrt(7e9 / 2000), rel=1e-12)
    assert dropin.christoffel_vmax(_row(iso, 2000.0), (1, 0, 1)) == pytest.approx(
        np.sqrt((7e9 + 3e9 + 2e9) / 2 / 2000), rel=1e-12)


def _objects(seismic=True):
    """Duck-typed Model / Domain after Model.build: three materials in the
    table, two painted (labels 0 and 2); material 1, absent from the
    geometry, is the fastest and must not set sigma_max."""
    import pandas as pd
    from types import SimpleNamespace as NS
    nx0, nz0, cpml_ = 9, 7, 3
    geom = np.zeros((nx0, nz0), dtype=int)
    geom[:, 4:] = 2
    geom[3, 1] = 2
    if seismic:
        rows = []
        for c11, c66, c16, rho in ((4e9, 1e9, 0.0, 2000.0), (9e10, 1e9, 0.0, 2000.0),
                                   (2e9, 7e9, 1.5e9, 1000.0)):
            r = {k: 0.0 for k in problem.STIFFNESS}
            r.update(c11=c11, c22=c11, c33=c11, c12=0.3 * c11, c13=0.3 * c11, c23=0.3 * c11,
                     c44=c66, c55=c66, c66=c66, c16=c16, rho=rho)
            rows.append(r)
        tab = dict(stiffness_coefficients=pd.DataFrame(rows), attenuation_coefficients=None)
    else:
        e = pd.DataFrame([{k: v for k, v in zip(problem.PERMITTIVITY, (3.2, 0, 0, 3.2, 0, 3.1 + m))}
                          for m in range(3)])
        c = pd.DataFrame([{k: v for k, v in zip(problem.CONDUCTIVITY, (1e-4 * (m + 1), 0, 0, 1e-4, 0, 2e-4))}
                          for m in range(3)])
        tab = dict(permittivity_coefficients=e, conductivity_coefficients=c)
    nt = 12
    model = NS(is_seismic=seismic, time_steps=nt, f0=100.0, dt=1e-4 if seismic else 1e-11,
               velocity_scaling_factor=1.0, x=2.0, y=0.0, z=1.5,
               sourcefunction_x=np.linspace(0, 1, nt), sourcefunction_y=np.zeros(nt),
               sourcefunction_z=np.linspace(1, 0, nt), density_method="harmonic", **tab)
    domain = NS(dim=2.0, cpml=cpml_, nx=nx0, nz=nz0, ny=1, dx=0.5, dz=0.25, dy=1.0, geometry=geom,
                nmats=1, NP=2.0, NPA=1.0, Rcoef=0.001, kappa_max=5.0, alpha_max_scalar=1.0)
    return model, domain


def test_problem_from_objects_seismic():
    """The object route (Model.run replacement): tables painted on the
    geometry and edge-extended into the C-PML (tensor2dat), sigma_max from
    the largest axis speed of the painted materials in the reference's
    Christoffel convention (cpmlcompute, definitions.py:612-652), the
    model's dt, source series and source cell."""
    model, domain = _objects(True)
    p = dropin.problem_from_objects(model, domain)
    cp_, geom = domain.cpml, domain.geometry
    tab = model.stiffness_coefficients
    assert (p.nx, p.nz, p.dt, p.nt) == (9 + 6, 7 + 6, model.dt, 12)
    for k in ("c11", "c16", "c55", "rho"):
        exp = tab[k].values[geom[np.clip(np.arange(p.nx) - cp_, 0, 8)][:, np.clip(np.arange(p.nz) - cp_, 0, 6)]]
        np.testing.assert_array_equal(p.coef[k], exp.T, err_msg=k)
    # x: material 2 along x, Gamma block [[c11, c16/sqrt2], [c16/sqrt2, c66/2]] / rho;
    # z: max(c33 / rho) over materials 0 and 2 (c55 / 2 < c33 for both)
    a, b, c = 2e9, 3.5e9, 1.5e9 / np.sqrt(2)
    vx = np.sqrt(0.5 * (a + b + np.sqrt((a - b) ** 2 + 4 * c * c)) / 1000.0)
    vx = max(vx, np.sqrt(4e9 / 2000.0))
    vz = max(np.sqrt(4e9 / 2000.0), np.sqrt(3.5e9 / 1000.0), np.sqrt(2e9 / 1000.0))
    for ax, n, d, v in (("x", p.nx, domain.dx, vx), ("z", p.nz, domain.dz, vz)):
        exp = cpml.axis_profile(n, cp_, d, cpml.sigma_max_seismic(v, cp_, 2.0, 0.001),
                                cpml.alpha_max_seismic(1.0, 100.0), 5.0, 2.0, 1.0, model.dt, True)
        for f in ("kappa", "a", "b", "kappa_h", "a_h", "b_h"):
            np.testing.assert_allclose(getattr(p.prof[ax], f), getattr(exp, f), rtol=1e-14, atol=0,
                                       err_msg=(ax, f))
    sh = p.shots[0]
    np.testing.assert_array_equal(sh.force[0], model.sourcefunction_x)
    np.testing.assert_array_equal(sh.force[2], model.sourcefunction_z)
    assert sh.force[1] is None
    assert sh.src == dropin.source_index(int(2.0 / 0.5), 0, int(1.5 / 0.25), cp_, 2.0, "legacy")
    assert p.density_method == "harmonic"


def test_problem_from_objects_em():
    """EM: permittivity and conductivity painted; sigma_max from the EM
    formula with deltamin = min(dx, dz) and c (definitions.py:653-689)."""
    model, domain = _objects(False)
    p = dropin.problem_from_objects(model, domain)
    np.testing.assert_array_equal(p.coef["e33"][domain.cpml, domain.cpml:-domain.cpml],
                                  3.1 + domain.geometry[:, 0])
    sm = cpml.sigma_max_em(domain.cpml, 0.25, 2.0, 0.001)
    for ax, n, d in (("x", p.nx, 0.5), ("z", p.nz, 0.25)):
        exp = cpml.axis_profile(n, domain.cpml, d, sm, cpml.alpha_max_em(1.0, 100.0), 5.0, 2.0, 1.0,
                                model.dt, False)
        for f in ("kappa", "a", "b", "kappa_h", "a_h", "b_h"):
            np.testing.assert_allclose(getattr(p.prof[ax], f), getattr(exp, f), rtol=1e-14, atol=0,
                                       err_msg=(ax, f))


# -------------------------------------------------------------- file formats
def test_fortran_record_roundtrip_and_errors(tmp_path):
    a = np.arange(12.0).reshape(3, 4)
    f = str(tmp_path / "x.dat")
    fortranio.write_record(f, a)
    np.testing.assert_array_equal(fortranio.read_record(f, shape=(3, 4)), a)
    raw = open(f, "rb").read()
    assert raw[:4] == raw[-4:] == (96).to_bytes(4, "little")
    with pytest.raises(ValueError, match="expected shape"):
        fortranio.read_record(f, shape=(5, 4))
    open(f, "wb").write(raw[:-6])
    with pytest.raises(ValueError, match="truncated"):
        fortranio.read_record(f)


def _em25_bytes(nx, ny, nz, first, recs):
    """An EM25 block built byte by byte (definitions.py:154-190 layout),
    without the package's writer: recs = [(step, channel, lo, hi, raw)]."""
    out = b"SEIDART_EM25B1  " + np.array([1, nx, ny, nz, first, len({r[0] for r in recs}),
                                           len({r[1] for r in recs})], "<i4").tobytes()
    for step, ch, lo, hi, raw in recs:
        out += np.array([step], "<i4").tobytes() + ch.encode() + np.array([lo, hi], "<f4").tobytes()
        out += np.asarray(raw, "<i2").tobytes()
    return out


def test_em25_block_reader_on_handmade_file(tmp_path):
    """Reader against a hand-built block: the end points of the int16 range
    map to min and max, 0 to the midpoint, values land in Fortran (x, y, z)
    order, max == min gives a constant, gzip is read transparently."""
    import gzip
    nx, ny, nz = 3, 2, 2
    raw = np.array([-32767, 0, 32767, 100, -100, 5, 6, 7, 8, 9, 10, 11], dtype=np.int16)
    recs = [(4, "Ex", -2.0, 6.0, raw), (4, "Ez", 1.5, 1.5, np.zeros(12, np.int16)),
            (5, "Ex", 0.0, 1.0, raw[::-1].copy()), (5, "Ez", 0.0, 2.0, raw)]
    blob = _em25_bytes(nx, ny, nz, 4, recs)
    for name, data in (("b.blk", blob), ("b.blk.gz", gzip.compress(blob))):
        pth = tmp_path / name
        pth.write_bytes(data)
        r = fortranio.read_em25_block(str(pth))
        assert r["metadata"] == {"version": 1, "nx": 3, "ny": 2, "nz": 2, "first_step": 4,
                                 "step_count": 2, "component_count": 2}
        ex = r["fields"]["Ex"]
        assert ex.shape == (2, 3, 2, 2)
        assert ex[0, 0, 0, 0] == np.float32(-2.0)       # raw -32767 -> min
        assert ex[0, 2, 0, 0] == np.float32(6.0)        # raw +32767 -> max
        assert abs(ex[0, 1, 0, 0] - 2.0) < 1e-6         # raw 0 -> midpoint
        # element (x=0, y=1, z=0) is raw[0 + 3*(1 + 2*0)] = raw[3]
        assert abs(ex[0, 0, 1, 0] - ((100 + 32767) / 65534 * 8.0 - 2.0)) < 1e-5
        assert np.all(r["fields"]["Ez"][0] == np.float32(1.5))
        assert r["scales"]["Ex"] == [(4, -2.0, 6.0), (5, 0.0, 1.0)]
        only = fortranio.read_em25_block(str(pth), channels=["Ez"])
        assert list(only["fields"]) == ["Ez"]


def test_em25_block_writer_roundtrip(tmp_path):
    rng = np.random.default_rng(7)
    ex = rng.standard_normal((3, 5, 4, 6)).astype(np.float32) * 1e-3
    ey = np.full((3, 5, 4, 6), 0.25, np.float32)
    pth = tmp_path / fortranio.em25_block_name(11, (5, 2, 3))
    assert pth.name == "EM25.000011.5.2.3.blk"
    fortranio.write_em25_block(str(pth), {"Ex": ex, "Ey": ey}, first_step=11)
    r = fortranio.read_em25_block(str(pth))
    assert r["metadata"]["first_step"] == 11 and r["metadata"]["step_count"] == 3
    got = np.transpose(r["fields"]["Ex"], (0, 3, 2, 1))   # back to (t, z, y, x)
    for t in range(3):
        step = (ex[t].max() - ex[t].min()) / 65534
        assert np.abs(got[t] - ex[t]).max() <= 0.5 * step * 1.001 + 1e-12
    assert np.all(r["fields"]["Ey"] == np.float32(0.25))
    with pytest.raises(ValueError):
        fortranio.write_em25_block(str(pth), {"Exx": ex}, first_step=0)


def test_snapshot_names():
    assert fortranio.snapshot_name("Vx", 7, (128, 30)) == "Vx.000007.128.30.dat"
    assert fortranio.snapshot_name("Vx", 7, (128, 30), legacy_2d_dots=True) == "Vx.000007.128.30..dat"
    assert fortranio.snapshot_name("Ez", 123456, (1, 2, 3)) == "Ez.123456.1.2.3.dat"


def test_trace_csv_name_and_format(tmp_path):
    """Array.save's CSV (arraybuild.py:1031-1056): '{base}-{channel}-{x}-{y}-{z}.csv',
    nt x nrec, no header, no index, floats in their shortest round-trip form."""
    tr = np.array([[0.1, -2.5e-7], [1.0 / 3.0, 0.0], [1e300, -0.0]])
    fn = dropin.write_trace_csv(tr, str(tmp_path / "proj"), "Vz", np.array([12.5, 0.0, 3.0]))
    assert fn.endswith("proj-Vz-12.5-0.0-3.0.csv")
    text = open(fn).read().splitlines()
    assert text == ["0.1,-2.5e-07", "0.3333333333333333,0.0", "1e+300,-0.0"]
    np.testing.assert_array_equal(np.loadtxt(fn, delimiter=","), tr)


# ------------------------------------------------------------ kernel coefs
def test_kernel_coef_averages_match_loops():
    """The vectorised staggered-node averages equal the legacy loop bodies
    (seismicfdtd.f95:269-278, :307-310, :342) evaluated cell by cell."""
    from helpers import tiny_problem
    p = tiny_problem("elastic2", seed=3, nx=12, nz=9, npml=2, nt=4, gamma=1e-3)
    K = problem.kernel_coefs(p)
    c11, c55, rho = p.coef["c11"], p.coef["c55"], p.coef["rho"]
    for j in range(p.nz):
        for i in range(p.nx):
            n11 = (c11[j, i + 1] + 2 * c11[j, i] + c11[j - 1, i]) / 4 if (j >= 1 and i <= p.nx - 2) else 0
            s55 = (c55[j + 1, i] + 2 * c55[j, i] + c55[j, i - 1]) / 4 if (j <= p.nz - 2 and i >= 1) else 0
            dtz = p.dt / ((2 * rho[j, i] + rho[j, i + 1] + rho[j + 1, i]) / 4) \
                if (j <= p.nz - 2 and i <= p.nx - 2) else 0
            assert K[0][j, i] == n11 and K[7][j, i] == s55 and K[9][j, i] == dtz


def test_node_density_methods():
    a, b, c = np.array([2.0]), np.array([3.0]), np.array([5.0])
    assert problem.node_density([a, b, c], [2, 1, 1], "legacy")[0] == (2 * 2.0 + 3.0 + 5.0) / 4
    assert problem.node_density([a, b, c], [2, 1, 1], "geometric")[0] == pytest.approx(
        (2.0 * 2.0 * 3.0 * 5.0) ** 0.25, rel=1e-14)
    assert problem.node_density([a, b, c], [2, 1, 1], "harmonic")[0] == pytest.approx(
        4 / (2 / 2.0 + 1 / 3.0 + 1 / 5.0), rel=1e-14)
    assert problem.node_density([a, b, c], [2, 1, 1], "none")[0] == 2.0
    with pytest.raises(ValueError):
        problem.node_density([a, b, c], [2, 1, 1], "bogus")


# ------------------------------------------------------------- file loader
def _write_project(tmp, seismic=True):
    """A small project the way Model.build leaves it on disk."""
    nx0, nz0, c = 30, 20, 5
    nx, nz = nx0 + 2 * c, nz0 + 2 * c
    rng = np.random.default_rng(4)
    prj = {"Domain": {"dim": 2, "nx": nx0, "ny": 1, "nz": nz0, "dx": 0.5, "dy": 1, "dz": 0.5,
                      "cpml": c, "nmats": 1, "alpha_max_scalar": 1.0, "kappa_max": 5, "NP": 2,
                      "NPA": 2, "Rcoef": 0.001, "image_file": None},
           "Seismic": {"Source": {"dt": 5e-5, "time_steps": 40, "x": 5.0, "y": 0.0, "z": 3.0,
                                  "xind": 10, "yind": 0, "zind": 6, "source_frequency": 200.0}},
           "Electromagnetic": {"Source": {"dt": 1e-11, "time_steps": 40, "x": 5.0, "y": 0.0,
                                          "z": 3.0, "xind": 10, "yind": 0, "zind": 6}}}
    json.dump(prj, open(os.path.join(tmp, "p.json"), "w"))
    vals = {}
    for n in problem.STIFFNESS:
        v = rng.uniform(1e9, 2e9, size=(nz, nx)) if n != "rho" else rng.uniform(900, 1000, size=(nz, nx))
        vals[n] = v
        fortranio.write_record(os.path.join(tmp, f"{n}.dat"), v)
    for a in ("x", "z"):
        fortranio.write_record(os.path.join(tmp, f"gamma_{a}.dat"), np.full((nz, nx), 1e-4))
    px = cpml.axis_profile(nx, c, 0.5, 900.0, 600.0, 5.0, 2, 2, 5e-5, True)
    pz = cpml.axis_profile(nz, c, 0.5, 800.0, 600.0, 5.0, 2, 2, 5e-5, True)
    for q, h in (("kappa", ""), ("acoef", ""), ("bcoef", ""), ("kappa", "_half"), ("acoef", "_half"),
                 ("bcoef", "_half")):
        attr = {"kappa": "kappa", "acoef": "a", "bcoef": "b"}[q] + ("_h" if h else "")
        ax, az = getattr(px, attr), getattr(pz, attr)
        f = np.ones((nz, nx)) if q != "acoef" else np.zeros((nz, nx))
        f[:, :c], f[:, nx - c:] = ax[None, :c], ax[None, nx - c:]
        f[:c, c:nx - c], f[nz - c:, c:nx - c] = az[:c, None], az[nz - c:, None]
        fortranio.write_record(os.path.join(tmp, f"{q}{h}_cpml.dat"), f)
    src = sources.point_force(40, 5e-5, 200.0, xz_rotation=90.0)
    for a, s in zip("xyz", src):
        fortranio.write_record(os.path.join(tmp, f"seismicsource{a}.dat"), s)
    return vals, px, pz


def test_problem_from_files(tmp_path):
    vals, px, pz = _write_project(str(tmp_path))
    rec = np.array([[1.0, 0.0, 1.0], [7.5, 0.0, 2.0]])
    p = dropin.problem_from_files(str(tmp_path / "p.json"), seismic=True, receivers=rec)
    assert (p.kind, p.nx, p.nz, p.nt) == ("elastic2", 40, 30, 40)
    np.testing.assert_array_equal(p.coef["c33"], vals["c33"])
    np.testing.assert_array_equal(p.prof["x"].a_h, px.a_h)
    np.testing.assert_array_equal(p.prof["z"].b, pz.b)
    # legacy source placement: Fortran index xind + cpml -> 0-based xind + cpml - 1
    assert p.shots[0].src == (10 + 5 - 1, 0, 6 + 5 - 1)
    # receivers: int(xyz / d) + cpml (arraybuild.py:185-192)
    np.testing.assert_array_equal(p.shots[0].rec, [[7, 0, 7], [20, 0, 9]])
    assert p.gamma["x"][0, 0] == 1e-4
    assert p.shots[0].force[1] is None


def test_air_density_gradient_closed_forms():
    """airsurf (definitions.py:475-557): 0.15 on the solid cell next to air,
    then 0.575, 0.7875 (halving what is left) one and two cells further in,
    across either axis; diagonal neighbours and air cells stay 1; an air
    label of 0 or None (air listed first) gives no gradient."""
    g = np.ones((6, 8), int)
    g[:, :3] = 2                                   # air in z < 3 of a (nx, nz) image
    got = dropin.air_density_gradient(g, 2, 3)
    assert got.shape == (6, 8)
    np.testing.assert_allclose(got[0], [1, 1, 1, 0.15, 0.575, 0.7875, 1, 1])
    assert np.all(got == got[0])
    g = np.ones((7, 7), int)
    g[3, 3] = 2
    got = dropin.air_density_gradient(g, 2, 2)
    for (i, k), v in {(2, 3): 0.15, (4, 3): 0.15, (3, 2): 0.15, (3, 4): 0.15, (1, 3): 0.575,
                      (3, 5): 0.575, (2, 2): 1.0, (3, 3): 1.0, (0, 0): 1.0}.items():
        assert got[i, k] == v, (i, k, got[i, k])
    assert np.all(dropin.air_density_gradient(g, 0) == 1) and np.all(dropin.air_density_gradient(g, None) == 1)


def test_initial_condition_files(tmp_path):
    """initialconditionV{x,z}.dat (Model.run writes them as .T records,
    classes.py:882-898) become the run's initial fields; all-zero files are
    ignored."""
    _write_project(str(tmp_path))
    base = dropin.problem_from_files(str(tmp_path / "p.json"), seismic=True)
    assert not base.init
    vx0 = np.zeros((base.nz, base.nx))
    vx0[12, 20] = 1e-3
    fortranio.write_record(str(tmp_path / "initialconditionVx.dat"), vx0)
    fortranio.write_record(str(tmp_path / "initialconditionVz.dat"), np.zeros((base.nz, base.nx)))
    p = dropin.problem_from_files(str(tmp_path / "p.json"), seismic=True)
    assert list(p.init[0]) == ["vx"]
    np.testing.assert_array_equal(p.init[0]["vx"], vx0)


def test_density_gradient_file_scales_rho(tmp_path):
    """density_gradient.dat ((nx, nz) interior) multiplies rho; the C-PML
    rows and columns repeat the scaled edge."""
    _write_project(str(tmp_path))
    base = dropin.problem_from_files(str(tmp_path / "p.json"), seismic=True)
    nx0, nz0, c = base.nx - 10, base.nz - 10, 5
    grad = np.ones((nx0, nz0))
    grad[:, 0] = 0.15
    grad[3, 4] = 0.5
    fortranio.write_record(str(tmp_path / "density_gradient.dat"), grad)
    p = dropin.problem_from_files(str(tmp_path / "p.json"), seismic=True)
    inner = base.coef["rho"][c:-c, c:-c] * grad.T
    np.testing.assert_allclose(p.coef["rho"][c:-c, c:-c], inner, rtol=0, atol=0)
    np.testing.assert_array_equal(p.coef["rho"][:c, c:-c], np.broadcast_to(inner[0], (c, nx0)))
    np.testing.assert_array_equal(p.coef["c11"], base.coef["c11"])


def _co_files(tmp):
    src, rec = os.path.join(tmp, "src.csv"), os.path.join(tmp, "rec.csv")
    np.savetxt(src, [[2.0, 0, 1.5], [4.5, 0, 1.5], [7.25, 0, 2.0]], delimiter=",", header="x,y,z",
               comments="")
    np.savetxt(rec, [[3.0, 0, 0.5], [5.5, 0, 0.5], [8.0, 0, 1.0]], delimiter=",", header="x,y,z",
               comments="")
    return src, rec


def test_common_offset_problem(tmp_path):
    """CommonOffset (common_offset.py:14-283): one shot per source row, its
    receiver int(xyz / d + cpml), the project's source series and model."""
    _write_project(str(tmp_path))
    src, rec = _co_files(str(tmp_path))
    co = survey.CommonOffset(src, "Vz", str(tmp_path / "p.json"), rec)
    q = co.problem()
    assert len(q.shots) == 3 and co.component == 1 and co.time_steps == 40
    assert q.shots[0].src == (int(2.0 / 0.5) + 5 - 1, 0, int(1.5 / 0.5) + 5 - 1)
    np.testing.assert_array_equal(q.shots[2].rec, [[int(8.0 / 0.5 + 5), 0, int(1.0 / 0.5 + 5)]])
    np.testing.assert_array_equal(q.shots[1].force[2], co.base.shots[0].force[2])
    with pytest.raises(ValueError):
        survey.CommonOffset(src, "Vq", str(tmp_path / "p.json"), rec)
    with pytest.raises(ValueError):   # a 2-D run has no y component
        survey.CommonOffset(src, "Vy", str(tmp_path / "p.json"), rec)


def test_dropin_cli_reports_missing_project(capsys):
    assert dropin.main(["/nonexistent/p.json", "seismic=true"]) == 1
    assert "seidart_amd.run" in capsys.readouterr().err
    assert dropin.main(["x.json", "density_method=bogus"]) == 2
