"""The project-file route (seidart_amd.project): loadproject, Model.build's
dt, sources, C-PML and painted files, from the JSON and geometry alone.

Pinned by closed forms (dt of media with known speeds, the source series and
moment tensors, the label image) and by equality with the file route: the
files this route writes, read back by the reader of the reference's files
(seidart_amd.run.problem_from_files), give the same Problem bit for bit."""
import json
import os

import numpy as np
import pytest

from seidart_amd import fortranio, project, sources
from seidart_amd.cpml import CLIGHT
from seidart_amd.problem import STIFFNESS

ICE = (917.0, 3840.0, 1920.0)
SNOW = (400.0, 1800.0, 900.0)


def _iso_row(rho, vp, vs, i):
    c11 = rho * vp * vp
    c55 = rho * vs * vs
    row = {"id": i}
    row.update({n: 0.0 for n in STIFFNESS})
    for n in ("c11", "c22", "c33"):
        row[n] = c11
    for n in ("c12", "c13", "c23"):
        row[n] = c11 - 2 * c55
    for n in ("c44", "c55", "c66"):
        row[n] = c55
    row["rho"] = rho
    return row


def write_project(tmp, nx=60, nz=40, cpml=8, dim=2, ny=1, time_steps=120, source_type="ac",
                  wavelet="gaus1", dt=1e-12, air=True, rotations=(0.0, 90.0, 0.0)):
    """A three-material project as prjbuild and save_to_json leave it: the
    JSON with its material tables, and geometry.dat (the label image as
    image2int writes it: one int64 record of the (nz, nx) labels).  Labels:
    0 snow, 1 ice, 2 air (on top when ``air``).  ``rotations``: the source's
    x-y, x-z and y-z rotations (strike, dip, rake of a double couple)."""
    geom = np.zeros((nz, nx), dtype=np.int64)
    z = np.arange(nz)[:, None]
    x = np.arange(nx)[None, :]
    geom[z > 15 + 3 * np.sin(x / 7.0)] = 1
    if air:
        geom[:4, :] = 2
    fortranio.write_record(os.path.join(tmp, "geometry.dat"), geom, dtype=np.int64)
    rows = [_iso_row(*SNOW, 0), _iso_row(*ICE, 1), _iso_row(1.2, 343.0, 0.0, 2)]
    att = [{"id": i, "gamma_x": g, "gamma_y": g, "gamma_z": 2 * g, "gamma_yz": 0.0, "gamma_xz": 0.0,
            "gamma_xy": 0.0, "reference_frequency": 1.0} for i, g in enumerate((2e-6, 1e-6, 0.0))]
    perm = [{"id": i, "e11": e, "e12": 0.0, "e13": 0.1 * i, "e22": e, "e23": 0.0, "e33": e + 0.05}
            for i, e in enumerate((1.8, 3.15, 1.0))]
    cond = [{"id": i, "s11": s, "s12": 0.0, "s13": 0.0, "s22": s, "s23": 0.0, "s33": s}
            for i, s in enumerate((1e-5, 5e-5, 0.0))]
    src = {"dt": dt, "time_steps": time_steps, "x": 12.5, "y": 0.0, "z": 4.0, "xind": 0, "yind": 0,
           "zind": 0, "source_frequency": 200.0, "x-z_rotation": rotations[1],
           "x-y_rotation": rotations[0], "y-z_rotation": rotations[2], "amplitude": 1.0,
           "source_wavelet": wavelet, "source_type": source_type}
    esrc = dict(src, source_frequency=4e8, source_type="ps")
    prj = {"Domain": {"dim": dim, "nx": nx, "ny": ny, "nz": nz, "dx": 0.5, "dy": 0.5, "dz": 0.5,
                      "cpml": cpml, "nmats": 3, "alpha_max_scalar": 1.0, "kappa_max": 5, "NP": 2,
                      "NPA": 2, "Rcoef": 0.001, "image_file": "none.png"},
           "Materials": [{"id": i, "name": n, "rgb": f"{i}/{i}/{i}", "temperature": -10,
                          "density": 0, "porosity": 0, "water_content": 0, "is_anisotropic": False,
                          "euler_angles": None} for i, n in enumerate(("snow", "ice1h", "air"))],
           "Seismic": {"Source": src, "Attenuation": att, "Stiffness_Coefficients": rows},
           "Electromagnetic": {"Source": esrc, "Permittivity_Coefficients": perm,
                               "Conductivity_Coefficients": cond}}
    jf = os.path.join(tmp, "prj.json")
    with open(jf, "w") as fh:
        json.dump(prj, fh)
    return jf, geom.T


def test_load_project_reads_tables_and_geometry(tmp_path):
    jf, geom = write_project(str(tmp_path))
    d, mat, sm, em = project.load_project(jf)
    assert (d.nx, d.nz, d.cpml, d.nmats, d.dim) == (60, 40, 8, 3, 2.0)
    np.testing.assert_array_equal(d.geometry, geom)
    assert list(mat.material) == ["snow", "ice1h", "air"]
    assert sm.stiffness_coefficients["c11"].values[1] == 917.0 * 3840.0 ** 2
    assert sm.attenuation_coefficients["gamma_z"].values[0] == 4e-6
    assert em.permittivity_coefficients["e13"].values[2] == 0.2
    assert sm.is_seismic and not em.is_seismic and sm.f0 == 200.0


def test_geometry_mask_marks_fluid_cells(tmp_path):
    """writemask's file (classes.py:549-606): 1 on air, water and oil cells,
    0 elsewhere, edges repeated into the C-PML, the transpose of the padded
    (nx, nz) array in one int32 record."""
    jf, geom = write_project(str(tmp_path), nx=30, nz=20, cpml=5)
    d, mat, sm, em = project.load_project(jf)
    m = project.geometry_mask(d, mat)
    assert m.shape == (40, 30)
    np.testing.assert_array_equal(m[5:35, 5:25], (geom == 2).astype(int))  # label 2 is air
    np.testing.assert_array_equal(m[:5, 5:25], np.repeat(m[5:6, 5:25], 5, axis=0))
    np.testing.assert_array_equal(m[5:35, :5], np.repeat(m[5:35, 5:6], 5, axis=1))
    assert m[0, 0] == m[5, 5] and m[-1, -1] == m[34, 24]
    assert set(np.unique(m)) == {0, 1}
    project.build(sm, d, mat, write=False)
    files = project.write_build_files(sm, d, mat, str(tmp_path / "build"))
    path = str(tmp_path / "build" / "geometry_mask.dat")
    assert path in files
    got = fortranio.read_record(path, dtype=np.int32).reshape(30, 40)
    np.testing.assert_array_equal(got, m.T)


def test_dt_closed_forms(tmp_path):
    """dt = CFL / (vmax sqrt(sum 1/d^2)) (classes.py:469-485): isotropic media
    have vmax = the largest Vp along the axes (ice, 3840 m/s); EM vmax =
    c / sqrt(min e_ii) (air, e = 1); 2.5-D caps the CFL at 1/sqrt(3)."""
    jf, _ = write_project(str(tmp_path))
    d, _, sm, em = project.load_project(jf)
    assert project.model_dt(sm, d) == pytest.approx((1 / np.sqrt(2)) / (3840.0 * np.sqrt(2 / 0.25)), rel=1e-13)
    assert project.model_dt(em, d) == pytest.approx((1 / np.sqrt(2)) / (CLIGHT * np.sqrt(2 / 0.25)), rel=1e-13)
    d.dim = 2.5
    project.build(sm, d)
    assert sm.CFL == pytest.approx(1 / np.sqrt(3))
    assert sm.dt == pytest.approx((1 / np.sqrt(3)) / (3840.0 * np.sqrt(3 / 0.25)), rel=1e-13)
    # an anisotropic medium: vmax along the axes is sqrt(c11 / rho) or sqrt(c33 / rho)
    tab = sm.stiffness_coefficients
    tab._a[1, tab.columns.index("c33")] *= 1.21          # vp_z = 1.1 * 3840
    d.dim = 2
    sm.CFL = 1 / np.sqrt(2)
    assert project.model_dt(sm, d) == pytest.approx((1 / np.sqrt(2)) / (1.1 * 3840.0 * np.sqrt(8.0)), rel=1e-12)


def test_source_series_closed_forms():
    nt, dt, f = 400, 2e-5, 500.0
    t = np.arange(nt) * dt
    # Ricker: peak at t = 1/f, zero crossings at 1/f +- 1/(sqrt(2) pi f)
    r = sources.wavelet(t, f, "gaus1")
    assert np.argmax(r) == round(1 / f / dt) and r.max() == 1.0
    # linear chirp: cos(2 pi (f0 t + (f1 - f0) t^2 / (2 t1)) - pi/2), f0 = 10 f, f1 = f, t1 = 1/f
    c = sources.wavelet(t, f, "chirp")
    ph = 2 * np.pi * (10 * f * t + (f - 10 * f) * t * t / (2 / f)) - np.pi / 2
    ref = np.cos(ph)
    np.testing.assert_allclose(c, ref / np.abs(ref).max(), rtol=0, atol=1e-12)
    cl = sources.wavelet(t, f, "chirplet")
    assert np.abs(cl).max() == 1.0 and abs(cl[-1]) < 1e-6
    # multimode: 4 n - 1 octave-spaced wavelets, normalised
    m = sources.multimodesrc(t, f, "gaus1", num_octaves=2)
    freqs = f * 2 ** np.linspace(-2, 0, 7)
    s = sum(sources.wavelet(t, q, "gaus1") for q in freqs)
    np.testing.assert_allclose(m, s / np.abs(s).max(), rtol=0, atol=1e-15)
    # broadband: causal minimum-phase band-pass; passes mid-band, blocks DC
    b = sources.broadbandsrc(np.arange(4096) * dt, 3000.0, 9000.0)
    H = np.abs(np.fft.rfft(b))
    fr = np.fft.rfftfreq(4096, dt)
    mid = H[np.argmin(abs(fr - 6000.0))]
    assert mid > 100 * H[0] and mid > 100 * H[np.argmin(abs(fr - 20000.0))]
    assert np.argmax(np.abs(b)) < 30 and b[201:].max() == 0.0
    with pytest.raises(ValueError):
        sources.wavelet(t, f, "boxcar")


def test_moment_tensors_and_source_types():
    M = sources.double_couple_tensor(30.0, 60.0, 45.0, M0=2.0)
    np.testing.assert_allclose(M, M.T, atol=0)
    assert abs(np.trace(M)) < 1e-15
    np.testing.assert_allclose(np.linalg.eigvalsh(M), [-2.0, 0.0, 2.0], atol=1e-14)
    nt, dt = 50, 1e-4
    s = sources.wavelet(np.arange(nt) * dt, 100.0, "gaus1")
    f, m = sources.pointsource(nt, dt, 100.0, 3.0, "gaus1", "tnt")
    assert all(not np.any(v) for v in f)
    for q in range(3):
        np.testing.assert_array_equal(m[q], 3.0 * s)
    f, m = sources.pointsource(nt, dt, 100.0, 1.0, "gaus1", "omni_ps", 20.0, 35.0, 10.0)
    np.testing.assert_allclose(m[0] + m[1] + m[2], 3.0 * s, rtol=0, atol=1e-14)
    f, m = sources.pointsource(nt, dt, 100.0, 1.0, "gaus1", "dc", 30.0, 60.0, 45.0)
    np.testing.assert_allclose(m[4], s * M[0, 2] / 2.0, rtol=0, atol=1e-15)
    f, m = sources.pointsource(nt, dt, 100.0, 1.0, "gaus1", "ac", 0.0, 90.0, 0.0)
    np.testing.assert_allclose(f[2], s, atol=1e-15)
    assert np.abs(f[0]).max() < 1e-15
    with pytest.raises(ValueError):
        sources.pointsource(nt, dt, 100.0, 1.0, "gaus1", "mystery")


def test_image2int_orders_colours(tmp_path):
    """Labels follow ascending 65536 R + 255 G + B (prjbuild.py:303-339)."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    img = np.zeros((6, 9, 3))
    img[:, :3] = (1.0, 0.0, 0.0)        # key 65536
    img[:, 3:6] = (0.0, 1.0, 0.0)       # key 255
    img[:, 6:] = (0.0, 0.0, 1.0)        # key 1
    fn = str(tmp_path / "g.png")
    plt.imsave(fn, img)
    lab, rgb = project.image2int(fn)
    assert lab.shape == (6, 9)
    assert (lab[:, :3] == 2).all() and (lab[:, 3:6] == 1).all() and (lab[:, 6:] == 0).all()
    np.testing.assert_array_equal(rgb, [[0, 0, 255], [0, 255, 0], [255, 0, 0]])


@pytest.mark.parametrize("seismic,source_type", [(True, "ac"), (True, "dc"), (False, "ac")])
def test_project_route_equals_file_route(tmp_path, seismic, source_type):
    """build(write=True) writes the files Model.build writes (coefficients,
    attenuation, density gradient, C-PML fields, sources) and the dt into the
    JSON; the file reader gives the same Problem as the in-memory route, bit
    for bit (coefficients, gamma, C-PML profiles, dt, source cell and series)."""
    from seidart_amd import run as dropin
    rot = (90.0, 45.0, 90.0) if source_type == "dc" else (0.0, 90.0, 0.0)
    jf, _ = write_project(str(tmp_path), source_type=source_type, rotations=rot)
    d, mat, sm, em = project.load_project(jf)
    model = sm if seismic else em
    project.build(model, d, mat, write=True)
    rec = np.array([[5.0, 0.0, 3.0], [20.0, 0.0, 3.0], [27.5, 0.0, 10.0]])
    a = project.problem_from_project(jf, seismic, receivers=rec)
    b = dropin.problem_from_files(jf, seismic, receivers=rec)
    assert a.kind == b.kind and a.dt == b.dt == model.dt and a.nt == b.nt
    assert set(a.coef) == set(b.coef)
    for k in a.coef:
        np.testing.assert_array_equal(a.coef[k], b.coef[k], err_msg=k)
    assert set(a.gamma) == set(b.gamma)
    for k in a.gamma:
        np.testing.assert_array_equal(a.gamma[k], b.gamma[k], err_msg=k)
    for ax in ("x", "z"):
        for q in ("kappa", "a", "b", "kappa_h", "a_h", "b_h"):
            np.testing.assert_array_equal(getattr(a.prof[ax], q), getattr(b.prof[ax], q), err_msg=(ax, q))
    sa, sb = a.shots[0], b.shots[0]
    assert sa.src == sb.src
    np.testing.assert_array_equal(sa.rec, sb.rec)
    for u, v in zip(sa.force, sb.force):
        assert (u is None) == (v is None)
        if u is not None:
            np.testing.assert_array_equal(u, v)
    for u, v in zip(sa.moment, sb.moment):
        assert (u is None) == (v is None)
        if u is not None:
            np.testing.assert_array_equal(u, v)
    if seismic:
        # the density gradient near air reached rho (air is label 2)
        assert a.coef["rho"][d.cpml + 4, d.cpml + 10] == 0.15 * 400.0
    if source_type == "dc":
        assert sa.moment[0] is not None and sa.moment[4] is not None
