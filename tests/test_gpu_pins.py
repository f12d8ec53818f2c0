"""The HIP path against closed forms and invariants, where the legacy
Fortran leaves the behaviour open (attenuation of vz in 2-D and of every
component in 2.5-D, the 2.5-D stress-rate source, the split C-PML variant).
The same properties pin the oracle (tests/test_oracle.py); here the device
results must satisfy them directly, not just match the oracle."""
import numpy as np
import pytest

from helpers import rel_l2, tiny_problem
from test_oracle import (REFL_BOUND, _homogeneous_decay, _homogeneous_iso, _interior,
                         reflection_case)

pytestmark = pytest.mark.gpu


def _sim(p, precision="fp64"):
    from seidart_amd.solver import Simulation
    return Simulation(p, precision=precision)


@pytest.mark.parametrize("kind", ["elastic2", "elastic25"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_attenuation_closed_form(kind, precision):
    nt = 8
    p, expect = _homogeneous_decay(kind, ny=41 if kind == "elastic25" else 1, nt=nt)
    sim = _sim(p, precision).run().check()
    m = 2 * nt + 4
    tol = 1e-14 if precision == "fp64" else 1e-6
    for c, e in expect.items():
        got = _interior(sim.field(0, c), m)
        assert got.size > 0
        np.testing.assert_allclose(got, e, rtol=tol, atol=0, err_msg=c)


def test_explosive_moment_source_mirror_symmetry_2d():
    p, c = _homogeneous_iso("elastic2")
    sim = _sim(p).run().check()
    vx, vz, sxx = (sim.field(0, n) for n in ("vx", "vz", "sxx"))
    assert np.abs(vx).max() > 0 and np.abs(vz).max() > 0
    i = np.arange(2, p.nx - 2)
    np.testing.assert_allclose(vx[:, 2 * c + 1 - i], -vx[:, i], rtol=0, atol=1e-12 * np.abs(vx).max())
    np.testing.assert_allclose(vz[2 * c - 1 - i, :], -vz[i, :], rtol=0, atol=1e-12 * np.abs(vz).max())
    np.testing.assert_allclose(sxx[:, 2 * c - i], sxx[:, i], rtol=0, atol=1e-12 * np.abs(sxx).max())
    one = _sim(p).run(1).check()
    r0 = p.shots[0].moment[0][0] * p.dt
    assert one.field(0, "sxx")[c, c] == r0 and one.field(0, "szz")[c, c] == r0


def test_explosive_moment_source_mirror_symmetry_25d():
    p, c = _homogeneous_iso("elastic25", n=41, ny=41, nt=24)
    cy = p.ny // 2
    sim = _sim(p).run().check()
    vx, vy, vz, sxx = (sim.field(0, n) for n in ("vx", "vy", "vz", "sxx"))
    for a in (vx, vy, vz, sxx):
        assert np.abs(a).max() > 0
    i = np.arange(2, p.nx - 2)
    j = np.arange(2, p.ny - 2)
    tol = lambda a: 1e-12 * np.abs(a).max()
    np.testing.assert_allclose(vx[:, :, 2 * c + 1 - i], -vx[:, :, i], rtol=0, atol=tol(vx))
    np.testing.assert_allclose(vy[:, 2 * cy - 1 - j, :], -vy[:, j, :], rtol=0, atol=tol(vy))
    np.testing.assert_allclose(vz[2 * c - 1 - i, :, :], -vz[i, :, :], rtol=0, atol=tol(vz))
    np.testing.assert_allclose(sxx[:, :, 2 * c - i], sxx[:, :, i], rtol=0, atol=tol(sxx))
    one = _sim(p).run(1).check()
    r0 = p.shots[0].moment[0][0] * p.dt
    for n in ("sxx", "syy", "szz"):
        assert one.field(0, n)[c, cy, c] == r0, n


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_split_cpml_without_memory_equals_legacy(precision):
    from seidart_amd.cpml import AxisProfile
    p = tiny_problem("elastic2", seed=71, nx=40, nz=36, npml=6, nt=60, gamma=1e-3, nrec=6)
    for a in ("x", "z"):
        pr = p.prof[a]
        p.prof[a] = AxisProfile(pr.kappa, np.zeros_like(pr.a), pr.b, pr.kappa_h, np.zeros_like(pr.a_h), pr.b_h)
    a = _sim(p, precision).run().check()
    p.cpml_split = True
    b = _sim(p, precision).run().check()
    np.testing.assert_array_equal(a.traces()[0], b.traces()[0])
    for n in ("vx", "vz", "sxx", "szz", "sxz"):
        np.testing.assert_array_equal(a.field(0, n), b.field(0, n))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_split_cpml_reflection_below_bound(precision):
    small, big = reflection_case()
    ref = _sim(big, "fp64").run(graph_block=50).check().traces()[0]
    small.cpml_split = True
    r = rel_l2(_sim(small, precision).run(graph_block=50).check().traces()[0], ref)
    print(f"split C-PML reflection ({precision}): {r:.3e} = {20 * np.log10(r):.1f} dB")
    assert r < REFL_BOUND


def test_split_cpml_energy_bounded_on_fwi_model():
    """The FWI model (VTI ice under snow and firn, 1024 x 512 + C-PML 20), one
    shot, FP32, all 6000 steps: with the split C-PML the kinetic energy
    (sum of v^2) falls after the source has passed and never returns to its
    peak; the legacy sharing of the memory arrays on the same model blows up
    (the device's 1e25 check trips)."""
    from seidart_amd._hip import UnstableError
    from seidart_amd import configs
    _, q = configs.fwi_vti()
    p = q.with_shots([q.shots[8]])
    assert p.cpml_split
    sim = _sim(p, "fp32")
    e = []
    for _ in range(12):
        sim.run(500, graph_block=50)
        e.append(sum(float(np.sum(sim.field(0, n) ** 2)) for n in ("vx", "vz")))
    sim.check()
    print("split energy per 500 steps:", " ".join(f"{v:.2e}" for v in e))
    peak = int(np.argmax(e))
    assert peak <= 3 and e[-1] < 1e-2 * e[peak]
    assert max(e[peak + 2:]) < e[peak]
    sim.close()
    p.cpml_split = False
    leg = _sim(p, "fp32").run(graph_block=50)
    with pytest.raises(UnstableError):
        leg.check()
    print("legacy C-PML: unstable from step", leg.unstable[0])
