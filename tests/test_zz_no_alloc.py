"""The allocation-free entry point (sdt_run_noalloc) against the plan path
(sdt_plan_create / sdt_plan_run / sdt_plan_check) it varies: the same
kernels, so bit-identical traces and states, through the C ABI, on two calls
in a row, the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest

from helpers import tiny_problem

pytestmark = pytest.mark.gpu


def _near_source_receivers(probs):
    """Receivers a few cells around each shot's source, so that short runs
    give non-zero traces."""
    for p in probs:
        for sh in p.shots:
            i, j, k = sh.src
            pts = [(i + a, j + (b if p.is3d else 0), k + c)
                   for a, b, c in ((0, 0, 0), (2, 1, 1), (-3, -1, 2), (1, 0, -2), (-2, 1, -1))]
            sh.rec = np.array([[min(max(x, 0), p.nx - 1), min(max(y, 0), p.ny - 1),
                                min(max(z, 0), p.nz - 1)] for x, y, z in pts], dtype=np.int32)
    return probs


def _pair(probs, precision):
    from seidart_amd.solver import Simulation
    a = Simulation(probs, precision=precision)
    b = Simulation(probs, precision=precision)
    return a, b


def _same(a, b):
    for x, y in zip(a.traces(), b.traces()):
        assert np.abs(x).max() > 0
        np.testing.assert_array_equal(x, y)
    for (_, _, ba), (_, _, bb) in zip(a.shots, b.shots):
        assert a.kind == b.kind
        import torch
        torch.cuda.synchronize()
        assert torch.equal(ba["state"], bb["state"])


@pytest.mark.parametrize("kind", ["elastic2", "em2", "elastic25", "em25"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_equals_plan_path_twice(kind, precision):
    is3d = kind.endswith("25")
    big = [tiny_problem(kind, seed=300 + m, npml=5, nt=40, nshots=2, gamma=1e-3 if "elastic" in kind else 0.0,
                        **(dict(nx=70, ny=12, nz=30) if is3d else dict(nx=180, nz=90)))
           for m in range(2)]
    small = [tiny_problem(kind, seed=310, npml=4, nt=24, nshots=1,
                          **(dict(nx=30, ny=9, nz=20) if is3d else dict(nx=50, nz=40)))]
    for probs in (big, small):
        _near_source_receivers(probs)
        a, b = _pair(probs, precision)
        a.run(graph_block=0).check()
        b.run_noalloc(13)
        b.run_noalloc()
        _same(a, b)
        a.close()
        b.close()


def test_noalloc_two_steps_per_launch_and_split():
    """A plan large enough for two steps per launch (its default), the split
    C-PML variant and an odd run: still equal to the plan path."""
    probs = [tiny_problem("elastic2", seed=320, npml=6, nt=31, nshots=4, gamma=1e-3, nx=1200, nz=900)]
    for p in probs:
        p.cpml_split = True
    _near_source_receivers(probs)
    a, b = _pair(probs, "fp32")
    assert a.steps_per_launch == 2
    a.run(graph_block=0).check()
    b.run_noalloc()
    _same(a, b)


def test_noalloc_reports_instability_and_rejects_oversize():
    from seidart_amd import _hip
    p = tiny_problem("elastic2", seed=53, nt=200)
    p.dt *= 8
    a, b = _pair([p], "fp64")
    a.run()
    with pytest.raises(_hip.UnstableError):
        a.check()
    with pytest.raises(_hip.UnstableError):
        b.run_noalloc()
    assert b.unstable == a.unstable and b.unstable[0] >= 0
    assert _hip.lib().sdt_noalloc_scratch_bytes() == 8 << 20
    # more receivers than the scratch holds the receivers-by-row index for
    many = tiny_problem("elastic2", seed=54, nt=4, nx=64, nz=64, nrec=1_100_000)
    c, _ = _pair([many], "fp32")
    with pytest.raises(_hip.SeidartHipError, match="scratch"):
        c.run_noalloc()
