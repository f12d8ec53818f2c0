"""Two time steps per launch (ELASTIC2 and EM2): the state of every shot
(fields and C-PML memory) and every receiver sample are bit-identical to one
step per launch, for both precisions, both C-PML variants, every tile shape,
graph blocks, odd run lengths and runs that start at odd steps; and the FP64
result matches the oracle."""
import numpy as np
import pytest
import torch

from helpers import rel_l2, tiny_problem

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _sim(p, precision, spl, **kw):
    from seidart_amd.solver import Simulation
    s = Simulation(p, precision=precision, **kw)
    s.steps_per_launch = spl
    assert s.steps_per_launch == spl
    return s


def _snap(sim):
    torch.cuda.synchronize()
    return ([b["state"].cpu().numpy().copy() for _, _, b in sim.shots], sim.traces())


def _assert_same(a, b):
    for x, y in zip(a[0], b[0]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a[1], b[1]):
        np.testing.assert_array_equal(x, y)


def _problems(nt, split=False, full21=True, seed=300):
    d = [dict(nx=150, nz=75, npml=6), dict(nx=61, nz=37, npml=5), dict(nx=40, nz=130, npml=7)]
    probs = [tiny_problem("elastic2", seed=seed + m, nt=nt, gamma=1e-3, nshots=2, nrec=9,
                          full21=full21, **kw) for m, kw in enumerate(d)]
    for p in probs:
        p.cpml_split = split
        # receivers on the outer ring and on one row as well as scattered
        p.shots[0].rec = np.concatenate([p.shots[0].rec, np.array(
            [[0, 0, 0], [p.nx - 1, 0, p.nz - 1], [5, 0, p.nz - 1], [p.nx - 2, 0, 1]], dtype=np.int32)])
        row = np.stack([np.arange(1, p.nx - 1, 3), np.zeros((p.nx - 2 + 2) // 3, int),
                        np.full((p.nx - 2 + 2) // 3, p.nz // 3)], axis=1).astype(np.int32)
        p.shots[1].rec = np.concatenate([p.shots[1].rec, row])
    return probs


@pytest.mark.parametrize("precision,split", [("fp32", False), ("fp64", False),
                                             ("fp32", True), ("fp64", True)])
def test_two_step_equals_one_step(precision, split):
    probs = _problems(61, split=split)
    ref = _snap(_sim(probs, precision, 1).run(graph_block=10).check())
    for gb in (10, 0):
        got = _snap(_sim(probs, precision, 2).run(graph_block=gb).check())
        _assert_same(ref, got)


def test_two_step_isotropic_without_attenuation():
    """The body without c15/c35 terms and without gamma (the headline's)."""
    probs = _problems(40, full21=False, seed=320)
    for p in probs:
        p.gamma = {}
    ref = _snap(_sim(probs, "fp32", 1).run().check())
    got = _snap(_sim(probs, "fp32", 2).run(graph_block=8).check())
    _assert_same(ref, got)


@pytest.mark.parametrize("precision,split", [("fp32", False), ("fp64", False),
                                             ("fp32", True), ("fp64", True)])
def test_one_row_tiles_equal_taller_tiles(precision, split):
    """Tiles of one row run the one-step kernel instance that issues all of a
    tile's loads before its first row step; the state and traces equal those
    of 8-row tiles bit for bit, for blocks of 1, 2 and 4 wavefronts, three
    models, receivers on the outer ring, a source in a C-PML corner and a
    stress-rate (moment) source."""
    probs = _problems(41, split=split, seed=340)
    probs[1].shots[1].src = (probs[1].nx - 3, 0, 2)
    p = probs[0]
    t = np.arange(p.nt) * p.dt
    w = np.exp(-((t - 15 * p.dt) / (5 * p.dt)) ** 2)
    p.shots[0].moment = (1e3 * w, None, 2e3 * w, None, -5e2 * w, None)
    ref = _snap(_sim(probs, precision, 1).set_tiles(4, 8).run(graph_block=8).check())
    for wb in (1, 2, 4):
        got = _snap(_sim(probs, precision, 1).set_tiles(wb, 1).run(graph_block=5).check())
        _assert_same(ref, got)


def test_two_step_tile_shapes():
    probs = _problems(30, seed=330)
    ref = _snap(_sim(probs, "fp32", 1).run().check())
    for wb, rows in ((1, 2), (2, 3), (4, 16), (1, 64), (4, 256), (2, 32), (1, 1), (4, 1)):
        got = _snap(_sim(probs, "fp32", 2).set_tiles(wb, rows).run(graph_block=6).check())
        _assert_same(ref, got)


def test_two_step_segments_and_mode_switches():
    """Runs of odd length that start at odd steps, odd graph blocks (one step
    per launch), and switching the mode between runs."""
    probs = _problems(57, seed=340)
    ref = _snap(_sim(probs, "fp32", 1).run().check())
    s = _sim(probs, "fp32", 2)
    for n, gb in ((1, 0), (3, 0), (6, 4), (5, 5), (9, 2)):
        s.run(n, graph_block=gb)
    s.steps_per_launch = 1
    s.run(7)
    s.steps_per_launch = 2
    s.run(graph_block=4)
    _assert_same(ref, _snap(s.check()))


def test_two_step_moment_source_and_oracle(oracle_lib):
    probs = _problems(45, seed=350)
    p = probs[0]
    nt = p.nt
    t = np.arange(nt) * p.dt
    w = np.exp(-((t - 20 * p.dt) / (6 * p.dt)) ** 2)
    sh = p.shots[0]
    sh.moment = (1e3 * w, None, 2e3 * w, None, -5e2 * w, None)
    sh.src = (p.nx - 8, 0, 4)   # source in the C-PML corner region
    ref = _snap(_sim(probs, "fp64", 1).run().check())
    sim = _sim(probs, "fp64", 2).run(graph_block=6).check()
    _assert_same(ref, _snap(sim))
    tr = sim.traces()
    for s in range(len(p.shots)):
        o = oracle_lib.OracleRun(p, shot=s).run()
        assert rel_l2(tr[s], o.trace) < FP64_TOL
        for n in o.names:
            exp = o.field(n)
            if np.abs(exp).max() > 0:
                assert rel_l2(sim.field(s, n), exp) < FP64_TOL, n


def test_two_step_reports_the_same_unstable_step():
    from seidart_amd._hip import UnstableError
    p = tiny_problem("elastic2", seed=360, nt=120, nshots=2)
    p.dt *= 8
    steps = []
    for spl in (1, 2):
        s = _sim(p, "fp32", spl).run(graph_block=10)
        with pytest.raises(UnstableError):
            s.check()
        steps.append(list(s.unstable))
    assert steps[0] == steps[1] and min(steps[0]) >= 0


def _em_problems(nt, seed=400):
    d = [dict(nx=150, nz=75, npml=6), dict(nx=61, nz=37, npml=5), dict(nx=40, nz=130, npml=7)]
    probs = [tiny_problem("em2", seed=seed + m, nt=nt, nshots=2, nrec=9, **kw)
             for m, kw in enumerate(d)]
    for p in probs:
        p.shots[0].rec = np.concatenate([p.shots[0].rec, np.array(
            [[0, 0, 0], [p.nx - 1, 0, p.nz - 1], [p.nx - 2, 0, p.nz - 2], [1, 0, 1]], dtype=np.int32)])
    return probs


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_em2_two_step_equals_one_step(precision):
    probs = _em_problems(61)
    probs[1].shots[1].src = (probs[1].nx - 3, 0, 2)   # source in the C-PML corner
    ref = _snap(_sim(probs, precision, 1).run(graph_block=10).check())
    for gb, tiles in ((10, None), (0, (1, 2)), (6, (2, 16)), (4, (4, 64)), (6, (2, 1)), (4, (4, 4)),
                      (0, (2, 4))):
        s = _sim(probs, precision, 2)
        if tiles:
            s.set_tiles(*tiles)
        _assert_same(ref, _snap(s.run(graph_block=gb).check()))


def test_em2_two_step_segments_and_oracle(oracle_lib):
    probs = _em_problems(47, seed=410)
    ref = _snap(_sim(probs, "fp64", 1).run().check())
    s = _sim(probs, "fp64", 2)
    for n, gb in ((1, 0), (3, 0), (6, 4), (5, 5), (9, 2)):
        s.run(n, graph_block=gb)
    s.steps_per_launch = 1
    s.run(7)
    s.steps_per_launch = 2
    s.run(graph_block=4).check()
    _assert_same(ref, _snap(s))
    p = probs[0]
    tr = s.traces()
    for k in range(len(p.shots)):
        o = oracle_lib.OracleRun(p, shot=k).run()
        assert rel_l2(tr[k], o.trace) < FP64_TOL
        for n in o.names:
            exp = o.field(n)
            if np.abs(exp).max() > 0:
                assert rel_l2(s.field(k, n), exp) < FP64_TOL, n
