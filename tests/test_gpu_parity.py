"""HIP path (through the C ABI) against the oracle and the legacy goldens.

Bars: FP64 device traces and fields within 1e-10 relative L2 of the FP64
oracle (itself bit-identical to the compiled legacy Fortran); FP32 device
traces within the FP32 tolerance below."""
import os

import numpy as np
import pytest
import torch

from golden.make_golden import CASES, case_problem
from helpers import rel_l2, tiny_problem

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-5   # relative L2 of FP32 traces against FP64, the bar of every configuration
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

SMALL = {
    "elastic2": dict(seed=31, nx=48, nz=40, npml=6, nt=90, gamma=1e-3),
    "em2": dict(seed=32, nx=48, nz=40, npml=6, nt=90),
    "elastic25": dict(seed=33, nx=28, ny=14, nz=22, npml=4, nt=45, gamma=1e-3),
    "em25": dict(seed=34, nx=28, ny=14, nz=22, npml=4, nt=45),
}


def _sim(p, precision="fp64", **kw):
    from seidart_amd.solver import Simulation
    return Simulation(p, precision=precision, **kw)


def _compare(sim, p, oracle_lib, tol, shot_offset=0):
    tr = sim.traces()
    for s in range(len(p.shots)):
        o = oracle_lib.OracleRun(p, shot=s).run()
        assert rel_l2(tr[shot_offset + s], o.trace) < tol, ("trace", s, rel_l2(tr[shot_offset + s], o.trace))
        for n in o.names:
            got = sim.field(shot_offset + s, n)
            exp = o.field(n)
            if np.abs(exp).max() == 0:
                continue
            assert rel_l2(got, exp) < tol, (n, s, rel_l2(got, exp))


@pytest.mark.parametrize("kind", list(SMALL))
def test_fp64_matches_oracle(oracle_lib, kind):
    p = tiny_problem(kind, **SMALL[kind])
    sim = _sim(p, "fp64").run().check()
    _compare(sim, p, oracle_lib, FP64_TOL)


@pytest.mark.parametrize("kind", list(SMALL))
def test_fp32_close_to_oracle(oracle_lib, kind):
    p = tiny_problem(kind, **SMALL[kind])
    sim = _sim(p, "fp32").run().check()
    tr = sim.traces()[0]
    o = oracle_lib.OracleRun(p).run()
    err = rel_l2(tr, o.trace)
    print(f"{kind} fp32 trace rel L2 vs fp64 oracle: {err:.3e}")
    assert err < FP32_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_matches_legacy_golden(name):
    """Straight against the legacy Fortran's own traces (tests/golden)."""
    g = np.load(os.path.join(GOLDEN, f"legacy_{name}.npz"))
    p = case_problem(name)
    sim = _sim(p, "fp64").run().check()
    tr = sim.traces()[0]
    chans = {"elastic2": ["Vx", "Vz"], "em2": ["Ex", "Ez"], "elastic25": ["Vx", "Vy", "Vz"],
             "em25": ["Ex", "Ey", "Ez"]}[p.kind]
    for c, ch in enumerate(chans):
        assert rel_l2(tr[:, c, :], g[f"trace_{ch}"]) < FP64_TOL, ch


@pytest.mark.parametrize("kind", ["elastic2", "em2", "elastic25", "em25"])
def test_batched_uneven_models(oracle_lib, kind):
    """Several models of different size, several shots each, one plan: every
    shot equals its own single-shot oracle run."""
    is3d = kind.endswith("25")
    dims = [dict(nx=40, nz=30), dict(nx=57, nz=45), dict(nx=33, nz=50)] if not is3d else \
        [dict(nx=24, ny=10, nz=18), dict(nx=31, ny=12, nz=22)]
    probs = [tiny_problem(kind, seed=40 + m, npml=4, nt=40, nshots=3, **d)
             for m, d in enumerate(dims)]
    sim = _sim(probs, "fp64").run().check()
    off = 0
    for p in probs:
        _compare(sim, p, oracle_lib, FP64_TOL, shot_offset=off)
        off += len(p.shots)


def test_el25_fp32_tiles_chunks_segments(oracle_lib):
    """FP32 2.5-D elastic against the FP64 oracle: traces near the source and
    every final field, on a grid several tiles wide in x, several chunks long
    in y and with an odd number of planes, run in odd-length segments."""
    p = tiny_problem("elastic25", seed=59, nx=140, ny=70, nz=27, npml=5, nt=31, nrec=9, gamma=1e-3)
    si, sj, sk = p.shots[0].src      # receivers the wave reaches within 31 steps (FP32 range)
    p.shots[0].rec = np.array([[si + dx, sj + dy, sk + dz] for dx, dy, dz in
                               ((0, 0, 0), (3, 0, 0), (-3, 1, 0), (0, 3, 1), (2, -2, -2), (1, 1, 3))],
                              dtype=np.int32)
    sim = _sim(p, "fp32")
    for n in (7, 11, 13):
        sim.run(n)
    sim.check()
    o = oracle_lib.OracleRun(p).run()
    assert rel_l2(sim.traces()[0], o.trace) < FP32_TOL
    for n in o.names:
        exp = o.field(n)
        if np.abs(exp).max() > 0:
            assert rel_l2(sim.field(0, n), exp) < FP32_TOL, n


@pytest.mark.parametrize("kind", ["elastic25", "em25"])
def test_3d_tiles_and_y_chunks(oracle_lib, kind):
    """A 2.5-D grid several tiles wide in x, with more y rows than one work
    item walks and a z extent that is not a multiple of the block's planes."""
    p = tiny_problem(kind, seed=58, nx=140, ny=70, nz=23, npml=5, nt=30, nrec=9,
                     **({"gamma": 1e-3} if kind == "elastic25" else {}))
    sim = _sim(p, "fp64").run().check()
    _compare(sim, p, oracle_lib, FP64_TOL)


@pytest.mark.parametrize("kind", ["elastic25", "em25"])
def test_3d_y_rows_above_32767(oracle_lib, kind):
    """A 2.5-D grid 32,832 rows long in y: the work items whose y chunk ends
    at or above row 32768 update their rows (the chunk bounds travel as two
    16-bit halves of one int).  The source sits in the last chunks."""
    p = tiny_problem(kind, seed=67, nx=12, ny=32832, nz=12, npml=2, nt=8, nrec=4,
                     src=(6, 32800, 6), **({"gamma": 1e-3} if kind == "elastic25" else {}))
    p.shots[0].rec = np.array([[6, 32800, 6], [5, 32803, 7], [7, 32795, 5], [6, 32829, 6]],
                              dtype=np.int32)
    sim = _sim(p, "fp64").run().check()
    _compare(sim, p, oracle_lib, FP64_TOL)
    assert np.abs(sim.field(0, "Ex" if kind == "em25" else "vx")[:, 32768:, :]).max() > 0


def test_em2_tile_shapes_keep_results(oracle_lib):
    """The single-pass EM 2-D kernel gives bit-identical results for every
    tile shape, and matches the oracle on several models and shots."""
    probs = [tiny_problem("em2", seed=95 + m, nt=60, nshots=2, **d)
             for m, d in enumerate([dict(nx=300, nz=75, npml=6), dict(nx=57, nz=45, npml=5)])]
    ref = _sim(probs, "fp32").run(graph_block=10).check().traces()
    for wb, rows in ((1, 16), (2, 64), (4, 8), (1, 2), (4, 4), (1, 1), (4, 1)):
        got = _sim(probs, "fp32").set_tiles(wb, rows).run(graph_block=10).check().traces()
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b)
    s = _sim(probs, "fp64")
    s.run(7)
    s.autotune(steps=4)
    s.run(graph_block=10).check()
    off = 0
    for p in probs:
        _compare(s, p, oracle_lib, FP64_TOL, shot_offset=off)
        off += len(p.shots)


def test_em25_block_snapshots_match_record_snapshots(tmp_path):
    """2.5-D EM snapshots written as one int16 block equal the per-step
    records of the same run to the block's quantisation step."""
    import glob
    from seidart_amd import fortranio, run as dropin
    p = tiny_problem("em25", seed=61, nx=20, ny=8, nz=16, npml=3, nt=12)
    tr, unstable, nsnap = dropin._execute(p, "fp64", 4, True, str(tmp_path), 0, False, "em25block")
    assert unstable < 0 and nsnap == 1
    blk = glob.glob(str(tmp_path / "EM25.*.blk"))
    assert len(blk) == 1
    r = fortranio.read_em25_block(blk[0])
    assert r["metadata"]["step_count"] == 3 and sorted(r["fields"]) == ["Ex", "Ey", "Ez"]
    _, _, n2 = dropin._execute(p, "fp64", 4, False, str(tmp_path), 0, False, "dat")
    assert n2 == 9
    si = p.shots[0].src
    for t, it in enumerate((4, 8, 12)):
        for ch in ("Ex", "Ey", "Ez"):
            rec = fortranio.read_snapshot(str(tmp_path / fortranio.snapshot_name(ch, it, tuple(si))),
                                          (p.nz, p.ny, p.nx), single=False)
            got = np.transpose(r["fields"][ch][t], (2, 1, 0))
            q = (rec.max() - rec.min()) / 65534
            tol = 0.5 * q + 4 * np.finfo(np.float32).eps * np.abs(rec).max() + 1e-30
            assert np.abs(got - rec).max() <= tol, (ch, it)


def test_common_offset_batched_matches_oracle(tmp_path, oracle_lib):
    """The batched common-offset survey (all source positions as shots of
    one plan) against each shot's own oracle run, and its CSV."""
    from seidart_amd import survey
    from test_host import _co_files, _write_project
    _write_project(str(tmp_path))
    src, rec = _co_files(str(tmp_path))
    co = survey.CommonOffset(src, "Vz", str(tmp_path / "p.json"), rec, precision="fp64")
    img = co.co_run(graph_block=0)
    assert img.shape == (40, 3)
    q = co.problem()
    for s in range(3):
        o = oracle_lib.OracleRun(q, shot=s).run()
        assert rel_l2(img[:, s], o.trace[:, 1, 0]) < FP64_TOL, s
    assert np.allclose(np.loadtxt(co.csvfilename, delimiter=","), img)


@pytest.mark.parametrize("kind,precision", [("elastic2", "fp64"), ("elastic2", "fp32"),
                                             ("elastic25", "fp64"), ("elastic25", "fp32")])
def test_moment_tensor_source_matches_oracle(oracle_lib, kind, precision):
    """Stress-rate (moment-tensor) sources with a force source beside them,
    on heterogeneous grids, against the oracle; 2-D runs use xx, zz, xz."""
    kw = dict(nx=70, ny=20, nz=40) if kind == "elastic25" else dict(nx=90, nz=60)
    p = tiny_problem(kind, seed=66, npml=5, nt=50, nrec=8, gamma=1e-3, **kw)
    rng = np.random.default_rng(3)
    f = np.asarray(p.shots[0].force[2], dtype=np.float64)
    p.shots[0].moment = tuple(f * rng.uniform(-2e6, 2e6) for _ in range(6))
    si, sj, sk = p.shots[0].src
    p.shots[0].rec = np.array([[si + d, sj, sk + e] for d, e in ((0, 0), (2, 1), (-3, 2), (1, -2))],
                              dtype=np.int32)
    sim = _sim(p, precision).run().check()
    if precision == "fp64":
        _compare(sim, p, oracle_lib, FP64_TOL)
    else:
        o = oracle_lib.OracleRun(p).run()
        assert rel_l2(sim.traces()[0], o.trace) < FP32_TOL


def test_graph_capture_equals_direct_launch():
    p = tiny_problem("elastic2", seed=51, nt=64, nshots=2)
    a = _sim(p, "fp32").run(graph_block=0).check().traces()
    b = _sim(p, "fp32").run(graph_block=8).check().traces()
    c = _sim(p, "fp32").run(nsteps=13, graph_block=5).run(graph_block=5).check().traces()
    for x, y, z in zip(a, b, c):
        np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(x, z)


def test_segments_equal_full_run():
    p = tiny_problem("em2", seed=52, nt=60)
    a = _sim(p).run().check().traces()[0]
    s = _sim(p)
    for n in (1, 10, 20, 29):
        s.run(n)
    b = s.check().traces()[0]
    np.testing.assert_array_equal(a, b)


def test_unstable_run_is_reported():
    """dt far above the CFL limit: the device flag reports the step, the run
    raises instead of the reference's silent process stop."""
    from seidart_amd._hip import UnstableError
    p = tiny_problem("elastic2", seed=53, nt=200)
    p.dt *= 8
    sim = _sim(p).run()
    with pytest.raises(UnstableError):
        sim.check()
    assert sim.unstable[0] >= 0


def test_receivers_on_edges_and_source_on_boundary(oracle_lib):
    p = tiny_problem("elastic2", seed=54, nt=40)
    p.shots[0].rec = np.array([[0, 0, 0], [p.nx - 1, 0, p.nz - 1], [5, 0, p.nz - 1]], dtype=np.int32)
    p.shots[0].src = (1, 0, 1)
    sim = _sim(p).run().check()
    _compare(sim, p, oracle_lib, FP64_TOL)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_split_cpml_variant_matches_oracle(oracle_lib, precision):
    """The split C-PML variant (SDT_SPLIT_CPML) against the oracle's split
    variant, several shots and models in one plan."""
    probs = [tiny_problem("elastic2", seed=70 + m, nt=90, gamma=1e-3, nshots=2, **d)
             for m, d in enumerate([dict(nx=48, nz=40, npml=6), dict(nx=61, nz=37, npml=5)])]
    for p in probs:
        p.cpml_split = True
    sim = _sim(probs, precision).run(graph_block=16).check()
    if precision == "fp64":
        off = 0
        for p in probs:
            _compare(sim, p, oracle_lib, FP64_TOL, shot_offset=off)
            off += len(p.shots)
    else:
        tr = sim.traces()
        o = oracle_lib.OracleRun(probs[0], shot=0).run()
        assert rel_l2(tr[0], o.trace) < FP32_TOL


def test_tile_shapes_and_autotune_keep_results(oracle_lib):
    """Every tile shape gives bit-identical results; autotune restores the
    state it times from, so a run after it equals a run without it."""
    probs = [tiny_problem("elastic2", seed=90 + m, nt=60, gamma=1e-3, nshots=2, **d)
             for m, d in enumerate([dict(nx=300, nz=75, npml=6), dict(nx=57, nz=45, npml=5)])]
    ref = _sim(probs, "fp32").run(graph_block=10).check().traces()
    for wb, rows in ((1, 16), (2, 64), (4, 8), (1, 2), (4, 4), (1, 1), (4, 1)):
        got = _sim(probs, "fp32").set_tiles(wb, rows).run(graph_block=10).check().traces()
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b)
    s = _sim(probs, "fp32")
    s.run(7)
    res = s.autotune(steps=4)
    n = 2 * len(type(s).TILE_ORDERS) * len(type(s).TILE_CANDIDATES)
    assert len(res) == n and (s.steps_per_launch, *s.tiles(), s.tile_order) in res
    got = s.run(graph_block=10).check().traces()
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("kind", ["elastic2", "em2"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_tile_orders_keep_results(kind, precision):
    """Both work-list orders (tiles dealt to the XCDs in turn, or column
    strips per XCD) give bit-identical states and traces, for one and two
    steps per launch, several tile shapes, two models and two shots each;
    an order other than 0 or 1 is refused."""
    probs = [tiny_problem(kind, seed=120 + m, nt=44, nshots=2, **d)
             for m, d in enumerate([dict(nx=300, nz=75, npml=6), dict(nx=57, nz=45, npml=5)])]

    def snap(s):
        torch.cuda.synchronize()
        return [b["state"].cpu().numpy().copy() for _, _, b in s.shots], s.traces()

    ref = snap(_sim(probs, precision).run(graph_block=10).check())
    for spl, wb, rows in ((1, 4, 1), (1, 2, 8), (2, 4, 16), (2, 1, 4)):
        s = _sim(probs, precision)
        s.steps_per_launch = spl
        s.set_tiles(wb, rows)
        s.tile_order = 1
        assert s.tile_order == 1
        got = snap(s.run(graph_block=10).check())
        for a, b in zip(ref[0] + ref[1], got[0] + got[1]):
            np.testing.assert_array_equal(a, b)
    with pytest.raises(RuntimeError):
        s.tile_order = 2


@pytest.mark.parametrize("kind", ["elastic2", "em2", "elastic25", "em25"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_initial_conditions_match_oracle(oracle_lib, kind, precision):
    """Non-zero initial velocity (elastic) or E (EM) fields, the
    initialcondition{V,E}{x,y,z}.dat inputs of the current solver
    (classes.py:888-898), with the source switched off: traces and every
    final field against the oracle, which starts from the same fields."""
    kw = SMALL[kind]
    p = tiny_problem(kind, **kw)
    rng = np.random.default_rng(77)
    names = ("vx", "vy", "vz") if kind.startswith("elastic") else ("Ex", "Ey", "Ez")
    from seidart_amd.problem import FIELDS, field_extent
    init = {}
    for n in names:
        if n in FIELDS[kind]:
            # zero outside the field's legacy extent, as the device keeps it
            f = np.zeros(p.field_shape)
            ext = field_extent(kind, n)
            f[ext] = rng.standard_normal(p.field_shape)[ext] * (1e-3 if kind.startswith("elastic") else 1.0)
            init[n] = f
    p.init = [init]
    sh = p.shots[0]
    sh.force = tuple(None if f is None else np.zeros_like(f) for f in sh.force)
    sim = _sim(p, precision).run().check()
    if precision == "fp64":
        _compare(sim, p, oracle_lib, FP64_TOL)
    else:
        o = oracle_lib.OracleRun(p).run()
        assert rel_l2(sim.traces()[0], o.trace) < FP32_TOL
        for n in o.names:
            if np.abs(o.field(n)).max() > 0:
                assert rel_l2(sim.field(0, n), o.field(n)) < FP32_TOL, n
