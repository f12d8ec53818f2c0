"""The drop-in boundary end to end on the GPU: ``python -m seidart_amd.run``
(the reference binary's command line, classes.py:910-924) on projects in the
reference's file contract, its snapshot records against the legacy writer's
(readwrite_routines.f95:473-497) and its trace CSVs against the oracle; and
the project-file route (JSON and geometry only) against the file route."""
import glob
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from golden.make_golden import CASES, case_problem
from helpers import rel_l2, write_reference_project

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cli(jf, seismic, *args):
    cmd = [sys.executable, "-u", "-m", "seidart_amd.run", jf,
           f"seismic={'true' if seismic else 'false'}", "density_method=legacy", *args]
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-3000:]
    return res


def _record(path):
    raw = open(path, "rb").read()
    (m,) = struct.unpack("<i", raw[:4])
    assert m == len(raw) - 8 and struct.unpack("<i", raw[-4:])[0] == m, path
    return raw


@pytest.mark.parametrize("name", ["elastic2", "em2"])
def test_cli_snapshots_match_legacy_records_and_csv_matches_oracle(tmp_path, oracle_lib, name):
    p = case_problem(name)
    seismic = p.kind.startswith("elastic")
    jf, rf = write_reference_project(p, str(tmp_path), cpml=CASES[name]["npml"])
    # legacy layout, FP64 records, legacy names, every 20 steps
    _cli(jf, seismic, "snapshot_every=20", "single=false", f"receivers={rf}",
         "snapshot_layout=legacy", "legacy_names=true")
    g = np.load(os.path.join(GOLDEN, f"legacy_snapshots_{name}.npz"))
    n = len([k for k in g.files if k.startswith("bytes_")])
    assert n == 8
    for q in range(n):
        ch, it, ref = str(g[f"channel_{q}"]), int(g[f"step_{q}"]), g[f"bytes_{q}"].tobytes()
        files = glob.glob(str(tmp_path / f"{ch}.{it:06d}.*..dat"))
        assert len(files) == 1, (ch, it, files)
        raw = _record(files[0])
        assert len(raw) == len(ref) and raw[:4] == ref[:4], (ch, it)
        got = np.frombuffer(raw[4:-4], dtype=np.float64)
        exp = np.frombuffer(ref[4:-4], dtype=np.float64)
        err = rel_l2(got, exp)
        print(f"{name} {ch} step {it}: record {len(raw)} B, rel L2 vs legacy record {err:.3e}")
        assert err < 1e-10, (ch, it, err)
    # trace CSVs (Array.save naming) against the oracle
    o = oracle_lib.OracleRun(p).run()
    csvs = sorted(glob.glob(str(tmp_path / "golden-*.csv")))
    assert len(csvs) == 2
    chans = ["Vx", "Vz"] if seismic else ["Ex", "Ez"]
    for c, ch in enumerate(chans):
        f = [x for x in csvs if f"-{ch}-" in os.path.basename(x)]
        assert len(f) == 1
        tr = np.loadtxt(f[0], delimiter=",", ndmin=2)
        assert tr.shape == (p.nt, len(p.shots[0].rec))
        assert rel_l2(tr, o.trace[:, c, :]) < 1e-10, ch


@pytest.mark.parametrize("name", ["elastic2", "em2"])
def test_cli_default_snapshots_are_padded_fp32_records(tmp_path, oracle_lib, name):
    """The default output: every channel on the full padded grid as FP32
    (what read_dat reshapes), named {ch}.{it:06d}.{xind}.{zind}.dat."""
    p = case_problem(name)
    seismic = p.kind.startswith("elastic")
    cpml = CASES[name]["npml"]
    jf, _ = write_reference_project(p, str(tmp_path), cpml=cpml)
    _cli(jf, seismic, "snapshot_every=40")
    i, _, k = p.shots[0].src
    o = oracle_lib.OracleRun(p)
    done = 0
    for it in (40, 80):
        o.run(done, it - done)
        done = it
        for ch, fld in (("Vx", "vx"), ("Vz", "vz")) if seismic else (("Ex", "Ex"), ("Ez", "Ez")):
            fn = tmp_path / f"{ch}.{it:06d}.{i - cpml + 1}.{k - cpml + 1}.dat"
            raw = _record(str(fn))
            assert len(raw) == 8 + 4 * p.nx * p.nz
            got = np.frombuffer(raw[4:-4], dtype=np.float32).reshape(p.nz, p.nx)
            exp = o.field(fld).astype(np.float32)
            assert np.abs(got.astype(np.float64) - exp).max() <= 2 * np.spacing(np.abs(exp).max()), (ch, it)


@pytest.mark.parametrize("seismic", [True, False])
def test_project_route_traces_equal_file_route(tmp_path, oracle_lib, seismic):
    """run_project (JSON and geometry only: dt, sources, C-PML, painting,
    density gradient computed here) gives the same traces, bit for bit, as
    run_files on the files the same build wrote, and matches the oracle."""
    from test_project import write_project
    from seidart_amd import project, run as dropin
    jf, _ = write_project(str(tmp_path), time_steps=160)
    d, mat, sm, em = project.load_project(jf)
    model = sm if seismic else em
    project.build(model, d, mat, write=True)
    rec = np.array([[5.0, 0.0, 3.0], [12.0, 0.0, 2.5], [20.0, 0.0, 3.0], [27.5, 0.0, 10.0]])
    a = project.run_project(jf, seismic, receivers=rec)
    b = dropin.run_files(jf, seismic, receivers=rec)
    assert np.abs(a.traces).max() > 0
    np.testing.assert_array_equal(a.traces, b.traces)
    o = oracle_lib.OracleRun(a.problem).run()
    assert rel_l2(a.traces, o.trace) < 1e-10
