"""Write the golden fixtures: receiver traces and final fields of the legacy
Fortran solver (oracle/_ref, compiled unmodified from the reference's
legacy/fortran sources) on small seeded problems.

Run in a container holding the reference tree:
    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import legacy_ref  # noqa: E402
from helpers import tiny_problem  # noqa: E402

# (kind, tiny_problem kwargs); dt is always the legacy solver's own dt
CASES = {
    "elastic2": dict(seed=11, nx=40, nz=36, npml=6, nt=80),
    "elastic2_gamma": dict(seed=12, nx=40, nz=36, npml=6, nt=80, gamma=2e-3),
    "em2": dict(seed=13, nx=40, nz=36, npml=6, nt=80),
    "elastic25": dict(seed=14, nx=24, ny=12, nz=20, npml=4, nt=40),
    "em25": dict(seed=15, nx=24, ny=12, nz=20, npml=4, nt=40),
}


def case_problem(name):
    kind = name.split("_")[0]
    kw = dict(CASES[name])
    p = tiny_problem(kind, **kw)
    p.dt = legacy_ref.legacy_dt(kind, p.coef, p.dx, p.dy, p.dz)
    # the legacy seismic2 applies gamma_x to vx only (gamma_z unset): keep z at 0
    if "z" in p.gamma:
        p.gamma = {"x": p.gamma["x"]}
    # profiles depend on dt: rebuild them for the legacy dt
    p2 = tiny_problem(kind, dt=p.dt, **kw)
    p2.gamma = p.gamma
    return p2


# snapshot records of the legacy writer (write_image2, FP64 records of each
# channel at its legacy extent) kept as raw bytes for these cases and steps
SNAPSHOT_CASES = ("elastic2", "em2")
SNAPSHOT_STEPS = (20, 40, 60, 80)


def snapshot_fixtures():
    """legacy_snapshots_{case}.npz: for each kept record, its channel, step
    and the raw bytes of the file the legacy solver wrote (record markers
    included)."""
    import glob
    import tempfile
    for name in SNAPSHOT_CASES:
        p = case_problem(name)
        tmp = tempfile.mkdtemp(prefix="legacy_snap_")
        legacy_ref.run_legacy(p, keep_dir=tmp)
        out = {}
        q = 0
        for ch in ("Vx", "Vz") if name.startswith("elastic") else ("Ex", "Ez"):
            for it in SNAPSHOT_STEPS:
                files = glob.glob(os.path.join(tmp, f"{ch}.{it:06d}.*"))
                assert len(files) == 1, files
                with open(files[0], "rb") as fh:
                    raw = fh.read()
                out[f"channel_{q}"] = np.array(ch)
                out[f"step_{q}"] = np.int64(it)
                out[f"bytes_{q}"] = np.frombuffer(raw, dtype=np.uint8)
                q += 1
        for f in glob.glob(os.path.join(tmp, "*")):
            os.remove(f)
        os.rmdir(tmp)
        np.savez_compressed(os.path.join(HERE, f"legacy_snapshots_{name}.npz"), **out)
        print("snapshots", name, q)


def main():
    snapshot_fixtures()
    for name in CASES:
        p = case_problem(name)
        leg = legacy_ref.run_legacy(p)
        rec = np.asarray(p.shots[0].rec)
        out = {"dt": np.float64(p.dt), "rec": rec}
        for ch, arr in leg.items():
            if arr.ndim == 3:
                out[f"trace_{ch}"] = arr[:, rec[:, 2], rec[:, 0]]
            else:
                out[f"trace_{ch}"] = arr[:, rec[:, 2], rec[:, 1], rec[:, 0]]
            out[f"final_{ch}"] = arr[-1]
        np.savez_compressed(os.path.join(HERE, f"legacy_{name}.npz"), **out)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
