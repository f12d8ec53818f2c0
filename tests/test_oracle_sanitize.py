"""The oracle's C restatement built with AddressSanitizer and UBSan
(`make -C oracle san`) and driven by oracle/san_check.c over small grids:
sources and receivers on the first and last cells of every axis, grids with
nz > nx (where the restatement clamps the legacy's out-of-range b_x_half read,
seismicfdtd.f95:255), segmented runs over one psi workspace, the split C-PML
variant, attenuation and stress-rate sources, for all four routines."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(__file__), "..", "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_oracle_clean_under_asan_ubsan(tmp_path):
    build = subprocess.run(["make", "-C", ORACLE, "san", f"SANDIR={tmp_path}"],
                           capture_output=True, text=True)
    if build.returncode and "asan" in build.stderr.lower():
        pytest.skip("gcc has no AddressSanitizer runtime here")
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(tmp_path / "san_check")], capture_output=True, text=True,
                         timeout=120)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "runtime error" not in run.stderr
    lines = [ln for ln in run.stdout.splitlines() if ln.startswith("kind ")]
    # 4 routines x 3 grids x 3 source cells, elastic2 twice (split C-PML)
    assert len(lines) == 45
    assert all(": rc 0 blowup -1 " in ln for ln in lines)
    # the centred sources reach the receivers beside them
    assert all("trace sum 0.000000e+00" not in ln for ln in lines
               if "source centre" in ln and " 3x" not in ln)
    assert run.stdout.strip().endswith("ok")
