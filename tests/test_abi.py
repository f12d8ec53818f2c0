"""The C ABI: the library loads, exports every symbol include/seidart_hip.h
declares, and its host-side argument checks reject bad input without a GPU."""
import ctypes as C
import os
import re

import pytest

from seidart_amd import _hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "seidart_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sdt_[a-z_]+)\s*\(", txt)))


def test_header_declares_the_bound_entry_points():
    assert declared_symbols() == sorted(_hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _hip.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.sdt_version()


def test_struct_layout_matches_header(tmp_path):
    # 6 ints, 4 doubles, scal[4], coef[40], prof[3][6], aux
    assert C.sizeof(_hip.SdtModel) == 24 + 32 + 32 + 8 * 40 + 8 * 18 + 8
    assert C.sizeof(_hip.SdtShot) == 16 + 8 * 9 + 8 + 8 * 9 + 8 + 8 * 3 + 8 + 8 + 8 + 8 * 6
    # and against the header itself, compiled by the host C compiler
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no host C compiler")
    inc = os.path.join(os.path.dirname(__file__), "..", "include")
    src = tmp_path / "probe.c"
    structs = {"sdt_model": _hip.SdtModel, "sdt_shot": _hip.SdtShot, "sdt_adj_shot": _hip.SdtAdjShot}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "seidart_hip.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, cls in structs.items():
        assert int(out[cname]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


@pytest.mark.parametrize("kind,dims,expect", [
    (_hip.SDT_ELASTIC2, (100, 1, 80, 10, 0, 12), 4 * 80 * 20 + 4 * 24 * 100),
    (_hip.SDT_EM2, (100, 1, 80, 10, 0, 12), 2 * 80 * 20 + 2 * 24 * 100),
    (_hip.SDT_ELASTIC25, (30, 20, 40, 5, 4, 6), 6 * (40 * 20 * 10 + 40 * 8 * 30 + 12 * 20 * 30)),
    (_hip.SDT_EM25, (30, 20, 40, 5, 4, 6), 4 * (40 * 20 * 10 + 40 * 8 * 30 + 12 * 20 * 30)),
    (_hip.SDT_ELASTIC2 | _hip.SDT_SPLIT_CPML, (100, 1, 80, 10, 0, 12), 6 * 80 * 20 + 6 * 24 * 100),
])
def test_psi_elems(kind, dims, expect):
    assert _hip.psi_elems(kind, *dims) == expect


def test_psi_elems_bad_kind():
    with pytest.raises(_hip.SeidartHipError, match="unknown kind"):
        _hip.psi_elems(9, 10, 1, 10, 1, 0, 1)
    with pytest.raises(_hip.SeidartHipError, match="unknown kind"):
        _hip.psi_elems(_hip.SDT_EM2 | _hip.SDT_SPLIT_CPML, 10, 1, 10, 1, 0, 1)


def test_state_and_adjoint_sizes():
    dims = (100, 1, 80, 10, 0, 12)
    for k in (_hip.SDT_ELASTIC2, _hip.SDT_ELASTIC2 | _hip.SDT_SPLIT_CPML):
        assert _hip.state_elems(k, *dims) == 5 * 8000 + _hip.psi_elems(k, *dims)
        assert _hip.adjoint_work_elems(k, *dims) == 12 * 8000
    assert _hip.state_elems(_hip.SDT_EM25, 30, 20, 40, 5, 4, 6) == \
        6 * 30 * 20 * 40 + _hip.psi_elems(_hip.SDT_EM25, 30, 20, 40, 5, 4, 6)
    assert _hip.adjoint_work_elems(_hip.SDT_EM2, *dims) == 4 * 8000
    with pytest.raises(_hip.SeidartHipError, match="ELASTIC2 and EM2 only"):
        _hip.adjoint_work_elems(_hip.SDT_EM25, 30, 20, 40, 5, 4, 6)
    # trail[2], trail_len (+pad), lam[2], work, resid, grad, gsrc
    assert C.sizeof(_hip.SdtAdjShot) == 16 + 8 + 16 + 8 * 4


def test_adjoint_calls_reject_null_plan():
    L = _hip.lib()
    assert L.sdt_adjoint_setup(None, None, 0, 0xfff, 0) == _hip.SDT_EINVAL
    assert L.sdt_plan_run_trail(None, 0, 1, 0, None) == _hip.SDT_EINVAL
    assert L.sdt_adjoint_run(None, 1, 0, 0, 0, None) == _hip.SDT_EINVAL
    assert L.sdt_plan_set_tiles(None, 4, 32) == _hip.SDT_EINVAL
    assert L.sdt_plan_set_tile_order(None, 1) == _hip.SDT_EINVAL
    assert L.sdt_plan_set_steps_per_launch(None, 2) == _hip.SDT_EINVAL
    v = C.c_int(0)
    assert L.sdt_plan_get_steps_per_launch(None, C.byref(v)) == _hip.SDT_EINVAL
    assert L.sdt_plan_get_tile_order(None, C.byref(v)) == _hip.SDT_EINVAL


def _model(nx=20, nz=20):
    M = _hip.SdtModel()
    M.nx, M.ny, M.nz = nx, 1, nz
    M.lx, M.lz = 3, 3
    M.dx = M.dy = M.dz = 1.0
    M.dt = 1e-4
    for q in range(12):
        M.coef[q] = 0x1000  # never dereferenced: validation fails first
    for a in (0, 2):
        for q in range(6):
            M.prof[a][q] = 0x1000
    M.aux = 0x1000
    return M


def _shot(model=0, src=(5, 0, 5)):
    S = _hip.SdtShot()
    S.model = model
    S.isrc, S.jsrc, S.ksrc = src
    for f in range(5):
        S.field[f] = 0x1000
        S.field_b[f] = 0x1000
    S.psi = 0x1000
    S.psi_b = 0x1000
    return S


@pytest.mark.parametrize("mutate,msg", [
    (lambda M, S: setattr(M, "nx", 2), "grid too small"),
    (lambda M, S: setattr(M, "lx", 50), "slab width"),
    (lambda M, S: setattr(M, "dt", 0.0), "must be positive"),
    (lambda M, S: setattr(S, "isrc", 99), "source outside"),
    (lambda M, S: setattr(S, "model", 3), "bad model index"),
    (lambda M, S: M.coef.__setitem__(3, None), "coefficient slot 3"),
    (lambda M, S: M.coef.__setitem__(4, None), "must be all set or all null"),
    (lambda M, S: setattr(M, "aux", None), "aux"),
    (lambda M, S: setattr(S, "nrec", 4), "receivers need"),
])
def test_plan_create_rejects_bad_input(mutate, msg):
    M, S = _model(), _shot()
    mutate(M, S)
    models = (_hip.SdtModel * 1)(M)
    shots = (_hip.SdtShot * 1)(S)
    plan = C.c_void_p()
    rc = _hip.lib().sdt_plan_create(_hip.SDT_ELASTIC2, _hip.SDT_FP64, models, 1, shots, 1, 10,
                                    C.byref(plan))
    assert rc == _hip.SDT_EINVAL
    assert msg in _hip.lib().sdt_last_error().decode()
    assert not plan.value


def test_plan_run_rejects_null_plan():
    rc = _hip.lib().sdt_plan_run(None, 0, 1, 0, None)
    assert rc == _hip.SDT_EINVAL
