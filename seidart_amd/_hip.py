"""ctypes binding of libseidart_hip.so (C ABI: include/seidart_hip.h).

The library is built in-tree by :func:`seidart_amd.build.build_library`.
There is no fallback: if the library is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libseidart_hip.so")

SDT_ELASTIC2, SDT_ELASTIC25, SDT_EM2, SDT_EM25 = 0, 1, 2, 3
SDT_SPLIT_CPML = 0x100
SDT_ADJ_REFERENCE = 1
SDT_FP64, SDT_FP32 = 0, 1
SDT_OK, SDT_EINVAL, SDT_EDEVICE, SDT_EUNSTABLE = 0, 1, 2, 3


class SdtModel(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_int), ("ly", C.c_int), ("lz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("scal", C.c_double * 4),
        ("coef", C.c_void_p * 40),
        ("prof", (C.c_void_p * 6) * 3),
        ("aux", C.c_void_p),
    ]


class SdtShot(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("isrc", C.c_int), ("jsrc", C.c_int), ("ksrc", C.c_int),
        ("field", C.c_void_p * 9),
        ("psi", C.c_void_p),
        ("field_b", C.c_void_p * 9),
        ("psi_b", C.c_void_p),
        ("src", C.c_void_p * 3),
        ("nrec", C.c_int),
        ("rec", C.c_void_p),
        ("trace", C.c_void_p),
        ("msrc", C.c_void_p * 6),
    ]


class SdtAdjShot(C.Structure):
    _fields_ = [
        ("trail", C.c_void_p * 2),
        ("trail_len", C.c_int),
        ("lam", C.c_void_p * 2),
        ("work", C.c_void_p),
        ("resid", C.c_void_p),
        ("grad", C.c_void_p),
        ("gsrc", C.c_void_p),
    ]


class SeidartHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[sdt error {code}] {msg}")
        self.code = code


class UnstableError(SeidartHipError):
    """A field exceeded 1e25 or became non-finite (the reference stops with
    'code became unstable and blew up', seismicfdtd.f95:378-379)."""


_lib = None

EXPORTS = ["sdt_psi_elems", "sdt_plan_create", "sdt_plan_run", "sdt_plan_check",
           "sdt_plan_profile", "sdt_plan_destroy", "sdt_last_error", "sdt_version",
           "sdt_state_elems", "sdt_adjoint_work_elems", "sdt_adjoint_setup",
           "sdt_plan_run_trail", "sdt_adjoint_run", "sdt_plan_set_tiles", "sdt_plan_get_tiles",
           "sdt_plan_set_tile_order", "sdt_plan_get_tile_order", "sdt_plan_set_steps_per_launch", "sdt_plan_get_steps_per_launch", "sdt_run_noalloc",
           "sdt_noalloc_scratch_bytes", "sdt_hbm_probe"]


def lib():
    """Load the library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is not built; run seidart_amd.build.build_library() "
                           "(or __graft_entry__.build()) first")
    L = C.CDLL(LIBPATH)
    L.sdt_psi_elems.argtypes = [C.c_int] * 7 + [C.POINTER(C.c_longlong)]
    L.sdt_psi_elems.restype = C.c_int
    L.sdt_plan_create.argtypes = [C.c_int, C.c_int, C.POINTER(SdtModel), C.c_int,
                                  C.POINTER(SdtShot), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.sdt_plan_create.restype = C.c_int
    L.sdt_plan_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.sdt_plan_run.restype = C.c_int
    L.sdt_plan_check.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.sdt_plan_check.restype = C.c_int
    L.sdt_plan_profile.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                   C.POINTER(C.c_double)]
    L.sdt_plan_profile.restype = C.c_int
    L.sdt_plan_destroy.argtypes = [C.c_void_p]
    L.sdt_plan_destroy.restype = None
    L.sdt_last_error.argtypes = []
    L.sdt_last_error.restype = C.c_char_p
    L.sdt_version.argtypes = []
    L.sdt_version.restype = C.c_char_p
    L.sdt_state_elems.argtypes = [C.c_int] * 7 + [C.POINTER(C.c_longlong)]
    L.sdt_state_elems.restype = C.c_int
    L.sdt_adjoint_work_elems.argtypes = [C.c_int] * 7 + [C.POINTER(C.c_longlong)]
    L.sdt_adjoint_work_elems.restype = C.c_int
    L.sdt_adjoint_setup.argtypes = [C.c_void_p, C.POINTER(SdtAdjShot), C.c_int, C.c_uint, C.c_int]
    L.sdt_adjoint_setup.restype = C.c_int
    L.sdt_plan_run_trail.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.sdt_plan_run_trail.restype = C.c_int
    L.sdt_adjoint_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.sdt_adjoint_run.restype = C.c_int
    L.sdt_plan_set_tiles.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.sdt_plan_set_tiles.restype = C.c_int
    L.sdt_plan_get_tiles.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.sdt_plan_get_tiles.restype = C.c_int
    L.sdt_plan_set_tile_order.argtypes = [C.c_void_p, C.c_int]
    L.sdt_plan_set_tile_order.restype = C.c_int
    L.sdt_plan_get_tile_order.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.sdt_plan_get_tile_order.restype = C.c_int
    L.sdt_plan_set_steps_per_launch.argtypes = [C.c_void_p, C.c_int]
    L.sdt_plan_set_steps_per_launch.restype = C.c_int
    L.sdt_plan_get_steps_per_launch.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.sdt_plan_get_steps_per_launch.restype = C.c_int
    L.sdt_run_noalloc.argtypes = [C.c_int, C.c_int, C.POINTER(SdtModel), C.c_int, C.POINTER(SdtShot),
                                  C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.sdt_run_noalloc.restype = C.c_int
    L.sdt_noalloc_scratch_bytes.argtypes = []
    L.sdt_noalloc_scratch_bytes.restype = C.c_longlong
    L.sdt_hbm_probe.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p]
    L.sdt_hbm_probe.restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc == SDT_OK:
        return
    msg = lib().sdt_last_error().decode(errors="replace")
    if rc == SDT_EUNSTABLE:
        raise UnstableError(rc, msg)
    raise SeidartHipError(rc, msg)


def psi_elems(kind: int, nx, ny, nz, lx, ly, lz) -> int:
    out = C.c_longlong(0)
    check(lib().sdt_psi_elems(kind, nx, ny, nz, lx, ly, lz, C.byref(out)))
    return int(out.value)


def state_elems(kind: int, nx, ny, nz, lx, ly, lz) -> int:
    out = C.c_longlong(0)
    check(lib().sdt_state_elems(kind, nx, ny, nz, lx, ly, lz, C.byref(out)))
    return int(out.value)


def adjoint_work_elems(kind: int, nx, ny, nz, lx, ly, lz) -> int:
    out = C.c_longlong(0)
    check(lib().sdt_adjoint_work_elems(kind, nx, ny, nz, lx, ly, lz, C.byref(out)))
    return int(out.value)
