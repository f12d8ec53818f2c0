"""Fortran file contract of SeidarT's solver boundary.

Every coefficient, source and snapshot file the reference hands to (or gets
back from) its solver is one Fortran unformatted sequential record: a 4-byte
little-endian length, the payload, the same length again.  2-D arrays are
written column-major (x fastest): ``FortranFile.write_record(arr.T)`` on an
``(nx, nz)`` array (src/seidart/routines/classes.py:608-679), read back by
``material_rw`` (legacy/fortran/readwrite_routines.f95:451-470).  The legacy
1-D C-PML profiles are raw stream files of float64
(readwrite_routines.f95:26-37; written with ``ndarray.tofile`` in
legacy/prjrun_101424.py:244-257).

Arrays in this package are held C-ordered as (nz, nx) or (nz, ny, nx), so
their bytes are already the Fortran column-major layout of (nx, nz) /
(nx, ny, nz).
"""
from __future__ import annotations

import os
import struct

import numpy as np

__all__ = [
    "write_record", "read_record", "write_stream", "read_stream",
    "snapshot_name", "write_snapshot", "read_snapshot",
]


def write_record(path: str, arr, dtype=np.float64) -> None:
    """Write one Fortran sequential record (classes.py:677-679)."""
    a = np.ascontiguousarray(np.asarray(arr, dtype=dtype))
    n = a.nbytes
    if n >= 2**31:
        raise ValueError(f"{path}: record of {n} bytes exceeds the 4-byte marker")
    with open(path, "wb") as f:
        f.write(struct.pack("<i", n))
        f.write(a.tobytes())
        f.write(struct.pack("<i", n))


def read_record(path: str, dtype=np.float64, shape=None) -> np.ndarray:
    """Read one Fortran sequential record; raise on a truncated file."""
    with open(path, "rb") as f:
        head = f.read(4)
        if len(head) != 4:
            raise ValueError(f"{path}: missing record marker")
        (n,) = struct.unpack("<i", head)
        if n < 0:
            raise ValueError(f"{path}: negative record length")
        body = f.read(n)
        tail = f.read(4)
    if len(body) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
        raise ValueError(f"{path}: truncated or corrupt record")
    a = np.frombuffer(body, dtype=dtype)
    if shape is not None:
        if int(np.prod(shape)) != a.size:
            raise ValueError(f"{path}: record holds {a.size} values, expected shape {shape}")
        a = a.reshape(shape)
    return a.copy()


def write_stream(path: str, arr) -> None:
    """Raw float64 stream file (the legacy 1-D C-PML profile format)."""
    np.ascontiguousarray(np.asarray(arr, dtype=np.float64)).tofile(path)


def read_stream(path: str, n: int | None = None) -> np.ndarray:
    a = np.fromfile(path, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"{path}: {a.size} values, expected {n}")
    return a


def snapshot_name(channel: str, it: int, src_idx, legacy_2d_dots: bool = False) -> str:
    """Snapshot file name.

    ``{ch}.{it:06d}.{x}.{z}.dat`` (2-D) or ``{ch}.{it:06d}.{x}.{y}.{z}.dat``
    (2.5-D), with the source indices *without* C-PML padding and ``it``
    1-based (legacy/fortran/readwrite_routines.f95:486-487, :510-511), which
    is what the trace reader globs for (src/seidart/routines/arraybuild.py:
    208-216).  The legacy 2-D writer emits an extra '.' before ``dat``
    (``a1, a4`` in its format); ``legacy_2d_dots`` reproduces that.
    """
    idx = ".".join(str(int(v)) for v in src_idx)
    tail = "..dat" if (legacy_2d_dots and len(src_idx) == 2) else ".dat"
    return f"{channel}.{int(it):06d}.{idx}{tail}"


def write_snapshot(path: str, field, single: bool = True) -> None:
    """Full padded field as one record, FP32 unless ``single`` is False
    (``sngl(image_data)``, readwrite_routines.f95:489-494)."""
    write_record(path, field, dtype=np.float32 if single else np.float64)


def read_snapshot(path: str, shape, single: bool = True) -> np.ndarray:
    return read_record(path, dtype=np.float32 if single else np.float64, shape=shape)


def exists_all(paths) -> bool:
    return all(os.path.exists(p) for p in paths)
