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
    "EM25_MAGIC", "write_em25_block", "read_em25_block", "em25_block_name",
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


# ---------------------------------------------------------------------------
# 2.5-D EM block files (src/seidart/routines/definitions.py:140-212, the
# reader of the blocks the current solver writes): a 16-byte ASCII magic,
# int32 version (1), nx, ny, nz, first step, step count, component count,
# then step_count * component_count records of int32 step, a 2-character
# channel, float32 min and max, and nx*ny*nz int16 values in Fortran order
# of (nx, ny, nz); value = (raw + 32767) / 65534 * (max - min) + min, and a
# record with max == min holds min everywhere.
# ---------------------------------------------------------------------------
EM25_MAGIC = "SEIDART_EM25B1"


def em25_block_name(first_step: int, src_idx) -> str:
    """File name this package gives a block: ``EM25.{first:06d}.{x}.{y}.{z}.blk``
    (source indices without C-PML padding, like the snapshot names).  The
    current solver's own block names are not in the reference sources."""
    return f"EM25.{int(first_step):06d}." + ".".join(str(int(v)) for v in src_idx) + ".blk"


def _quantize(v: np.ndarray):
    """int16 block-floating form of one field: (min, max, raw)."""
    v = np.asarray(v, dtype=np.float32)
    lo, hi = np.float32(v.min()), np.float32(v.max())
    if hi == lo:
        return lo, hi, np.zeros(v.size, dtype="<i2")
    q = (v.astype(np.float64).ravel() - float(lo)) / (float(hi) - float(lo)) * 65534.0 - 32767.0
    return lo, hi, np.clip(np.rint(q), -32767, 32767).astype("<i2")


def write_em25_block(path: str, fields: dict, first_step: int) -> None:
    """Write one block.  ``fields`` maps a 2-character channel ("Ex", ...) to
    an array (step_count, nz, ny, nx) in this package's C order (whose bytes
    are the Fortran order of (nx, ny, nz)); steps are first_step,
    first_step + 1, ...  Records go step by step, channels in dict order."""
    chans = list(fields)
    if not chans:
        raise ValueError("no channels")
    for ch in chans:
        if len(ch) != 2:
            raise ValueError(f"channel names have two characters: {ch!r}")
    arr0 = np.asarray(fields[chans[0]])
    if arr0.ndim != 4:
        raise ValueError("each channel is (step_count, nz, ny, nx)")
    nt, nz, ny, nx = arr0.shape
    with open(path, "wb") as f:
        f.write(EM25_MAGIC.ljust(16).encode("ascii"))
        f.write(np.array([1, nx, ny, nz, first_step, nt, len(chans)], dtype="<i4").tobytes())
        for t in range(nt):
            for ch in chans:
                a = np.asarray(fields[ch])
                if a.shape != arr0.shape:
                    raise ValueError("all channels must have the same shape")
                lo, hi, raw = _quantize(a[t])
                f.write(np.array([first_step + t], dtype="<i4").tobytes())
                f.write(ch.encode("ascii"))
                f.write(np.array([lo, hi], dtype="<f4").tobytes())
                f.write(raw.tobytes())


def read_em25_block(path: str, channels=None, dtype=np.float32) -> dict:
    """Read a block (plain or gzip-compressed) back: {"metadata": {...},
    "fields": {channel: (step_count, nx, ny, nz) array}, "scales":
    {channel: [(step, min, max), ...]}}, the arrays indexed like the
    reference reader's (x, y, z)."""
    import gzip
    with open(path, "rb") as fh:
        head = fh.read(2)
    opener = gzip.open if head == b"\x1f\x8b" else open
    want = None if channels is None else {c.strip() for c in channels}
    dt = np.dtype(dtype).type
    with opener(path, "rb") as f:
        if f.read(16).decode("ascii").strip() != EM25_MAGIC:
            raise ValueError(f"{path} is not an EM25 block file")
        version, nx, ny, nz, first, count, ncomp = (int(v) for v in np.frombuffer(f.read(28), "<i4"))
        if version != 1:
            raise ValueError(f"unsupported EM25 block version {version}")
        n = nx * ny * nz
        meta = {"version": version, "nx": nx, "ny": ny, "nz": nz, "first_step": first,
                "step_count": count, "component_count": ncomp}
        fields, scales = {}, {}
        for _ in range(count * ncomp):
            step = int(np.frombuffer(f.read(4), "<i4")[0])
            ch = f.read(2).decode("ascii")
            lo, hi = (dt(v) for v in np.frombuffer(f.read(8), "<f4"))
            raw = np.frombuffer(f.read(2 * n), "<i2")
            if want is not None and ch not in want:
                continue
            if ch not in fields:
                fields[ch] = np.empty((count, nx, ny, nz), dtype=dt)
                scales[ch] = []
            t = step - first
            if hi == lo:
                fields[ch][t] = lo
            else:
                vals = (raw.astype(dt) + dt(32767.0)) / dt(65534.0) * (hi - lo) + lo
                fields[ch][t] = vals.reshape((nx, ny, nz), order="F")
            scales[ch].append((step, float(lo), float(hi)))
    return {"metadata": meta, "fields": fields, "scales": scales}
