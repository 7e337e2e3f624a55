"""Binary case and state files read and written by the Fortran driver.

Case file (stream, native endianness), read by fortran/isph_main.f90:
    int32 magic 0x48505349 ('ISPH' in little-endian bytes), int32 version 1
    isph_params as the C struct
    int64 n, int64 ns
    float64 x y u v p m rho rho0 vol0 c   (n each)
    int32 tid (n), uint8 flags (n)
    float64 sx sy sm srho                 (ns each)
    float64 time, displaytime
State file, written by the driver at the end of its run:
    int32 magic, int32 version, int64 n,
    the ten float64 arrays, tid, flags (n each), float64 t, dt
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

MAGIC = 0x48505349
VERSION = 1
PART_FIELDS = ("x", "y", "u", "v", "p", "m", "rho", "rho0", "vol0", "c")
SOLID_FIELDS = ("x", "y", "m", "rho")


def _params_struct():
    from . import Params
    return Params


def write_case(path, case, time: float, displaytime: float, params: dict | None = None) -> None:
    """Write `case` (particles, solids, run parameters) and the run controls."""
    Params = _params_struct()
    P = Params.from_dict(dict(case.params if params is None else params))
    with open(path, "wb") as f:
        f.write(np.array([MAGIC, VERSION], np.int32).tobytes())
        f.write(bytes(P))
        f.write(np.array([case.n, case.ns], np.int64).tobytes())
        for k in PART_FIELDS:
            f.write(np.ascontiguousarray(case.parts[k], np.float64).tobytes())
        f.write(np.ascontiguousarray(case.parts["tid"], np.int32).tobytes())
        f.write(np.ascontiguousarray(case.parts["flags"], np.uint8).tobytes())
        for k in SOLID_FIELDS:
            if case.ns:
                f.write(np.ascontiguousarray(case.solids[k], np.float64).tobytes())
        f.write(np.array([time, displaytime], np.float64).tobytes())


def read_case(path) -> tuple[dict, dict, dict, float, float]:
    """(params, parts, solids, time, displaytime) of a case file."""
    Params = _params_struct()
    raw = pathlib.Path(path).read_bytes()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    magic, version = take(np.int32, 2)
    if magic != MAGIC or version != VERSION:
        raise ValueError(f"{path}: not an ISPH case file (version {VERSION})")
    P = Params.from_buffer_copy(raw[off:off + C.sizeof(Params)])
    off += C.sizeof(Params)
    params = {name: getattr(P, name) for name, _ in Params._fields_}
    n, ns = (int(v) for v in take(np.int64, 2))
    parts = {k: take(np.float64, n) for k in PART_FIELDS}
    parts["tid"] = take(np.int32, n)
    parts["flags"] = take(np.uint8, n)
    solids = {k: take(np.float64, ns) for k in SOLID_FIELDS}
    time, displaytime = take(np.float64, 2)
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    return params, parts, solids, float(time), float(displaytime)


def read_state(path) -> dict:
    """Final particle state written by the driver (cell-sorted order)."""
    raw = pathlib.Path(path).read_bytes()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    magic, version = take(np.int32, 2)
    if magic != MAGIC or version != VERSION:
        raise ValueError(f"{path}: not an ISPH state file (version {VERSION})")
    n = int(take(np.int64, 1)[0])
    out = {k: take(np.float64, n) for k in PART_FIELDS}
    out["tid"] = take(np.int32, n)
    out["flags"] = take(np.uint8, n)
    out["t"], out["dt"] = (float(v) for v in take(np.float64, 2))
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    return out
