"""ctypes wrapper of the CPU oracle (oracle/isph_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, to check (or time as a baseline) the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
NBR_CAP = 64
RCAP = NBR_CAP + 2


class IsphParams(C.Structure):
    _fields_ = [
        ("cellx", C.c_int32), ("celly", C.c_int32), ("fixed_dt", C.c_int32),
        ("precision", C.c_int32), ("max_iter", C.c_int32), ("max_outer", C.c_int32),
        ("h", C.c_double), ("dl", C.c_double), ("lam", C.c_double), ("rho", C.c_double),
        ("rel_den", C.c_double), ("rhomin", C.c_double), ("g", C.c_double),
        ("mu", C.c_double), ("por", C.c_double), ("d50", C.c_double), ("lamk", C.c_double),
        ("csh", C.c_double), ("maxshift", C.c_double), ("fs_thresh", C.c_double),
        ("sig1", C.c_double), ("tschmidt", C.c_double), ("xlcutoff", C.c_double),
        ("ytcutoff", C.c_double), ("buf_xmin", C.c_double), ("buf_xmax", C.c_double),
        ("dt", C.c_double), ("tol", C.c_double), ("inner_tol", C.c_double),
    ]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer", C.c_int32), ("converged", C.c_int32),
                ("breakdown", C.c_int32), ("rel_residual", C.c_double), ("bnorm", C.c_double),
                ("nnz", C.c_double)]


class Scalars(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("dtsol", C.c_double),
                ("umax", C.c_double), ("numax", C.c_double), ("solsteps", C.c_int32),
                ("n", C.c_int32)]


def make_params(d: dict) -> IsphParams:
    p = IsphParams()
    for name, _ in IsphParams._fields_:
        setattr(p, name, d[name])
    return p


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_libs: dict = {}


def lib(path=None):
    """The oracle library (the default build, or the one at `path`)."""
    path = str(path or LIB_PATH)
    if path not in _libs:
        if path == str(LIB_PATH) and not LIB_PATH.exists():
            build()
        L = C.CDLL(path)
        vp, i64, d = C.c_void_p, C.c_int64, C.c_double
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(IsphParams), i64, i64]
        L.or_destroy.argtypes = [vp]
        L.or_n.restype = i64
        L.or_n.argtypes = [vp]
        L.or_ns.restype = i64
        L.or_ns.argtypes = [vp]
        L.or_set_params.argtypes = [vp, C.POINTER(IsphParams)]
        L.or_upload.argtypes = [vp, i64] + [vp] * 12
        L.or_upload_solids.argtypes = [vp, i64] + [vp] * 4
        L.or_array.restype = vp
        L.or_array.argtypes = [vp, C.c_char_p]
        for f in ("or_projection", "or_cellshift", "or_cellshiftalt", "or_effpor",
                  "or_eddyvis", "or_int_vel", "or_comp_vel", "or_comp_pos", "or_freesurf",
                  "or_opt2_shift", "or_densityupdate", "or_timestep", "or_scalart",
                  "or_presstoden", "or_dentopress", "or_calmasscons", "or_updateden"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = None
        for f in ("or_neighbour", "or_ppe_assemble", "or_init_state"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = C.c_int
        L.or_ppe_solve.argtypes = [vp, C.POINTER(SolveStats)]
        L.or_step.argtypes = [vp, C.c_int, C.POINTER(SolveStats)]
        L.or_scalarupdate.argtypes = [vp, d]
        L.or_get_scalars.argtypes = [vp, C.POINTER(Scalars)]
        L.or_set_scalars.argtypes = [vp, C.POINTER(Scalars)]
        L.or_wab.restype = d
        L.or_wab.argtypes = [d, d]
        L.or_wo.restype = d
        L.or_wo.argtypes = [d]
        L.or_wabd.restype = d
        L.or_wabd.argtypes = [d, d, d]
        L.or_vis.restype = d
        L.or_vis.argtypes = [d]
        L.or_inv2.argtypes = [vp]
        L.or_reduce.restype = d
        L.or_reduce.argtypes = [i64, vp]
        L.or_threads.restype = C.c_int
        L.or_solve_system.argtypes = [vp, i64, vp, vp, vp, vp, vp, C.POINTER(SolveStats)]
        L.or_set_buffer.argtypes = [vp, C.c_int, i64] + [vp] * 5 + [d] * 7
        L.or_buffer_update.argtypes = [vp, C.c_int]
        L.or_set_wcsph.argtypes = [vp, d, d, d, d, C.c_int]
        L.or_wavemaker_motion.argtypes = [d, d, d, d, d, C.POINTER(d), C.POINTER(d)]
        L.or_set_wavemaker.argtypes = [vp, d, d, d, d]
        L.or_wavemaker.argtypes = [vp]
        _libs[path] = L
    return _libs[path]


_DTYPES = {
    "tid": (np.int32, 1), "cell": (np.int32, 1), "flags": (np.uint8, 1),
    "ncnt": (np.int32, 1), "nbr": (np.int32, NBR_CAP), "nr": (np.float64, NBR_CAP),
    "klt1": (np.float64, NBR_CAP), "klt2": (np.float64, NBR_CAP), "coff": (np.float64, 4),
    "alen": (np.int32, 1), "acol": (np.int32, RCAP), "aval": (np.float64, RCAP),
    "cstart": (np.int32, 1),
}

PART_FIELDS = ("x", "y", "u", "v", "p", "m", "rho", "rho0", "vol0", "c")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One scenario held by the CPU oracle."""

    def __init__(self, case, cap: int | None = None, params: dict | None = None,
                 library=None):
        self.L = L = library if library is not None else lib()  # noqa
        self.case = case
        self.pdict = dict(case.params if params is None else params)
        self.P = make_params(self.pdict)
        self.cap = cap or case.n
        self.s = L.or_create(C.byref(self.P), self.cap, max(case.ns, 1))
        self._keep = []
        parts = {k: np.ascontiguousarray(case.parts[k], dtype=np.float64) for k in PART_FIELDS}
        parts["tid"] = np.ascontiguousarray(case.parts["tid"], dtype=np.int32)
        parts["flags"] = np.ascontiguousarray(case.parts["flags"], dtype=np.uint8)
        self._keep = parts
        args = [_ptr(parts[k]) for k in PART_FIELDS] + [_ptr(parts["tid"]), _ptr(parts["flags"])]
        rc = L.or_upload(self.s, case.n, *args)
        if rc != 0:
            raise RuntimeError(f"or_upload failed: {rc}")
        if case.ns:
            sol = {k: np.ascontiguousarray(v) for k, v in case.solids.items()}
            rc = L.or_upload_solids(self.s, case.ns, _ptr(sol["x"]), _ptr(sol["y"]),
                                    _ptr(sol["m"]), _ptr(sol["rho"]))
            if rc != 0:
                raise RuntimeError(f"or_upload_solids failed: {rc}")
        wave = case.meta.get("wavemaker")
        if wave:
            rc = L.or_set_wavemaker(self.s, wave["stroke"], wave["period"], wave["ramp"],
                                    wave["t0"])
            if rc != 0:
                raise RuntimeError(f"or_set_wavemaker failed: {rc}")

    def __del__(self):
        try:
            if getattr(self, "s", None):
                self.L.or_destroy(self.s)
                self.s = None
        except Exception:
            pass

    @property
    def n(self) -> int:
        return int(self.L.or_n(self.s))

    def arr(self, name: str, count: int | None = None) -> np.ndarray:
        """Copy of an internal array over the live particles."""
        dt, w = _DTYPES.get(name, (np.float64, 1))
        cnt = self.n if count is None else count
        ptr = self.L.or_array(self.s, name.encode())
        if not ptr:
            raise KeyError(name)
        buf = (C.c_char * (cnt * w * np.dtype(dt).itemsize)).from_address(ptr)
        a = np.frombuffer(buf, dtype=dt).copy()
        return a.reshape(cnt, w) if w > 1 else a

    def put(self, name: str, values) -> None:
        """Overwrite an internal per-particle array (live particles) in place."""
        dt, w = _DTYPES.get(name, (np.float64, 1))
        ptr = self.L.or_array(self.s, name.encode())
        if not ptr:
            raise KeyError(name)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=dt),
                                                 (self.n, w) if w > 1 else (self.n,)))
        C.memmove(ptr, a.ctypes.data, a.nbytes)

    def cstart(self) -> np.ndarray:
        ncell = self.P.cellx * self.P.celly
        ptr = self.L.or_array(self.s, b"cstart")
        buf = (C.c_int32 * (ncell + 1)).from_address(ptr)
        return np.frombuffer(buf, dtype=np.int32).copy()

    def scalars(self) -> Scalars:
        s = Scalars()
        self.L.or_get_scalars(self.s, C.byref(s))
        return s

    def set_scalars(self, s: Scalars) -> None:
        self.L.or_set_scalars(self.s, C.byref(s))

    def neighbours(self):
        cnt = self.arr("ncnt")
        nb = self.arr("nbr")
        return [nb[i, :cnt[i]].copy() for i in range(self.n)]

    def call(self, name: str, *args):
        return getattr(self.L, "or_" + name)(self.s, *args)

    def ppe_solve(self):
        st = SolveStats()
        rc = self.L.or_ppe_solve(self.s, C.byref(st))
        return rc, st

    def set_buffer(self, zone: int, x, y, u, v, p, mass: float, rho: float, vol0: float,
                   c: float, rv: float, interval: float = 0.0, time_shift: float = 0.0) -> int:
        t = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, u, v, p)]
        return self.L.or_set_buffer(self.s, zone, t[0].size, *[_ptr(a) for a in t], mass, rho,
                                   vol0, c, rv, interval, time_shift)

    def buffer_update(self, zone: int) -> int:
        return self.L.or_buffer_update(self.s, zone)

    def set_wavemaker(self, stroke: float, period: float, ramp: float = 0.0,
                      t0: float = 0.0) -> int:
        return self.L.or_set_wavemaker(self.s, stroke, period, ramp, t0)

    def step(self, nsteps: int = 1):
        st = SolveStats()
        rc = self.L.or_step(self.s, nsteps, C.byref(st))
        return rc, st

    def state(self) -> dict:
        out = {k: self.arr(k) for k in PART_FIELDS + ("tid", "flags")}
        return out


def build_variant(out_dir, mutant: int) -> pathlib.Path:
    """Compile the oracle with planted mistake `mutant` (-DOR_MUTANT=k, see the
    MUT(k) marks in isph_oracle.c) into out_dir; returns the library path.
    Pin tests use it to show that a check fails on a plausible mistake."""
    out = pathlib.Path(out_dir) / f"liboracle_m{mutant}.so"
    cc = os.environ.get("CC", "gcc")
    cmd = [cc, "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-std=c11",
           f"-DOR_MUTANT={int(mutant)}", "-shared", "-o", str(out), str(HERE / "isph_oracle.c"),
           "-lm"]
    subprocess.run(cmd, check=True)
    return out


def threads() -> int:
    return int(lib().or_threads())


def wavemaker_motion(stroke: float, period: float, ramp: float, t0: float, t: float):
    """(X, U) of the oracle's paddle motion at time t."""
    x, u = C.c_double(), C.c_double()
    rc = lib().or_wavemaker_motion(stroke, period, ramp, t0, t, C.byref(x), C.byref(u))
    if rc != 0:
        raise ValueError(f"or_wavemaker_motion: {rc}")
    return x.value, u.value
