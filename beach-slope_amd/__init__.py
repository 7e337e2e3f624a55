"""MI355X-native ISPH per-step pipeline: Python host side.

The reference drives its per-step pipeline through argument-less Fortran
subroutines acting on module state (main.f90:199-261).  `Scenario` mirrors that
interface one method per subroutine -- projection, cellshift, resetid,
neighbour, effpor, eddyvis, int_vel, ppesolve, comp_vel, comp_pos,
cellshiftalt, freesurf, opt2_shift, densityupdate, timestep,
scalart_anisotropic, scalarupdate(stime) -- over the C ABI of
`libisph_hip.so` (include/isph_abi.h).  State is device resident; downloads
copy it back.

The library is built by `build.py` next to this file; importing this package
without it raises an error: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libisph_hip.so"

NBR_CAP = 64
ROW_CAP = NBR_CAP + 2
NOALLOC_MAX_ROWS = 1 << 20   # ISPH_NOALLOC_MAX_ROWS
# kernels the autotuner sizes (ISPH_TUNE_N, in isph_autotune's order)
TUNED_KERNELS = ("neighbour", "effpor", "eddyvis", "int_vel_varts", "int_vel",
                 "ppe_assemble", "comp_vel", "opt2_shift", "scalart")
# launch shapes (isph_autotune): one thread per particle in blocks of 64, 128
# or 256 threads; for the neighbour search also one workgroup per tile of 4
# or 8 cells of a cell column, its neighbourhood staged in LDS
TUNE_SHAPES = ("p64", "p128", "p256", "tile4", "tile8")
PPE_FP64, PPE_FP32_REFINE = 0, 1

ERRORS = {-1: "HIP error", -2: "bad argument", -3: "capacity overflow",
          -4: "non-finite", -5: "solver not converged", -6: "call out of order"}


class IsphError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = ""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({code}) {msg}".strip())


class Params(C.Structure):
    """`isph_params` of include/isph_abi.h."""
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

    @classmethod
    def from_dict(cls, d: dict) -> "Params":
        p = cls()
        for name, _ in cls._fields_:
            setattr(p, name, d[name])
        return p


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer", C.c_int32), ("converged", C.c_int32),
                ("breakdown", C.c_int32), ("rel_residual", C.c_double), ("bnorm", C.c_double),
                ("nnz", C.c_double)]


class Scalars(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("dtsol", C.c_double),
                ("umax", C.c_double), ("numax", C.c_double), ("solsteps", C.c_int32),
                ("n", C.c_int32)]


# entry points taking only the context (one per reference subroutine)
PHASES = ("projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis", "int_vel",
          "comp_vel", "comp_pos", "cellshiftalt", "freesurf", "opt2_shift",
          "densityupdate", "timestep", "scalart_anisotropic", "ppe_assemble", "init_state",
          "presstoden", "dentopress", "calmasscons", "updateden")

EXPORTS = ("isph_create", "isph_destroy", "isph_last_error", "isph_set_params",
           "isph_upload", "isph_upload_solids", "isph_download", "isph_get_scalars",
           "isph_set_scalars", "isph_ppesolve", "isph_ppe_solve", "isph_scalarupdate",
           "isph_step", "isph_dump_cells", "isph_dump_neighbours", "isph_dump_props",
           "isph_dump_ppe", "isph_ppe_set_rhs", "isph_ppe_get_solution",
           "isph_load_system", "isph_batch_create", "isph_batch_step", "isph_batch_stats",
           "isph_batch_destroy", "isph_set_buffer", "isph_buffer_update", "isph_set_wcsph",
           "isph_footprint_bytes", "isph_prof_enable", "isph_prof_read",
           "isph_wavemaker_motion", "isph_set_wavemaker", "isph_wavemaker",
           "isph_ppe_solve_noalloc", "isph_batch_launches", "isph_batch_count",
           "isph_batch_set_graph", "isph_autotune", "isph_batch_autotune", "isph_set_shape",
           "isph_batch_set_shape") + tuple("isph_" + p for p in PHASES)

_lib = None


def library_path() -> pathlib.Path:
    return LIB_PATH


def lib():
    """Load libisph_hip.so (built by build.py); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run beach-slope_amd/build.py (or "
                          "__graft_entry__.build()) first; there is no CPU fallback")
    L = C.CDLL(str(LIB_PATH))
    vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int32, C.c_double
    L.isph_create.argtypes = [C.POINTER(Params), i64, i64, C.POINTER(vp)]
    L.isph_destroy.argtypes = [vp]
    L.isph_last_error.restype = C.c_char_p
    L.isph_last_error.argtypes = [vp]
    L.isph_set_params.argtypes = [vp, C.POINTER(Params)]
    L.isph_upload.argtypes = [vp, i64] + [vp] * 12
    L.isph_upload_solids.argtypes = [vp, i64] + [vp] * 4
    L.isph_download.argtypes = [vp, C.POINTER(i64)] + [vp] * 12
    L.isph_get_scalars.argtypes = [vp, C.POINTER(Scalars)]
    L.isph_set_scalars.argtypes = [vp, C.POINTER(Scalars)]
    for p in PHASES:
        getattr(L, "isph_" + p).argtypes = [vp]
    L.isph_ppesolve.argtypes = [vp, C.POINTER(SolveStats)]
    L.isph_ppe_solve.argtypes = [vp, C.POINTER(SolveStats)]
    L.isph_scalarupdate.argtypes = [vp, d]
    L.isph_step.argtypes = [vp, i32, C.POINTER(SolveStats)]
    L.isph_dump_cells.argtypes = [vp, vp, vp]
    L.isph_dump_neighbours.argtypes = [vp, vp, vp]
    L.isph_dump_props.argtypes = [vp] + [vp] * 13
    L.isph_dump_ppe.argtypes = [vp, vp, vp, vp, vp]
    L.isph_ppe_set_rhs.argtypes = [vp, vp, vp]
    L.isph_ppe_get_solution.argtypes = [vp, vp]
    L.isph_load_system.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.isph_batch_create.argtypes = [i32, C.POINTER(vp), C.POINTER(vp)]
    L.isph_batch_step.argtypes = [vp, i32]
    L.isph_batch_stats.argtypes = [vp, i32, C.POINTER(SolveStats)]
    L.isph_batch_destroy.argtypes = [vp]
    L.isph_batch_launches.restype = i64
    L.isph_batch_launches.argtypes = [vp]
    L.isph_batch_count.argtypes = [vp, i32, C.POINTER(i64)]
    L.isph_batch_set_graph.argtypes = [vp, i32]
    L.isph_autotune.argtypes = [vp, i32, vp, vp]
    L.isph_batch_autotune.argtypes = [vp, i32, vp, vp]
    L.isph_set_shape.argtypes = [vp, i32, i32]
    L.isph_batch_set_shape.argtypes = [vp, i32, i32]
    L.isph_set_buffer.argtypes = [vp, i32, i64] + [vp] * 5 + [d] * 7
    L.isph_buffer_update.argtypes = [vp, i32]
    L.isph_set_wcsph.argtypes = [vp, d, d, d, d, i32]
    L.isph_wavemaker_motion.argtypes = [d, d, d, d, d, C.POINTER(d), C.POINTER(d)]
    L.isph_set_wavemaker.argtypes = [vp, d, d, d, d]
    L.isph_wavemaker.argtypes = [vp]
    L.isph_prof_enable.argtypes = [vp, i32]
    L.isph_prof_read.argtypes = [vp, C.c_char_p, C.POINTER(d), C.POINTER(i64)]
    L.isph_footprint_bytes.restype = i64
    L.isph_footprint_bytes.argtypes = [i64, i64, i32, i32, i32]
    L.isph_ppe_solve_noalloc.argtypes = [i32, i32] + [vp] * 5 + [d, i32, i32, vp,
                                                                C.POINTER(SolveStats)]
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


PART_FIELDS = ("x", "y", "u", "v", "p", "m", "rho", "rho0", "vol0", "c")
PROP_FIELDS = ("porosity", "resist", "gradvx", "coff", "nut", "tke", "varts", "cdiff", "vel",
               "us", "vs", "xs", "ys")


class Scenario:
    """One device-resident scenario behind the C ABI."""

    def __init__(self, case, params: dict | None = None, cap: int | None = None):
        L = lib()
        self.case = case
        self.params = dict(case.params if params is None else params)
        self._P = Params.from_dict(self.params)
        self.cap = int(cap or case.n)
        self._ctx = C.c_void_p()
        self._check(L.isph_create(C.byref(self._P), self.cap, max(case.ns, 1),
                                  C.byref(self._ctx)), "isph_create", ctx=False)
        parts = {k: np.ascontiguousarray(case.parts[k], dtype=np.float64) for k in PART_FIELDS}
        parts["tid"] = np.ascontiguousarray(case.parts["tid"], dtype=np.int32)
        parts["flags"] = np.ascontiguousarray(case.parts["flags"], dtype=np.uint8)
        self._keep = parts  # the arrays must outlive the call that reads them
        args = [_ptr(parts[k]) for k in PART_FIELDS] + [_ptr(parts["tid"]), _ptr(parts["flags"])]
        self._check(L.isph_upload(self._ctx, case.n, *args), "isph_upload")
        if case.ns:
            s = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in case.solids.items()}
            self._check(L.isph_upload_solids(self._ctx, case.ns, _ptr(s["x"]), _ptr(s["y"]),
                                             _ptr(s["m"]), _ptr(s["rho"])),
                        "isph_upload_solids")
        wave = case.meta.get("wavemaker")
        if wave:
            self.set_wavemaker(wave["stroke"], wave["period"], wave["ramp"], wave["t0"])

    # -- plumbing --------------------------------------------------------
    def _check(self, rc: int, where: str, ctx: bool = True):
        if rc != 0:
            # without a context (failed isph_create) the library keeps the
            # message of the calling thread's last create
            handle = self._ctx if (ctx and self._ctx) else None
            msg = (lib().isph_last_error(handle) or b"").decode()
            raise IsphError(rc, where, msg)

    def close(self):
        if getattr(self, "_ctx", None):
            lib().isph_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._ctx

    def last_error(self) -> str:
        if not self._ctx:
            return ""
        return (lib().isph_last_error(self._ctx) or b"").decode()

    def set_params(self, **kw):
        params = dict(self.params, **kw)
        P = Params.from_dict(params)
        self._check(lib().isph_set_params(self._ctx, C.byref(P)), "isph_set_params")
        self.params, self._P = params, P

    # -- one method per reference subroutine ------------------------------
    def _phase(self, name):
        self._check(getattr(lib(), "isph_" + name)(self._ctx), "isph_" + name)

    def projection(self): self._phase("projection")
    def cellshift(self): self._phase("cellshift")
    def resetid(self): self._phase("resetid")
    def neighbour(self): self._phase("neighbour")
    def effpor(self): self._phase("effpor")
    def eddyvis(self): self._phase("eddyvis")
    def int_vel(self): self._phase("int_vel")
    def comp_vel(self): self._phase("comp_vel")
    def comp_pos(self): self._phase("comp_pos")
    def cellshiftalt(self): self._phase("cellshiftalt")
    def freesurf(self): self._phase("freesurf")
    def opt2_shift(self): self._phase("opt2_shift")
    def densityupdate(self): self._phase("densityupdate")
    def timestep(self): self._phase("timestep")
    def scalart_anisotropic(self): self._phase("scalart_anisotropic")
    def ppe_assemble(self): self._phase("ppe_assemble")
    def init_state(self): self._phase("init_state")

    def scalarupdate(self, stime: float):
        self._check(lib().isph_scalarupdate(self._ctx, float(stime)), "isph_scalarupdate")

    def ppe_solve(self, check: bool = True) -> SolveStats:
        st = SolveStats()
        rc = lib().isph_ppe_solve(self._ctx, C.byref(st))
        if check:
            self._check(rc, "isph_ppe_solve")
        return st

    def ppesolve(self) -> SolveStats:
        st = SolveStats()
        self._check(lib().isph_ppesolve(self._ctx, C.byref(st)), "isph_ppesolve")
        return st

    def step(self, nsteps: int = 1, check: bool = True):
        st = SolveStats()
        rc = lib().isph_step(self._ctx, int(nsteps), C.byref(st))
        if check:
            self._check(rc, "isph_step")
        return rc, st

    # -- state -------------------------------------------------------------
    def scalars(self) -> Scalars:
        s = Scalars()
        self._check(lib().isph_get_scalars(self._ctx, C.byref(s)), "isph_get_scalars")
        return s

    def set_scalars(self, s: Scalars):
        self._check(lib().isph_set_scalars(self._ctx, C.byref(s)), "isph_set_scalars")

    @property
    def n(self) -> int:
        return int(self.scalars().n)

    def download(self) -> dict:
        n = self.n
        out = {k: np.empty(n) for k in PART_FIELDS}
        out["tid"] = np.empty(n, np.int32)
        out["flags"] = np.empty(n, np.uint8)
        nn = C.c_int64()
        args = [_ptr(out[k]) for k in PART_FIELDS] + [_ptr(out["tid"]), _ptr(out["flags"])]
        self._check(lib().isph_download(self._ctx, C.byref(nn), *args), "isph_download")
        return out

    def cells(self):
        n = self.n
        ncell = self.params["cellx"] * self.params["celly"]
        cell = np.empty(n, np.int32)
        cs = np.empty(ncell + 1, np.int32)
        self._check(lib().isph_dump_cells(self._ctx, _ptr(cell), _ptr(cs)), "isph_dump_cells")
        return cell, cs

    def neighbours(self):
        n = self.n
        cnt = np.empty(n, np.int32)
        lst = np.empty((n, NBR_CAP), np.int32)
        self._check(lib().isph_dump_neighbours(self._ctx, _ptr(cnt), _ptr(lst)),
                    "isph_dump_neighbours")
        return cnt, lst

    def props(self) -> dict:
        n = self.n
        out = {k: np.empty(n) for k in PROP_FIELDS}
        out["coff"] = np.empty((n, 4))
        self._check(lib().isph_dump_props(self._ctx, *[_ptr(out[k]) for k in PROP_FIELDS]),
                    "isph_dump_props")
        return out

    def ppe(self):
        n = self.n
        ln = np.empty(n, np.int32)
        col = np.empty((n, ROW_CAP), np.int32)
        val = np.empty((n, ROW_CAP))
        rhs = np.empty(n)
        self._check(lib().isph_dump_ppe(self._ctx, _ptr(ln), _ptr(col), _ptr(val), _ptr(rhs)),
                    "isph_dump_ppe")
        return ln, col, val, rhs

    def set_buffer(self, zone: int, x, y, u, v, p, mass: float, rho: float, vol0: float,
                   c: float, rv: float, interval: float = 0.0, time_shift: float = 0.0):
        """Buffer zone 1 or 2 (isph_set_buffer): template particles and the
        values inserted particles take, removal coordinate, insertion clock."""
        t = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, u, v, p)]
        self._keep_buf = t
        self._check(lib().isph_set_buffer(self._ctx, zone, t[0].size, *[_ptr(a) for a in t],
                                          mass, rho, vol0, c, rv, interval, time_shift),
                    "isph_set_buffer")

    def set_wcsph(self, c0: float, gamma: float = 7.0, delta: float = 0.1, swl: float = 0.0,
                  in_step: bool = True):
        """Weakly compressible form (isph_set_wcsph)."""
        self._check(lib().isph_set_wcsph(self._ctx, c0, gamma, delta, swl, int(in_step)),
                    "isph_set_wcsph")

    def set_wavemaker(self, stroke: float, period: float, ramp: float = 0.0, t0: float = 0.0):
        """Piston wavemaker (isph_set_wavemaker): the ISPH_F_PADDLE walls move by
        X(t) from their current positions, which are taken as rest + X(now)."""
        self._check(lib().isph_set_wavemaker(self._ctx, stroke, period, ramp, t0),
                    "isph_set_wavemaker")

    def wavemaker(self):
        """Move the paddle to X, U at the context's current time (isph_wavemaker)."""
        self._check(lib().isph_wavemaker(self._ctx), "isph_wavemaker")

    def presstoden(self): self._phase("presstoden")
    def dentopress(self): self._phase("dentopress")
    def calmasscons(self): self._phase("calmasscons")
    def updateden(self): self._phase("updateden")

    def buffer_update(self, zone: int):
        """remove_buffer then insert_buffer of the zone (isph_buffer_update)."""
        self._check(lib().isph_buffer_update(self._ctx, zone), "isph_buffer_update")

    def load_system(self, length, col, val, rhs, x0):
        length = np.ascontiguousarray(length, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        rhs = np.ascontiguousarray(rhs, np.float64)
        x0 = np.ascontiguousarray(x0, np.float64)
        self._check(lib().isph_load_system(self._ctx, len(rhs), _ptr(length), _ptr(col),
                                           _ptr(val), _ptr(rhs), _ptr(x0)), "isph_load_system")

    def autotune(self, trials: int = 5) -> dict:
        """Choose each particle kernel's launch shape by timing its shapes on
        the current state (isph_autotune; call at a step boundary).
        Returns {kernel: (chosen shape, {shape: mean us per launch})}."""
        return _tune(lib().isph_autotune, self._ctx, trials, self._check, "isph_autotune")

    def set_shape(self, kernel: str, shape: str):
        """Launch `kernel` (one of TUNED_KERNELS) in `shape` (one of
        TUNE_SHAPES) from now on (isph_set_shape)."""
        self._check(lib().isph_set_shape(self._ctx, TUNED_KERNELS.index(kernel),
                                         TUNE_SHAPES.index(shape)), "isph_set_shape")

    def prof_enable(self, on: bool = True):
        self._check(lib().isph_prof_enable(self._ctx, int(on)), "isph_prof_enable")

    def prof_read(self, kernel: str):
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(lib().isph_prof_read(self._ctx, kernel.encode(), C.byref(ms), C.byref(cnt)),
                    "isph_prof_read")
        return ms.value, cnt.value

    def solution(self) -> np.ndarray:
        x = np.empty(self.n)
        self._check(lib().isph_ppe_get_solution(self._ctx, _ptr(x)), "isph_ppe_get_solution")
        return x


class Batch:
    """Independent scenarios stepped together (isph_batch_*).

    While the batch exists its scenarios share one stream, and the PPE solves
    of the precision-form scenarios run as one set of launches over all of
    them.  Close the batch before closing its scenarios."""

    def __init__(self, scenarios):
        self.scenarios = list(scenarios)
        handles = (C.c_void_p * len(self.scenarios))(*[s.handle for s in self.scenarios])
        self._b = C.c_void_p()
        rc = lib().isph_batch_create(len(self.scenarios), handles, C.byref(self._b))
        if rc != 0:
            raise IsphError(rc, "isph_batch_create", self.scenarios[0].last_error())

    def step(self, nsteps: int = 1, check: bool = True) -> int:
        rc = lib().isph_batch_step(self._b, nsteps)
        if check and rc != 0:
            msgs = "; ".join(f"{k}: {s.last_error()}" for k, s in enumerate(self.scenarios)
                             if s.last_error())
            raise IsphError(rc, "isph_batch_step", msgs)
        return rc

    def set_shape(self, kernel: str, shape: str):
        """isph_batch_set_shape: the launch shape of a batched particle phase."""
        rc = lib().isph_batch_set_shape(self._b, TUNED_KERNELS.index(kernel),
                                        TUNE_SHAPES.index(shape))
        if rc != 0:
            raise IsphError(rc, "isph_batch_set_shape", self.scenarios[0].last_error())

    def autotune(self, trials: int = 3) -> dict:
        """isph_batch_autotune: launch shapes of the batched particle phases."""
        def check(rc, where):
            if rc != 0:
                raise IsphError(rc, where, self.scenarios[0].last_error())
        return _tune(lib().isph_batch_autotune, self._b, trials, check, "isph_batch_autotune")

    def set_graph(self, on: bool = True):
        """Replay the batched phase sequences as HIP graphs (isph_batch_set_graph)."""
        rc = lib().isph_batch_set_graph(self._b, int(on))
        if rc != 0:
            raise IsphError(rc, "isph_batch_set_graph", self.scenarios[0].last_error())

    def launches(self) -> int:
        """Kernel launches of the last batched step (0: phases per scenario)."""
        return int(lib().isph_batch_launches(self._b))

    def count(self, s: int) -> int:
        """Live particles of scenario s after the last batched step."""
        n = C.c_int64()
        rc = lib().isph_batch_count(self._b, s, C.byref(n))
        if rc != 0:
            raise IsphError(rc, "isph_batch_count")
        return int(n.value)

    def stats(self, s: int) -> SolveStats:
        st = SolveStats()
        rc = lib().isph_batch_stats(self._b, s, C.byref(st))
        if rc != 0:
            raise IsphError(rc, "isph_batch_stats")
        return st

    def close(self):
        if getattr(self, "_b", None):
            lib().isph_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tune(fn, handle, trials, check, where) -> dict:
    shape = np.zeros(len(TUNED_KERNELS), np.int32)
    nc = len(TUNE_SHAPES)
    us = np.zeros(nc * len(TUNED_KERNELS))
    check(fn(handle, int(trials), _ptr(shape), _ptr(us)), where)
    return {k: (TUNE_SHAPES[shape[q]],
                {t: v for t, v in zip(TUNE_SHAPES, us[nc * q:nc * q + nc].tolist()) if v >= 0})
            for q, k in enumerate(TUNED_KERNELS)}


def wavemaker_motion(stroke: float, period: float, ramp: float, t0: float, t: float):
    """(X, U) of the paddle motion at time t (isph_wavemaker_motion; host code,
    no device needed)."""
    x, u = C.c_double(), C.c_double()
    rc = lib().isph_wavemaker_motion(stroke, period, ramp, t0, t, C.byref(x), C.byref(u))
    if rc != 0:
        raise IsphError(rc, "isph_wavemaker_motion")
    return x.value, u.value


def footprint_bytes(n_cap: int, n_solid_cap: int, cellx: int, celly: int,
                    precision: int = 0) -> int:
    return int(lib().isph_footprint_bytes(n_cap, n_solid_cap, cellx, celly, precision))


def ppe_solve_noalloc(n: int, ld: int, row_len: int, col: int, val: int, rhs: int, x: int,
                      tol: float, max_iter: int, max_outer: int = 0, stream: int | None = None,
                      check: bool = True):
    """FP64 BiCGStab on a system already in device memory, with no allocation
    (isph_ppe_solve_noalloc).  row_len, col, val, rhs and x are device
    addresses (integers); col/val are ELL slot-major with leading dimension
    ld; x holds the initial guess and receives the solution.  Returns
    (return code, SolveStats)."""
    st = SolveStats()
    rc = lib().isph_ppe_solve_noalloc(int(n), int(ld), row_len, col, val, rhs, x, float(tol),
                                      int(max_iter), int(max_outer), stream, C.byref(st))
    if check and rc != 0:
        raise IsphError(rc, "isph_ppe_solve_noalloc",
                        (lib().isph_last_error(None) or b"").decode())
    return rc, st
