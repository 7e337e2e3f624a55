"""Synthetic case generators for the benchmark configurations.

The reference hard-codes one coastal-aquifer scenario in its generators
(boundary.f90:12-606, internal.f90:10-445, setup.f90:10-150).  The
configurations measured here are different scenarios built on the same
conventions:

* three wall layers at the fluid spacing, the innermost layer "fixed"
  (tid 1) and the outer two "ghost" (tid 2)            -- boundary.f90:47-86
* fluid particles on a regular lattice, mass rho*dx^2, volume dx^2,
  hydrostatic initial pressure                          -- internal.f90:37-91
* static porous-solid particles of bulk density 1800 whose kernel sum gives
  the porosity                                          -- boundary.f90:523-593
* the scalar is stored as the reference does, c_ref in [-0.5, 0.5]
  (main.f90:169-179, output.f90:203): a reported concentration c in {0, 1}
  maps to c_ref = c - 0.5, so densityupdate gives rho0 = 1000 + 25 c
  (scalar.f90:414, setup.f90:19).
* the cell grid has one halo ring; interior cell (ci, cj) covers
  x in [(cj-1)*4h, cj*4h), y in [(celly-2-ci)*4h, (celly-1-ci)*4h)
  (setup.f90:120-123), so every particle is placed at x, y >= 0.

Each generator returns a `Case`: the run parameters (the fields of the C
struct `isph_params`), the particle SoA arrays and the solid SoA arrays.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

TID_FIXED, TID_GHOST, TID_FLUID = 1, 2, 3
F_BUFFER, F_FREE, F_VICINITY, F_PADDLE = 1, 2, 4, 8

RHO = 1000.0
REL_DEN = 1.025
G = -9.81
BULKDEN = 1800.0


@dataclasses.dataclass
class Case:
    name: str
    params: dict
    parts: dict       # x y u v p m rho rho0 vol0 c (f64), tid (i32), flags (u8)
    solids: dict      # x y m rho (f64)
    steps: int
    meta: dict

    @property
    def n(self) -> int:
        return int(self.parts["x"].shape[0])

    @property
    def ns(self) -> int:
        return int(self.solids["x"].shape[0])


def default_params(dx: float, h_fac: float = 1.3) -> dict:
    """Run parameters with the reference's constants (initialize.f90:21-83,
    porous.f90:111, turbulence.f90:16) for a lattice spacing dx."""
    h = h_fac * dx
    return dict(
        cellx=0, celly=0, fixed_dt=1, precision=0, max_iter=20000, max_outer=20,
        h=h, dl=dx, lam=0.0010 * (h * h), rho=RHO, rel_den=REL_DEN, rhomin=1000.0,
        g=G, mu=0.0010, por=0.380, d50=0.001020, lamk=600.0, csh=0.50,
        maxshift=0.1, fs_thresh=0.80 * 2.0, sig1=0.20, tschmidt=1.0,
        xlcutoff=-1.0e30, ytcutoff=-1.0e30, buf_xmin=1.0e30, buf_xmax=-1.0e30,
        dt=1.0e-4, tol=1.0e-10, inner_tol=0.0,
    )


def _grid(params: dict, xmax: float, ymax: float) -> None:
    cs = 4.0 * params["h"]
    params["cellx"] = int(math.ceil(xmax / cs)) + 3
    params["celly"] = int(math.ceil(ymax / cs)) + 3


def _walls_box(x0: float, x1: float, y0: float, ytop: float, dx: float):
    """Three wall layers below [x0,x1] and outside x0 / x1 up to ytop.
    Returns (x, y, tid) with the innermost layer tid 1."""
    xs, ys, ts = [], [], []
    nx = int(round((x1 - x0) / dx))
    ny = int(round((ytop - y0) / dx))
    for k in range(3):
        t = TID_FIXED if k == 0 else TID_GHOST
        # bottom, including the corner columns of the side walls
        xb = x0 + (np.arange(-k - 1, nx + k + 1) + 0.5) * dx
        xs.append(xb)
        ys.append(np.full(xb.shape, y0 - (k + 0.5) * dx))
        ts.append(np.full(xb.shape, t))
        yv = y0 + (np.arange(-k, ny) + 0.5) * dx
        for xw in (x0 - (k + 0.5) * dx, x1 + (k + 0.5) * dx):
            xs.append(np.full(yv.shape, xw))
            ys.append(yv)
            ts.append(np.full(yv.shape, t))
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ts).astype(np.int32)


def wave_number(period: float, depth: float, g: float = abs(G)) -> float:
    """k of the linear dispersion relation w^2 = g k tanh(k d), by Newton's
    method (f(k) is increasing and concave, so the iteration from the
    shallow-water guess w / sqrt(g d), which lies above the root, falls
    monotonically onto it)."""
    w = 2.0 * math.pi / period
    k = w / math.sqrt(g * depth)
    for _ in range(100):
        th = math.tanh(k * depth)
        f = g * k * th - w * w
        df = g * th + g * k * depth * (1.0 - th * th)
        dk = f / df
        k -= dk
        if abs(dk) <= 1e-16 * k:
            break
    return k


def piston_stroke(wave_h: float, period: float, depth: float, g: float = abs(G)) -> float:
    """Piston-wavemaker stroke S for a target wave height H by linear
    wavemaker (Biesel) theory: H/S = 2 (cosh 2kd - 1) / (sinh 2kd + 2kd)
    (Dean & Dalrymple, Water Wave Mechanics for Engineers and Scientists,
    1991, sec. 6.2).  H/S tends to kd in shallow water and to 2 in deep."""
    kd2 = 2.0 * wave_number(period, depth, g) * depth
    return wave_h * (math.sinh(kd2) + kd2) / (2.0 * (math.cosh(kd2) - 1.0))


def _walls_paddle_flume(x0: float, x1: float, y0: float, ytop: float, dx: float, back: float):
    """Walls of a flume whose left end is a piston paddle: three bottom layers
    from x0 - back to the right wall, a three-layer right wall as in
    `_walls_box`, and a three-layer paddle at x0 - (k + 1/2) dx from y0 + dx/2
    up (above the bottom, so it slides over it) carrying F_PADDLE.  back
    covers the paddle's travel behind its rest position.
    Returns (x, y, tid, flags)."""
    xs, ys, ts, fs = [], [], [], []
    nback = int(math.ceil(back / dx))
    nx = int(round((x1 - x0) / dx))
    ny = int(round((ytop - y0) / dx))
    for k in range(3):
        t = TID_FIXED if k == 0 else TID_GHOST
        xb = x0 + (np.arange(-nback, nx + k + 1) + 0.5) * dx
        xs.append(xb)
        ys.append(np.full(xb.shape, y0 - (k + 0.5) * dx))
        ts.append(np.full(xb.shape, t))
        fs.append(np.zeros(xb.shape, np.uint8))
        yv = y0 + (np.arange(-k, ny) + 0.5) * dx
        xs.append(np.full(yv.shape, x1 + (k + 0.5) * dx))
        ys.append(yv)
        ts.append(np.full(yv.shape, t))
        fs.append(np.zeros(yv.shape, np.uint8))
        yp = y0 + (np.arange(0, ny) + 0.5) * dx
        xs.append(np.full(yp.shape, x0 - (k + 0.5) * dx))
        ys.append(yp)
        ts.append(np.full(yp.shape, t))
        fs.append(np.full(yp.shape, F_PADDLE, np.uint8))
    return (np.concatenate(xs), np.concatenate(ys), np.concatenate(ts).astype(np.int32),
            np.concatenate(fs))


def kernel_porosity(params: dict, solids: dict, px: np.ndarray, py: np.ndarray) -> np.ndarray:
    """Porosity at the points (px, py) from the solids' kernel sum, as effpor
    computes it for a fluid particle there (porous.f90:52-100, 182-188):
    n = 1 - min(sum over solids within 2h of (1 - por) W(r) m_s / rho_s, 1 - por)."""
    from scipy.spatial import cKDTree

    n = np.ones(px.size)
    if solids["x"].size == 0 or px.size == 0:
        return n
    h, por = params["h"], params["por"]
    pts = cKDTree(np.c_[px, py])
    sol = cKDTree(np.c_[solids["x"], solids["y"]])
    pairs = pts.sparse_distance_matrix(sol, 2 * h, output_type="coo_matrix")
    i, j, r = pairs.row, pairs.col, pairs.data
    keep = r < 2 * h
    i, j, r = i[keep], j[keep], r[keep]
    k1 = 10.0 / (float(np.float32(28.0) * np.float32(3.14)) * h * h)
    k2 = 1.0 / (float(np.float32(3.0) * np.float32(3.14)) * h * h)
    q = r / h
    a = 2.0 - q
    t1 = np.where(r < h, 4.0 - 6.0 * q * q + 3.0 * q ** 3, a ** 3)
    t2 = np.where(r < h, q ** 3 - 6.0 * q + 6.0, a ** 3)
    w = 0.1 * t1 * k1 + 0.9 * t2 * k2
    sf = np.bincount(i, weights=(1.0 - por) * w * solids["m"][j] / solids["rho"][j],
                     minlength=px.size)
    return 1.0 - np.minimum(sf, 1.0 - por)


def pore_volumes(c: Case) -> Case:
    """Give every fluid particle the bulk volume of its lattice site,
    m / rho = dx^2, where rho = n rho0 is the density effpor sets
    (porous.f90:188-190): volume vol0 = n dx^2 and mass m = rho0 vol0, with n
    the porosity at the particle.  A particle in the pores then represents
    the pore water of one lattice site, and the particle volume m / rho that
    every sum of the pipeline weights by is the lattice area on both sides
    of the beach surface and at the walls.  The reference reaches the same
    state differently: its pore water fills the medium on a lattice of
    spacing 2 pr / sqrt(por) with the open-water mass (internal.f90:37-39,
    setup.f90:71-72), so that m / rho = V0 / n is again the area per
    particle.  Without it a pore particle would carry the open-water volume
    V0 and m / rho = V0 / n would jump at the beach surface; comp_vel's
    pressure term n_i sum_j V_j (p_i + p_j) gradW~ / rho0 then holds a
    2 p_i sum_j V_j gradW~ that grows with depth, and at 0.3-0.4 m depth
    the flume diverges within a few steps."""
    fl = c.parts["tid"] == TID_FLUID
    dx = c.meta["dx"]
    n = kernel_porosity(c.params, c.solids, c.parts["x"][fl], c.parts["y"][fl])
    vol0 = c.parts["vol0"].copy()
    vol0[fl] = n * dx * dx
    rho = c.parts["rho"].copy()
    rho[fl] = n * c.parts["rho0"][fl]
    c.parts["vol0"] = vol0
    c.parts["m"] = c.parts["rho0"] * vol0
    c.parts["rho"] = rho
    c.meta["pore_volumes"] = True
    return c


def _assemble(params, name, fx, fy, fc, wx, wy, wt, dx, depth_top, sx=None, sy=None,
              steps=1, meta=None, wflags=None) -> Case:
    n_f, n_w = fx.size, wx.size
    x = np.concatenate([wx, fx])
    y = np.concatenate([wy, fy])
    tid = np.concatenate([wt, np.full(n_f, TID_FLUID, np.int32)]).astype(np.int32)
    c_ref = np.concatenate([np.full(n_w, -0.5), fc - 0.5])
    rhomax = params["rho"] * params["rel_den"]
    rho0 = 0.5 * (rhomax + params["rhomin"]) + c_ref * (rhomax - params["rhomin"])
    rho0[:n_w] = RHO
    vol0 = np.full(x.size, dx * dx)
    m = rho0 * vol0
    p = np.maximum(RHO * abs(G) * (depth_top - y), 0.0)
    flags = np.zeros(x.size, np.uint8)
    if wflags is not None:
        flags[:n_w] = wflags
    parts = dict(
        x=x, y=y, u=np.zeros_like(x), v=np.zeros_like(x), p=p, m=m, rho=rho0.copy(),
        rho0=rho0, vol0=vol0, c=c_ref, tid=tid, flags=flags,
    )
    if sx is None:
        sx = np.zeros(0)
        sy = np.zeros(0)
    vs = dx * dx
    solids = dict(x=sx.astype(np.float64), y=sy.astype(np.float64),
                  m=np.full(sx.size, BULKDEN * vs), rho=np.full(sx.size, BULKDEN))
    return Case(name=name, params=params, parts=parts, solids=solids, steps=steps,
                meta=meta or {})


def dam_break(dx: float = 0.005, jitter: float = 0.0, seed: int = 1234,
              tank=(2.0, 0.6), column=(0.5, 0.25), toe: float = 1.0,
              slope: float = 0.1, steps: int = 200, h_fac: float = 1.3) -> Case:
    """Dam break onto a 1:10 impermeable slope (FP64 configuration, and with a
    seeded uniform +-jitter*dx on fluid positions the precision-form one).
    Tank 2.0 x 0.6 m, water column 0.5 x 0.25 m, h = 1.3 dx (h_fac dx),
    dt = 1e-4 s."""
    params = default_params(dx, h_fac)
    params.update(dt=1.0e-4, tol=1.0e-10)
    off = 3.0 * dx + 0.5 * dx            # keeps the walls at x, y >= 0
    L, H = tank
    wx, wy, wt = _walls_box(off, off + L, off, off + H, dx)
    # 1:10 slope from the toe to the right wall: three layers under the line
    xs_ = off + toe + (np.arange(int(round((L - toe) / dx))) + 0.5) * dx
    sl_x, sl_y, sl_t = [], [], []
    for k in range(3):
        yk = off + (xs_ - off - toe) * slope - (k + 0.5) * dx
        keep = yk > off + 0.25 * dx
        sl_x.append(xs_[keep])
        sl_y.append(yk[keep])
        sl_t.append(np.full(int(keep.sum()), TID_FIXED if k == 0 else TID_GHOST, np.int32))
    wx = np.concatenate([wx] + sl_x)
    wy = np.concatenate([wy] + sl_y)
    wt = np.concatenate([wt] + sl_t)
    nx, ny = int(round(column[0] / dx)), int(round(column[1] / dx))
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    fx = off + (gx.ravel() + 0.5) * dx
    fy = off + (gy.ravel() + 0.5) * dx
    if jitter:
        rng = np.random.default_rng(seed)
        fx = fx + rng.uniform(-jitter * dx, jitter * dx, fx.size)
        fy = fy + rng.uniform(-jitter * dx, jitter * dx, fy.size)
    fc = np.zeros(fx.size)
    _grid(params, off + L + 4 * dx, off + H + 4 * dx)
    name = "dam_break_jitter" if jitter else "dam_break"
    return _assemble(params, name, fx, fy, fc, wx, wy, wt, dx, off + column[1],
                     steps=steps, meta=dict(dx=dx, jitter=jitter, seed=seed, off=off,
                                            toe=off + toe, slope=slope, tank=tank,
                                            column=column))


def porous_flume(dx: float = 0.0025, length: float = 10.0, depth: float = 0.4,
                 tank_h: float = 0.6, toe: float = 6.0, slope: float = 0.1,
                 porosity: float = 0.4, d50: float = 0.01,
                 patch=(0.2, 0.1), dt: float = 5.0e-5, steps: int = 2000,
                 wave_h: float = 0.05, wave_t: float = 1.5, grade_h: float = 0.0,
                 wavemaker: bool = True, ramp: float | None = None,
                 name: str = "porous_flume", pore_volume: bool = True) -> Case:
    """Wave-runup flume on a porous beach (the 650k-particle configuration at
    the defaults: 4000 x 160 fluid lattice, 3 wall layers, 128,000 beach
    solids).  The 1:10 beach under the line y = (x - toe)*slope is filled
    with static solid particles at the fluid spacing (bulk density 1800,
    boundary.f90:523-593); the fluid lattice runs through the beach, and
    `pore_volumes` gives each fluid particle the bulk volume of its lattice
    site (m / rho = dx^2 at the porosity effpor computes there).  A scalar
    patch of c = 1 (0.2 m x 0.1 m) sits at the toe.

    The left end is a piston wavemaker (with `wavemaker`): a three-layer
    paddle of F_PADDLE walls whose stroke gives waves of height wave_h and
    period wave_t by linear wavemaker theory (`piston_stroke`), started with a
    linear ramp over `ramp` seconds (one period by default).  The bottom
    extends behind the paddle's rest position by half the stroke plus 4 dx,
    and the flume sits that much further right, rounded up to an even number
    of cells (meta "off" is the paddle's rest x, "offy" the bottom's y).
    meta["wavemaker"] holds the motion (stroke, period, ramp, t0) that
    `Scenario` and the oracle's wrapper apply on upload.  Without it the
    left end is an ordinary three-layer wall.

    The length is rounded to a whole number of lattice spacings dx, so the
    right wall sits on the fluid lattice.

    The beach surface is sharp by default: the layer has porosity `porosity`
    right up to the slope line (the kernel sum smooths n over 2h).  With
    grade_h > 0 the solid volume instead ramps linearly from 0 at the beach
    surface to full at a depth of grade_h*h.

    pore_volume=False leaves every fluid particle the open-water volume dx^2
    (and mass rho0 dx^2), so m / rho = dx^2 / n jumps at the beach surface:
    at 0.3-0.4 m depth that flume diverges within a few steps."""
    params = default_params(dx)
    params.update(dt=dt, por=porosity, d50=d50, tol=1.0e-10)
    # a whole number of lattice spacings: the right wall then continues the
    # fluid and bottom lattice one spacing beyond the last fluid column (off
    # the lattice it would leave a gap of 0.5-1.5 dx and a misaligned bottom
    # corner, where the pressure solve goes unstable within tens of steps)
    length = int(round(length / dx)) * dx
    wflags = None
    if wavemaker:
        stroke = piston_stroke(wave_h, wave_t, depth)
        back = 0.5 * stroke + 4.0 * dx
        # the flume moves right by an even number of cells, so the lattice sits
        # on the cell grid, and cells pair into the first coarse multigrid
        # level, as without the paddle; it stays at the same height
        cs = 4.0 * params["h"]
        off = 3.0 * dx + 0.5 * dx + 2.0 * math.ceil(back / (2.0 * cs)) * cs
        offy = 3.0 * dx + 0.5 * dx
        wx, wy, wt, wflags = _walls_paddle_flume(off, off + length, offy, offy + tank_h, dx,
                                                 back)
    else:
        off = offy = 3.0 * dx + 0.5 * dx
        wx, wy, wt = _walls_box(off, off + length, off, off + tank_h, dx)
    nx, ny = int(round(length / dx)), int(round(depth / dx))
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    fx = off + (gx.ravel() + 0.5) * dx
    fy = offy + (gy.ravel() + 0.5) * dx
    lx = fx - off
    ly = fy - offy
    fc = ((np.abs(lx - toe) <= 0.5 * patch[0]) & (ly <= patch[1])).astype(np.float64)
    # porous beach solids
    sxi = np.arange(int(math.floor(toe / dx)), nx)
    sgx, sgy = np.meshgrid(sxi, np.arange(ny + 40), indexing="ij")
    sx = (sgx.ravel() + 0.5) * dx
    sy = (sgy.ravel() + 0.5) * dx
    below = (sx - toe) * slope - sy
    keep = below > 0.0
    weight = np.clip(below[keep] / (grade_h * params["h"]), 0.0, 1.0) if grade_h > 0 \
        else np.ones(int(keep.sum()))
    sx = off + sx[keep]
    sy = offy + sy[keep]
    _grid(params, off + length + 4 * dx, offy + tank_h + 4 * dx)
    meta = dict(dx=dx, length=length, depth=depth, toe=toe, slope=slope, porosity=porosity,
                d50=d50, wave_h=wave_h, wave_t=wave_t, grade_h=grade_h, off=off, offy=offy)
    if wavemaker:
        meta["wavemaker"] = dict(stroke=stroke, period=wave_t,
                                 ramp=wave_t if ramp is None else ramp, t0=0.0)
    c = _assemble(params, name, fx, fy, fc, wx, wy, wt, dx, offy + depth, sx, sy,
                  steps=steps, meta=meta, wflags=wflags)
    c.solids["m"] = c.solids["m"] * weight
    return pore_volumes(c) if pore_volume else c


def flume_650k(steps: int = 2000, dt: float = 5.0e-5) -> Case:
    """The 650k-particle wave-runup workload: the 10 m x 0.4 m flume of
    `porous_flume` at dx = 0.0025 (4000 x 160 fluid lattice, 3 wall layers,
    piston wavemaker for H = 0.05 m, T = 1.5 s, scalar patch at the toe) on
    its 1:10 porous beach of porosity 0.4 and d50 = 0.01 m (128,000 solids)."""
    return porous_flume(steps=steps, dt=dt, name="flume_650k")


def scenario_mix(n_scen: int = 64, seed: int = 2024, n_min: int = 20_000,
                 n_max: int = 300_000, log_uniform: bool = False, steps: int = 1000,
                 beach_solids: bool = True):
    """Independent porous-beach flume scenarios of uneven particle count with
    parameters drawn as the batched configuration states: slope 1:8-1:20,
    porosity 0.3-0.5, d50 0.005-0.02 m, H 0.02-0.08 m, T 1.0-2.5 s.  Each
    flume is 6 m x 0.3 m; the lattice spacing is chosen so the fluid count
    matches the drawn N.  beach_solids=False leaves the beach solids out
    (an impermeable-free flume with the drawn porosity and d50 unused)."""
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(n_scen):
        if log_uniform:
            n = int(math.exp(rng.uniform(math.log(n_min), math.log(n_max))))
        else:
            n = int(rng.integers(n_min, n_max + 1))
        slope = 1.0 / rng.uniform(8.0, 20.0)
        por = rng.uniform(0.3, 0.5)
        d50 = rng.uniform(0.005, 0.02)
        wh = rng.uniform(0.02, 0.08)
        wt = rng.uniform(1.0, 2.5)
        length, depth = 6.0, 0.3
        dx = math.sqrt(length * depth / n)
        c = porous_flume(dx=dx, length=length, depth=depth, tank_h=0.45,
                         toe=length - depth / slope if depth / slope < length else 0.5,
                         slope=slope, porosity=por, d50=d50, dt=0.2 * dx / 10.0,
                         steps=steps, wave_h=wh, wave_t=wt, name=f"scenario{k}")
        c.params["tol"] = 1.0e-8
        if not beach_solids:
            c.solids = {kk: v[:0] for kk, v in c.solids.items()}
            c.meta["beach_solids"] = False
        cases.append(c)
    return cases


def tiny_box(nx: int = 12, ny: int = 10, dx: float = 0.01, jitter: float = 0.05,
             seed: int = 7, solids: bool = True, patch: bool = True) -> Case:
    """A few-hundred-particle tank for oracle-speed parity tests: jittered
    fluid, three wall layers, an optional porous corner and scalar patch."""
    params = default_params(dx)
    params.update(dt=2.0e-4, tol=1.0e-12, por=0.4, d50=0.01)
    off = 3.0 * dx + 0.5 * dx
    L, H = (nx + 4) * dx, (ny + 6) * dx
    wx, wy, wt = _walls_box(off, off + L, off, off + H, dx)
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rng = np.random.default_rng(seed)
    fx = off + (gx.ravel() + 0.5) * dx + rng.uniform(-jitter, jitter, gx.size) * dx
    fy = off + (gy.ravel() + 0.5) * dx + rng.uniform(-jitter, jitter, gy.size) * dx
    fc = ((gx.ravel() < nx // 3) & (gy.ravel() < ny // 3)).astype(np.float64) if patch \
        else np.zeros(fx.size)
    sx = sy = None
    if solids:
        sgx, sgy = np.meshgrid(np.arange(nx // 2, nx + 4), np.arange(ny // 2), indexing="ij")
        sx = off + (sgx.ravel() + 0.5) * dx
        sy = off + (sgy.ravel() + 0.5) * dx
    _grid(params, off + L + 4 * dx, off + H + 4 * dx)
    c = _assemble(params, "tiny_box", fx, fy, fc, wx, wy, wt, dx, off + ny * dx, sx, sy,
                  steps=3, meta=dict(nx=nx, ny=ny, dx=dx, jitter=jitter, seed=seed))
    # a little motion so every term of the pipeline is exercised
    c.parts["u"][wx.size:] = 0.05 * np.sin(fy * 40.0)
    c.parts["v"][wx.size:] = 0.03 * np.cos(fx * 30.0)
    return c


def buffered_box(nx: int = 12, ny: int = 8, dx: float = 0.01, nbuf: int = 3,
                 entry_vel: float = 0.05, insert_steps: int = 3, seed: int = 7) -> Case:
    """A small tank with an inflow buffer zone at the left (zone 1), for
    parity tests of the buffer pass.  The first fluid column is left empty;
    the next `nbuf` columns are buffer particles moving at entry_vel; the
    zone's template is the empty column (hydrostatic pressure, entry_vel), as
    the reference's buffer1 (internal.f90:105-136).  meta["buffer1"] holds the
    arguments of Scenario.set_buffer: every insert_steps steps the template
    column is inserted and buffer particles of the entry cells (the cell
    column holding the template) at x >= off + 1.5 dx are removed.  The release window of comp_pos (buf_xmin,
    buf_xmax) lies past the zone."""
    params = default_params(dx)
    params.update(dt=2.0e-4, tol=1.0e-12, por=0.4, d50=0.01)
    off = 3.0 * dx + 0.5 * dx
    L, H = (nx + 4) * dx, (ny + 6) * dx
    wx, wy, wt = _walls_box(off, off + L, off, off + H, dx)
    gx, gy = np.meshgrid(np.arange(1, nx), np.arange(ny), indexing="ij")
    fx = off + (gx.ravel() + 0.5) * dx
    fy = off + (gy.ravel() + 0.5) * dx
    fc = ((gx.ravel() >= nx // 2) & (gy.ravel() < ny // 3)).astype(np.float64)
    _grid(params, off + L + 4 * dx, off + H + 4 * dx)
    rv = off + 1.5 * dx
    params.update(buf_xmin=off + (nbuf + 0.5) * dx, buf_xmax=off + (nx - 1) * dx)
    c = _assemble(params, "buffered_box", fx, fy, fc, wx, wy, wt, dx, off + ny * dx,
                  steps=2 * insert_steps, meta=dict(nx=nx, ny=ny, dx=dx, nbuf=nbuf))
    fl = c.parts["tid"] == TID_FLUID
    col = np.floor((c.parts["x"] - off) / dx + 1e-9).astype(int)
    buf = fl & (col >= 1) & (col <= nbuf)
    c.parts["flags"] = np.where(buf, F_BUFFER, 0).astype(np.uint8)
    c.parts["u"] = np.where(buf, entry_vel, 0.02 * np.sin(c.parts["y"] * 40.0) * fl)
    ty = off + (np.arange(ny) + 0.5) * dx
    tx = np.full(ny, off + 0.5 * dx)
    tp = RHO * abs(G) * (off + ny * dx - ty)
    c.meta["buffer1"] = dict(x=tx, y=ty, u=np.full(ny, entry_vel), v=np.zeros(ny), p=tp,
                             mass=RHO * dx * dx, rho=RHO, vol0=dx * dx, c=-0.5, rv=rv,
                             interval=insert_steps * params["dt"] * (1.0 - 1e-9))
    return c
