"""Gauges on a downloaded particle state (host post-processing).

The reference writes whole-field snapshots only (output.f90); these gauges
reduce a state to the quantities the configurations compare: point
pressures, free-surface elevation at fixed stations, the surge front, the
runup on the slope, and a particle-by-particle comparison of two states of
the same scenario.  Two runs of one scenario have no shared particle ids
after rebinning (the state is held in cell order), so the field comparison
pairs each particle of one state with the nearest particle of the other.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

TID_FLUID = 3
F_FREE = 2


def _wab(r: np.ndarray, h: float) -> np.ndarray:
    """The reference's kernel W(r, h) (kernel.f90:34-66): a blend, weights
    0.1 / 0.9, of two cubic profiles with real4 constants 28*3.14, 3*3.14."""
    k1 = 10.0 / (float(np.float32(28.0) * np.float32(3.14)) * h * h)
    k2 = 1.0 / (float(np.float32(3.0) * np.float32(3.14)) * h * h)
    q = r / h
    near = r < h
    t1 = np.where(near, 4.0 - 6.0 * q * q + 3.0 * q ** 3, (2.0 - q) ** 3)
    t2 = np.where(near, q ** 3 - 6.0 * q + 6.0, (2.0 - q) ** 3)
    w = 0.1 * t1 * k1 + 0.9 * t2 * k2
    return np.where(r < 2.0 * h, w, 0.0)


def fluid(state: dict) -> np.ndarray:
    return state["tid"] == TID_FLUID


def probe_pressure(state: dict, pts: np.ndarray, h: float) -> np.ndarray:
    """Shepard (normalised kernel-weighted) pressure of the fluid particles at
    the probe points; NaN where no fluid particle lies within 2h."""
    f = fluid(state)
    xy = np.stack([state["x"][f], state["y"][f]], axis=1)
    vol = state["m"][f] / state["rho"][f]
    tree = cKDTree(xy)
    out = np.full(len(pts), np.nan)
    for k, nb in enumerate(tree.query_ball_point(pts, 2.0 * h)):
        if not nb:
            continue
        nb = np.asarray(nb)
        w = _wab(np.hypot(xy[nb, 0] - pts[k, 0], xy[nb, 1] - pts[k, 1]), h) * vol[nb]
        if w.sum() > 0:
            out[k] = float(np.dot(w, state["p"][f][nb]) / w.sum())
    return out


def surface_elevation(state: dict, stations: np.ndarray, half_width: float) -> np.ndarray:
    """Highest fluid particle within +-half_width of each station (NaN if
    none): the free-surface elevation there."""
    f = fluid(state)
    x, y = state["x"][f], state["y"][f]
    out = np.full(len(stations), np.nan)
    for k, xs in enumerate(stations):
        sel = np.abs(x - xs) <= half_width
        if sel.any():
            out[k] = float(y[sel].max())
    return out


def front_position(state: dict) -> float:
    """Largest x reached by a fluid particle (the surge front)."""
    return float(state["x"][fluid(state)].max())


def runup(state: dict, toe: float, bed: float) -> float:
    """Highest elevation above the bed level `bed` reached by fluid on the
    slope (x >= toe); 0 while no fluid has reached the toe."""
    f = fluid(state) & (state["x"] >= toe)
    return float(state["y"][f].max() - bed) if f.any() else 0.0


def pair_states(a: dict, b: dict) -> dict:
    """Pair every fluid particle of `a` with the nearest fluid particle of
    `b`; report the largest pairing distance, the relative L2 difference of
    the paired pressures and velocities, and the fraction of pairs whose
    free-surface flags differ."""
    fa, fb = fluid(a), fluid(b)
    pa = np.stack([a["x"][fa], a["y"][fa]], axis=1)
    pb = np.stack([b["x"][fb], b["y"][fb]], axis=1)
    d, j = cKDTree(pb).query(pa)

    def rel(ua, ub):
        den = np.linalg.norm(ub)
        return float(np.linalg.norm(ua - ub) / den) if den > 0 else float(np.linalg.norm(ua))

    free_a = (a["flags"][fa] & F_FREE) != 0
    free_b = (b["flags"][fb] & F_FREE) != 0
    return dict(
        fluid_a=int(fa.sum()), fluid_b=int(fb.sum()),
        max_pair_distance=float(d.max()),
        pressure_rel_l2=rel(a["p"][fa], b["p"][fb][j]),
        velocity_rel_l2=rel(np.hypot(a["u"][fa], a["v"][fa]),
                            np.hypot(b["u"][fb][j], b["v"][fb][j])),
        free_flag_mismatch=float(np.mean(free_a != free_b[j])),
    )
