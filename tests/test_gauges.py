"""Gauges (host post-processing of a state): the kernel equals the oracle's
restatement of Wab, Shepard probes reproduce a linear pressure field at
lattice-symmetric points, and the surface / front / runup / pairing gauges
give the values fixed by construction on synthetic states."""
import numpy as np


def _lattice(nx=40, ny=20, dx=0.01, p0=100.0, b=-9810.0):
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = (gx.ravel() + 0.5) * dx
    y = (gy.ravel() + 0.5) * dx
    n = x.size
    return dict(x=x, y=y, u=np.zeros(n), v=np.zeros(n), p=p0 + b * y,
                m=np.full(n, 1000 * dx * dx), rho=np.full(n, 1000.0),
                tid=np.full(n, 3, np.int32), flags=np.zeros(n, np.uint8))


def test_kernel_matches_oracle(pkg, oracle):
    from beach_slope_amd import gauges
    L = oracle.lib()
    h = 0.013
    r = np.linspace(0.0, 2.2 * h, 97)
    ref = np.array([L.or_wab(float(v), h) if v < 2 * h else 0.0 for v in r])
    assert np.allclose(gauges._wab(r, h), ref, rtol=1e-13, atol=0.0)


def test_probe_pressure_linear_field(pkg):
    from beach_slope_amd import gauges
    dx = 0.01
    s = _lattice(dx=dx)
    pts = np.array([[0.20, 0.10], [0.10, 0.05], [0.30, 0.12]])  # lattice nodes' corners
    got = gauges.probe_pressure(s, pts, h=1.3 * dx)
    assert np.allclose(got, 100.0 - 9810.0 * pts[:, 1], rtol=1e-12)
    assert np.isnan(gauges.probe_pressure(s, np.array([[5.0, 5.0]]), h=1.3 * dx))[0]


def test_surface_front_runup_and_pairing(pkg):
    from beach_slope_amd import gauges
    s = _lattice(nx=40, ny=20, dx=0.01)
    assert np.allclose(gauges.surface_elevation(s, np.array([0.105, 0.305]), 0.005), 0.195)
    assert abs(gauges.front_position(s) - 0.395) < 1e-12
    assert abs(gauges.runup(s, toe=0.30, bed=0.0) - 0.195) < 1e-12
    assert gauges.runup(s, toe=1.0, bed=0.0) == 0.0
    perm = np.random.default_rng(3).permutation(s["x"].size)
    t = {k: v[perm] for k, v in s.items()}
    r = gauges.pair_states(s, t)
    assert r["max_pair_distance"] == 0.0 and r["pressure_rel_l2"] == 0.0
    assert r["free_flag_mismatch"] == 0.0
    t["p"] = t["p"] * (1 + 1e-6)
    assert abs(gauges.pair_states(t, s)["pressure_rel_l2"] - 1e-6) < 1e-9
