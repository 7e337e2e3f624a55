"""The porous beach at full depth, on the CPU oracle.

Pore-water particles carry the bulk volume of their lattice site
(`case.pore_volumes`: vol0 = n dx^2, m = rho0 vol0, so m / rho = dx^2 with
rho = n rho0 from effpor, porous.f90:188-190).  A 0.4 m-deep flume on a 1:10
porous beach (porosity 0.4) then stays bounded; with the open-water volume in
the pores (m / rho = dx^2 / n) the same flume diverges within a few steps,
from the volume jump at the beach surface.  The reference's seepage case
reaches m / rho = area per particle by spacing its pore lattice 2 pr / sqrt(por)
(internal.f90:37-39, setup.f90:71-72)."""
import numpy as np

DX = 0.02


def _flume(cases, pore_volume):
    return cases.porous_flume(dx=DX, length=3.0, depth=0.4, toe=1.0, slope=0.1,
                              porosity=0.4, d50=0.01, dt=4.0e-4, pore_volume=pore_volume)


def _max_speed(o):
    fl = o.arr("tid") == 3
    return float(np.hypot(o.arr("u"), o.arr("v"))[fl].max())


def test_pore_volume_is_the_lattice_area(oracle, cases):
    # the generator's porosity is effpor's: after init_state every fluid
    # particle has m / rho = dx^2, inside the beach as outside
    c = _flume(cases, True)
    o = oracle.Oracle(c)
    assert o.call("init_state") == 0
    fl = o.arr("tid") == 3
    n = o.arr("por")[fl]
    assert np.sum(n < 0.999) > 200 and n.min() < 0.45
    V = (o.arr("m") / o.arr("rho"))[fl]
    assert np.allclose(V, DX * DX, rtol=1e-12, atol=0)


def test_full_depth_porous_flume_bounded_with_pore_volumes(oracle, cases):
    o = oracle.Oracle(_flume(cases, True))
    assert o.call("init_state") == 0
    worst = 0.0
    for _ in range(100):
        rc, st = o.step(1)
        assert rc == 0 and st.converged
        worst = max(worst, _max_speed(o))
    # the paddle's orbital velocities (H 0.05 m, T 1.5 s) are ~0.1 m/s
    assert np.isfinite(worst) and worst < 0.2


def test_full_depth_porous_flume_diverges_with_open_water_volume(oracle, cases):
    o = oracle.Oracle(_flume(cases, False))
    assert o.call("init_state") == 0
    worst = 0.0
    for _ in range(10):
        rc, _ = o.step(1)
        s = _max_speed(o)
        worst = max(worst, s if np.isfinite(s) else np.inf)
        if rc != 0 or worst > 10.0:
            break
    assert worst > 10.0
