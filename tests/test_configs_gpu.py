"""The batched and served configurations at their stated sizes, through
properties that hold at any size (the oracle takes minutes per step here).

Batched: 64 porous-beach flumes, N ~ U[20k, 300k] (seeded, ~10M particles),
slope 1:8-1:20, porosity 0.3-0.5, d50 0.005-0.02 m, H 0.02-0.08 m,
T 1.0-2.5 s, precision form at 1e-8, stepped together.
Served: a seeded Poisson stream of flumes, N log-uniform on [10k, 1M],
admitted against the 240 GB budget and continuously batched.

After the steps every solve met its tolerance, every particle is still there
with its volume, fluid masses are rho0 vol0, the fields are finite, cell keys
ascend with their cell starts, and the beach's pore water holds its porosity
(n >= por)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check_scenario(pkg, c, before, after, cell, cs):
    n = before["x"].size
    assert after["x"].size == n == c.n
    for k in ("x", "y", "u", "v", "p", "rho", "c"):
        assert np.isfinite(after[k]).all(), k
    assert np.array_equal(np.sort(before["vol0"]), np.sort(after["vol0"]))
    fl = after["tid"] == 3
    assert np.array_equal(after["m"][fl], after["rho0"][fl] * after["vol0"][fl])
    assert np.array_equal(np.bincount(before["tid"], minlength=4),
                          np.bincount(after["tid"], minlength=4))
    ncell = c.params["cellx"] * c.params["celly"]
    assert cs[0] == 0 and cs[-1] == n and np.all(np.diff(cell) >= 0)
    assert np.array_equal(np.searchsorted(cell, np.arange(ncell + 1)), cs)


def test_batched_configuration_64_flumes(pkg, cases):
    mix = cases.scenario_mix()
    assert len(mix) == 64 and all(c.ns > 0 for c in mix)
    for c in mix:
        c.params.update(precision=1, tol=1e-8, max_outer=20)
    total = sum(c.n for c in mix)
    print(f"batched: 64 flumes, {total} particles, {sum(c.ns for c in mix)} solids", flush=True)
    assert 9_000_000 < total < 11_500_000
    scen = [pkg.Scenario(c) for c in mix]
    for s in scen:
        s.init_state()
    before = [s.download() for s in scen]
    b = pkg.Batch(scen)
    for k in range(3):
        b.step(1)
        its = []
        for q in range(len(scen)):
            st = b.stats(q)
            assert st.converged and st.rel_residual <= 1e-8, (k, q, st.rel_residual)
            its.append(st.iterations)
        print(f"batched step {k}: iterations {min(its)}-{max(its)}, "
              f"{b.launches()} particle-phase launches", flush=True)
    assert b.launches() < 80
    b.close()
    for q, s in enumerate(scen):
        after = s.download()
        cell, cs = s.cells()
        _check_scenario(pkg, mix[q], before[q], after, cell, cs)
        por = s.props()["porosity"][after["tid"] == 3]
        assert np.all(por >= mix[q].params["por"] - 1e-12)
        s.close()


def test_served_configuration_sizes_10k_to_1m(pkg, cases):
    from beach_slope_amd import serve
    mix = cases.scenario_mix(n_scen=6, seed=25, n_min=10_000, n_max=1_000_000,
                             log_uniform=True, steps=3)
    sizes = [c.n for c in mix]
    print("served sizes:", sizes, flush=True)
    assert min(sizes) < 30_000 and max(sizes) > 900_000
    for c in mix:
        c.params.update(precision=1, tol=1e-8, max_outer=20)
    arrivals = serve.poisson_arrivals(len(mix), rate=2.0, seed=25)
    reqs = [serve.Request(k, c, steps=3, arrival=float(a)) for k, (c, a) in
            enumerate(zip(mix, arrivals))]
    srv = serve.Server(pkg, hbm_budget=240 * 10**9, max_slots=64, clock=serve.StepClock())
    srv.run(reqs)
    for r in reqs:
        assert r.finished is not None and len(r.iterations) == 3
        res = r.result
        n = r.case.n
        assert res["x"].size == n
        for k in ("x", "y", "u", "v", "p", "rho", "c"):
            assert np.isfinite(res[k]).all(), (r.rid, k)
        assert np.array_equal(np.sort(res["vol0"]), np.sort(r.case.parts["vol0"]))
        fl = res["tid"] == 3
        assert np.array_equal(res["m"][fl], res["rho0"][fl] * res["vol0"][fl])
    assert srv.particle_steps == 3 * sum(sizes)
