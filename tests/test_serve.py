"""Served form: the request loop's admission, budget and slot logic on the
CPU (with a stand-in for the device package), and on the GPU the result of
every request equals the same scenario stepped alone although the active set
changes between steps."""
import types

import numpy as np
import pytest


class _FakeScenario:
    live = 0

    def __init__(self, case):
        self.case = case
        self.steps = 0
        _FakeScenario.live += 1

    def init_state(self):
        pass

    def download(self):
        return {"x": np.full(3, float(self.steps))}

    def close(self):
        _FakeScenario.live -= 1


class _FakeBatch:
    log = []

    def __init__(self, scen):
        self.scen = scen
        _FakeBatch.log.append(tuple(s.case.name for s in scen))

    def step(self, n):
        for s in self.scen:
            s.steps += n

    def stats(self, q):
        return types.SimpleNamespace(iterations=7)

    def count(self, q):
        return self.scen[q].case.n

    def close(self):
        pass


def _fake_pkg(footprints):
    return types.SimpleNamespace(
        Scenario=_FakeScenario, Batch=_FakeBatch,
        footprint_bytes=lambda n, ns, cx, cy, prec: footprints[n])


def _case(name, n):
    return types.SimpleNamespace(name=name, n=n, ns=0,
                                 params=dict(cellx=4, celly=4, precision=1))


def test_admission_budget_and_slots(pkg):
    from beach_slope_amd import serve
    _FakeBatch.log.clear()
    # footprints keyed by particle count: 100 -> 40 B, 200 -> 70 B, 300 -> 50 B
    pkg = _fake_pkg({100: 40, 200: 70, 300: 50})
    reqs = [serve.Request(0, _case("a", 100), steps=3, arrival=0.0),
            serve.Request(1, _case("b", 200), steps=2, arrival=0.0),
            serve.Request(2, _case("c", 300), steps=2, arrival=1.0)]
    srv = serve.Server(pkg, hbm_budget=120, max_slots=8, clock=serve.StepClock())
    srv.run(reqs)
    # a and b fit (110 B); c (50 B) waits until b leaves at t = 2
    assert [r.admitted for r in reqs] == [0.0, 0.0, 2.0]
    assert [r.finished for r in reqs] == [3.0, 2.0, 4.0]
    assert [r.latency for r in reqs] == [3.0, 2.0, 3.0]
    assert all(np.all(r.result["x"] == r.steps) for r in reqs)
    assert _FakeBatch.log == [("a", "b"), ("a", "c"), ("c",)]  # rebuilt on membership change
    assert _FakeScenario.live == 0
    assert srv.steps_run == 4 and srv.particle_steps == 3 * 100 + 2 * 200 + 2 * 300
    with pytest.raises(ValueError):
        serve.Server(pkg, hbm_budget=60, clock=serve.StepClock()).run(
            [serve.Request(9, _case("big", 200), steps=1, arrival=0.0)])


def test_slot_limit_and_idle_gap(pkg):
    from beach_slope_amd import serve
    pkg = _fake_pkg({100: 1})
    reqs = [serve.Request(k, _case(f"s{k}", 100), steps=2, arrival=a)
            for k, a in enumerate([0.0, 0.0, 0.0, 10.0])]
    serve.Server(pkg, hbm_budget=100, max_slots=2, clock=serve.StepClock()).run(reqs)
    assert [r.admitted for r in reqs] == [0.0, 0.0, 2.0, 10.0]
    assert [r.finished for r in reqs] == [2.0, 2.0, 4.0, 12.0]


def test_poisson_arrivals_seeded(pkg):
    from beach_slope_amd import serve
    a = serve.poisson_arrivals(2000, rate=2.0, seed=5)
    assert np.all(np.diff(a) > 0)
    assert abs(np.mean(np.diff(a)) - 0.5) < 0.05
    assert np.array_equal(a, serve.poisson_arrivals(2000, rate=2.0, seed=5))


@pytest.mark.gpu
def test_served_results_equal_alone(pkg, cases):
    from beach_slope_amd import serve
    mix = cases.scenario_mix(n_scen=4, seed=11, n_min=2500, n_max=7000, steps=3)
    for c in mix:
        c.params.update(precision=1, tol=1e-8)
    steps = [3, 2, 4, 2]
    arrivals = [0.0, 1.0, 1.0, 3.0]
    reqs = [serve.Request(k, c, steps=s, arrival=a)
            for k, (c, s, a) in enumerate(zip(mix, steps, arrivals))]
    srv = serve.Server(pkg, max_slots=3, clock=serve.StepClock())
    srv.run(reqs)
    print("served:", [(r.rid, r.admitted, r.finished) for r in reqs], flush=True)
    for r in reqs:
        s = pkg.Scenario(r.case)
        s.init_state()
        its = []
        for _ in range(r.steps):
            _, st = s.step(1)
            its.append(st.iterations)
        ref = s.download()
        s.close()
        assert r.iterations == its, (r.rid, r.iterations, its)
        for f in ("x", "y", "u", "v", "p", "c", "rho", "tid", "flags"):
            assert np.array_equal(r.result[f], ref[f]), (r.rid, f)
