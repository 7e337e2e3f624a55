"""Batched form: independent scenarios stepped together give, scenario by
scenario, the same results as each scenario stepped alone (bit for bit: a
scenario's sums run over its own rows in its own block order, whatever else
shares the launch)."""
import numpy as np
import pytest

FIELDS = ("x", "y", "u", "v", "p", "c", "rho", "rho0", "m", "vol0", "tid", "flags")


def _mix(cases, precisions):
    mix = cases.scenario_mix(n_scen=len(precisions), seed=7, n_min=2500, n_max=9000, steps=3)
    for c, prec in zip(mix, precisions):
        c.params.update(precision=prec, tol=1e-8 if prec else 1e-10)
    return mix


def _alone(pkg, c, k):
    s = pkg.Scenario(c)
    s.init_state()
    its = []
    for _ in range(k):
        _, st = s.step(1)
        its.append(st.iterations)
    out = s.download()
    s.close()
    return out, its


@pytest.mark.gpu
def test_batch_results_equal_alone(pkg, cases):
    k = 3
    precisions = (1, 1, 0, 1)
    mix = _mix(cases, precisions)
    print("batch mix n =", [c.n for c in mix], "precision =", precisions, flush=True)
    alone = [_alone(pkg, c, k) for c in mix]
    scen = [pkg.Scenario(c) for c in mix]
    for s in scen:
        s.init_state()
    b = pkg.Batch(scen)
    its = [[] for _ in scen]
    for _ in range(k):
        b.step(1)
        for q in range(len(scen)):
            its[q].append(b.stats(q).iterations)
    b.close()
    for q, s in enumerate(scen):
        got = s.download()
        ref, ref_its = alone[q]
        assert its[q] == ref_its, (q, its[q], ref_its)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (q, f)
        s.close()
    print("batch: per-scenario iterations", its, flush=True)


@pytest.mark.gpu
def test_batch_phases_are_batched_and_form_is_fixed(pkg, cases):
    """The particle phases of a batch run as one launch per phase: the
    launches of a batched step do not grow with the number of scenarios.
    While a context belongs to a batch its solve form cannot change."""
    counts = {}
    for S in (2, 6):
        mix = _mix(cases, (1,) * S)
        scen = [pkg.Scenario(c) for c in mix]
        for s in scen:
            s.init_state()
        b = pkg.Batch(scen)
        b.step(2)
        counts[S] = b.launches()
        live = [b.count(q) for q in range(S)]
        assert live == [s.n for s in scen]
        with pytest.raises(pkg.IsphError) as e:
            scen[0].set_params(precision=0)
        assert e.value.code == -6
        with pytest.raises(pkg.IsphError) as e:
            scen[1].set_wcsph(1.0)
        assert e.value.code == -6
        b.close()
        scen[0].set_params(precision=0)     # free again once the batch is gone
        for s in scen:
            s.close()
    print("particle-phase launches per batched step:", counts, flush=True)
    assert 0 < counts[2] < 80 and 0 < counts[6] < 80


@pytest.mark.gpu
def test_batch_graph_replays_equal_eager_steps(pkg, cases):
    """The batched phases replayed as HIP graphs: every step's state, scenario
    by scenario, equals the scenario stepped alone without graphs (the graph
    is captured at the first step and replayed at the next three, with the
    solver's own launches, host polls and downloads between replays)."""
    k = 4
    mix = _mix(cases, (1, 1, 1))
    ref = []
    for c in mix:
        s = pkg.Scenario(c)
        s.init_state()
        states = []
        for _ in range(k):
            s.step(1)
            states.append(s.download())
        s.close()
        ref.append(states)
    scen = [pkg.Scenario(c) for c in mix]
    for s in scen:
        s.init_state()
    b = pkg.Batch(scen)
    b.set_graph(True)
    for step in range(k):
        b.step(1)
        print(f"graph step {step}: {b.launches()} particle-phase launches", flush=True)
        for q, s in enumerate(scen):
            got = s.download()
            for f in FIELDS:
                assert np.array_equal(got[f], ref[q][step][f]), (step, q, f)
    b.close()
    for s in scen:
        s.close()


@pytest.mark.gpu
def test_autotuned_launches_give_the_same_results(pkg, cases):
    """The autotuner times the particle kernels at 64, 128 and 256 threads per
    block and keeps the fastest; results do not depend on the block size, so
    a tuned scenario and a tuned batch step exactly as untuned ones."""
    k = 3
    mix = _mix(cases, (1, 1))
    ref = [_alone(pkg, c, k)[0] for c in mix]
    s = pkg.Scenario(mix[0])
    s.init_state()
    tuned = s.autotune(trials=2)
    print("autotune (alone):", {q: v[0] for q, v in tuned.items()}, flush=True)
    assert set(tuned) == set(pkg.TUNED_KERNELS)
    for _ in range(k):
        s.step(1)
    got = s.download()
    for f in FIELDS:
        assert np.array_equal(got[f], ref[0][f]), f
    s.close()
    scen = [pkg.Scenario(c) for c in mix]
    for sc in scen:
        sc.init_state()
    b = pkg.Batch(scen)
    tb = b.autotune(trials=2)
    print("autotune (batch):", {q: v[0] for q, v in tb.items()}, flush=True)
    b.step(k)
    b.close()
    for q, sc in enumerate(scen):
        got = sc.download()
        for f in FIELDS:
            assert np.array_equal(got[f], ref[q][f]), (q, f)
        sc.close()
