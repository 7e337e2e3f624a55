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
