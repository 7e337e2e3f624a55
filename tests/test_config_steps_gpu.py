"""The dam-break configurations at their stated step count, on the HIP path
and the CPU oracle from the same seeded input: 200 fused FP64 steps
(BiCGStab to 1e-10, dt 1e-4, shifting on) of the dam break onto the 1:10
slope, and of the same dam break with seeded +-0.01 dx jitter.  Every
step's iteration count and restarts agree, and the particle state and the
time-stepping scalars are bit-identical at steps 50, 100 and 200."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(name, a, b):
    a, b = np.asarray(a), np.asarray(b)
    if not np.array_equal(a, b, equal_nan=True):
        bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
        i = int(bad[0])
        raise AssertionError(f"{name}: {bad.size} entries differ, first at {i}: "
                             f"gpu={a[i]!r} oracle={b[i]!r}")


def _run(pkg, oracle, c, steps=200, checks=(50, 100, 200)):
    g = pkg.Scenario(c)
    o = oracle.Oracle(c)
    g.init_state()
    assert o.call("init_state") == 0
    its = []
    for k in range(1, steps + 1):
        _, st = g.step(1)
        rc, sto = o.step(1)
        assert rc == 0, (k, rc)
        assert (st.iterations, st.outer, st.converged) == \
            (sto.iterations, sto.outer, sto.converged), k
        its.append(st.iterations)
        if k in checks:
            gs, os_ = g.download(), o.state()
            assert gs["x"].size == os_["x"].size, k
            for f in ("x", "y", "u", "v", "p", "m", "rho", "rho0", "vol0", "c", "tid", "flags"):
                _same(f"step {k}: {f}", gs[f], os_[f])
            a, b = g.scalars(), o.scalars()
            for f in ("t", "dt", "dtsol", "umax", "numax", "solsteps", "n"):
                assert getattr(a, f) == getattr(b, f), (k, f)
            print(f"{c.name}: step {k} bit-identical (n={gs['x'].size}, "
                  f"iterations {min(its)}-{max(its)})", flush=True)
    g.close()
    return its


def test_dam_break_200_steps_bit_exact(pkg, oracle, cases):
    c = cases.dam_break()
    assert c.params["dt"] == 1e-4 and c.params["tol"] == 1e-10
    _run(pkg, oracle, c)


def test_jittered_dam_break_fp64_200_steps_bit_exact(pkg, oracle, cases):
    c = cases.dam_break(jitter=0.01)
    _run(pkg, oracle, c)
