"""The 653k-particle wave flume on its 1:10 porous beach (porosity 0.4,
d50 0.01 m, 128,000 beach solids; piston wavemaker, H 0.05 m, T 1.5 s,
dx 2.5 mm) at full size, in the plain FP64 form and in the precision form.
The oracle takes minutes per step here, so the checks are properties that
hold at any size: every solve meets the 1e-8 relative residual, particles
and volumes are conserved, fixed walls stay put and the paddle moves rigidly
by the piston's displacement, the counting sort leaves every particle in the
cell range of its own cell, the beach's pore water keeps its porosity and
volume m / rho = dx^2, and on the same assembled system the precision form's
pressure lies within 1e-6 (relative) of a FP64 solve converged to 1e-11."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _order(s, mask):
    idx = np.flatnonzero(mask)
    return idx[np.lexsort((s["x"][idx], s["y"][idx]))]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "precision"])
def test_flume_650k_properties(pkg, cases, precision):
    c = cases.flume_650k()
    assert c.ns == 128_000
    print(f"flume_650k n={c.n} ns={c.ns} precision={precision}", flush=True)
    prm = dict(c.params, precision=precision, tol=1e-8, max_outer=20)
    g = pkg.Scenario(c, params=prm)
    g.init_state()
    before = g.download()
    t0 = g.scalars().t
    for k in range(3):
        _, st = g.step(1)
        print(f"step {k}: {st.iterations} iterations, rel residual {st.rel_residual:.2e}",
              flush=True)
        assert st.converged and st.rel_residual <= 1e-8
    after = g.download()
    t1 = g.scalars().t
    n = before["x"].size
    assert after["x"].size == n == c.n
    for k in ("x", "y", "u", "v", "p", "rho", "c"):
        assert np.isfinite(after[k]).all(), k
    # volumes travel with the particles; densityupdate (scalar.f90:399-426)
    # sets a fluid particle's mass to rho0 vol0, walls keep theirs
    assert np.array_equal(np.sort(before["vol0"]), np.sort(after["vol0"]))
    fl = after["tid"] == 3
    assert np.array_equal(after["m"][fl], after["rho0"][fl] * after["vol0"][fl])
    assert np.array_equal(np.sort(before["m"][before["tid"] != 3]), np.sort(after["m"][~fl]))
    assert np.array_equal(np.bincount(before["tid"], minlength=4),
                          np.bincount(after["tid"], minlength=4))
    # the beach: pore water particles keep their volume fraction, and the
    # porosity effpor gives them keeps m / rho at the lattice area
    pr = g.props()["porosity"]
    dx = c.meta["dx"]
    assert np.sum(pr[fl] < 0.5) > 100_000
    inside = fl & (after["vol0"] < 0.99 * dx * dx)
    assert inside.sum() > 100_000
    assert np.all(pr[fl] >= c.params["por"] - 1e-12)

    # fixed walls keep their positions exactly; paddle walls move together
    fixed = lambda s: (s["tid"] != 3) & ((s["flags"] & cases.F_PADDLE) == 0)
    paddle = lambda s: (s["tid"] != 3) & ((s["flags"] & cases.F_PADDLE) != 0)
    fb, fa = _order(before, fixed(before)), _order(after, fixed(after))
    assert fb.size == fa.size > 0
    assert np.array_equal(before["x"][fb], after["x"][fa])
    assert np.array_equal(before["y"][fb], after["y"][fa])
    pb, pa = _order(before, paddle(before)), _order(after, paddle(after))
    assert pb.size == pa.size > 0
    d = after["x"][pa] - before["x"][pb]
    w = c.meta["wavemaker"]
    x_t1, _ = pkg.wavemaker_motion(w["stroke"], w["period"], w["ramp"], w["t0"], t1)
    x_t0, _ = pkg.wavemaker_motion(w["stroke"], w["period"], w["ramp"], w["t0"], t0)
    assert abs(x_t1 - x_t0) > 0
    assert np.abs(d - (x_t1 - x_t0)).max() <= 1e-12
    assert np.array_equal(before["y"][pb], after["y"][pa])

    # counting sort: cell keys ascend, cstart brackets each cell, positions lie
    # in their cell (to half a particle spacing: shifting follows the rebin)
    cell, cs = g.cells()
    ncell = c.params["cellx"] * c.params["celly"]
    assert cs[0] == 0 and cs[-1] == n and np.all(np.diff(cs) >= 0)
    assert np.all(np.diff(cell) >= 0)
    assert np.array_equal(np.searchsorted(cell, np.arange(ncell + 1)), cs)
    size = 4.0 * c.params["h"]
    cj, ci = cell // c.params["celly"], cell % c.params["celly"]
    tol = 0.5 * c.meta.get("dx", 0.0025)
    x, y = after["x"], after["y"]
    assert np.all(x >= (cj - 1) * size - tol) and np.all(x <= cj * size + tol)
    assert np.all(y >= (c.params["celly"] - 2 - ci) * size - tol)
    assert np.all(y <= (c.params["celly"] - 1 - ci) * size + tol)

    if not precision:
        g.close()
        return
    # the precision form's pressure against a tight FP64 solve of the same rows
    g.ppe_assemble()
    st32 = g.ppe_solve()
    x32 = g.solution()
    g.set_params(precision=0, tol=1e-11, max_outer=20)
    st64 = g.ppe_solve()
    x64 = g.solution()
    rel = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    print(f"precision form {st32.iterations} iterations ({st32.rel_residual:.2e}); "
          f"FP64 {st64.iterations} iterations ({st64.rel_residual:.2e}); "
          f"pressure error {rel:.2e}", flush=True)
    assert st32.converged and st32.rel_residual <= 1e-8
    assert st64.converged and st64.rel_residual <= 1e-11
    assert rel <= 1e-6
    g.close()
