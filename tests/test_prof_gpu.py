"""Kernel timers of the precision-form solve: every timed kernel reports its
launches, and the counts follow the iteration's launch sequence (per
iteration: one of each SpMV, two of each level-0 pass and of the level-1
restriction and coarse workgroup, 2 x the multi-workgroup cell levels of
each cell pass, two finalizes), counted over the iterations that did work."""
import pytest

TIMED = ("spmv_p32", "spmv_s32", "mg_pre0", "mg_post0", "mg_restrict0", "mg_down",
         "mg_coarse", "mg_up", "pfinalize")
PER_ITER = {"spmv_p32": 1, "spmv_s32": 1, "mg_pre0": 2, "mg_post0": 2, "mg_restrict0": 2,
            "mg_coarse": 2, "pfinalize": 2}


def _multi_workgroup_levels(case):
    """Cell levels run by the many-workgroup passes: those before the first
    level with at most 1024 unknowns (the coarse workgroup's first)."""
    ny, nx, lv = case.params["celly"], case.params["cellx"], []
    while True:
        lv.append(ny * nx)
        if ny * nx <= 64:
            break
        ny, nx = (ny + 1) // 2, (nx + 1) // 2
    return next(i for i, m in enumerate(lv) if m <= 1024)


@pytest.mark.gpu
def test_solver_timers_count_every_working_launch(pkg, cases):
    c = cases.porous_flume(length=4.0, depth=0.4, dx=0.005, toe=2.5, steps=3)
    c.params.update(precision=1, tol=1e-8)
    nmb = _multi_workgroup_levels(c)
    print("timers: n =", c.n, "multi-workgroup cell levels", nmb, flush=True)
    assert nmb >= 1
    s = pkg.Scenario(c)
    s.init_state()
    s.step(1)
    s.prof_enable(True)
    its = 0
    for _ in range(2):
        _, st = s.step(1)
        assert st.converged
        its += st.iterations
    got = {k: s.prof_read(k) for k in TIMED}
    s.prof_enable(False)
    s.close()
    print("timers:", {k: (round(ms, 3), n) for k, (ms, n) in got.items()}, flush=True)
    for k, (ms, n) in got.items():
        assert n > 0 and ms > 0.0, k
    assert got["spmv_p32"][1] == its
    for k, m in PER_ITER.items():
        assert got[k][1] == m * its, k
    assert got["mg_down"][1] == got["mg_up"][1] == 2 * nmb * its
