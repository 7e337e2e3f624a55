"""The neighbour search in every launch shape gives the same lists, sums and
flags, bit for bit: one thread per particle in blocks of 64, 128 and 256
threads, and one workgroup per tile of 4 or 8 cells of a cell column with
the tile's neighbourhood staged in LDS.  A tiled workgroup whose
neighbourhood holds more particles than its LDS capacity (640 particles for
4-row tiles, 1152 for 8-row ones), or that holds a particle in an edge cell
of the grid, searches from global memory.  The dam break at h = 1.3 dx
stages every occupied tile in LDS; a 0.5 m deep one at h = 1.8 dx (52
particles to a cell) falls back on its dense tiles.  Each shape is checked
against the CPU oracle (neighbour lists after init_state, the whole state
after three steps), and the batched phases in the tiled shapes against the
scenarios stepped alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAP = {"tile4": (4, 640), "tile8": (8, 1152)}
FIELDS = ("x", "y", "u", "v", "p", "m", "rho", "rho0", "vol0", "c", "tid", "flags")


def _staged(cstart, cellx, celly, rows, cap):
    """Occupied tiles, and those of them the tiled kernel stages in LDS."""
    occ = staged = 0
    for cj in range(cellx):
        k = cj * celly
        for r0 in range(0, celly, rows):
            r1 = min(r0 + rows, celly)
            if cstart[k + r1] <= cstart[k + r0]:
                continue
            occ += 1
            if (cj < 1 or cj + 1 >= cellx or (r0 == 0 and cstart[k + 1] > cstart[k]) or
                    (r1 == celly and cstart[k + celly] > cstart[k + celly - 1])):
                continue
            lo, hi = max(r0 - 1, 0), min(r1 + 1, celly)
            tot = sum(int(cstart[c * celly + hi] - cstart[c * celly + lo])
                      for c in (cj - 1, cj, cj + 1))
            staged += tot <= cap
    return occ, staged


@pytest.mark.parametrize("h_fac,column", [(1.3, (0.5, 0.25)), (1.8, (0.5, 0.5))])
def test_neighbour_shapes_match_oracle(pkg, oracle, cases, h_fac, column):
    c = cases.dam_break(dx=0.005, h_fac=h_fac, column=column, steps=3)
    o = oracle.Oracle(c)
    assert o.call("init_state") == 0
    ocnt, onb = o.arr("ncnt"), o.arr("nbr")
    for _ in range(3):
        rc, _ = o.step(1)
        assert rc == 0
    ref = o.state()
    cellx, celly = c.params["cellx"], c.params["celly"]
    for shape in pkg.TUNE_SHAPES:
        g = pkg.Scenario(c)
        g.set_shape("neighbour", shape)
        g.init_state()
        if shape in CAP:
            _, cstart = g.cells()
            occ, staged = _staged(cstart, cellx, celly, *CAP[shape])
            print(f"h = {h_fac} dx, {shape}: {staged} of {occ} occupied tiles in LDS",
                  flush=True)
            if h_fac == 1.3:
                assert staged > occ // 2
            else:
                assert staged < occ
        cnt, lst = g.neighbours()
        assert np.array_equal(cnt, ocnt), shape
        for i in range(len(cnt)):
            assert np.array_equal(lst[i, :cnt[i]], onb[i, :cnt[i]]), (shape, i)
        g.step(3)
        got = g.download()
        assert got["x"].size == ref["x"].size, shape
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f], equal_nan=True), (shape, f)
        g.close()


@pytest.mark.parametrize("shape", ["tile4", "tile8"])
def test_batched_tiled_neighbour_equals_alone(pkg, cases, shape):
    k = 3
    mix = cases.scenario_mix(n_scen=3, seed=11, n_min=3000, n_max=9000, steps=k)
    for c, prec in zip(mix, (1, 0, 1)):
        c.params.update(precision=prec, tol=1e-8 if prec else 1e-10)
    ref = []
    for c in mix:
        s = pkg.Scenario(c)
        s.init_state()
        s.step(k)
        ref.append(s.download())
        s.close()
    scen = [pkg.Scenario(c) for c in mix]
    for s in scen:
        s.init_state()
    b = pkg.Batch(scen)
    b.set_shape("neighbour", shape)
    b.step(k)
    b.close()
    for q, s in enumerate(scen):
        got = s.download()
        for f in FIELDS:
            assert np.array_equal(got[f], ref[q][f], equal_nan=True), (shape, q, f)
        s.close()
