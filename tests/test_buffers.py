"""Inflow/outflow buffer zones (internal.f90:447-586, main.f90:217-235).

CPU: the oracle's buffer pass against the reference's semantics restated
independently here -- per cell, entry cells keep their non-removed particles
in order and append the zone's template particles in template order, other
cells are untouched, inserted particles carry the zone's values.  GPU: the
device pass and the insertion clock inside fused steps equal the oracle bit
for bit."""
import numpy as np
import pytest

FIELDS = ("x", "y", "u", "v", "p", "m", "rho", "rho0", "vol0", "c")


def _cell_of(case, x, y):
    """Cell key of points well inside a cell: column cj = floor(x/cs) + 1,
    row ci = celly - 2 - floor(y/cs) (setup.f90:120-123)."""
    p = case.params
    cs = 4.0 * p["h"]
    cj = np.floor(x / cs).astype(int) + 1
    ci = p["celly"] - 2 - np.floor(y / cs).astype(int)
    return cj * p["celly"] + ci


def _buffer_args(b):
    return (b["x"], b["y"], b["u"], b["v"], b["p"], b["mass"], b["rho"], b["vol0"], b["c"],
            b["rv"])


def test_oracle_buffer_pass_semantics(pkg, oracle, cases):
    c = cases.buffered_box()
    b = c.meta["buffer1"]
    o = oracle.Oracle(c, cap=c.n + 64)
    assert o.set_buffer(1, *_buffer_args(b)) == 0
    assert o.call("init_state") == 0
    # move the buffer particles so some pass the removal coordinate
    for _ in range(2):
        assert o.step(1)[0] == 0
    pre = o.state()
    pre_cell = o.arr("cell").copy()
    ncell = c.params["cellx"] * c.params["celly"]
    tcell = _cell_of(c, b["x"], b["y"])
    entry = np.zeros(ncell, bool)
    entry[tcell] = True
    # expected lists, restated from remove_buffer1 / insert_buffer1
    order = []
    for k in range(ncell):
        idx = np.flatnonzero(pre_cell == k)
        for i in idx:
            drop = entry[k] and (pre["flags"][i] & 1) and pre["x"][i] >= b["rv"]
            if not drop:
                order.append(("kept", int(i)))
        if entry[k]:
            order += [("new", int(t)) for t in np.flatnonzero(tcell == k)]
    removed = sum(1 for k in range(ncell) if entry[k]
                  for i in np.flatnonzero(pre_cell == k)
                  if (pre["flags"][i] & 1) and pre["x"][i] >= b["rv"])
    assert removed > 0
    assert o.buffer_update(1) == 0
    post = o.state()
    assert o.n == len(order) == pre["x"].size - removed + b["x"].size
    for d, (kind, q) in enumerate(order):
        if kind == "kept":
            for f in FIELDS + ("tid", "flags"):
                assert post[f][d] == pre[f][q], (d, f)
        else:
            want = dict(x=b["x"][q], y=b["y"][q], u=b["u"][q], v=b["v"][q], p=b["p"][q],
                        m=b["mass"], rho=b["rho"], rho0=b["rho"], vol0=b["vol0"], c=b["c"])
            for f, v in want.items():
                assert post[f][d] == v, (d, f)
            assert post["tid"][d] == 3 and post["flags"][d] == 1
    assert np.array_equal(o.arr("cell")[:o.n], np.array(
        [pre_cell[q] if kind == "kept" else tcell[q] for kind, q in order]))


def test_oracle_buffer_errors(pkg, oracle, cases):
    c = cases.buffered_box()
    b = c.meta["buffer1"]
    o = oracle.Oracle(c, cap=c.n)
    assert o.buffer_update(1) == -2                      # zone not set
    assert o.set_buffer(3, *_buffer_args(b)) == -2       # no zone 3
    assert o.set_buffer(1, *_buffer_args(b)) == 0
    assert o.buffer_update(1) == -3                      # no room for the template


@pytest.mark.gpu
def test_buffer_pass_and_clock_bit_exact(pkg, oracle, cases):
    from test_parity_gpu import compare_state, compare_scalars
    c = cases.buffered_box()
    b = c.meta["buffer1"]
    cap = c.n + 64
    g = pkg.Scenario(c, cap=cap)
    o = oracle.Oracle(c, cap=cap)
    g.set_buffer(1, *_buffer_args(b), interval=b["interval"])
    assert o.set_buffer(1, *_buffer_args(b), b["interval"]) == 0
    print(f"buffered_box n={c.n} template={b['x'].size}", flush=True)
    g.init_state()
    assert o.call("init_state") == 0
    for k in range(9):
        _, st = g.step(1)
        rc, sto = o.step(1)
        assert rc == 0 and st.iterations == sto.iterations, (k, rc)
        compare_state(g, o, f"buffered step {k}")
    compare_scalars(g, o, "buffered after 9 steps")
    assert g.n > c.n
    # an explicit pass (the clock's next insertion would come at step 12)
    g.buffer_update(1)
    assert o.buffer_update(1) == 0
    compare_state(g, o, "explicit buffer pass")
