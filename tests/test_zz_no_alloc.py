"""Allocation-free FP64 solve (isph_ppe_solve_noalloc) against the context
solve it is a variant of (isph_load_system + isph_ppe_solve, FP64 form) on
the same systems: the iterates must agree bit for bit.  Two calls in a row
share the library's one static scratch array, the second on a smaller
system, so a value the first call left behind would show."""
import ctypes
import json
import os

import numpy as np
import pytest


class DeviceArrays:
    """Device copies of host arrays through the HIP runtime the library itself
    links (libamdhip64.so.7): the test allocates, the entry point under test
    does not.  (torch carries its own HIP runtime, which cannot share the
    process with this one.)"""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p.value)
        assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for b in self.bufs:
            self.hip.hipFree(b)
        self.bufs = []


def test_noalloc_rejects_sizes_before_any_device_call(pkg):
    """Sizes beyond the maximum, a short leading dimension and null pointers
    are refused on the host (no HIP call is made: this runs without a GPU)."""
    big = pkg.NOALLOC_MAX_ROWS + 1
    rc, _ = pkg.ppe_solve_noalloc(big, big, 1, 1, 1, 1, 1, 1e-10, 10, check=False)
    assert rc == -2
    assert b"rows outside" in pkg.lib().isph_last_error(None)
    rc, _ = pkg.ppe_solve_noalloc(100, 99, 1, 1, 1, 1, 1, 1e-10, 10, check=False)
    assert rc == -2
    assert b"leading dimension" in pkg.lib().isph_last_error(None)
    rc, _ = pkg.ppe_solve_noalloc(100, 100, None, 1, 1, 1, 1, 1e-10, 10, check=False)
    assert rc == -2
    assert b"null" in pkg.lib().isph_last_error(None)


def _known_answer_5x5():
    with open(os.path.join(os.path.dirname(__file__), "golden", "solvertest_5x5.json")) as f:
        g = json.load(f)
    rowp = [r - 1 for r in g["row"]]
    n = len(rowp) - 1
    W = 66
    ln = np.array([rowp[i + 1] - rowp[i] for i in range(n)], np.int32)
    col = np.zeros((n, W), np.int32)
    val = np.zeros((n, W))
    for i in range(n):
        col[i, :ln[i]] = [c - 1 for c in g["col"][rowp[i]:rowp[i + 1]]]
        val[i, :ln[i]] = g["val"][rowp[i]:rowp[i + 1]]
    return ln, col, val, np.array(g["b"], np.float64), np.zeros(n), np.array(g["x"], float)


@pytest.mark.gpu
def test_noalloc_solve_equals_context_solve_twice(pkg, cases):
    # system 1: the dam break's assembled PPE (7.5k rows), from the HIP path
    c = cases.dam_break()
    prm = dict(c.params)
    prm.update(precision=0, tol=1e-10, max_iter=5000, max_outer=4)
    g = pkg.Scenario(c, params=prm)
    g.init_state()
    for ph in ("projection", "cellshift", "neighbour", "effpor", "eddyvis", "int_vel",
               "ppe_assemble"):
        getattr(g, ph)()
    ln, col, val, rhs = g.ppe()
    x0 = g.download()["p"]
    g.close()
    ln5, col5, val5, b5, x05, xk = _known_answer_5x5()
    systems = [("dam break", ln, col, val, rhs, x0), ("5x5", ln5, col5, val5, b5, x05)]

    ref = pkg.Scenario(c, params=prm)
    W = pkg.ROW_CAP
    for name, ln, col, val, rhs, x0 in systems:
        n = len(rhs)
        print(f"noalloc: {name}, n={n}", flush=True)
        ref.load_system(ln, col, val, rhs, x0)
        st_ref = ref.ppe_solve()
        x_ref = ref.solution()
        # the same system in device memory, ELL slot-major, leading dimension n + 3
        ld = n + 3
        cs = np.zeros((W, ld), np.int32)
        vs = np.zeros((W, ld))
        cs[:, :n] = col.T
        vs[:, :n] = val.T
        dv = DeviceArrays()
        d_x = dv.put(x0.astype(np.float64))
        rc, st = pkg.ppe_solve_noalloc(n, ld, dv.put(ln.astype(np.int32)), dv.put(cs),
                                       dv.put(vs), dv.put(rhs.astype(np.float64)), d_x,
                                       prm["tol"], prm["max_iter"], prm["max_outer"])
        x = dv.get(d_x, x0.astype(np.float64))
        dv.free()
        print(f"noalloc: {name}: {st.iterations} iterations ({st_ref.iterations} in the "
              f"context), rel residual {st.rel_residual:.3e}", flush=True)
        assert rc == 0 and st.converged
        assert (st.iterations, st.outer, st.breakdown) == \
            (st_ref.iterations, st_ref.outer, st_ref.breakdown)
        assert st.rel_residual == st_ref.rel_residual
        assert st.bnorm == st_ref.bnorm
        assert np.array_equal(x, x_ref), f"{name}: max |d| {np.max(np.abs(x - x_ref))}"
        if name == "5x5":
            assert np.allclose(x, xk, atol=1e-10)
    ref.close()


@pytest.mark.gpu
def test_noalloc_rejects_bad_columns_on_the_device(pkg):
    """A column outside [0, n) is caught by the device check before the first
    product (nothing is gathered out of range); the next call on the same
    scratch then solves a valid system."""
    n, W = 300, pkg.ROW_CAP
    # tridiagonal (-1, 4, -1): diagonally dominant, no breakdown
    ln = np.full(n, 3, np.int32)
    ln[0] = ln[-1] = 2
    cs = np.zeros((W, n), np.int32)
    vs = np.zeros((W, n))
    A = np.zeros((n, n))
    for i in range(n):
        cols = [j for j in (i - 1, i, i + 1) if 0 <= j < n]
        for k, j in enumerate(cols):
            cs[k, i] = j
            vs[k, i] = 4.0 if j == i else -1.0
            A[i, j] = vs[k, i]
    b = np.linspace(1.0, 2.0, n)
    bad = cs.copy()
    bad[1, 17] = n + 5
    dv = DeviceArrays()
    t = [dv.put(a) for a in (ln, bad, vs, b, np.zeros(n))]
    print("noalloc: bad column", flush=True)
    rc, _ = pkg.ppe_solve_noalloc(n, n, *t, 1e-12, 100, check=False)
    assert rc == -2
    assert b"column outside" in pkg.lib().isph_last_error(None)
    t[1] = dv.put(cs)
    print("noalloc: tridiagonal", flush=True)
    rc, st = pkg.ppe_solve_noalloc(n, n, *t, 1e-12, 100)
    assert rc == 0 and st.converged
    assert np.allclose(dv.get(t[4], np.zeros(n)), np.linalg.solve(A, b), rtol=0, atol=1e-10)
    dv.free()
