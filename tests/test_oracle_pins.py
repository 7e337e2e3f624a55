"""Pins of the CPU oracle against what the mathematics and the reference fix.

The reference cannot be compiled here (its modules include MKL interface
files the image lacks), so the oracle is pinned by closed forms, invariants,
brute force on tiny inputs and the one known answer the reference holds
(the 5x5 system of solvertest.f90:10-16, exact solution 1..5).  Each check is
chosen so that a plausible transcription mistake (a sign, an index, a
transposed operand, a dropped term) fails it.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def L(oracle):
    return oracle.lib()


H = 0.0065


# ---- kernel.f90 -----------------------------------------------------------
def test_kernel_normalisation_with_reference_constants(L):
    # 2-D integral of W over the support: the reference's real4 3.14
    # constants make it 1.000507, not 1 (kernel.f90:61-62)
    r = np.linspace(0.0, 2 * H, 40001)
    w = np.array([L.or_wab(x, H) for x in r])
    integral = np.trapezoid(2 * math.pi * r * w, r)
    assert abs(integral - 1.000507) < 2e-5


def test_kernel_continuity_and_support(L):
    assert abs(L.or_wab(H * (1 - 1e-12), H) - L.or_wab(H, H)) < 1e-6 * L.or_wab(H, H)
    assert L.or_wab(2 * H, H) == 0.0
    assert L.or_wo(H) == L.or_wab(0.0, H)


def test_kernel_gradient_is_derivative_of_kernel(L):
    # Wabx(p1=j, p2=i) = dW/dx_i: compare with a central difference of W
    for q in (0.3, 0.7, 0.95, 1.05, 1.4, 1.9):
        for ang in (0.0, 0.6, 2.0, 3.5):
            dx, dy = q * H * math.cos(ang), q * H * math.sin(ang)
            r = math.hypot(dx, dy)
            e = 1e-9
            num = (L.or_wab(math.hypot(dx + e, dy), H) - L.or_wab(math.hypot(dx - e, dy), H)) / (2 * e)
            ana = L.or_wabd(dx, r, H)
            assert abs(ana - num) <= 1e-5 * abs(num) + 1e-3, (q, ang, ana, num)
            # antisymmetric in i <-> j
            assert L.or_wabd(-dx, r, H) == -ana


def test_inverse_2x2_closed_form_and_fallback(L):
    import ctypes
    a = np.array([2.0, 0.5, -0.25, 3.0])
    m = a.copy()
    L.or_inv2(m.ctypes.data_as(ctypes.c_void_p))
    M = np.array([[a[0], a[1]], [a[2], a[3]]])
    Mi = np.array([[m[0], m[1]], [m[2], m[3]]])
    assert np.allclose(M @ Mi, np.eye(2), atol=1e-15)
    # |det| below the real4 1e-5 -> identity (functions.f90:75-88)
    s = np.array([1e-3, 0.0, 0.0, 1e-3])
    L.or_inv2(s.ctypes.data_as(ctypes.c_void_p))
    assert list(s) == [1.0, 0.0, 0.0, 1.0]
    s = np.array([math.nan, 0.0, 0.0, 1.0])
    L.or_inv2(s.ctypes.data_as(ctypes.c_void_p))
    assert list(s) == [1.0, 0.0, 0.0, 1.0]


def test_viscosity_law(L):
    assert abs(L.or_vis(1000.0) - 1.002e-3) < 1e-18
    s = 25.0
    assert abs(L.or_vis(1025.0) - (1.002 + 0.001652 * s + 0.0000083 * s * s) * 1e-3) < 1e-17


def test_reduction_tree_is_exact_on_integers_and_close_on_reals(L):
    import ctypes
    rng = np.random.default_rng(0)
    for n in (1, 255, 256, 257, 70000):
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        assert L.or_reduce(n, v.ctypes.data_as(ctypes.c_void_p)) == float(v.sum())
        x = rng.standard_normal(n)
        got = L.or_reduce(n, x.ctypes.data_as(ctypes.c_void_p))
        assert abs(got - math.fsum(x)) <= 1e-12 * np.abs(x).sum()


# ---- solver: the reference's own known answer ----------------------------
def test_solver_known_answer_5x5(oracle, cases):
    # solvertest.f90:10-16, kept as data in tests/golden/solvertest_5x5.json:
    # a 1-based CSR system with exact solution x = 1..5
    import ctypes
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "solvertest_5x5.json")) as f:
        g = json.load(f)
    rowp = [r - 1 for r in g["row"]]
    col = [c - 1 for c in g["col"]]
    n, W = len(rowp) - 1, oracle.RCAP
    ln = np.array([rowp[i + 1] - rowp[i] for i in range(n)], np.int32)
    C = np.zeros((n, W), np.int32)
    V = np.zeros((n, W))
    for i in range(n):
        C[i, :ln[i]] = col[rowp[i]:rowp[i + 1]]
        V[i, :ln[i]] = g["val"][rowp[i]:rowp[i + 1]]
    b = np.array(g["b"], dtype=np.float64)
    x = np.zeros(n)
    c = cases.tiny_box()
    p = dict(c.params)
    p.update(tol=1e-14, max_iter=200)
    o = oracle.Oracle(c, params=p)
    st = oracle.SolveStats()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = oracle.lib().or_solve_system(o.s, n, P(ln), P(C), P(V), P(b), P(x), ctypes.byref(st))
    assert rc == 0
    assert np.allclose(x, g["x"], atol=1e-10)


# ---- pipeline phases on a small jittered box ------------------------------
@pytest.fixture(scope="module")
def box(oracle, cases):
    c = cases.tiny_box(nx=14, ny=12, solids=False, patch=True)
    o = oracle.Oracle(c)
    assert o.call("init_state") == 0
    return c, o


def test_neighbour_sets_match_brute_force(box):
    c, o = box
    x, y = o.arr("x"), o.arr("y")
    h = c.params["h"]
    cnt, nb = o.arr("ncnt"), o.arr("nbr")
    for i in range(o.n):
        r = np.sqrt((x[i] - x) ** 2 + (y[i] - y) ** 2)
        want = np.flatnonzero((r < 2 * h) & (np.arange(o.n) != i))
        assert list(nb[i, :cnt[i]]) == list(want), i


def test_correction_matrix_reproduces_linear_fields(box):
    # sum_j V_j (x_j - x_i) gradW~ = I for every fluid particle: pins coff as
    # the inverse of M with the reference's index order (search.f90:77-85,
    # 272-278); a transposed M fails on anisotropic neighbourhoods
    c, o = box
    x, y, m, rho = o.arr("x"), o.arr("y"), o.arr("m"), o.arr("rho")
    tid, cf = o.arr("tid"), o.arr("coff")
    cnt, nb, k1, k2 = o.arr("ncnt"), o.arr("nbr"), o.arr("klt1"), o.arr("klt2")
    worst = 0.0
    for i in np.flatnonzero(tid == 3):
        j = nb[i, :cnt[i]]
        V = m[j] / rho[j]
        gx = cf[i, 0] * k1[i, :cnt[i]] + cf[i, 1] * k2[i, :cnt[i]]
        gy = cf[i, 2] * k1[i, :cnt[i]] + cf[i, 3] * k2[i, :cnt[i]]
        G = np.array([[np.sum(V * (x[j] - x[i]) * gx), np.sum(V * (y[j] - y[i]) * gx)],
                      [np.sum(V * (x[j] - x[i]) * gy), np.sum(V * (y[j] - y[i]) * gy)]])
        worst = max(worst, np.abs(G - np.eye(2)).max())
    assert worst < 1e-10


def test_porosity_without_solids_is_one(box):
    c, o = box
    por = o.arr("por")
    assert np.all(por == 1.0)
    assert np.all(o.arr("resist") == 0.0)


def test_still_water_intermediate_velocity(oracle, cases):
    # u = 0, nut = tke = 0: every viscous/SPS/tke term vanishes, so
    # u* = 0 and v* = g n dt exactly (integrator.f90:679-685)
    c = cases.tiny_box(nx=12, ny=10, solids=False)
    c.parts["u"][:] = 0.0
    c.parts["v"][:] = 0.0
    o = oracle.Oracle(c)
    o.call("init_state")
    o.call("int_vel")
    fl = o.arr("tid") == 3
    assert np.all(o.arr("us")[fl] == 0.0)
    g, dt = c.params["g"], c.params["dt"]
    assert np.all(o.arr("vs")[fl] == (g * o.arr("por")[fl]) * dt)


def test_eddy_viscosity_of_linear_shear(oracle, cases):
    # u = gamma*y: S = gamma, nu_t = (0.15 dl)^2 gamma, k = 0.08 (dl gamma)^2
    # (turbulence.f90:80-87) wherever the corrected gradient is exact
    c = cases.tiny_box(nx=14, ny=12, solids=False, jitter=0.0)
    gamma = 3.0
    c.parts["u"][:] = gamma * c.parts["y"]
    c.parts["v"][:] = 0.0
    # walls move with the same field so every neighbourhood sees it
    o = oracle.Oracle(c)
    o.call("init_state")
    nut, tke = o.arr("nut"), o.arr("tke")
    fl = np.flatnonzero(o.arr("tid") == 3)
    dl = c.params["dl"]
    assert np.allclose(nut[fl], (0.15 * dl) ** 2 * gamma, rtol=1e-9)
    assert np.allclose(tke[fl], 0.08 * (dl * gamma) ** 2, rtol=1e-9)


def test_ppe_rows_unit_diagonal_and_zero_sum(box, oracle):
    c, o = box
    assert o.call("ppe_assemble") == 0
    ln, col, val = o.arr("alen"), o.arr("acol"), o.arr("aval")
    tid, flags = o.arr("tid"), o.arr("flags")
    for i in range(o.n):
        cs, vs = col[i, :ln[i]], val[i, :ln[i]]
        assert list(cs) == sorted(cs)           # columns ascending
        assert vs[list(cs).index(i)] == 1.0      # Jacobi-scaled diagonal
        if tid[i] == 3 and not (flags[i] & 2):  # non-free fluid: lamp = 1
            assert abs(vs.sum()) < 1e-12


def test_ppe_solution_matches_direct_solve(box, oracle):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    c, o = box
    assert o.call("ppe_assemble") == 0
    n = o.n
    ln, col, val, b = o.arr("alen"), o.arr("acol"), o.arr("aval"), o.arr("b")
    rows = np.repeat(np.arange(n), ln)
    A = sp.csr_matrix((np.concatenate([val[i, :ln[i]] for i in range(n)]),
                       (rows, np.concatenate([col[i, :ln[i]] for i in range(n)]))), shape=(n, n))
    x_direct = spl.spsolve(A.tocsc(), b)
    rc, st = o.ppe_solve()
    assert rc == 0 and st.rel_residual <= c.params["tol"]
    p = o.arr("p")
    assert np.linalg.norm(p - x_direct) <= 1e-6 * np.linalg.norm(x_direct)


def test_uniform_scalar_does_not_diffuse(oracle, cases):
    c = cases.tiny_box(nx=12, ny=10, solids=False, patch=False)
    o = oracle.Oracle(c)
    o.call("init_state")
    o.call("scalart")
    fl = o.arr("tid") == 3
    assert np.all(o.arr("cdiff")[fl] == 0.0)


def test_density_update_closed_form(oracle, cases):
    c = cases.tiny_box(nx=12, ny=10, solids=False, patch=True)
    o = oracle.Oracle(c)
    o.call("init_state")
    o.call("densityupdate")
    fl = o.arr("tid") == 3
    cref = o.arr("c")[fl]
    rho0 = o.arr("rho0")[fl]
    assert np.allclose(rho0, 1012.5 + 25.0 * cref, rtol=0, atol=1e-12)
    assert set(np.round(rho0, 9)) <= {1000.0, 1025.0}


def test_time_step_substeps_closed_form(oracle, cases):
    c = cases.tiny_box(nx=12, ny=10, solids=False)
    o = oracle.Oracle(c)
    o.call("init_state")
    s = o.scalars()
    dl = c.params["dl"]
    dtsol = (0.2 * dl * dl) / (float(np.float32(1e-9)) + s.numax / 1.0)
    want_steps = 1 if dtsol >= s.dt else math.ceil(s.dt / dtsol)
    assert s.solsteps == want_steps
    assert s.dtsol == (s.dt / want_steps if want_steps > 1 else s.dt)


def test_rebin_preserves_particles_and_orders_cells(oracle, cases):
    # after a rebin: a permutation of the particles (none dropped inside the
    # tank), cell keys non-decreasing, walls first in every cell, and every
    # particle inside its cell's half-open bounds (setup.f90:120-123)
    c = cases.tiny_box(nx=14, ny=12, solids=False)
    o = oracle.Oracle(c)
    o.call("init_state")
    before = sorted(zip(o.arr("x"), o.arr("y")))
    o.call("projection")
    o.call("cellshift")
    cell = o.arr("cell")
    assert np.all(np.diff(cell) >= 0)
    tid = o.arr("tid")
    cs = o.cstart()
    for k in np.unique(cell):
        seg = tid[cs[k]:cs[k + 1]]
        nwall = int(np.sum(seg != 3))
        assert np.all(seg[:nwall] != 3)
    celly = c.params["celly"]
    h = c.params["h"]
    x, y = o.arr("x"), o.arr("y")
    cj, ci = cell // celly, cell % celly
    assert np.all((x >= (cj - 1) * 4 * h) & (x < cj * 4 * h))
    assert np.all((y >= (celly - 2 - ci) * 4 * h) & (y < (celly - 1 - ci) * 4 * h))
    assert o.n == len(before)


# ---- pins of the flow phases, each shown to fail on a planted mistake ------
# A planted mistake is a build of the oracle with -DOR_MUTANT=k (the MUT(k)
# marks in oracle/isph_oracle.c): every pin below passes on the oracle and
# fails on the mutant named next to it.
MUTANTS = {
    1: "comp_vel: factor 2 of the fluid pressure term dropped",
    2: "comp_vel: n_i^2 written as n_i",
    3: "opt2_shift: shift taken up the concentration gradient",
    4: "opt2_shift: x and y gradient terms swapped in the velocity correction",
    5: "effpor: lower-left quadrant walks the right-hand column",
    6: "effpor: Forchheimer term with n^2 in place of n^3",
    7: "int_vel: factor 2 of the viscous term dropped",
    8: "int_vel: sign of the tke term flipped",
    9: "int_vel: Darcy-Forchheimer resistance explicit instead of implicit",
    10: "PPE rhs: u*_i - u*_j in place of u*_j - u*_i",
    11: "PPE rhs: x velocity paired with the y gradient",
    12: "scalar: arithmetic mean of the diffusivities without the 1/2",
    13: "scalar: (c_j - c_i) in place of (c_i - c_j)",
    14: "projection: u times n in place of u / n",
    15: "comp_pos: projected position in place of the saved one",
    16: "freesurf: a surface particle flags itself, not its neighbours",
    17: "int_vel: the y term dropped from the free-surface velocity gradient",
    18: "PPE: ghost rows take entries for fluid neighbours",
    19: "int_vel: factor 2 of the SPS harmonic mean dropped",
}


@pytest.fixture(scope="session")
def mutant(oracle, tmp_path_factory):
    """mutant(k): the oracle library with planted mistake k."""
    import concurrent.futures as cf
    d = tmp_path_factory.mktemp("oracle_mutants")
    with cf.ThreadPoolExecutor(8) as ex:
        paths = dict(zip(MUTANTS, ex.map(lambda k: oracle.build_variant(d, k), MUTANTS)))
    return lambda k: oracle.lib(paths[k])


def regular_box(cases, nx=24, ny=20, dx=0.01, jitter=0.0, seed=3, solids=False):
    """Fluid filling a walled tank [off, off + nx dx] x [off, off + ny dx] (the
    walls go 4 dx higher), on a regular lattice unless jitter is set.  With
    `solids`, porous solids on the same lattice cover every interior cell, so
    each fluid particle's kernel sum exceeds 1 - por and the clamp gives the
    uniform porosity n = por (porous.f90:182-190)."""
    P = cases.default_params(dx)
    P.update(dt=2.0e-4, tol=1e-12, por=0.4, d50=0.01)
    off = 3.5 * dx
    wx, wy, wt = cases._walls_box(off, off + nx * dx, off, off + (ny + 4) * dx, dx)
    gx, gy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    fx = off + (gx.ravel() + 0.5) * dx
    fy = off + (gy.ravel() + 0.5) * dx
    if jitter:
        rng = np.random.default_rng(seed)
        fx = fx + rng.uniform(-jitter, jitter, fx.size) * dx
        fy = fy + rng.uniform(-jitter, jitter, fy.size) * dx
    cases._grid(P, off + (nx + 4) * dx, off + (ny + 8) * dx)
    sx = sy = None
    if solids:
        cs = 4 * P["h"]
        g1, g2 = np.meshgrid(np.arange(int((P["cellx"] - 2) * cs / dx)),
                             np.arange(int((P["celly"] - 2) * cs / dx)), indexing="ij")
        sx = (g1.ravel() + 0.5) * dx
        sy = (g2.ravel() + 0.5) * dx
    return cases._assemble(P, "regular_box", fx, fy, np.zeros(fx.size), wx, wy, wt, dx,
                           off + ny * dx, sx, sy, meta=dict(nx=nx, ny=ny, dx=dx, off=off))


def interior(c, o, away_from_walls):
    """Fluid particles whose 2h support lies below the free surface (and, with
    away_from_walls, holds no wall particle)."""
    x, y, tid = o.arr("x"), o.arr("y"), o.arr("tid")
    m, h = c.meta, c.params["h"]
    top = m["off"] + m["ny"] * m["dx"]
    k = (tid == 3) & (y < top - 2 * h - m["dx"])
    if away_from_walls:
        k &= (x > m["off"] + 2 * h + m["dx"]) & (x < m["off"] + m["nx"] * m["dx"] - 2 * h - m["dx"])
        k &= y > m["off"] + 2 * h + m["dx"]
    return k


def porous_state(oracle, c, L):
    """init_state, then neighbour and effpor again, so the correction matrix
    is built with the porous particle volumes m/rho = V0/n (the order of a
    production step: effpor's density feeds the next neighbour search)."""
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    if c.ns:
        assert o.call("neighbour") == 0
        o.call("effpor")
    return o


# comp_vel, integrator.f90:705-843 ------------------------------------------
def hydrostatic_residual(oracle, cases, L, solids):
    """Still water: int_vel gives v* = g n dt; with the hydrostatic pressure
    p = rho0 |g| (H - y) comp_vel must return u = v = 0, since the pressure
    term is (n/rho0) grad p for a linear p and uniform n."""
    c = regular_box(cases, solids=solids)
    o = porous_state(oracle, c, L)
    y = o.arr("y")
    top = c.meta["off"] + c.meta["ny"] * c.meta["dx"]
    o.put("p", c.params["rho"] * abs(c.params["g"]) * (top - y))
    o.put("nut", 0.0)
    o.put("tke", 0.0)
    o.call("int_vel")
    o.call("comp_vel")
    k = interior(c, o, away_from_walls=solids)
    assert k.sum() > 100
    if solids:
        assert np.allclose(o.arr("por")[k], o.arr("por")[k][0], rtol=1e-12)
        assert o.arr("por")[k][0] < 0.5
    return max(np.abs(o.arr("u")[k]).max(), np.abs(o.arr("v")[k]).max())


def test_comp_vel_keeps_hydrostatic_water_at_rest(oracle, cases, mutant):
    # n = 1 (walls included in the support) and uniform n < 1 (away from the
    # walls, whose volume m/rho stays V0): |u| <= 1e-12 against g dt ~ 2e-3
    L = oracle.lib()
    assert hydrostatic_residual(oracle, cases, L, False) < 1e-12
    assert hydrostatic_residual(oracle, cases, L, True) < 1e-12
    assert hydrostatic_residual(oracle, cases, mutant(1), False) > 1e-4
    assert hydrostatic_residual(oracle, cases, mutant(2), True) > 1e-4


# projection isph.f90:30-37 and comp_pos integrator.f90:859-863 --------------
def kinematics_error(oracle, cases, L, solids):
    """projection with u_old, then comp_pos with u_new = u_old + a dt: the
    trapezoid rule integrates a constant acceleration exactly, and particles
    move at the pore velocity u/n: x = x_old + (u_old dt + a dt^2 / 2) / n."""
    c = regular_box(cases, nx=16, ny=12, jitter=0.05, solids=solids)
    o = porous_state(oracle, c, L)
    fl = o.arr("tid") == 3
    x0, y0, n = o.arr("x"), o.arr("y"), o.arr("por")
    U, V, ax, ay, dt = 0.2, -0.1, 3.0, -5.0, c.params["dt"]
    o.put("u", U)
    o.put("v", V)
    o.call("projection")
    o.put("u", U + ax * dt)
    o.put("v", V + ay * dt)
    o.call("comp_pos")
    x1, y1 = o.arr("x"), o.arr("y")
    assert np.all(x1[~fl] == x0[~fl]) and np.all(y1[~fl] == y0[~fl])
    ex = np.abs(x1 - (x0 + (U * dt + 0.5 * ax * dt * dt) / n))[fl]
    ey = np.abs(y1 - (y0 + (V * dt + 0.5 * ay * dt * dt) / n))[fl]
    return max(ex.max(), ey.max())


def test_projection_and_comp_pos_integrate_constant_acceleration(oracle, cases, mutant):
    L = oracle.lib()
    for solids in (False, True):
        assert kinematics_error(oracle, cases, L, solids) < 1e-15
    assert kinematics_error(oracle, cases, mutant(14), True) > 1e-6
    assert kinematics_error(oracle, cases, mutant(15), False) > 1e-6


# freesurf, isph.f90:49-86 ------------------------------------------------------
def vicinity_mismatches(oracle, cases, L):
    """The vicinity flags equal a brute-force search: every particle within
    2h of a non-buffer fluid particle with gradvx <= fs_thresh."""
    c = cases.tiny_box(nx=14, ny=12, solids=False, patch=False)
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    before = o.arr("flags").copy()
    o.call("freesurf")
    x, y, tid, gv = o.arr("x"), o.arr("y"), o.arr("tid"), o.arr("gradvx")
    surf = np.flatnonzero((tid == 3) & ((before & 1) == 0) & (gv <= c.params["fs_thresh"]))
    assert surf.size > 10
    want = (before & 4) > 0
    for i in surf:
        near = np.hypot(x[i] - x, y[i] - y) < 2 * c.params["h"]
        near[i] = False
        want |= near
    return int(np.sum(((o.arr("flags") & 4) > 0) != want))


def test_freesurf_vicinity_matches_brute_force(oracle, cases, mutant):
    assert vicinity_mismatches(oracle, cases, oracle.lib()) == 0
    assert vicinity_mismatches(oracle, cases, mutant(16)) > 0


# opt2_shift, part_shift.f90:12-191 ---------------------------------------------
def shift_errors(oracle, cases, L):
    """The shift is -csh h^2 frac grad C, C(x) = sum_j V_j W(|x - x_j|) the
    particle concentration (grad C by central differences of the kernel sum,
    an independent route to the analytic gradient sum), capped at
    maxshift dl; and for a linear velocity field the velocity correction is
    exactly grad u . shift.  Returns the worst shift error relative to the
    cap and the worst correction error relative to |shift|."""
    c = regular_box(cases, nx=16, ny=12, jitter=0.08)
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    x, y, tid, flags = o.arr("x"), o.arr("y"), o.arr("tid"), o.arr("flags")
    V = o.arr("m") / o.arr("rho")
    a, b, cc, d = 0.3, -0.2, 0.15, 0.25
    o.put("u", a * x + b * y)
    o.put("v", cc * x + d * y)
    o.call("opt2_shift")
    xs, ys, us, vs = o.arr("xs"), o.arr("ys"), o.arr("us"), o.arr("vs")
    P = c.params
    h, cap, e = P["h"], P["maxshift"] * P["dl"], 1e-6 * P["h"]
    ws = wl = 0.0
    for i in np.flatnonzero(tid == 3):
        js = np.flatnonzero((np.hypot(x[i] - x, y[i] - y) < 2 * h) & (np.arange(x.size) != i))

        def conc(px, py):
            return math.fsum(V[j] * L.or_wab(math.hypot(px - x[j], py - y[j]), h) for j in js)
        gx = (conc(x[i] + e, y[i]) - conc(x[i] - e, y[i])) / (2 * e)
        gy = (conc(x[i], y[i] + e) - conc(x[i], y[i] - e)) / (2 * e)
        frac = 0.01 if flags[i] & 2 else 1.0
        dx, dy = -P["csh"] * h * h * frac * gx, -P["csh"] * h * h * frac * gy
        ws = max(ws, abs(xs[i] - math.copysign(min(abs(dx), cap), dx)) / cap,
                 abs(ys[i] - math.copysign(min(abs(dy), cap), dy)) / cap)
        if abs(dx) < cap and abs(dy) < cap:
            sc = abs(xs[i]) + abs(ys[i])
            wl = max(wl, abs(us[i] - (a * xs[i] + b * ys[i])) / sc,
                     abs(vs[i] - (cc * xs[i] + d * ys[i])) / sc)
    return ws, wl


def test_shift_follows_concentration_gradient_and_corrects_velocity(oracle, cases, mutant):
    ws, wl = shift_errors(oracle, cases, oracle.lib())
    assert ws < 1e-7 and wl < 1e-12
    assert shift_errors(oracle, cases, mutant(3))[0] > 0.1
    assert shift_errors(oracle, cases, mutant(4))[1] > 0.1


# effpor / porsearch, porous.f90:32-224 -----------------------------------------
def porosity_errors(oracle, cases, L, lamk):
    """Porosity against a brute-force sum over every solid within 2h (no
    cell walk): n = 1 - min(sum (1 - por) W m_s / rho_s, 1 - por); and with
    lamk = 150 the resistance is Ergun's equation per unit mass (Ergun 1952):
    -(150 nu (1-n)^2 / (n^3 d^2) + 1.75 (1-n) |u| / (n^3 d)), nu = mu(rho0)/rho0.
    Returns the worst porosity difference, the worst relative resistance
    difference, and the counts of porous and clamped particles."""
    c = cases.tiny_box(nx=14, ny=12, solids=True, patch=False)
    c.params["lamk"] = lamk
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    x, y = o.arr("x"), o.arr("y")
    fl = np.flatnonzero(o.arr("tid") == 3)
    S = c.solids
    h, por, d = c.params["h"], c.params["por"], c.params["d50"]
    nb = np.empty(fl.size)
    for q, i in enumerate(fl):
        r = np.hypot(x[i] - S["x"], y[i] - S["y"])
        sf = math.fsum((1.0 - por) * L.or_wab(r[j], h) * S["m"][j] / S["rho"][j]
                       for j in np.flatnonzero(r < 2 * h))
        nb[q] = 1.0 - min(sf, 1.0 - por)
    n = o.arr("por")[fl]
    u, v, rho0 = o.arr("u")[fl], o.arr("v")[fl], o.arr("rho0")[fl]
    nu = np.array([L.or_vis(r0) for r0 in rho0]) / rho0
    ergun = -(150.0 * nu * (1 - n) ** 2 / (n ** 3 * d * d) +
              1.75 * (1 - n) * np.hypot(u, v) / (n ** 3 * d))
    pm = n < 1.0
    er = np.abs(o.arr("resist")[fl][pm] / ergun[pm] - 1).max()
    return np.abs(n - nb).max(), er, int(pm.sum()), int(np.sum(n == por))


def test_effpor_matches_brute_force_and_ergun(oracle, cases, mutant):
    en, er, npor, nclamp = porosity_errors(oracle, cases, oracle.lib(), 150.0)
    assert npor > 20 and nclamp > 5 and en < 1e-14 and er < 1e-12
    assert porosity_errors(oracle, cases, mutant(5), 150.0)[0] > 1e-3
    assert porosity_errors(oracle, cases, mutant(6), 150.0)[1] > 1e-2


# int_vel, integrator.f90:459-703 ------------------------------------------------
def viscous_ratio(oracle, cases, L, nut):
    """u = 3 x^2, v = -2 y^2 (Laplacians 6 and -4), uniform nu_t: away from
    walls and surface, (u* - u) / (dt (nu + nu_t)) is the Laplacian, to the
    operator's lattice error (the 1e-3 h^2 regulariser costs ~0.1%).
    Returns the range of the ratio to the exact Laplacian."""
    c = regular_box(cases, nx=30, ny=24)
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    x, y = o.arr("x"), o.arr("y")
    u0, v0 = 3.0 * x * x, -2.0 * y * y
    o.put("u", u0)
    o.put("v", v0)
    o.put("nut", nut)
    o.put("tke", 0.0)
    o.call("int_vel")
    k = interior(c, o, away_from_walls=True)
    nu = L.or_vis(1000.0) / 1000.0
    dt, g = c.params["dt"], c.params["g"]
    r = np.concatenate([(o.arr("us")[k] - u0[k]) / (dt * (nu + nut)) / 6.0,
                        (o.arr("vs")[k] - v0[k] - g * dt) / (dt * (nu + nut)) / -4.0])
    return r.min(), r.max()


def test_int_vel_viscous_and_sps_terms_are_the_laplacian(oracle, cases, mutant):
    # nu_t below the real4 1e-6 threshold (arithmetic mean) and above it
    # (harmonic mean) both reduce to nu_t for a uniform field
    for nut in (0.0, 5e-7, 2e-5):
        lo, hi = viscous_ratio(oracle, cases, oracle.lib(), nut)
        assert 0.99 < lo <= hi < 1.01, (nut, lo, hi)
    assert viscous_ratio(oracle, cases, mutant(7), 0.0)[0] < 0.9
    assert viscous_ratio(oracle, cases, mutant(19), 2e-5)[0] < 0.9


def tke_error(oracle, cases, L):
    """u = v = 0, k = 0.7 x - 0.4 y: the tke term -(2/3) n grad k is exact
    for a linear k (corrected gradient), so u* = -(2/3)(0.7) dt and
    v* = -(2/3)(-0.4) dt + g dt for every fluid particle."""
    c = regular_box(cases, jitter=0.05)
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    x, y = o.arr("x"), o.arr("y")
    o.put("u", 0.0)
    o.put("v", 0.0)
    o.put("nut", 0.0)
    o.put("tke", 0.7 * x - 0.4 * y)
    o.call("int_vel")
    fl = o.arr("tid") == 3
    dt, g = c.params["dt"], c.params["g"]
    eu = o.arr("us")[fl] / (-(2.0 / 3.0) * 0.7 * dt) - 1
    ev = (o.arr("vs")[fl] - g * dt) / (-(2.0 / 3.0) * -0.4 * dt) - 1
    return max(np.abs(eu).max(), np.abs(ev).max())


def test_int_vel_tke_term_is_the_gradient(oracle, cases, mutant):
    assert tke_error(oracle, cases, oracle.lib()) < 1e-9
    assert tke_error(oracle, cases, mutant(8)) > 0.5


def darcy_error(oracle, cases, L):
    """Uniform flow U through uniform porosity: every flow term vanishes and
    int_vel is the implicit Euler step of du/dt = R n u, u* = U / (1 - dt R n)."""
    c = regular_box(cases, solids=True)
    o = porous_state(oracle, c, L)
    U = 0.3
    o.put("u", U)
    o.put("v", 0.0)
    o.put("nut", 0.0)
    o.put("tke", 0.0)
    o.call("effpor")  # resistance at the new speed
    o.call("int_vel")
    fl = o.arr("tid") == 3
    R, n, dt = o.arr("resist")[fl], o.arr("por")[fl], c.params["dt"]
    assert R.max() < -100.0
    return np.abs(o.arr("us")[fl] * (1.0 - dt * R * n) / U - 1).max()


def test_int_vel_darcy_forchheimer_is_implicit(oracle, cases, mutant):
    assert darcy_error(oracle, cases, oracle.lib()) < 1e-14
    assert darcy_error(oracle, cases, mutant(9)) > 1e-4


def varts_error(oracle, cases, L):
    """Free-surface particles with u = 0.4 x, v = -0.3 y (n = 1): the
    corrected velocity gradient is exact, so
    varts = dl^2 |(0.4, -0.3)| |u_i| / ((umax + 1e-6) n)."""
    c = regular_box(cases, jitter=0.05)
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    x, y = o.arr("x"), o.arr("y")
    o.put("u", 0.4 * x)
    o.put("v", -0.3 * y)
    o.put("nut", 0.0)
    o.put("tke", 0.0)
    umax = o.scalars().umax
    o.call("int_vel")
    free = (o.arr("flags") & 2) > 0
    assert free.sum() > 10
    dl = c.params["dl"]
    want = dl * dl * 0.5 * np.hypot(0.4 * x, -0.3 * y) / (
        (umax + float(np.float32(1e-6))) * o.arr("por"))
    return np.abs(o.arr("varts")[free] / want[free] - 1).max()


def test_int_vel_varts_is_the_velocity_gradient(oracle, cases, mutant):
    assert varts_error(oracle, cases, oracle.lib()) < 1e-12
    assert varts_error(oracle, cases, mutant(17)) > 0.1


# PPE right-hand side and wall rows, isph.f90:208-410 ---------------------------------
def _row_apply(ln, col, val, q):
    return np.array([np.dot(val[i, :ln[i]], q[col[i, :ln[i]]]) for i in range(len(ln))])


def ppe_rhs_checks(oracle, cases, L, jitter):
    """u* linear, n = 1: the raw source sum_j V_j (u*_j - u*_i) rho / dt . grad W~
    is rho div(u*) / dt exactly, so after the Jacobi scaling every non-free
    fluid row's rhs is proportional to div(u*) (ratio of two fields exact,
    zero for a divergence-free field).  The scale: the same row applied to
    q = x^2 + y^2 gives lap(q) / diag = 4 / diag, so rhs / (A q) =
    rho div / (4 dt) to the Laplacian's lattice error (regular lattice).
    Returns (ratio error, divergence-free residual, scale range)."""
    c = regular_box(cases, nx=30, ny=24, jitter=jitter)
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    x, y, tid = o.arr("x"), o.arr("y"), o.arr("tid")
    dt, rho = c.params["dt"], c.params["rho"]
    rhs = {}
    for name, (a, b, cc, d) in {"f1": (0.5, 0.2, -0.3, 0.9), "f2": (0.25, 0.7, 0.1, 0.0),
                                "free": (0.6, 0.2, -0.3, -0.6)}.items():
        o.put("us", a * x + b * y)
        o.put("vs", cc * x + d * y)
        assert o.call("ppe_assemble") == 0
        rhs[name] = o.arr("b").copy()
    nf = (tid == 3) & ((o.arr("flags") & 2) == 0)
    ratio = np.abs(rhs["f1"][nf] * 0.25 / (rhs["f2"][nf] * 1.4) - 1).max()
    zero = np.abs(rhs["free"][nf]).max() / np.abs(rhs["f1"][nf]).max()
    Aq = _row_apply(o.arr("alen"), o.arr("acol"), o.arr("aval"), x * x + y * y)
    k = interior(c, o, away_from_walls=False) & nf
    sc = rhs["f1"][k] / Aq[k] / (rho * 1.4 / (4 * dt))
    return ratio, zero, (sc.min(), sc.max())


def test_ppe_rhs_is_the_scaled_divergence(oracle, cases, mutant):
    ratio, zero, _ = ppe_rhs_checks(oracle, cases, oracle.lib(), 0.05)
    assert ratio < 1e-12 and zero < 1e-12
    _, _, (lo, hi) = ppe_rhs_checks(oracle, cases, oracle.lib(), 0.0)
    assert 0.99 < lo <= hi < 1.01
    lo, hi = ppe_rhs_checks(oracle, cases, mutant(10), 0.0)[2]
    assert not (0.99 < lo <= hi < 1.01)
    r11 = ppe_rhs_checks(oracle, cases, mutant(11), 0.05)
    assert r11[0] > 0.1 and r11[1] > 0.1


def wall_row_checks(oracle, cases, L):
    """Ghost rows (isph.f90:272-293) couple only to wall particles, sum to
    zero and have no source; fixed-wall rows (isph.f90:302-410) sum to zero
    (the diagonal is the sum of the off-diagonal terms in both)."""
    c = regular_box(cases, jitter=0.05)
    o = oracle.Oracle(c, library=L)
    assert o.call("init_state") == 0
    x, y = o.arr("x"), o.arr("y")
    o.put("us", 0.3 * x - 0.1 * y)
    o.put("vs", 0.2 * y)
    assert o.call("ppe_assemble") == 0
    tid, ln, col, val, b = (o.arr(k) for k in ("tid", "alen", "acol", "aval", "b"))
    gh, fx = np.flatnonzero(tid == 2), np.flatnonzero(tid == 1)
    only_walls = all(np.all(tid[col[i, :ln[i]]] <= 2) for i in gh)
    sums = max(abs(val[i, :ln[i]].sum()) for i in np.concatenate([gh, fx]))
    return only_walls, sums, np.abs(b[gh]).max()


def test_ppe_wall_rows(oracle, cases, mutant):
    only_walls, sums, ghost_rhs = wall_row_checks(oracle, cases, oracle.lib())
    assert only_walls and sums < 1e-13 and ghost_rhs == 0.0
    assert not wall_row_checks(oracle, cases, mutant(18))[0]


# scalart_anisotropic, scalar.f90:242-362 ---------------------------------------
def scalar_ratio(oracle, cases, L, solids, quadratic=True):
    """u = v = 0.  Without solids (n = 1 >= 0.7) the coefficient is the
    molecular Dm (arithmetic mean, nu_t = 0); with solids everywhere
    (n = por < 0.7) the anisotropic tensor is 0/0 at zero velocity and the
    cap 0.12 dl^2 / dt applies (merge, scalar.f90:311).  For c = 3(x^2 + y^2)
    cdiff n / (D lap c) is 1 to the lattice error of the uncorrected kernel
    Laplacian (returns its range); for a linear c cdiff is 0 to round-off
    (returns max |cdiff| / (D |grad c| / h))."""
    c = regular_box(cases, nx=30, ny=24, solids=solids)
    o = porous_state(oracle, c, L)
    x, y, tid = o.arr("x"), o.arr("y"), o.arr("tid")
    o.put("u", 0.0)
    o.put("v", 0.0)
    o.put("nut", 0.0)
    o.put("c", 3.0 * (x * x + y * y) if quadratic else 0.7 * x - 0.4 * y)
    o.call("scalart")
    k = interior(c, o, away_from_walls=True) & (tid == 3)
    P = c.params
    D = 0.12 * P["dl"] ** 2 / P["dt"] if solids else float(np.float32(1e-9))
    cd, n = o.arr("cdiff")[k], o.arr("por")[k]
    if quadratic:
        r = cd * n / (D * 12.0)
        return r.min(), r.max()
    return np.abs(cd).max() / (D * 0.81 / P["h"])


def test_scalar_transport_is_the_laplacian_in_both_branches(oracle, cases, mutant):
    L = oracle.lib()
    for solids in (False, True):
        lo, hi = scalar_ratio(oracle, cases, L, solids)
        assert 0.95 < lo <= hi < 1.02, (solids, lo, hi)
        assert scalar_ratio(oracle, cases, L, solids, quadratic=False) < 1e-12
    assert scalar_ratio(oracle, cases, mutant(12), False)[1] > 1.5
    assert scalar_ratio(oracle, cases, mutant(13), True)[1] < 0


# comp_vel's per-cell maxima and timestep, integrator.f90:831-841, 914-924 --------
def test_emptied_cell_keeps_its_velocity_maximum(oracle, cases):
    """comp_vel refreshes dpcell%maxvel only for cells that hold particles and
    timestep takes the maximum over every cell, so a cell a fast particle
    has left still holds its speed.  A lone particle above the water moves
    at 5 m/s; after it is moved one cell on and slowed to 0.1 m/s, umax
    stays 5 while its old cell is empty.  Left in place, it gives 0.1."""
    c = cases.tiny_box(nx=14, ny=8, solids=False, patch=False)
    dx, h = c.meta["dx"], c.params["h"]
    top = 3.5 * dx + (c.meta["ny"] + 5.5) * dx     # inside the walls, above the water
    for k, val in dict(x=3.5 * dx + 9.0 * dx, y=top, u=0.0, v=0.0, p=0.0, m=1000 * dx * dx,
                       rho=1000.0, rho0=1000.0, vol0=dx * dx, c=-0.5).items():
        c.parts[k] = np.append(c.parts[k], val)
    c.parts["tid"] = np.append(c.parts["tid"], np.int32(3)).astype(np.int32)
    c.parts["flags"] = np.append(c.parts["flags"], np.uint8(0)).astype(np.uint8)

    def run(move):
        o = oracle.Oracle(c)
        assert o.call("init_state") == 0
        d = int(np.argmax(np.where(o.arr("tid") == 3, o.arr("y"), -1.0)))
        cell0 = int(o.arr("cell")[d])
        cs = o.cstart()
        assert cs[cell0 + 1] - cs[cell0] == 1          # alone in its cell
        for name in ("p", "us", "vs"):
            o.put(name, 0.0)
        us = o.arr("us")
        us[d] = 5.0
        o.put("us", us)
        o.call("comp_vel")
        o.call("timestep")
        assert o.scalars().umax == 5.0
        if move:
            x = o.arr("x")
            x[d] += 4.0 * h
            o.put("x", x)
            o.call("cellshiftalt")
            assert o.call("neighbour") == 0
            o.call("effpor")
            o.call("eddyvis")
        d = int(np.argmax(np.where(o.arr("tid") == 3, o.arr("y"), -1.0)))
        for name in ("p", "us", "vs"):
            o.put(name, 0.0)
        us = o.arr("us")
        us[d] = 0.1
        o.put("us", us)
        o.call("comp_vel")
        o.call("timestep")
        cs = o.cstart()
        return o.scalars().umax, int(cs[cell0 + 1] - cs[cell0]), o.arr("cmaxvel", c.params["cellx"] * c.params["celly"])[cell0]

    umax, left, kept = run(move=True)
    assert left == 0 and kept == 5.0 and umax == 5.0
    umax, left, kept = run(move=False)
    assert left == 1 and umax == kept == 0.1


# an isolated particle, isph.f90:425-440 ------------------------------------------------
def test_isolated_particle_gets_the_identity_row(oracle, cases):
    """A fluid particle with no neighbour (a droplet) has an all-zero PPE row;
    the reference divides it by its zero diagonal (isph.f90:432-440) and
    leaves the remedy, diagonal 1, commented out (isph.f90:425-427).  Here
    the row is the identity with rhs 0, so the droplet's pressure is 0 (the
    free surface's), the rest of the system is the same, and the step goes on:
    it falls freely, v = v0 + g dt per step."""
    c = cases.tiny_box(nx=14, ny=8, solids=False, patch=False)
    dx = c.meta["dx"]
    for k, val in dict(x=3.5 * dx + 7.0 * dx, y=3.5 * dx + 13.5 * dx, u=0.0, v=0.0, p=0.0,
                       m=1000 * dx * dx, rho=1000.0, rho0=1000.0, vol0=dx * dx,
                       c=-0.5).items():
        c.parts[k] = np.append(c.parts[k], val)
    c.parts["tid"] = np.append(c.parts["tid"], np.int32(3)).astype(np.int32)
    c.parts["flags"] = np.append(c.parts["flags"], np.uint8(0)).astype(np.uint8)
    o = oracle.Oracle(c)
    assert o.call("init_state") == 0
    for ph in ("projection", "cellshift"):
        o.call(ph)
    assert o.call("neighbour") == 0
    for ph in ("effpor", "eddyvis", "int_vel"):
        o.call(ph)
    assert o.call("ppe_assemble") == 0
    d = int(np.argmax(np.where(o.arr("tid") == 3, o.arr("y"), -1.0)))
    assert o.arr("ncnt")[d] == 0
    ln, col, val, b = o.arr("alen"), o.arr("acol"), o.arr("aval"), o.arr("b")
    assert ln[d] == 1 and col[d, 0] == d and val[d, 0] == 1.0 and b[d] == 0.0
    rc, st = o.ppe_solve()
    assert rc == 0 and st.converged
    assert o.arr("p")[d] == 0.0
    o2 = oracle.Oracle(c)
    assert o2.call("init_state") == 0
    v0 = None
    for k in range(3):
        rc, st = o2.step(1)
        assert rc == 0
        y, v, tid = o2.arr("y"), o2.arr("v"), o2.arr("tid")
        d = int(np.argmax(np.where(tid == 3, y, -1.0)))
        if v0 is not None:
            assert abs(v[d] - (v0 + c.params["g"] * c.params["dt"])) < 1e-12
        v0 = v[d]
