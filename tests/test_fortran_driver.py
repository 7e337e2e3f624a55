"""The Fortran side of the boundary: the iso_c_binding module and the driver
that runs the production loop of main.f90:199-261 through the C ABI.

CPU: the case file round trip and the driver's argument and format checks
(neither reaches a HIP call).  GPU: the driver's state after k steps, called
one reference subroutine at a time and fused, equals the Python host's
isph_step run bit for bit.
"""
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
DRIVER = ROOT / "beach-slope_amd" / "isph_driver"


def _need_driver():
    if not DRIVER.exists():
        pytest.skip("Fortran driver not built (no Fortran compiler)")


def test_case_file_round_trip(pkg, cases, tmp_path):
    from beach_slope_amd import caseio
    c = cases.tiny_box(nx=6, ny=5, solids=True, patch=True)
    f = tmp_path / "case.bin"
    caseio.write_case(f, c, time=1.5e-3, displaytime=5e-4)
    params, parts, solids, time, disp = caseio.read_case(f)
    assert time == 1.5e-3 and disp == 5e-4
    want = pkg.Params.from_dict(c.params)
    for name, _ in pkg.Params._fields_:
        assert params[name] == getattr(want, name), name
    for k in caseio.PART_FIELDS:
        assert np.array_equal(parts[k], c.parts[k])
    assert np.array_equal(parts["tid"], c.parts["tid"])
    assert np.array_equal(parts["flags"], c.parts["flags"])
    for k in caseio.SOLID_FIELDS:
        assert np.array_equal(solids[k], c.solids[k])


def test_driver_usage_and_bad_file(tmp_path):
    _need_driver()
    r = subprocess.run([str(DRIVER)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stdout
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\0" * 64)
    r = subprocess.run([str(DRIVER), str(bad), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "not an ISPH case file" in r.stdout


def test_create_failure_is_reported(pkg, cases, tmp_path):
    """A rejected isph_create leaves a null handle and its message readable
    through isph_last_error(NULL): the driver prints it and stops with code 1,
    the Python host raises it.  (The cell grid is invalid, so no HIP call is
    made.)"""
    _need_driver()
    from beach_slope_amd import caseio
    c = cases.tiny_box(nx=6, ny=5)
    bad = dict(c.params, cellx=2)
    f = tmp_path / "case.bin"
    caseio.write_case(f, c, time=1e-3, displaytime=1e-3, params=bad)
    r = subprocess.run([str(DRIVER), str(f), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert "error in isph_create: code -2" in r.stdout and "cell grid" in r.stdout
    with pytest.raises(pkg.IsphError, match="cell grid"):
        pkg.Scenario(c, params=bad)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["phases", "fused"])
def test_driver_matches_host_steps(pkg, cases, tmp_path, mode):
    _need_driver()
    from beach_slope_amd import caseio
    k = 3
    c = cases.tiny_box(nx=10, ny=8, solids=True, patch=True)
    dt = c.params["dt"]
    f = tmp_path / "case.bin"
    out = tmp_path / "state.bin"
    caseio.write_case(f, c, time=(k - 0.5) * dt, displaytime=dt)
    args = [str(DRIVER), str(f), str(out)] + (["fused"] if mode == "fused" else [])
    print(f"driver {mode}: tiny_box n={c.n} steps={k}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"# {k} steps" in r.stdout
    got = caseio.read_state(out)

    s = pkg.Scenario(c)
    s.init_state()
    s.step(k)
    ref = s.download()
    sc = s.scalars()
    s.close()
    assert got["t"] == sc.t and got["dt"] == sc.dt
    for key in caseio.PART_FIELDS + ("tid", "flags"):
        assert np.array_equal(got[key], ref[key]), key


@pytest.mark.gpu
def test_driver_snapshot_files(pkg, cases, tmp_path):
    """snapshots=DIR: the reference's output files (output.f90:11-222) with
    its names (1000 + iter), edit descriptors and line order; the last fluid
    file holds the final state to the printed precision."""
    _need_driver()
    from beach_slope_amd import caseio
    k = 2
    c = cases.tiny_box(nx=10, ny=8, solids=True, patch=True)
    dt = c.params["dt"]
    f = tmp_path / "case.bin"
    out = tmp_path / "state.bin"
    snap = tmp_path / "snap"
    for d in ("fluid", "buffer", "boundary", "ghost"):
        (snap / d).mkdir(parents=True)
    caseio.write_case(f, c, time=(k - 0.5) * dt, displaytime=dt)
    print(f"driver snapshots: tiny_box n={c.n} steps={k}", flush=True)
    r = subprocess.run([str(DRIVER), str(f), str(out), f"snapshots={snap}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    names = sorted(p.name for p in (snap / "fluid").iterdir())
    assert names == [f"fluid_{1000 + i}.txt" for i in range(1, k + 2)], names
    assert (snap / "boundary" / "boundary_1000.txt").exists()
    assert (snap / "ghost" / "ghost_1000.txt").exists()
    por = np.loadtxt(snap / "porous.txt", ndmin=2)
    assert por.shape == (c.ns, 2)
    nb = len((snap / "boundary" / "boundary_1000.txt").read_text().splitlines())
    assert nb == int(np.sum(c.parts["tid"] == 1))
    st = caseio.read_state(out)
    lines = (snap / "fluid" / f"fluid_{1000 + k + 1}.txt").read_text().splitlines()
    sel = (st["tid"] == 3) & ((st["flags"] & 1) == 0)
    assert len(lines) == int(sel.sum())
    assert all(len(ln) == 10 + 1 + 10 + 1 + 10 + 1 + 10 + 1 + 10 + 1 + 10 for ln in lines)
    vals = np.array([[float(ln[i:i + 10]) for i in range(0, 66, 11)] for ln in lines])
    want = np.stack([st["x"][sel], st["y"][sel], st["p"][sel], st["u"][sel], st["v"][sel],
                     st["c"][sel] + 0.5], axis=1)
    assert np.all(np.abs(vals[:, :3] - want[:, :3]) <= 5e-4 + 1e-12)
    for col in (3, 4):
        assert np.all(np.abs(vals[:, col] - want[:, col]) <= 5e-4 * np.abs(want[:, col]) + 1e-300)
    assert np.all(np.abs(vals[:, 5] - want[:, 5]) <= 5e-6 + 1e-12)
