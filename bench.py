"""Benchmark: particle-steps/s of the per-step ISPH pipeline at ~650k particles
on one MI355X.

A step is one production-loop iteration of the reference (main.f90:199-261)
over the 650k-particle flume on its porous beach (128,000 solids) with its
piston wavemaker: scalar transport,
projection, the paddle's motion, rebin, neighbour
search, porosity, SPS, intermediate velocity, PPE assembly and solve,
pressure-gradient update, trapezoid positions, second rebin and neighbour
search, free surface, porosity, shifting, density, time step, SPS.  State is
device resident before the timed region starts.

Two forms are measured on the same input:
  plain   FP64 BiCGStab (the parity path);
  chosen  BiCGStab right-preconditioned by a cell-multigrid V-cycle: FP16/FP32
          V-cycle, FP32 search directions, FP64 outer products, solution,
          residual and dots, with a final FP64 residual check.
Both solve to the same relative residual (1e-8) so their throughputs compare
like for like.  `value` is the chosen form.

Each form also reports the whole step against the HBM roofline: its
algorithmic bytes per particle-step (bytes_model.step_bytes, at the measured
iteration count), the rate that gives at the measured speed, and the
fraction of the 8 TB/s peak.

The roofline object is for the headline form's dominant kernel: the
solver's kernels (the two SpMVs, the V-cycle's level-0 passes, level-1
restriction, cell-level down and up passes and coarse workgroup, and the
dot-product finalizes; together ~90% of a step's kernel time) are timed with
HIP events on the context's stream over as many steps again, right after
the timed region (the events cost ~10% of a step, so the timed region runs
without them), and the kernel function with the most time per step (all
its launches: both applications of the V-cycle, every cell level) is
reported, with its algorithmic bytes per launch (bytes_model) over its mean
launch time.  Its
`traffic` (HBM bytes per launch) comes from the FETCH_SIZE / WRITE_SIZE
counter passes summarised in profiles/traffic_flume_650k.md, when that file
holds the kernel.  The CPU baseline is the oracle (a C/OpenMP restatement of
the reference) on a bounded sample of the same workload: one untimed step,
then three timed steps.

Three more legs run by default, at bounded sizes, and go into the same line:
  precision_errors  the jittered dam break, 200 steps in each form: pressure,
           free-surface, front and runup differences of the precision form
           against the FP64 run (gauges.py); and, since the runup does not
           develop in those 200 steps, the same dam break at dx 0.01 m with
           the toe 5 cm beyond the column for 2,000 steps (runup, front and
           slope-surface differences);
  batched  64 independent porous-beach flumes, N ~ U[20k, 300k] (about 10M
           particles), precision form, stepped together (one launch per
           particle phase and one set of solver launches for all), 20 steps,
           against the same scenarios stepped one after another alone; then
           20 more batched steps with the phase sequences replayed as HIP
           graphs (graph_*);
  served   a seeded Poisson stream (2 requests/s) of 8 porous-beach flumes
           with N log-uniform on [10k, 1M], 2,000 steps each, admitted
           against a 240 GB HBM budget and continuously batched; p50/p99
           completion latency, scenarios/h and particle-steps/s.

The particle kernels run in the launch shapes the autotuner picks on the
warmed-up state (isph_autotune; --no-tune keeps the defaults): tuning sits
outside every timed region, and results do not depend on the shape.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode both|plain|chosen]
                       [--no-batch] [--no-serve] [--no-cpu] [--no-tune]
       python bench.py --batch-full 1000   (the 64 batched flumes for their configured
                                             1,000 steps, every solve checked; not
                                             part of the default run)
       python bench.py --flume-full 2000   (the 650k flume for its configured
                                             2,000 steps: FP64 at 1e-10, FP64 and the
                                             precision form at --tol, compared at the
                                             end; not part of the default run)
       python bench.py --copy-calibration   (a 1 GiB device copy, for the
                                             counter calibration only)
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

import numpy as np

import __graft_entry__ as entry

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 650k ISPH; scenarios/h + p99 latency vs plain GPU FP64"


def _shapes(tuned: dict) -> dict:
    return {k: v[0] for k, v in tuned.items()}


def run_gpu(pkg, case, precision: int, steps: int, warmup: int, tol: float, tune: bool):
    params = dict(case.params)
    params.update(precision=precision, tol=tol, max_outer=20)
    g = pkg.Scenario(case, params=params)
    g.init_state()
    for _ in range(warmup):
        g.step(1)
    # launch shapes of the particle kernels, timed on the warmed-up state
    shapes = _shapes(g.autotune(trials=3)) if tune else None
    # the timed region runs without the per-kernel HIP-event timers: their
    # events cost ~10% of a precision-form step at 650k
    iters, outer = [], []
    nnz = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = g.step(1)
        iters.append(st.iterations)
        outer.append(st.outer)
        nnz = st.nnz
    t1 = time.perf_counter()
    # kernel durations: as many further steps with the timers on
    g.prof_enable(True)
    prof_iters = []
    for _ in range(steps):
        _, st = g.step(1)
        prof_iters.append(st.iterations)
    from beach_slope_amd import bytes_model
    n = g.n
    timed = {}
    for kname in (TIMED_PRECISION if precision else TIMED_PLAIN):
        ms, launches = g.prof_read(kname)
        kern = KERNEL_OF[kname]
        if kern == "kb_mg_coarse":
            kb = bytes_model.coarse_bytes(case.params["cellx"], case.params["celly"])
        elif kern in CELL_PASSES:
            kb = bytes_model.cell_pass_bytes(kern, n, case.params["cellx"], case.params["celly"])
        else:
            kb = bytes_model.kernel_bytes(kern, n, nnz)
        avg = 1e3 * ms / max(launches, 1)
        timed[kname] = dict(kernel=kern, ms_per_step=ms / steps, launches=launches,
                            avg_us=avg, bytes=kb,
                            gbs=kb / (avg * 1e-6) / 1e9 if launches else 0.0)
    g.prof_enable(False)
    g.close()
    wall = t1 - t0
    # the dominant kernel: the most time per step
    kname = max(timed, key=lambda k: timed[k]["ms_per_step"])
    avg_us, launches = timed[kname]["avg_us"], timed[kname]["launches"]
    bytes_per, achieved = timed[kname]["bytes"], timed[kname]["gbs"]
    # the whole step against the HBM roofline: algorithmic bytes per step
    # (bytes_model.step_bytes) at the measured rate, over the 8 TB/s peak
    step_b = bytes_model.step_bytes("precision" if precision else "plain", n, nnz,
                                    float(np.mean(iters)))
    step_gbs = step_b * steps / wall / 1e9
    return dict(
        n=n, seconds=wall, ms_per_step=1e3 * wall / steps,
        particle_steps_per_s=n * steps / wall, iters_mean=float(np.mean(iters)),
        outer_mean=float(np.mean(outer)), nnz_per_row=nnz / n,
        timer_steps_iters_mean=float(np.mean(prof_iters)),
        kernel=kname, kernel_avg_us=avg_us, kernel_launches=launches,
        kernel_bytes=bytes_per, kernel_gbs=achieved,
        timed_kernels={k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                           for kk, vv in v.items()} for k, v in timed.items()},
        launch_shapes=shapes,
        step_bytes_per_particle=step_b / n, step_gbs=step_gbs,
        step_hbm_frac=step_gbs / HBM_PEAK_GBS,
        hbm_bound_particle_steps_per_s=n * HBM_PEAK_GBS * 1e9 / step_b,
    )


def run_batch(pkg, cases, n_scen: int, steps: int, warmup: int, tune: bool) -> dict:
    """The batched configuration: n_scen independent flumes, N ~ U[20k, 300k]
    (seeded), precision form at 1e-8, stepped together; the same scenarios
    stepped one after another alone give the comparison."""
    mix = cases.scenario_mix(n_scen=n_scen, seed=2024, n_min=20_000, n_max=300_000)
    for c in mix:
        c.params.update(precision=1, tol=1e-8, max_outer=20)
    scen = [pkg.Scenario(c) for c in mix]
    for sc in scen:
        sc.init_state()
    b = pkg.Batch(scen)
    if warmup:
        b.step(warmup)
    shapes = _shapes(b.autotune(trials=2)) if tune else None
    t0 = time.perf_counter()
    b.step(steps)
    wall = time.perf_counter() - t0
    its = [b.stats(q).iterations for q in range(len(scen))]
    launches = b.launches()
    # the same batch's next steps with the fixed-shape phase sequences
    # replayed as HIP graphs (captured in one untimed step)
    b.set_graph(True)
    b.step(1)
    t0 = time.perf_counter()
    b.step(steps)
    wall_graph = time.perf_counter() - t0
    its_graph = [b.stats(q).iterations for q in range(len(scen))]
    b.close()
    n_live = [sc.n for sc in scen]
    for sc in scen:
        sc.close()
    alone = 0.0
    for c in mix:
        sc = pkg.Scenario(c)
        sc.init_state()
        if warmup:
            sc.step(warmup)
        if tune:
            sc.autotune(trials=2)
        t1 = time.perf_counter()
        sc.step(steps)
        alone += time.perf_counter() - t1
        sc.close()
    ps = sum(n_live) * steps / wall
    ps_alone = sum(n_live) * steps / alone
    return dict(scenarios=n_scen, particles=int(sum(n_live)), steps=steps, warmup=warmup,
                seconds=wall, particle_steps_per_s=ps, alone_particle_steps_per_s=ps_alone,
                speedup_vs_alone=ps / ps_alone,
                scenarios_per_h_at_1000_steps=3600.0 * n_scen / (1000.0 * wall / steps),
                particle_phase_launches_per_step=launches, launch_shapes=shapes,
                solids=int(sum(c.ns for c in mix)),
                iterations_last_step_mean=float(np.mean(its)),
                iterations_last_step_max=int(max(its)),
                graph_seconds=wall_graph,
                graph_particle_steps_per_s=sum(n_live) * steps / wall_graph,
                graph_iterations_last_step_mean=float(np.mean(its_graph)))


def run_batch_full(pkg, cases, n_scen: int, steps: int, tune: bool) -> dict:
    """The batched configuration for its configured step count: the 64 flumes
    of the batched leg stepped together `steps` steps after one warm-up step
    (and the launch-shape tuning), with every scenario's solve checked at
    every step; scenarios/h is measured, not extrapolated."""
    mix = cases.scenario_mix(n_scen=n_scen, seed=2024, n_min=20_000, n_max=300_000)
    for c in mix:
        c.params.update(precision=1, tol=1e-8, max_outer=20)
    scen = [pkg.Scenario(c) for c in mix]
    for sc in scen:
        sc.init_state()
    b = pkg.Batch(scen)
    b.step(1)
    shapes = _shapes(b.autotune(trials=2)) if tune else None
    n0 = [b.count(q) for q in range(n_scen)]
    its_sum, its_max, worst, unconv = 0, 0, 0.0, 0
    t0 = time.perf_counter()
    for k in range(steps):
        b.step(1)
        for q in range(n_scen):
            st = b.stats(q)
            its_sum += st.iterations
            its_max = max(its_max, st.iterations)
            worst = max(worst, st.rel_residual)
            unconv += 0 if st.converged else 1
        if (k + 1) % 100 == 0:
            print(f"bench: batch step {k + 1}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    n1 = [b.count(q) for q in range(n_scen)]
    b.close()
    for sc in scen:
        sc.close()
    return dict(scenarios=n_scen, particles=int(sum(n0)), solids=int(sum(c.ns for c in mix)),
                steps=steps, seconds=wall, particle_steps_per_s=sum(n0) * steps / wall,
                scenarios_per_h=3600.0 * n_scen / wall, launch_shapes=shapes,
                iterations_mean=its_sum / (steps * n_scen), iterations_max=its_max,
                worst_rel_residual=worst, unconverged_solves=unconv,
                particles_kept=n0 == n1)


def run_serve(pkg, cases, n_req: int, steps: int, rate: float) -> dict:
    """The served configuration: a seeded Poisson stream of porous-beach flume
    requests, N log-uniform on [10k, 1M], admitted against a 240 GB HBM
    budget, each kept device resident for its steps (the configuration's
    2,000), continuous batching between steps.  The request count is reduced
    to bound the run."""
    from beach_slope_amd import serve
    mix = cases.scenario_mix(n_scen=n_req, seed=99, n_min=10_000, n_max=1_000_000,
                             log_uniform=True, steps=steps)
    for c in mix:
        c.params.update(precision=1, tol=1e-8, max_outer=20)
    arr = serve.poisson_arrivals(n_req, rate=rate, seed=99)
    reqs = [serve.Request(k, c, steps=steps, arrival=float(arr[k])) for k, c in enumerate(mix)]
    srv = serve.Server(pkg, hbm_budget=240 * 10**9, max_slots=64, clock=serve.WallClock(),
                       keep_results=False)
    srv.run(reqs)
    wall = srv.clock.now()
    out = serve.summary(reqs, wall, srv.particle_steps)
    out.update(rate_per_s=rate, steps_per_request=steps, hbm_budget_gb=240,
               particles_total=int(sum(c.n for c in mix)), batched_steps=srv.steps_run,
               gpu_busy_s=srv.busy)
    return out


def run_precision_errors(pkg, cases, steps: int) -> dict:
    """The precision-form configuration: the dam break with seeded +-0.01 dx
    jitter run `steps` steps in the FP64 form (tolerance 1e-10) and in the
    precision form (1e-8) from the same state; pressure, free-surface, front
    and runup differences through the gauges."""
    from beach_slope_amd import gauges
    c = cases.dam_break(jitter=0.01)
    m = c.meta
    h = c.params["h"]
    out = {"workload": f"dam_break_jitter n={c.n}", "steps": steps}
    states = {}
    for form, prec, tol in (("fp64", 0, 1e-10), ("precision", 1, 1e-8)):
        prm = dict(c.params, precision=prec, tol=tol, max_outer=20)
        sc = pkg.Scenario(c, params=prm)
        sc.init_state()
        t0 = time.perf_counter()
        its = []
        for _ in range(steps):
            _, st = sc.step(1)
            its.append(st.iterations)
        out[f"{form}_seconds"] = time.perf_counter() - t0
        out[f"{form}_iterations_mean"] = float(np.mean(its))
        states[form] = sc.download()
        sc.close()
    a, b = states["precision"], states["fp64"]
    off, col_h = m["off"], m["column"][1]
    probes = np.array([[off + 0.05, off + 0.02], [off + 0.25, off + 0.02],
                       [off + 0.05, off + 0.12], [off + 0.25, off + 0.10]])
    pa, pb = gauges.probe_pressure(a, probes, h), gauges.probe_pressure(b, probes, h)
    stations = off + np.array([0.05, 0.15, 0.25, 0.35, 0.45])
    ea = gauges.surface_elevation(a, stations, 0.5 * m["dx"])
    eb = gauges.surface_elevation(b, stations, 0.5 * m["dx"])
    fa, fb = gauges.front_position(a), gauges.front_position(b)
    ra, rb = gauges.runup(a, m["toe"], off), gauges.runup(b, m["toe"], off)
    pair = gauges.pair_states(a, b)
    out.update(
        probe_pressure_rel_l2=float(np.linalg.norm(pa - pb) / np.linalg.norm(pb)),
        surface_max_abs_m=float(np.nanmax(np.abs(ea - eb))),
        surface_max_rel_to_column=float(np.nanmax(np.abs(ea - eb)) / col_h),
        front_abs_m=abs(fa - fb), front_m=fb,
        runup_abs_m=abs(ra - rb), runup_m=rb,
        paired_pressure_rel_l2=pair["pressure_rel_l2"],
        paired_velocity_rel_l2=pair["velocity_rel_l2"],
        paired_max_distance_m=pair["max_pair_distance"],
        free_flag_mismatch=pair["free_flag_mismatch"],
    )
    return out


def run_runup_errors(pkg, cases, steps: int) -> dict:
    """Runup against the FP64 run where it develops: the C3 dam break has its
    toe 0.5 m beyond the column, out of reach in its 200 steps, so the same
    dam break at dx = 0.01 m with the toe 5 cm beyond the column and seeded
    +-0.01 dx jitter is run `steps` steps in both forms (the front reaches
    the slope by step ~700).  A run that stops with an error is reported
    instead of compared (an isolated fluid particle no longer stops it: its
    PPE row is the identity)."""
    from beach_slope_amd import gauges
    c = cases.dam_break(dx=0.01, jitter=0.01, toe=0.55)
    m = c.meta
    out = {"workload": f"dam_break_jitter dx=0.01 toe=0.55 n={c.n}", "steps": steps}
    states = {}
    for form, prec, tol in (("fp64", 0, 1e-10), ("precision", 1, 1e-8)):
        prm = dict(c.params, precision=prec, tol=tol, max_outer=20)
        sc = pkg.Scenario(c, params=prm)
        sc.init_state()
        t0 = time.perf_counter()
        try:
            for k in range(steps):
                sc.step(1)
        except pkg.IsphError as e:
            out[f"{form}_stopped"] = f"step {k}: {e}"
            sc.close()
            return out
        out[f"{form}_seconds"] = time.perf_counter() - t0
        states[form] = sc.download()
        sc.close()
    a, b = states["precision"], states["fp64"]
    stations = m["toe"] + np.array([0.02, 0.05, 0.10, 0.15])
    ea = gauges.surface_elevation(a, stations, 0.5 * m["dx"])
    eb = gauges.surface_elevation(b, stations, 0.5 * m["dx"])
    ra, rb = gauges.runup(a, m["toe"], m["off"]), gauges.runup(b, m["toe"], m["off"])
    out.update(runup_m=rb, runup_abs_m=abs(ra - rb), runup_rel=abs(ra - rb) / rb if rb else None,
               front_m=gauges.front_position(b),
               front_abs_m=abs(gauges.front_position(a) - gauges.front_position(b)),
               slope_surface_max_abs_m=float(np.nanmax(np.abs(ea - eb))))
    return out


def run_flume_full(pkg, cases, steps: int, tol: float) -> dict:
    """The 650k porous-beach flume for its configured step count, from the
    same initial state, three times: the FP64 form at 1e-10 (the reference
    run), the FP64 form at `tol` and the precision form at `tol`.  Reports
    whether every solve met its tolerance, each run's wall time, and at the
    end each `tol` run's differences from the reference run (probe pressures
    along the flume, paired particles, free-surface elevation at stations):
    the FP64 run at `tol` shows how far a solve tolerance of `tol` alone
    carries a run over these steps, the precision run adds its own form."""
    from beach_slope_amd import gauges
    c = cases.flume_650k()
    m = c.meta
    h = c.params["h"]
    out = {"workload": f"flume_650k n={c.n} ns={c.ns}", "steps": steps, "tol": tol}
    states = {}
    for form, prec, tl in (("fp64_ref", 0, 1e-10), ("fp64", 0, tol), ("precision", 1, tol)):
        prm = dict(c.params, precision=prec, tol=tl, max_outer=20)
        sc = pkg.Scenario(c, params=prm)
        sc.init_state()
        its, worst, unconv = [], 0.0, 0
        t0 = time.perf_counter()
        for k in range(steps):
            _, st = sc.step(1)
            its.append(st.iterations)
            worst = max(worst, st.rel_residual)
            unconv += 0 if st.converged else 1
            if (k + 1) % 250 == 0:
                print(f"bench: flume {form} step {k + 1}: {st.iterations} iterations",
                      file=sys.stderr, flush=True)
        out[f"{form}_seconds"] = time.perf_counter() - t0
        out[f"{form}_iterations_mean"] = float(np.mean(its))
        out[f"{form}_worst_rel_residual"] = worst
        out[f"{form}_unconverged_solves"] = unconv
        out[f"{form}_t_end_s"] = sc.scalars().t
        states[form] = sc.download()
        sc.close()
    ref = states["fp64_ref"]
    off, offy = m["off"], m["offy"]
    probes = np.array([[off + x, offy + y] for x, y in
                       ((0.5, 0.2), (2.0, 0.2), (4.0, 0.2), (5.5, 0.2), (7.0, 0.25))])
    stations = off + np.array([0.3, 1.0, 2.0, 4.0, 5.5, 6.5, 7.5])
    p_ref = gauges.probe_pressure(ref, probes, h)
    e_ref = gauges.surface_elevation(ref, stations, 0.5 * m["dx"])
    fl_r = gauges.fluid(ref)
    out.update(speedup_precision_vs_fp64=out["fp64_seconds"] / out["precision_seconds"],
               probe_pressure_pa=[float(v) for v in p_ref], surface_m=[float(v) for v in e_ref],
               max_speed_m_s=float(np.hypot(ref["u"][fl_r], ref["v"][fl_r]).max()))
    for form in ("fp64", "precision"):
        a = states[form]
        pair = gauges.pair_states(a, ref)
        fl_a = gauges.fluid(a)
        pa = gauges.probe_pressure(a, probes, h)
        ea = gauges.surface_elevation(a, stations, 0.5 * m["dx"])
        out[f"{form}_vs_ref"] = dict(
            probe_pressure_rel_l2=float(np.linalg.norm(pa - p_ref) / np.linalg.norm(p_ref)),
            surface_max_abs_m=float(np.nanmax(np.abs(ea - e_ref))),
            paired_pressure_rel_l2=pair["pressure_rel_l2"],
            paired_velocity_rel_l2=pair["velocity_rel_l2"],
            paired_max_distance_m=pair["max_pair_distance"],
            free_flag_mismatch=pair["free_flag_mismatch"],
            particles=int(a["x"].size),
            fluid_volume_m2=float(np.sum(a["m"][fl_a] / a["rho"][fl_a])))
    out["ref_fluid_volume_m2"] = float(np.sum(ref["m"][fl_r] / ref["rho"][fl_r]))
    return out


KERNEL_OF = {"spmv_p": "k_bicg_1", "spmv_s": "k_bicg_2", "spmv_p32": "kb_pb_1",
             "spmv_s32": "kb_pb_2", "mg_pre0": "kb_mg_pre0", "mg_post0": "kb_mg_post0",
             "mg_coarse": "kb_mg_coarse", "mg_restrict0": "kb_mg_restrict0",
             "mg_down": "kb_mg_down", "mg_up": "kb_mg_up", "pfinalize": "kb_pfinalize"}
TIMED_PLAIN = ("spmv_p", "spmv_s")
TIMED_PRECISION = ("spmv_p32", "spmv_s32", "mg_pre0", "mg_post0", "mg_restrict0", "mg_down",
                   "mg_coarse", "mg_up", "pfinalize")
CELL_PASSES = ("kb_mg_restrict0", "kb_mg_down", "kb_mg_up", "kb_pfinalize")
TRAFFIC_FILE = pathlib.Path(__file__).resolve().parent / "profiles" / "traffic_flume_650k.md"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the counter summary, or None."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        m = re.match(r"^traffic\s+(\S+)\s+([0-9.eE+]+)\s*$", line.strip())
        if m and m.group(1) == KERNEL_OF.get(kernel, kernel):
            return float(m.group(2))
    return None


def copy_calibration() -> int:
    """Three launches of a plain 1 GiB device-to-device copy (torch's copy
    kernel): the FETCH_SIZE and WRITE_SIZE counters of one launch should read
    1 GiB each; the ratio calibrates the counters of the solver kernels."""
    import torch
    n = 1 << 28
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}))
    return 0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(oracle, case, tol: float, steps: int = 3):
    """The oracle on the same workload: one untimed step after its init, then
    `steps` timed steps."""
    params = dict(case.params)
    params.update(precision=0, tol=tol)
    o = oracle.Oracle(case, params=params)
    o.call("init_state")
    o.step(1)
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        rc, st = o.step(1)
        its.append(st.iterations)
    dt = time.perf_counter() - t0
    return dict(value=o.n * steps / dt, unit="particle-steps/s", cores=oracle.threads(),
                kind="port", cpu=cpu_model(),
                sample=f"{steps} steps of the same {o.n}-particle workload after one "
                       f"untimed step ({its} FP64 BiCGStab iterations, {dt:.1f} s)")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="both", choices=["both", "plain", "chosen"])
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (no autotuning after warmup)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--error-steps", type=int, default=200)
    ap.add_argument("--runup-steps", type=int, default=2000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-scenarios", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=8)
    ap.add_argument("--serve-steps", type=int, default=2000)
    ap.add_argument("--batch-full", type=int, default=0, metavar="STEPS",
                    help="only run the batched configuration (64 flumes) for STEPS steps, "
                         "every solve checked, and print that line (1,000 steps: ~2.5 min)")
    ap.add_argument("--flume-full", type=int, default=0, metavar="STEPS",
                    help="only run the 650k flume for STEPS steps (FP64 at 1e-10, FP64 and "
                         "the precision form at --tol) and print that comparison (the "
                         "configuration's 2,000 steps take ~7 min)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(one flume runs on one GPU; independent scenarios would be placed one per GPU)",
              file=sys.stderr)
        return 2
    pkg = entry.load_package()
    from beach_slope_amd import case as cases
    if args.flume_full > 0:
        print(json.dumps(run_flume_full(pkg, cases, args.flume_full, args.tol)), flush=True)
        return 0
    if args.batch_full > 0:
        print(json.dumps(run_batch_full(pkg, cases, args.batch_scenarios, args.batch_full,
                                        not args.no_tune)), flush=True)
        return 0
    c = cases.flume_650k()
    print(f"bench: workload flume_650k n={c.n} steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)
    res = {}
    if args.mode in ("both", "plain"):
        res["plain"] = run_gpu(pkg, c, 0, args.steps, args.warmup, args.tol,
                               not args.no_tune)
        print(f"bench: plain {res['plain']}", file=sys.stderr, flush=True)
    if args.mode in ("both", "chosen"):
        res["chosen"] = run_gpu(pkg, c, 1, args.steps, args.warmup, args.tol,
                                not args.no_tune)
        print(f"bench: chosen {res['chosen']}", file=sys.stderr, flush=True)
    head = res.get("chosen") or res["plain"]
    batched = served = errors = None
    if not args.no_errors:
        errors = run_precision_errors(pkg, cases, args.error_steps)
        if args.runup_steps > 0:
            errors["runup"] = run_runup_errors(pkg, cases, args.runup_steps)
        print(f"bench: precision errors {errors}", file=sys.stderr, flush=True)
    if not args.no_batch:
        batched = run_batch(pkg, cases, args.batch_scenarios, args.batch_steps, 1,
                            not args.no_tune)
        print(f"bench: batched {batched}", file=sys.stderr, flush=True)
    if not args.no_serve:
        served = run_serve(pkg, cases, args.serve_requests, args.serve_steps, 2.0)
        print(f"bench: served {served}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(entry.load_oracle(), c, args.tol)
        print(f"bench: cpu {cpu}", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": head["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if "chosen" in res else "f64",
        "data": "synthetic",
        "config": {"workload": "flume_650k: 10 m x 0.4 m flume on a 1:10 porous beach "
                               "(porosity 0.4, d50 0.01 m, 128,000 solids), dx 0.0025, "
                               "4000x160 fluid + 3 wall layers, piston wavemaker (H 0.05 m, "
                               "T 1.5 s), scalar patch, dt 5e-5",
                   "particles": head["n"], "ppe_rel_tol": args.tol,
                   "form": "chosen" if "chosen" in res else "plain"},
        "roofline": {"bound": "hbm", "achieved": head["kernel_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["kernel_gbs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(head["kernel"]),
                     "kernel": KERNEL_OF[head["kernel"]],
                     "avg_launch_us": head["kernel_avg_us"],
                     "ms_per_step": head["timed_kernels"][head["kernel"]]["ms_per_step"]},
        "cpu_baseline": cpu,
        "precision_errors": errors,
        "batched": batched,
        "served": served,
    }
    for k, v in res.items():
        out[k] = {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
    if "plain" in res and "chosen" in res:
        out["speedup_chosen_vs_plain"] = (res["chosen"]["particle_steps_per_s"] /
                                          res["plain"]["particle_steps_per_s"])
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
