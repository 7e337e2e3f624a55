"""Served form: a request loop with continuous batching over the C ABI.

Requests (a scenario and a step count) arrive over time.  The loop admits a
waiting request when its device footprint (isph_footprint_bytes) fits in the
HBM budget next to the scenarios already resident and a slot is free; it
uploads the scenario once and keeps its state on the device until the
request's last step.  Between steps the active set changes: arrivals join,
finished scenarios are downloaded and leave.  Every step advances all active
scenarios together (isph_batch_step: the PPE solves of the set run as one set
of launches).  A scenario's results do not depend on the set it shares steps
with, so a request's result is the same as when it runs alone.

The reference is a single-scenario batch program (main.f90); this loop has no
counterpart there.  The clock is injectable: `WallClock` for measurements,
`StepClock` (time advances by one unit per batched step) for deterministic
tests.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    case: object
    steps: int
    arrival: float
    footprint: int = 0
    admitted: float | None = None
    finished: float | None = None
    result: dict | None = None
    iterations: list = dataclasses.field(default_factory=list)

    @property
    def latency(self) -> float | None:
        return None if self.finished is None else self.finished - self.arrival


class WallClock:
    def __init__(self):
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def wait_until(self, t: float) -> None:
        dt = t - self.now()
        if dt > 0:
            time.sleep(dt)

    def tick(self) -> None:
        pass


class StepClock:
    """Time advances by `dt` per batched step and jumps to the next arrival
    when nothing is active."""

    def __init__(self, dt: float = 1.0):
        self.t = 0.0
        self.dt = dt

    def now(self) -> float:
        return self.t

    def wait_until(self, t: float) -> None:
        self.t = max(self.t, t)

    def tick(self) -> None:
        self.t += self.dt


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times of a seeded Poisson stream (exponential gaps, mean 1/rate)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


class Server:
    def __init__(self, pkg, hbm_budget: int = 240 * 10**9, max_slots: int = 64,
                 clock=None, keep_results: bool = True):
        self.pkg = pkg
        self.budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.clock = clock or WallClock()
        self.keep_results = keep_results
        self.steps_run = 0
        self.particle_steps = 0
        self.busy = 0.0          # time spent in batched steps

    def footprint(self, case) -> int:
        p = case.params
        return self.pkg.footprint_bytes(case.n, max(case.ns, 1), p["cellx"], p["celly"],
                                        p["precision"])

    def run(self, requests: list[Request]) -> list[Request]:
        """Serve `requests` (sorted by arrival) to completion."""
        pkg = self.pkg
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            r.footprint = self.footprint(r.case)
            if r.footprint > self.budget:
                raise ValueError(f"request {r.rid} needs {r.footprint} B, above the budget")
        waiting: list[Request] = []
        active: list[tuple[Request, object, int]] = []   # (request, scenario, steps done)
        batch = None
        members: tuple = ()
        used = 0
        while pending or waiting or active:
            now = self.clock.now()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            # admission in arrival order while budget and slots allow
            while waiting and len(active) < self.max_slots and \
                    used + waiting[0].footprint <= self.budget:
                r = waiting.pop(0)
                s = pkg.Scenario(r.case)
                s.init_state()
                r.admitted = self.clock.now()
                used += r.footprint
                active.append((r, s, 0))
            if not active:
                if pending:
                    self.clock.wait_until(pending[0].arrival)
                continue
            ids = tuple(id(s) for _, s, _ in active)
            if ids != members:
                if batch is not None:
                    batch.close()
                batch = pkg.Batch([s for _, s, _ in active])
                members = ids
            t0 = time.perf_counter()
            batch.step(1)
            self.busy += time.perf_counter() - t0
            self.clock.tick()
            self.steps_run += 1
            still = []
            for q, (r, s, k) in enumerate(active):
                r.iterations.append(batch.stats(q).iterations)
                self.particle_steps += batch.count(q)   # live particles of the step
                k += 1
                if k >= r.steps:
                    r.finished = self.clock.now()
                    if self.keep_results:
                        r.result = s.download()
                    used -= r.footprint
                else:
                    still.append((r, s, k))
            if len(still) != len(active):
                batch.close()
                batch = None
                members = ()
                for r, s, k in active:
                    if r.finished is not None:
                        s.close()
            active = still
        if batch is not None:
            batch.close()
        return requests


def summary(requests: list[Request], wall: float, particle_steps: int) -> dict:
    lat = np.array([r.latency for r in requests], dtype=np.float64)
    return dict(
        requests=len(requests),
        p50_latency_s=float(np.percentile(lat, 50)),
        p99_latency_s=float(np.percentile(lat, 99)),
        mean_latency_s=float(lat.mean()),
        particle_steps_per_s=particle_steps / wall if wall > 0 else 0.0,
        scenarios_per_h=3600.0 * len(requests) / wall if wall > 0 else 0.0,
        wall_s=wall,
    )
