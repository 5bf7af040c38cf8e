"""Continuous-batching serving loop over the osb_pool C ABI.

A stream of solve requests arrives as a Poisson process; each request is
admitted into the pool as soon as a solve slot and the device-memory budget
allow it, stays resident on the GPU until its last step, and is then
fetched and released.  Solves of one shock-capturing signature advance
together through the batched launches; others overlap on side streams.

Request mix: 2-D four-quadrant Riemann problems on [0,1]^2 (N in
{256, 512, 1024}, seeded quadrant states rho ~ U(0.1, 1.5),
p ~ U(0.03, 1.5), |u| < 1.2, zero-gradient boundaries, WENO-Z, dt from
CFL 0.4) and viscous Taylor-Green cubes (N in 64..192 step 16,
Re ~ U(400, 3200), Ma 0.1, central Feiereisen + constant viscosity);
steps ~ U{100..1000}.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import native as nat
from . import problems as P


@dataclass
class Request:
    rid: int
    arrival: float                  # seconds after the start of the stream
    problem: P.Problem
    steps: int
    work: float                     # grid-point RK-stage updates
    admitted: Optional[float] = None
    finished: Optional[float] = None
    slot: Optional[int] = None
    result: Optional[np.ndarray] = None
    iters: int = 0
    finite: bool = True             # False: the solve's state went non-finite (failed alone)

    @property
    def time_to_solution(self) -> float:
        return self.finished - self.arrival


def riemann_request(rng: np.random.Generator, n: int, steps: int) -> P.Problem:
    states = []
    for _ in range(4):
        rho = float(rng.uniform(0.1, 1.5))
        p = float(rng.uniform(0.03, 1.5))
        speed = float(rng.uniform(0.0, 1.2))
        ang = float(rng.uniform(0.0, 2.0 * math.pi))
        states.append((rho, speed * math.cos(ang), speed * math.sin(ang), p))
    pr = P.riemann_2d(n=n, states=states, steps=steps)
    pr.desc.cfl = 0.4
    pr.desc.dt = 0.0
    return pr


def tgv_request(rng: np.random.Generator, n: int, steps: int) -> P.Problem:
    re = float(rng.uniform(400.0, 3200.0))
    return P.tgv(n=n, steps=steps, re=re)


def make_stream(count: int = 64, rate: float = 50.0, seed: int = 20261203, riemann_sizes=(256, 512, 1024),
                tgv_sizes=tuple(range(64, 193, 16)), steps_range=(100, 1000), tgv_fraction: float = 0.5) -> List[Request]:
    """Seeded request stream: exponential inter-arrival times (mean 1/rate)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        if rng.uniform() < tgv_fraction:
            pr = tgv_request(rng, int(rng.choice(np.array(tgv_sizes))), steps)
        else:
            pr = riemann_request(rng, int(rng.choice(np.array(riemann_sizes))), steps)
        pr.name = f"req{i}_{pr.name}"
        out.append(Request(i, t, pr, steps, pr.desc.npoints * 3.0 * steps))
    return out


@dataclass
class ServeReport:
    requests: List[Request]
    makespan: float
    p50: float
    p99: float
    throughput: float               # aggregate GPts*stage/s over the makespan
    peak_running: int
    peak_bytes: int
    extra: dict = field(default_factory=dict)

    @property
    def failed(self) -> List[Request]:
        """Requests whose state went non-finite (each fails alone: instances
        share launches but not data, so the others are unaffected)."""
        return [r for r in self.requests if not r.finite]


def serve(lib: nat.Library, requests: List[Request], hbm_budget_bytes: int, max_solves: int = 32,
          nstreams: int = 4, chunk: int = 25, keep_results: bool = False, realtime: bool = True,
          policy: str = "fifo") -> ServeReport:
    """Run the stream through an osb_pool.  With realtime=True a request can be
    admitted only once its arrival time has passed on the wall clock.

    ``policy`` orders the waiting requests for admission: "fifo" (arrival
    order) or "sjf" (least work first -- grid-point RK-stage updates -- which
    lowers the median time-to-solution when the queue is long).  A request that
    does not fit the remaining budget does not block smaller ones behind it
    under "sjf"."""
    if policy not in ("fifo", "sjf"):
        raise ValueError(f"unknown admission policy {policy!r}")
    pending = sorted(requests, key=lambda r: r.arrival)
    waiting: List[Request] = []
    running = {}
    done: List[Request] = []
    peak_running = 0
    peak_bytes = 0
    with lib.pool(hbm_budget_bytes, max_solves=max_solves, nstreams=nstreams) as pool:
        t0 = time.perf_counter()
        while pending or waiting or running:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                waiting.append(pending.pop(0))
            # admit in policy order while slots and memory allow
            if policy == "sjf":
                waiting.sort(key=lambda r: (r.work, r.arrival))
            k = 0
            while k < len(waiting):
                r = waiting[k]
                slot = pool.admit(r.problem.desc, r.problem.q0, r.steps)
                if slot is None:
                    if policy == "fifo" or len(running) + k >= max_solves:
                        break
                    k += 1                           # too big for the budget left: try a smaller one
                    continue
                r.slot = slot
                r.admitted = time.perf_counter() - t0
                running[slot] = r
                waiting.pop(k)
            used, nrun, _ = pool.usage()
            peak_running = max(peak_running, nrun)
            peak_bytes = max(peak_bytes, used)
            if running:
                pool.advance(chunk)
                for slot in pool.poll():
                    r = running.pop(slot)
                    q, st = pool.fetch(slot)
                    r.finished = time.perf_counter() - t0
                    r.iters = st.iter
                    # the analogue of the generated loop's ops_NaNcheck on rho
                    # (helperfunctions.py:201-214), over the whole state and
                    # failing this request only
                    r.finite = bool(np.isfinite(q).all())
                    if keep_results:
                        r.result = q
                    pool.release(slot)
                    done.append(r)
            if not running and waiting:
                raise RuntimeError(f"request {waiting[0].rid} does not fit the pool's budget on its own")
            if not running and pending and realtime:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
            elif running:
                time.sleep(0.0002)
        makespan = time.perf_counter() - t0
    tts = np.array([r.time_to_solution for r in done])
    work = sum(r.work for r in done)
    return ServeReport(done, makespan, float(np.percentile(tts, 50)), float(np.percentile(tts, 99)),
                       work / makespan / 1e9, peak_running, peak_bytes)
