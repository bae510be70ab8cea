"""Serving loop: a request stream of stealthy ground-state solves over the continuous-batching solver.

Requests arrive by a Poisson process; each is a seeded configuration (N lognormal, d in {2,3}, chi uniform,
optional soft core e0 (1 - r/sigma)^2 with sigma = 0.4 density^(-1/d)).  Arrived requests are submitted at
once; the solver admits them into free slots as HBM allows, so requests that arrive while others are on the
GPU join the running batch.  Latency = completion time - arrival time.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .optimize import BatchSolver
from .potential import GetKs, RandomConfiguration, getK


@dataclass
class Request:
    rid: int
    arrival: float
    N: int
    d: int
    chi: float
    seed: int
    epsilon: float = 0.0
    sigma: float = 0.0
    job: int = -1
    submitted: float = 0.0
    done: float = 0.0
    result: dict = field(default_factory=dict)


def make_stream(n_requests, rate, seed=20261203, median_N=1000, sigma_ln=0.6, N_clip=(100, 50000),
                chi_range=(0.1, 0.49), softcore_eps=1.0):
    """Poisson arrivals at `rate` requests/s; N ~ lognormal(median, sigma_ln) clipped; d in {2,3}; chi ~ U."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        N = int(np.clip(round(median_N * math.exp(sigma_ln * rng.standard_normal())), *N_clip))
        d = int(rng.choice([2, 3]))
        chi = float(rng.uniform(*chi_range))
        sig = 0.4 if softcore_eps else 0.0  # density 1: sigma = 0.4 rho^(-1/d)
        reqs.append(Request(i, t, N, d, chi, int(rng.integers(0, 2 ** 31)), softcore_eps, sig))
    return reqs


def make_batch(count=256, master_seed=20261203, N_range=(100, 4000), dims=(2, 3), chi_range=(0.2, 0.45)):
    """A batch of independent stealthy requests, all present at t = 0: N log-uniform on N_range (rounded), d a fair
    coin over `dims`, chi uniform on chi_range and one configuration seed each, drawn in that order per request from
    one MT19937 stream (numpy's RandomState) seeded with `master_seed`.  No soft core."""
    rng = np.random.RandomState(master_seed)
    reqs = []
    lo, hi = math.log(N_range[0]), math.log(N_range[1])
    for i in range(count):
        N = int(round(math.exp(rng.uniform(lo, hi))))
        d = dims[0] if rng.uniform() < 0.5 else dims[1]
        chi = float(rng.uniform(*chi_range))
        reqs.append(Request(i, 0.0, N, d, chi, int(rng.randint(0, 2 ** 31 - 1))))
    return reqs


def serve(reqs, Emin=1e-18, MaxStep=10000, max_slots=64, hbm_budget=200 << 30, step_evals=16, timeout=600,
          graph=False):
    """Run the stream in real time.  Returns the requests with results, and the wall time."""
    solver = BatchSolver(max_slots=max_slots, hbm_budget=hbm_budget, check_interval=step_evals, graph=graph)
    pending = sorted(reqs, key=lambda r: r.arrival)
    inflight = {}
    t0 = time.perf_counter()
    i = 0
    while i < len(pending) or inflight:
        now = time.perf_counter() - t0
        if now > timeout:
            raise TimeoutError("serving loop timeout")
        while i < len(pending) and pending[i].arrival <= now:
            r = pending[i]
            cfg = RandomConfiguration(r.seed, r.N, r.d)
            k = GetKs(cfg, getK(r.chi, 1.0, r.d))
            r.job = solver.submit(cfg, k, np.ones(len(k)), epsilon=r.epsilon, sigma=r.sigma, MaxStep=MaxStep, Emin=Emin)
            r.submitted = time.perf_counter() - t0
            inflight[r.job] = r
            i += 1
        if inflight:
            solver.step(step_evals)
            for job in list(inflight):
                res = solver.result(job)
                if res["state"] >= BatchSolver.DONE:
                    r = inflight.pop(job)
                    r.done = time.perf_counter() - t0
                    r.result = res
                    solver.release(job)
        elif i < len(pending):
            time.sleep(min(0.01, max(0.0, pending[i].arrival - now)))
    return reqs, time.perf_counter() - t0
