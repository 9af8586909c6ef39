"""Served request streams: Poisson arrivals of tracking requests against a pool of lattice
variants, run through the continuous-batching server and through one-job-at-a-time tracking.

The stream follows the served-stream workload: a pool of error variants of the 901-element ring
(relative strength errors sigma = 1e-3 on every quadrupole k1 and sextupole k2), particle counts
log-uniform in [n_lo, n_hi], turn counts log-uniform in [t_lo, t_hi], Gaussian beams with 3x the
base sigmas, exponential inter-arrival times. Everything is drawn from one seeded generator.
"""
from __future__ import annotations

import math
import sys
import threading
import time

import numpy as np

from . import lattices as L
from . import elements as E
from . import tracking as T


def make_stream(seed: int, n_requests: int, n_variants: int = 32, n_lo: int = 1000,
                n_hi: int = 1_000_000, t_lo: int = 10, t_hi: int = 10_000):
    """Requests as dicts {variant, n, turns, seed}, arrival gaps drawn later from the rate."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_requests):
        n = int(math.floor(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        t = int(math.floor(math.exp(rng.uniform(math.log(t_lo), math.log(t_hi)))))
        out.append({"variant": int(rng.integers(n_variants)), "n": n, "turns": t,
                    "seed": int(rng.integers(1 << 31))})
    gaps = rng.exponential(1.0, n_requests)  # in units of the mean inter-arrival time
    return out, gaps


class _Progress:
    """A line on stderr every `every` seconds while a long stream runs (completed requests)."""

    def __init__(self, what: str, total: int, every: float = 20.0):
        self.what, self.total, self.every = what, total, every
        self.t0 = self.last = time.perf_counter()

    def step(self, done: int):
        now = time.perf_counter()
        if now - self.last >= self.every or done == self.total:
            self.last = now
            print(f"{self.what}: {done}/{self.total} requests done, {now - self.t0:.1f} s",
                  file=sys.stderr, flush=True)


def request_beam(req):
    return L.gaussian_beam(req["n"], seed=req["seed"], sigma=3 * L.SIGMA0)


def run_served(ctx, lats, reqs, arrivals, hbm_budget, chunk_turns=100, compact_turns=10):
    """Submit each request at its arrival time (seconds from start) to a server; returns
    per-request latencies, the makespan and the server's pass count."""
    srv = T.Server(ctx, hbm_budget, chunk_turns, compact_turns)
    beams = [request_beam(r) for r in reqs]
    ids = [None] * len(reqs)
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        delay = arrivals[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids[i] = srv.submit(lats[r["variant"]], beams[i], r["turns"])
    lat_s = []
    prog = _Progress("served", len(ids))
    for j, i in enumerate(ids):
        _, st = srv.wait(i)
        lat_s.append(st["latency_s"])
        prog.step(j + 1)
    makespan = time.perf_counter() - t0
    tot = srv.totals()
    srv.close()
    return {"latency_s": lat_s, "makespan_s": makespan, "passes": tot["passes"],
            "chunks": tot["chunks"], "busy_s": tot["busy_s"]}


def run_plain(ctx, lats, reqs, arrivals):
    """One job at a time in arrival order (FIFO), each request one bmx_track_turns call on its
    own beam (no batching with other requests): per-request latencies (from arrival to
    completion), makespan and passes."""
    beams = [request_beam(r) for r in reqs]
    ctx.take_work()
    t0 = time.perf_counter()
    lat_s = []
    prog = _Progress("one job at a time", len(reqs))
    for i, r in enumerate(reqs):
        now = time.perf_counter() - t0
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        b = T.Beam.from_host(ctx, beams[i])
        T.track_beam(ctx, lats[r["variant"]], b, r["turns"])
        ctx.synchronize()
        b.download()
        b.close()
        lat_s.append(time.perf_counter() - t0 - arrivals[i])
        prog.step(i + 1)
    makespan = time.perf_counter() - t0
    return {"latency_s": lat_s, "makespan_s": makespan, "passes": ctx.take_work()}


def run_plain_elementwise(ctx, lats, reqs, arrivals):
    """The plain GPU path of the reference loop nest -- one launch per element per turn, state
    through HBM (bmx_track_turns_plain, one turn captured as a graph and replayed) -- one
    request at a time in arrival order, each tracked in full: per-request latencies (from
    arrival to completion), makespan and passes."""
    beams = [request_beam(r) for r in reqs]
    ctx.take_work()
    t0 = time.perf_counter()
    lat_s = []
    prog = _Progress("plain", len(reqs))
    for i, r in enumerate(reqs):
        now = time.perf_counter() - t0
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        b = T.Beam.from_host(ctx, beams[i])
        T.track_beam_plain(ctx, lats[r["variant"]], b, r["turns"])
        ctx.synchronize()
        b.download()
        b.close()
        lat_s.append(time.perf_counter() - t0 - arrivals[i])
        prog.step(i + 1)
    makespan = time.perf_counter() - t0
    return {"latency_s": lat_s, "makespan_s": makespan, "passes": ctx.take_work()}


def percentiles(v):
    a = np.asarray(v)
    return {"p50_s": float(np.percentile(a, 50)), "p99_s": float(np.percentile(a, 99)),
            "mean_s": float(a.mean()), "max_s": float(a.max())}
