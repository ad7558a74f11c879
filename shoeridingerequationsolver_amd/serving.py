"""Continuous batching of solve requests on one GPU.

Requests (seeded solves of mixed shapes, potentials, step counts and
precisions) arrive over time.  The server

* admits a waiting request as soon as its device footprint fits the HBM
  budget (``tdse_plan_bytes``; the context enforces the same cap) and builds
  its potential, mask and packet directly in HBM (``tdse_plan_build``; or,
  with ``device_build=False``, on host threads while the GPU works),
* keeps every admitted state resident in HBM (one plan each),
* overlaps the work of all active requests on the device, by one of two
  policies:

  - ``"streams"`` (default): each request advances on its own stream in
    chunks of about ``chunk_ms`` of device time; the loop polls the streams,
    queues the next chunk of a request as soon as its last one finished,
    retires finished requests and admits arrivals at once.  No request waits
    for another's chunk boundary, and nothing synchronises the device as a
    whole.
  - ``"batch"``: all active requests advance together in quanta, one
    ``tdse_batch_step`` of ``min(quantum, fewest remaining steps)`` steps, so
    requests that share a sweep shape share launches; retiring and admitting
    happen between quanta.

The reference has no serving layer; its CPU loop advances one scene per frame
(scenario/src/base_director.ixx:339-357).  The baseline ``run_sequential``
solves the same trace one request at a time, in plain FP64 or at each
request's own precision.
"""
from __future__ import annotations

import concurrent.futures as cf
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import Batch, Context, Propagator, plan_bytes
from .workloads import SplitMix64, Solve

STREAM_SHAPES = [(65536, 1, 1), (1024, 1024, 1), (64, 64, 64), (128, 128, 128),
                 (192, 192, 192), (256, 256, 256)]
STREAM_POTENTIALS = ["harmonic", "regularized_coulomb", "barrier"]


def request_stream(seed: int, count: int, rate: float, max_steps: int = 2000,
                   shapes=None) -> list:
    """`count` requests with Poisson arrivals at `rate` per second.

    Shape from STREAM_SHAPES (3D boxes 40 a.u. for <= 128^3, 80 a.u. above;
    2D 80; 1D 200), potential from harmonic / regularized Coulomb / barrier
    with an 8 a.u. absorbing mask, steps ~ U(100, max_steps), FP32 or FP64 with
    equal odds.  Returns [(arrival_time_s, Solve)]."""
    rng = SplitMix64(seed)
    shapes = shapes or STREAM_SHAPES
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        n = rng.choice(shapes)
        dims = sum(1 for v in n if v > 1)
        big = dims == 3 and n[0] > 128
        half = {3: 40.0 if big else 20.0, 2: 40.0, 1: 100.0}[dims]
        lo = tuple(-half if v > 1 else 0.0 for v in n)
        hi = tuple(half if v > 1 else 1.0 for v in n)
        kind = rng.choice(STREAM_POTENTIALS if dims > 1 else ["harmonic", "barrier"])
        mask = 0.0
        if kind == "harmonic":
            params, dt = {"omega": rng.uniform(0.05, 0.3)}, 0.01
        elif kind == "regularized_coulomb":
            params, dt = {"Z": rng.uniform(0.5, 2.0)}, 0.005
        else:
            params, dt = {"V0": rng.uniform(0.5, 2.0), "x_lo": 0.0, "x_hi": 2.0}, 0.01
            mask = 0.1 * half
        steps = int(rng.uniform(100, max_steps + 1))
        prec = "fp32" if rng.uniform() < 0.5 else "fp64"
        r0 = tuple((-0.3 * half if (a == 0 and kind == "barrier") else
                    rng.uniform(-0.1 * half, 0.1 * half)) if v > 1 else 0.0
                   for a, v in enumerate(n))
        sig = rng.uniform(1.0, 3.0)
        k0 = tuple((rng.uniform(0.8, 1.5) if (a == 0 and kind == "barrier") else
                    rng.uniform(-0.5, 0.5)) if v > 1 else 0.0 for a, v in enumerate(n))
        s = Solve(f"req{i}", n, lo, hi, kind, params, r0,
                  tuple(sig if v > 1 else 1.0 for v in n), k0, dt, steps, prec, mask_width=mask,
                  seed=seed)
        out.append((t, s))
    return out


@dataclass
class Result:
    name: str
    arrival: float
    admitted: float
    finished: float
    psi: np.ndarray | None
    steps: int
    precision: str
    shape: tuple
    obs: dict | None = None  # norm and <E> of the final state, when asked for

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Active:
    req: Solve
    arrival: float
    admitted: float
    prop: Propagator
    remaining: int
    inflight: int = 0       # steps queued and not yet seen finished
    chunk: int = 1          # steps per chunk (streams policy)


@dataclass
class ServeReport:
    results: list
    makespan: float
    quanta: int
    max_active: int
    peak_bytes: int
    extra: dict = field(default_factory=dict)

    def summary(self) -> dict:
        lat = np.array([r.latency for r in self.results])
        return {"solves": len(self.results), "makespan_s": self.makespan,
                "solves_per_s": len(self.results) / self.makespan if self.makespan > 0 else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)),
                "quanta": self.quanta, "max_active": self.max_active,
                "peak_bytes": self.peak_bytes}


_OBS = 1 | 4  # TDSE_OBS_NORM | TDSE_OBS_ENERGY


def _finish(a_req, prop, keep, observe, on_result, arrival, admitted, now):
    """Read a finished request back (and its observables), free its plan."""
    obs = None
    if observe:
        o = prop.observe(_OBS)
        obs = {"norm": o["norm"], "energy": o["energy"]}
    psi = prop.download() if (keep or on_result) else None
    r = Result(a_req.name, arrival, admitted, now(), psi, a_req.steps, a_req.precision, a_req.n,
               obs)
    prop.close()
    if on_result:
        on_result(r)
        if not keep:
            r.psi = None
    return r


def _build(req: Solve):
    return req.build_potential(), req.build_mask(), req.build_psi0()


class Server:
    """Continuous batching server over one device context with an HBM budget."""

    def __init__(self, hbm_budget_bytes: int, quantum: int = 200, device: int = 0,
                 keep_results: bool = True, builder_threads: int = 8,
                 device_build: bool = True, policy: str = "streams", chunk_ms: float = 4.0,
                 observe: bool = False, on_result=None):
        if policy not in ("streams", "batch"):
            raise ValueError("policy must be 'streams' or 'batch'")
        self.ctx = Context(device, hbm_budget_bytes)
        self.budget = hbm_budget_bytes
        self.quantum = quantum
        self.keep = keep_results
        self.device_build = device_build
        self.policy = policy
        self.chunk_ms = chunk_ms
        self.observe = observe
        self.on_result = on_result
        self.pool = cf.ThreadPoolExecutor(max_workers=builder_threads)

    def close(self):
        self.pool.shutdown(wait=True)
        self.ctx.close()

    def footprint(self, req: Solve) -> int:
        return plan_bytes(req.grid(), req.precision, with_mask=req.mask_width > 0)

    def _chunk(self, prop: Propagator) -> int:
        """Steps per chunk for about chunk_ms of device time: the step's bytes
        at 4 TB/s, but no less than 12 us per step (the latency of a small
        grid's launches)."""
        est_s = max(prop.step_bytes() / 4.0e12, 12e-6)
        return int(min(500, max(4, self.chunk_ms * 1e-3 / est_s)))

    def run(self, trace: list) -> ServeReport:
        """Serve [(arrival_s, Solve)] in real time; returns per-request results."""
        for _, req in trace:
            if self.footprint(req) > self.budget:
                raise ValueError(f"{req.name} ({req.n}) alone exceeds the HBM budget")
        pending = list(trace)           # not yet arrived or not yet admitted, by arrival
        building = {}                   # index in trace -> future of host inputs
        active: list[_Active] = []
        results = []
        reserved = 0                    # bytes of admitted plans
        t0 = time.perf_counter()
        quanta = 0
        max_active = 0
        peak = 0
        nxt = 0                         # next trace index whose inputs may be prebuilt
        batch, members = None, ()       # batch policy: the batch of the current active set
        streams = self.policy == "streams"

        def now():
            return time.perf_counter() - t0

        def retire(a):
            nonlocal reserved
            results.append(_finish(a.req, a.prop, self.keep, self.observe, self.on_result,
                                   a.arrival, a.admitted, now))
            reserved -= self.footprint(a.req)

        while pending or active:
            clock = now()
            # prebuild host inputs of requests that have arrived (bounded look-ahead)
            while (not self.device_build and nxt < len(trace) and trace[nxt][0] <= clock
                   and len(building) < 16):
                building[nxt] = self.pool.submit(_build, trace[nxt][1])
                nxt += 1
            # admit in arrival order while memory allows
            while pending and pending[0][0] <= clock:
                arr, req = pending[0]
                idx = len(trace) - len(pending)
                need = self.footprint(req)
                if reserved + need > self.budget:
                    break
                if self.device_build:
                    prop = Propagator(req.grid(), None, req.dt, precision=req.precision,
                                      context=self.ctx)
                    prop.build(req.recipe())
                else:
                    fut = building.pop(idx, None) or self.pool.submit(_build, req)
                    v, m, psi0 = fut.result()
                    prop = Propagator(req.grid(), v, req.dt, precision=req.precision, mask=m,
                                      context=self.ctx)
                    prop.upload(psi0)
                reserved += need
                a = _Active(req, arr, now(), prop, req.steps)
                if streams:
                    a.chunk = self._chunk(prop)
                    a.inflight = min(a.chunk, a.remaining)
                    prop.step(a.inflight)
                    quanta += 1
                active.append(a)
                pending.pop(0)
            peak = max(peak, reserved)
            if not active:
                # idle until the next arrival
                if pending:
                    wait = pending[0][0] - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            max_active = max(max_active, len(active))
            if streams:
                progressed = False
                still = []
                for a in active:
                    if a.inflight and a.prop.done():
                        a.remaining -= a.inflight
                        a.inflight = 0
                        progressed = True
                        if a.remaining == 0:
                            retire(a)
                            continue
                        a.inflight = min(a.chunk, a.remaining)
                        a.prop.step(a.inflight)
                        quanta += 1
                    still.append(a)
                active = still
                if not progressed:
                    time.sleep(0.0002)
                continue
            s = min(self.quantum, min(a.remaining for a in active))
            # a new batch (device tables, streams) only when the active set changed
            if tuple(id(a.prop) for a in active) != members:
                if batch is not None:
                    batch.close()
                batch = Batch([a.prop for a in active], context=self.ctx)
                members = tuple(id(a.prop) for a in active)
            batch.step(s)
            batch.sync()
            quanta += 1
            still = []
            for a in active:
                a.remaining -= s
                if a.remaining == 0:
                    if batch is not None:  # the batch holds the finished plan
                        batch.close()
                        batch, members = None, ()
                    retire(a)
                else:
                    still.append(a)
            active = still
        if batch is not None:
            batch.close()
        return ServeReport(results, now(), quanta, max_active, peak)


def run_sequential(trace: list, precision: str | None = "fp64", device: int = 0,
                   keep_results: bool = False, device_build: bool = True, observe: bool = False,
                   on_result=None) -> ServeReport:
    """Baseline: the same trace solved one request at a time, in plain FP64
    (or `precision`; None: each request's own precision), each plan created,
    stepped to the end and read back before the next request starts."""
    ctx = Context(device)
    t0 = time.perf_counter()
    results = []
    for arr, req in trace:
        wait = arr - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        start = time.perf_counter() - t0
        prec = precision or req.precision
        if device_build:
            p = Propagator(req.grid(), None, req.dt, precision=prec, context=ctx)
            p.build(req.recipe())
        else:
            v, m, psi0 = _build(req)
            p = Propagator(req.grid(), v, req.dt, precision=prec, mask=m, context=ctx)
            p.upload(psi0)
        p.step(req.steps)
        obs = None
        if observe:
            o = p.observe(_OBS)
            obs = {"norm": o["norm"], "energy": o["energy"]}
        psi = p.download() if (keep_results or on_result) else None
        p.close()
        r = Result(req.name, arr, start, time.perf_counter() - t0, psi, req.steps, prec, req.n,
                   obs)
        if on_result:
            on_result(r)
            if not keep_results:
                r.psi = None
        results.append(r)
    report = ServeReport(results, time.perf_counter() - t0, len(trace), 1, 0)
    ctx.close()
    return report
