"""Serving: a stream of scene-solve requests, continuously batched on one GPU.

A request is a scene (SceneConfig), a step budget and an arrival time.
Requests wait in arrival order (policy "fifo"), or ordered by their work,
cells x step budget, smallest first (policy "shortest": lower median latency
under load, at the price of making large requests wait behind later small
ones) or largest first (policy "longest": for a batch submitted at once, the
largest scenes start first, so that the end of the batch is not one large
scene running alone). The head of the queue is admitted when a slot
is free and its device footprint (``ffr_scene_footprint``) fits the context's
HBM budget beside the scenes already resident; the scene is then built (flags,
boundary lists, initial temperature: ``device.prepare``), created on the
device and submitted to the native serving loop (``ffr_submit``), which
overlaps the steps of every admitted scene on per-scene streams. When a
request finishes (its step budget, or FluidSolver::run's convergence test),
its fields are copied to the host if asked for and its device memory is
freed, which admits the next requests.

Latency is host time from a request's arrival to its result being on the
host; throughput is the sum of cells x completed steps over the wall time of
the stream. There is no counterpart in the reference (one scene per process).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Deque, Dict, List, Optional

import numpy as np

from . import autotune, capi, device
from .config import SceneConfig


@dataclass
class Request:
    config: SceneConfig
    max_steps: int
    arrival: float = 0.0  # seconds after the start of the stream


@dataclass
class Completion:
    index: int
    cells: int
    steps: int
    arrival: float
    admitted: float
    finished: float
    last: dict
    fields: Optional[Dict[str, np.ndarray]] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Active:
    index: int
    req: Request
    scene: "capi.DeviceScene"
    cells: int
    admitted: float


def percentile(xs: List[float], q: float) -> float:
    """q-th percentile (0..100), linear interpolation between order statistics."""
    if not xs:
        return float("nan")
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


class Server:
    """Admission, submission and completion around the native serving loop.

    ``overrides`` are ffr_params fields applied to every request (for example
    ``precision=capi.PREC_MIXED``). ``gangs`` >= 1 advances the active scenes
    in that many groups of merged launches (ffr_serve_gangs); 0 runs every
    scene on its own stream, run by ``threads`` host threads (ffr_serve_threads;
    by default one per slot)."""

    def __init__(self, ctx: capi.Context, slots: int = 6, hbm_budget: Optional[int] = None,
                 keep_fields: bool = False, tuned: bool = True, gangs: int = 0,
                 policy: str = "fifo", threads: Optional[int] = None, **overrides):
        if policy not in ("fifo", "shortest", "longest"):
            raise ValueError(f"unknown admission policy {policy!r} (fifo, shortest, longest)")
        self.ctx = ctx
        self.slots = slots
        self.policy = policy
        ctx.serve_config(slots)
        ctx.serve_threads(threads if threads is not None else max(slots, 1))
        ctx.serve_gangs(gangs)
        info = ctx.device_info()
        self.budget = hbm_budget if hbm_budget is not None else int(info["hbm_bytes"] * 0.8)
        self.keep_fields = keep_fields
        self.tuned = tuned
        # scenes advanced by worker threads launch eagerly: measured on
        # MI355X, replayed graphs of several scenes at once do not overlap
        # (about 960 scene-steps/s for four 72x49x58 rooms against 1600 eager)
        if threads is None or threads > 1:
            overrides.setdefault("graphs", 0)
        self.overrides = overrides
        self.pending: Deque = deque()   # (index, Request, PreparedScene or None)
        self.active: Dict[int, _Active] = {}
        self.done: List[Completion] = []
        self._next = 0
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request) -> int:
        i = self._next
        self._next += 1
        self.pending.append((i, req, None))
        return i

    @staticmethod
    def work(req: Request) -> int:
        """Cells x step budget: the admission key of policies "shortest" and "longest"."""
        g = req.config.room.grid
        return int(g[0]) * int(g[1]) * int(g[2]) * int(req.max_steps)

    def _admit(self):
        if self.policy != "fifo" and len(self.pending) > 1:
            sign = 1 if self.policy == "shortest" else -1
            self.pending = deque(sorted(self.pending, key=lambda e: (sign * self.work(e[1]), e[0])))
        while self.pending and len(self.active) < self.slots:
            i, req, ps = self.pending[0]
            if ps is None:
                # the cells are marked on the device: no host Grid per arrival
                ps = device.prepare(req.config, on_device=True, **self.overrides)
                if self.tuned:  # per-scene settings of the autotuner's table for this size class
                    entry = autotune.lookup(ps.nx, ps.ny, ps.nz, ps.params.precision)
                    if entry:
                        autotune.apply(ps.params, {k: v for k, v in entry.items()
                                                   if k != "max_blocks"})
                self.pending[0] = (i, req, ps)
            need = ps.footprint()
            if need == 0:
                raise ValueError(f"request {i}: invalid scene descriptor")
            if need > self.budget:
                raise MemoryError(f"request {i} needs {need} bytes, above the budget {self.budget}")
            if self.ctx.bytes_in_use() + need > self.budget:
                break  # wait for memory to be freed by finishing requests
            self.pending.popleft()
            scene = device.create(self.ctx, ps)
            ticket = self.ctx.submit(scene, req.max_steps)
            self.active[ticket] = _Active(i, req, scene, ps.cells, self.now())

    def pump(self) -> int:
        """Admits what fits, advances every active scene once and collects
        finished requests; returns the number of steps completed."""
        self._admit()
        n = self.ctx.progress() if self.active else 0
        for ticket in list(self.active):
            steps, fin = self.ctx.ticket_info(ticket)
            if not fin:
                continue
            a = self.active.pop(ticket)
            _, last = self.ctx.poll(ticket)
            fields = a.scene.download() if self.keep_fields else None
            a.scene.close()
            self.ctx.release(ticket)
            self.done.append(Completion(a.index, a.cells, steps, a.req.arrival, a.admitted,
                                        self.now(), last, fields))
        if self.done:
            self._admit()
        return n

    def idle(self) -> bool:
        return not self.pending and not self.active

    def run(self, requests: List[Request], realtime: bool = True) -> List[Completion]:
        """Feeds `requests` (sorted by arrival) and serves until all finished.
        With realtime, a request enters the queue when the wall clock reaches
        its arrival time; otherwise all enter at once."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        k = 0
        while k < len(reqs) or not self.idle():
            t = self.now()
            while k < len(reqs) and (not realtime or reqs[k].arrival <= t):
                r = reqs[k]
                if not realtime:
                    r = Request(r.config, r.max_steps, 0.0)
                self.submit(r)
                k += 1
            n = self.pump()
            if n == 0 and self.idle() and k < len(reqs):
                time.sleep(min(max(reqs[k].arrival - self.now(), 0.0), 0.001))
            elif n == 0:
                time.sleep(5e-5)  # the native loop's threads advance the scenes meanwhile
        return sorted(self.done, key=lambda c: c.index)


def summary(done: List[Completion]) -> dict:
    """p50/p99 latency and aggregate cell-updates/s of a served stream."""
    lat = [c.latency for c in done]
    start = min(c.arrival for c in done)
    end = max(c.finished for c in done)
    work = sum(c.cells * c.steps for c in done)
    return {"requests": len(done), "p50_latency_s": percentile(lat, 50),
            "p99_latency_s": percentile(lat, 99), "wall_s": end - start,
            "cell_updates_per_s": work / max(end - start, 1e-12),
            "steps": sum(c.steps for c in done), "cell_updates": work}
