"""Continuous-batching inversion server on one GPU.

Requests (``invert.Request``) arrive over time.  The server owns one plan
whose slots cover a fixed mix of grid sizes.  The plan's wavefields,
checkpoints and coefficient fields (every slot's, allocated together), and
the line-search fan-out plan when the machines evaluate line searches, are
reserved against the HBM budget when the server is built; a request is
admitted into a free slot of its size when its own state (LBFGS history,
observed data, vector temporaries) fits what is left.  Waiting requests are admitted first-fit in arrival order,
so one waiting for a busy size does not hold back the others.  Every loop iteration collects the pending propagator op
of every in-flight request and runs them as one batched launch per op kind;
requests that arrive meanwhile join at the next iteration (continuous
batching, not a queue of whole inversions).  Per-request state (model,
gradients, LBFGS pairs, observed data) stays on the device.

The reference has no serving layer: it inverts samples one by one, one
process per GPU (core/kaggle_support.py:482-495).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np
import torch

from ._lib import CAP_ADJOINT, CAP_FORWARD, CAP_TANGENT
from .geometry import Geometry
from .invert import Executor, Request, fanout_plan_bytes, lbfgs_machine
from .propagator import plan_bytes


@dataclass
class Served:
    rid: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    result: torch.Tensor | None = None
    stats: dict = field(default_factory=dict)
    failed: bool = False

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def request_footprint(g: Geometry, history: int) -> int:
    """Device bytes one request holds beyond the plans: its LBFGS history
    ring (history + 1 pairs of float64 vectors plus coefficients), about
    eight float64 vectors of its loop (x, d, g, previous g, trial points) and
    its float32 observed data."""
    n = g.n_params
    lbfgs = 8 * ((history + 1) * (2 * n + 2) + 1) + 8
    return lbfgs + 8 * n * 8 + 4 * g.n_data


class InversionServer:
    """Admits inversion requests into fixed slots and batches their
    propagator ops.  Each admitted request runs ``machine(req)``: by default
    LBFGS with the given settings; ``schedule.serving_machine()`` runs the
    reference's classified multi-stage chains instead (then
    ``history_size`` should bound the longest LBFGS stage's pairs, which
    the footprint reserves).  ``ckpt_every=None`` keeps the full forward
    history of every slot when all of them fit ``hbm_budget`` and
    checkpoints every 50 steps otherwise; results are bit-identical either
    way, the full history saves the adjoint's segment recomputes."""

    def __init__(self, slots: Sequence[Geometry], hbm_budget: int, ckpt_every: int | None = 50,
                 max_iter: int = 30, history_size: int = 10, tolerance_grad: float = 1e-7,
                 tolerance_change: float = 0.0, device="cuda", machine: Callable | None = None,
                 line_search: bool | None = None):
        self.slots = list(slots)
        self.budget = int(hbm_budget)
        caps = CAP_FORWARD | CAP_TANGENT | CAP_ADJOINT
        # the default machine (LBFGS) runs no line-search fan-out; a custom
        # machine (Gauss-Newton passes, chains) may
        self.line_search = (machine is not None) if line_search is None else bool(line_search)
        fan = fanout_plan_bytes(self.slots) if self.line_search else 0
        if ckpt_every is None:
            # keep every slot's full forward history when the plan fits the
            # budget (the adjoint then reads it instead of recomputing
            # checkpoint segments), else checkpoint every 50 steps
            full = plan_bytes(self.slots, ckpt_every=0, caps=caps)
            ckpt_every = 0 if full + fan <= self.budget else 50
        self.ckpt_every = ckpt_every
        self.ex = Executor(self.slots, ckpt_every=ckpt_every, device=device)
        self.plan_bytes = plan_bytes(self.slots, ckpt_every=ckpt_every, caps=caps)
        self.reserved = self.plan_bytes + fan
        if self.reserved > self.budget:
            raise ValueError(f"the plans of these slots need {self.reserved} bytes, over the HBM budget "
                             f"of {self.budget}")
        self.kw = dict(max_iter=max_iter, history_size=history_size, tolerance_grad=tolerance_grad,
                       tolerance_change=tolerance_change)
        self.machine = machine if machine is not None else (lambda req: lbfgs_machine(req, **self.kw))
        self._fp = {g: request_footprint(g, history_size) for g in set(self.slots)}
        self.used = self.reserved
        self.free: dict = {}
        for i, g in enumerate(self.slots):
            self.free.setdefault(g, []).append(i)
        self.log: list = []

    def footprint(self, req: Request) -> int:
        return self._fp[req.geom]

    def _execute(self, pending: dict) -> dict:
        """One loop iteration's batched launches; with ``timing`` set, HIP
        events on the launch stream bracket them (see ``gpu_busy_s``)."""
        if not getattr(self, "timing", False):
            return self.ex.execute(pending)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out = self.ex.execute(pending)
        e1.record()
        self._spans.append((e0, e1))
        return out

    def start_timing(self):
        """Record the device time of every loop iteration from now on."""
        self.timing = True
        self._spans = []

    def gpu_busy_s(self) -> float:
        """Device time spanned by the recorded iterations' launches (from the
        first to the last operation each iteration enqueued, so it includes
        gaps the host leaves between launches inside one iteration); the rest
        of the wall time is host work between iterations (machine steps,
        scalar reads, admission)."""
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in getattr(self, "_spans", [])) / 1e3

    def _check_admissible(self, req: Request):
        """Refuse a request no state of the server could ever admit: no slot
        of its size, or a footprint above the whole budget."""
        if req.geom not in self.free:
            raise ValueError(f"no slot of this server fits a {req.geom.nz}x{req.geom.nx} request")
        if getattr(self, "reserved", 0) + self.footprint(req) > self.budget:
            raise ValueError(f"a {req.geom.nz}x{req.geom.nx} request's footprint ({self.footprint(req)} bytes) "
                             f"exceeds the server's HBM budget ({self.budget} bytes) beyond its plans")

    def _try_admit(self, req: Request):
        free = self.free.get(req.geom)
        if not free:
            return None
        if self.used + self.footprint(req) > self.budget:
            return None
        slot = free.pop()
        self.used += self.footprint(req)
        return slot

    def _release(self, slot: int, req: Request):
        self.free[self.slots[slot]].append(slot)
        self.used -= self.footprint(req)

    @staticmethod
    def _finish(rec: Served, req: Request, value, t: float):
        rec.finished = t
        rec.result = value
        rec.stats = dict(req.meta.get("stats", {}))
        if "family" in req.meta:
            rec.stats["family"] = req.meta["family"]
            rec.stats["stages"] = list(req.meta.get("stages", []))
        rec.failed = bool(torch.isnan(value).any())

    def run(self, requests: Sequence[Request], arrivals: Sequence[float],
            on_event: Callable | None = None) -> list:
        """Serve `requests` arriving at `arrivals` (seconds from start).
        Returns one Served record per request (same order)."""
        order = np.argsort(arrivals, kind="stable")
        queue = deque((requests[i], float(arrivals[i]), int(i)) for i in order)
        out = [Served(rid=requests[i].rid, arrival=float(arrivals[i])) for i in range(len(requests))]
        waiting: deque = deque()
        active: dict = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while queue or waiting or active:
            t = now()
            while queue and queue[0][1] <= t:
                self._check_admissible(queue[0][0])
                waiting.append(queue.popleft())
            # first fit in arrival order: a request that cannot be admitted
            # (no free slot of its size, or footprint over budget) does not
            # hold back later requests that can
            blocked: deque = deque()
            while waiting:
                req, arr, idx = waiting.popleft()
                slot = self._try_admit(req)
                if slot is None:
                    blocked.append((req, arr, idx))
                    continue
                self.ex.set_target(slot, req.target)
                m = self.machine(req)
                out[idx].admitted = now()
                try:
                    active[slot] = [req, idx, m, next(m)]
                except StopIteration as e:
                    self._finish(out[idx], req, e.value, now())
                    self._release(slot, req)
                    if on_event:
                        on_event(out[idx])
            waiting = blocked
            if not active:
                if waiting:        # an idle server admits every admissible request
                    raise RuntimeError("waiting requests could not be admitted on an idle server")
                if queue:
                    time.sleep(max(0.0, queue[0][1] - now()))
                continue
            pending = {s: v[3] for s, v in active.items()}
            results = self._execute(pending)
            self.log.append((now(), len(pending)))
            for s, res in results.items():
                req, idx, m, _ = active[s]
                try:
                    active[s][3] = m.send(res)
                except StopIteration as e:
                    rec = out[idx]
                    self._finish(rec, req, e.value, now())
                    del active[s]
                    self._release(s, req)
                    if on_event:
                        on_event(rec)
        return out

    def run_stream(self, source, on_event: Callable | None = None) -> list:
        """Serve requests pulled from the iterable ``source`` as slots free
        up: the next request is pulled once the previous one is admitted, so
        a long request never holds a batch back (a data set streamed from
        disk keeps every slot busy).  Returns the Served records in
        admission order; ``on_event`` sees each one as it finishes."""
        it = iter(source)
        nxt = next(it, None)
        out: list = []
        active: dict = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while nxt is not None or active:
            while nxt is not None:
                self._check_admissible(nxt)
                slot = self._try_admit(nxt)
                if slot is None:
                    break
                rec = Served(rid=nxt.rid, arrival=now(), admitted=now())
                out.append(rec)
                self.ex.set_target(slot, nxt.target)
                m = self.machine(nxt)
                try:
                    active[slot] = [nxt, rec, m, next(m)]
                except StopIteration as e:
                    self._finish(rec, nxt, e.value, now())
                    self._release(slot, nxt)
                    if on_event:
                        on_event(rec)
                nxt = next(it, None)
            if not active:
                if nxt is not None:   # admissible, yet refused by an idle server
                    raise RuntimeError("a request could not be admitted on an idle server")
                break
            results = self._execute({s: v[3] for s, v in active.items()})
            self.log.append((now(), len(results)))
            for s, res in results.items():
                req, rec, m, _ = active[s]
                try:
                    active[s][3] = m.send(res)
                except StopIteration as e:
                    self._finish(rec, req, e.value, now())
                    del active[s]
                    self._release(s, req)
                    if on_event:
                        on_event(rec)
        return out


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    """Arrival times of a Poisson process with `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n)) - 0.0


def percentiles(latencies: Sequence[float]) -> dict:
    a = np.asarray(latencies, dtype=np.float64)
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)),
            "mean": float(a.mean()), "max": float(a.max())}
