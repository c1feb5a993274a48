"""A continuous-batching request loop on one GPU (no counterpart in the
reference, which runs one model per process).

Requests (a Simulation each) arrive over time.  The loop keeps up to
`max_slots` batches on the device at once, each on its own stream: when a
slot is free and the device memory allows, the queued requests that can
share a batch with the oldest one (same polynomial order, coordinate system,
azimuthal order, media and record cadence; at most `max_batch`) are admitted
together as one BatchSolver; requests that arrive while others run wait
only for a slot, not for the others to finish.  Every slot is kept busy by
queueing its next chunk of steps as soon as its stream is idle; a finished
batch returns each request's traces (bitwise those of the request run
alone) and, for borehole arrays, their slowness-time coherence.

Host work that does not touch the device (building a request's setup on
arrival, the coherence of finished traces) runs on worker threads, so the
loop keeps feeding the device meanwhile; every device call stays on the
loop's thread.
"""

from __future__ import annotations

import ctypes
import time
from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib, post
from .core import BatchSolver, SetupImage, Simulation


@dataclass
class Request:
    rid: int
    sim: Simulation
    arrival: float                      # s after the stream's start
    image: Optional[SetupImage] = None
    key: Tuple = ()
    footprint: int = 0                  # estimated device bytes
    admitted: float = float("nan")
    done: float = float("nan")
    dt: float = float("nan")
    n_steps: int = 0
    traces: Dict[Tuple[str, str], np.ndarray] = field(default_factory=dict)
    stc: Optional[post.STC] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def batch_key(img: SetupImage) -> Tuple:
    """Requests with equal keys can share a batch."""
    return (img.ngll, img.axisymmetric, img.azimuthal_order, tuple(d["medium"] for d in img.domains),
            img.time.record_every)


def footprint_estimate(img: SetupImage) -> int:
    """Upper estimate of the device bytes of one model (fields, mass, CSR,
    element data, force buffers), for admission."""
    npe = img.ngll * img.ngll
    total = 0
    for d in img.domains:
        c, ng, ne = d["ncomp"], d["n_global"], d["n_elements"]
        slots = ne * npe
        total += 4 * (5 * c * ng + 3 * ng)       # u, v, a, f, scratch; mass, 1/m, CSR offsets
        total += 4 * slots * (1 + 2 * c + 4 * c)  # CSR addresses, f_local twice, group rows (<= 4 c per slot)
        total += slots * 5                       # l2g + rank
        total += 4 * slots * (8 + 2 + 3)         # metric, coefficients, axisymmetric factors
    return int(total * 1.1) + (1 << 20)


def device_memory() -> Tuple[int, int]:
    """(free, total) bytes of the current device."""
    f, t = ctypes.c_int64(), ctypes.c_int64()
    _lib.check_hip(_lib.hip().insem_hip_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


@dataclass
class _Slot:
    batch: BatchSolver
    requests: List[Request]
    footprint: int

    def idle(self) -> bool:
        return self.batch.idle()


class Server:
    """Serves a stream of requests; see the module docstring.

    hbm_budget: device bytes the served models may hold together (default:
    90 % of the device memory free at start).  stc: keyword arguments of
    post.stc for each borehole request's first quantity over its receivers
    (slowness grid, window, ...; offsets are the receivers' z), or None."""

    def __init__(self, max_slots: int = 4, max_batch: int = 16, chunk: int = 1024,
                 hbm_budget: Optional[int] = None, stc: Optional[dict] = None):
        self.max_slots = max_slots
        self.max_batch = max_batch
        self.chunk = chunk
        self.hbm_budget = int(0.9 * device_memory()[0]) if hbm_budget is None else int(hbm_budget)
        self.stc = stc

    def _admit(self, queue: deque, now: float, used: int) -> Optional[_Slot]:
        head = queue[0]
        group, need = [], 0
        for r in list(queue):
            if r.key != head.key or len(group) == self.max_batch:
                continue
            if used + need + r.footprint > self.hbm_budget:
                break
            group.append(r)
            need += r.footprint
        if not group:
            return None
        for r in group:
            queue.remove(r)
            r.admitted = now
        images = [r.image for r in group]
        for r in group:
            r.image = None  # the batch takes the images over
        b = BatchSolver(images)
        for m, r in enumerate(group):
            r.dt, r.n_steps = b.dt(m), b.n_steps(m)
        return _Slot(b, group, need)

    def _coherence(self, r: Request, step: float) -> None:
        q = r.sim.receivers[0].quantities[0]
        recs = [rec for rec in r.sim.receivers if (rec.name, q) in r.traces]
        x = np.array([r.traces[(rec.name, q)][:, 0] for rec in recs])
        z = np.array([rec.position[1] for rec in recs])
        r.stc = post.stc(x, z, step, **self.stc)

    def _finish(self, slot: _Slot, pool: ThreadPoolExecutor, t0: float) -> List[Future]:
        """Reads the traces of a finished batch; the coherence of borehole
        arrays goes to the workers.  Returns their futures."""
        b = slot.batch
        jobs = []
        for m, r in enumerate(slot.requests):
            r.traces = b.traces(m)
            if self.stc is not None and b.models[m].axisymmetric and len(r.sim.receivers) > 1:
                step = b.models[m].time.record_every * r.dt

                def job(r=r, step=step):
                    self._coherence(r, step)
                    r.done = time.perf_counter() - t0

                jobs.append(pool.submit(job))
            else:
                r.done = time.perf_counter() - t0
        b.close()
        return jobs

    @staticmethod
    def _prepare(r: Request) -> Request:
        r.image = r.sim.build_setup()
        r.key = batch_key(r.image)
        r.footprint = footprint_estimate(r.image)
        return r

    def serve(self, arrivals: Sequence[Tuple[float, Simulation]]) -> List[Request]:
        """Runs the stream (arrival times in s from now) to completion; returns
        the requests in arrival order with their timings and results."""
        reqs = [Request(i, sim, float(t)) for i, (t, sim) in enumerate(sorted(arrivals, key=lambda a: a[0]))]
        pending = deque(reqs)
        building: deque = deque()   # futures of setups in arrival order
        queue: deque = deque()
        slots: List[_Slot] = []
        jobs: List[Future] = []
        used = 0
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=4) as pool:
            while pending or building or queue or slots or jobs:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    building.append(pool.submit(self._prepare, pending.popleft()))
                while building and building[0].done():
                    r = building.popleft().result()
                    if r.footprint > self.hbm_budget:
                        raise MemoryError(f"request {r.rid} needs about {r.footprint} device bytes, "
                                          f"more than the budget of {self.hbm_budget}")
                    queue.append(r)
                while queue and len(slots) < self.max_slots:
                    slot = self._admit(queue, time.perf_counter() - t0, used)
                    if slot is None:
                        break
                    slots.append(slot)
                    used += slot.footprint
                busy = False
                for slot in list(slots):
                    b = slot.batch
                    if b.step >= b.max_steps:
                        if slot.idle():
                            jobs += self._finish(slot, pool, t0)
                            slots.remove(slot)
                            used -= slot.footprint
                    elif slot.idle():
                        b.advance(self.chunk)
                        busy = True
                for j in [j for j in jobs if j.done()]:
                    j.result()
                    jobs.remove(j)
                if not busy:
                    if not slots and not queue and not building and not jobs and pending:
                        time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                    else:
                        time.sleep(2e-4)
        return reqs


def summary(reqs: Sequence[Request], wall: float) -> dict:
    """Solves/s and latency percentiles (ms) of a served stream."""
    lat = np.array([r.latency for r in reqs]) * 1e3
    wait = np.array([r.admitted - r.arrival for r in reqs]) * 1e3
    return {"solves": len(reqs), "solves_per_s": len(reqs) / wall, "latency_p50_ms": float(np.percentile(lat, 50)),
            "latency_p99_ms": float(np.percentile(lat, 99)), "queue_wait_p50_ms": float(np.percentile(wait, 50)),
            "queue_wait_p99_ms": float(np.percentile(wait, 99))}
