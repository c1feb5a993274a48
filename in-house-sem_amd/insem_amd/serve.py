"""A continuous-batching request loop on one GPU (no counterpart in the
reference, which runs one model per process).

Requests (a Simulation each) arrive over time and go through four stages:

1. setup: the request's host setup is built on a worker thread as soon as
   it arrives;
2. time step: its power iteration (the lambda_max that fixes its Δt and
   step count, as the request run alone would get them) is queued on the
   device on a short-lived context of its own, and the loop polls it
   without waiting, so that the running batches keep getting their steps.
   The context's set-up is staged: a worker lays it out in host memory for
   a device block the loop reserved, and the loop only copies it over and
   queues the iteration;
3. admission: requests whose time step is known wait for a slot.  Requests
   of at least `large_dofs` degrees of freedom (the 2D layered models) have
   a slot class of their own (`max_large_slots` slots, one request each),
   so that they neither hold up nor wait for the borehole batches; such a
   request runs as a plain Solver on the context its power iteration
   already built.  When a small slot is free, the waiting small requests
   that can share a batch (same polynomial order, coordinate system,
   azimuthal order, media and record cadence; at most `max_batch`) are
   admitted together as one BatchSolver, within the device-memory budget:
   the oldest key whose group is full, or whose oldest request has waited
   `batch_wait` seconds since its time step was known, or that can get no
   more requests (none are still arriving or being set up); among the
   group, the requests whose step counts are closest to the oldest one's
   come first (a batch runs until its longest model ends);
4. stepping: each slot keeps `lookahead` chunks of steps queued on its
   stream (never past the last step of its longest unfinished model), so
   that the device does not wait for the loop; progress is read from
   markers of the queued calls without blocking.  A request's traces are
   read as soon as the device has done its own last step, without waiting
   for the batch's queued steps (bitwise those of the request run alone;
   with their slowness-time coherence for borehole arrays), while the rest
   of its batch runs on.

Failures end one request, not the loop: a setup or power iteration that
raises, a request larger than the memory budget (refused), traces that are
not finite (a model that diverged) and a request still unfinished `timeout`
seconds after its arrival (its batch is stopped when none of its requests
is left) all set the request's `error` and its `done` time.

Host work that does not touch the device (setups, staged contexts, the
coherence of finished traces) runs on worker threads; every device call
stays on the loop's thread.
"""

from __future__ import annotations

import ctypes
import time
from contextlib import contextmanager
from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib, post
from .core import BatchSolver, PowerJob, SetupImage, Simulation, StagedPower, reserve_block, unreserve_block


@dataclass
class Request:
    rid: int
    sim: Simulation
    arrival: float                      # s after the stream's start
    image: Optional[SetupImage] = None
    key: Tuple = ()
    footprint: int = 0                  # estimated device bytes
    n_dofs: int = 0
    large: bool = False
    lam: float = float("nan")           # lambda_max of its power iteration
    ready: float = float("nan")         # when its time step was known
    job: object = None                  # a large request's finished power job (its context)
    stage_bytes: int = 0                # device bytes of its context (for the staged start)
    admitted: float = float("nan")
    done: float = float("nan")
    dt: float = float("nan")
    n_steps: int = 0
    traces: Dict[Tuple[str, str], np.ndarray] = field(default_factory=dict)
    stc: Optional[post.STC] = None
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    @property
    def ok(self) -> bool:
        return self.error is None


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


class _Solo:
    """A large request's Solver behind the slot interface of a batch of one
    model (the slot runs it as the request alone would run)."""

    def __init__(self, solver):
        self.solver = solver
        self.models = [solver.image]

    def dt(self, m: int) -> float:
        return self.solver.dt

    def n_steps(self, m: int) -> int:
        return self.solver.n_steps

    @property
    def step(self) -> int:
        return self.solver.step

    def idle(self) -> bool:
        flag = ctypes.c_int32()
        _lib.check_hip(_lib.hip().insem_hip_query(ctypes.c_void_p(self.solver.context), ctypes.byref(flag)))
        return bool(flag.value)

    def device_step(self) -> int:
        n = ctypes.c_int32()
        _lib.check_hip(_lib.hip().insem_hip_progress(ctypes.c_void_p(self.solver.context), ctypes.byref(n)))
        return n.value

    def advance(self, n: int) -> None:
        self.solver.advance(n)

    def traces(self, m: int):
        return self.solver.traces

    def traces_done(self, m: int):
        # the solver never has steps queued past its last one
        return self.solver.traces

    def close(self) -> None:
        self.solver.close()


@dataclass
class _Slot:
    batch: BatchSolver
    requests: List[Request]
    footprint: int
    large: bool
    open: List[int] = field(default_factory=list)   # models whose traces are still to be returned
    held: bool = False  # its bytes were counted in `used` before admission (a large request's context)

    def idle(self) -> bool:
        return self.batch.idle()


_FAILED = object()


class _StagedStart:
    """A power iteration's start in two halves (Server._start_power)."""

    def __init__(self, server: "Server", pool: ThreadPoolExecutor, r: Request):
        self.server = server
        self.block = reserve_block(r.stage_bytes + 65536)  # slack past the last array, as separate blocks have
        self.future = pool.submit(StagedPower, r.image, self.block)

    def poll(self):
        if not self.future.done():
            return None
        try:
            staged = self.future.result()
        except Exception:
            unreserve_block(self.block)
            raise
        with self.server._timed("power_start_s"):
            return PowerJob.from_staged(staged)


class Server:
    """Serves a stream of requests; see the module docstring.

    hbm_budget: device bytes the served models (and the contexts of running
    power iterations) may hold together (default: 90 % of the device memory
    free at start).  stc: keyword arguments of post.stc for each borehole
    request's first quantity over its receivers (slowness grid, window, ...;
    offsets are the receivers' z), or None.  timeout: seconds after its
    arrival by which a request must have finished, or None."""

    def __init__(self, max_slots: int = 3, max_batch: int = 16, chunk: int = 1024, lookahead: int = 2,
                 hbm_budget: Optional[int] = None, stc: Optional[dict] = None, max_large_slots: int = 1,
                 large_dofs: int = 2_000_000, max_power_jobs: int = 4, timeout: Optional[float] = None,
                 workers: int = 4, batch_wait: float = 0.0):
        if max_slots < 1 or max_batch < 1 or chunk < 1 or max_large_slots < 0 or max_power_jobs < 1:
            raise ValueError("slots, batch size, chunk and power jobs must be positive")
        if batch_wait < 0 or lookahead < 1:
            raise ValueError("batch_wait must be at least 0 and lookahead at least 1")
        self.lookahead = lookahead
        self.batch_wait = batch_wait
        self.max_slots = max_slots
        self.max_batch = max_batch
        self.chunk = chunk
        self.hbm_budget = int(0.9 * device_memory()[0]) if hbm_budget is None else int(hbm_budget)
        self.stc = stc
        # with no large slots, large requests go to the ordinary ones
        self.max_large_slots = max_large_slots
        self.large_dofs = large_dofs if max_large_slots > 0 else None
        self.max_power_jobs = max_power_jobs
        self.timeout = timeout
        self.workers = workers
        self.stats: Dict[str, float] = {}
        self.events: List[tuple] = []

    @contextmanager
    def _timed(self, name: str):
        """Adds the time of the block to stats[name] (the loop thread's time
        in each kind of device call)."""
        t = time.perf_counter()
        try:
            yield
        finally:
            self.stats[name] = self.stats.get(name, 0.0) + time.perf_counter() - t

    # The device-facing steps, one method each (a test replaces them).
    def _setup(self, r: Request) -> SetupImage:
        return r.sim.build_setup()

    def _power_job(self, img: SetupImage):
        return PowerJob(img)

    def _batch(self, images: List[SetupImage], lambdas: List[float]):
        return BatchSolver(images, lambdas)

    def _solo(self, job, image: SetupImage):
        return _Solo(job.into_solver(image))

    def _stage_size(self, image: SetupImage) -> int:
        return image.stage_size()

    def _start_power(self, pool: ThreadPoolExecutor, r: Request):
        """A power iteration's start, polled by the loop (poll() returns the
        job once started): the context's host set-up on a worker, which makes
        no device call, then its copy to the device and the iteration's
        launches on the loop thread (device calls stay on the thread that
        makes every other device call of the loop: contexts set up with
        device calls on worker threads while the loop replays graphs have
        stalled)."""
        return _StagedStart(self, pool, r)

    def _prepare(self, r: Request) -> Request:
        r.image = self._setup(r)
        r.stage_bytes = self._stage_size(r.image)
        r.key = batch_key(r.image)
        # the staged size is the exact device size of its context
        r.footprint = max(footprint_estimate(r.image), r.stage_bytes)
        r.n_dofs = r.image.n_dofs
        r.large = self.large_dofs is not None and r.n_dofs >= self.large_dofs
        return r

    def _admit(self, queue: deque, large: bool, t0: float, used: int, head: Optional[Request] = None):
        """A batch of the waiting requests of one class, led by `head` (default
        the oldest); None when none fits the memory left, _FAILED when the
        batch could not be set up (its requests then fail).  A large request
        (whose power iteration's context holds its model already: its bytes
        are in `used`) becomes a Solver on that context."""
        head = queue[0] if head is None else head
        if large:
            queue.remove(head)
            head.admitted = time.perf_counter() - t0
            img, job = head.image, head.job
            head.image = head.job = None
            try:
                with self._timed("batch_setup_s"):
                    b = self._solo(job, img)
            except Exception as e:  # noqa: BLE001 - the request fails, the loop goes on
                self._fail(head, f"set-up failed: {e}", t0)
                return _FAILED
            finally:
                job.close()
            head.dt, head.n_steps = b.dt(0), b.n_steps(0)
            self.events.append((time.perf_counter() - t0, "admit", [head.rid], [head.n_steps]))
            return _Slot(b, [head], head.footprint, True, [0], held=True)
        limit = self.max_batch
        # the oldest request, then those that can share its batch by how
        # close their step counts are to its own (a batch runs as long as
        # its longest model), oldest first among equals
        same = [r for r in queue if r is not head and r.key == head.key]
        same.sort(key=lambda r: abs(np.log(max(r.n_steps, 1) / max(head.n_steps, 1))))
        group, need = [], 0
        for r in [head] + same:
            if len(group) == limit:
                break
            if used + need + r.footprint > self.hbm_budget:
                if r is head:
                    return None
                continue
            group.append(r)
            need += r.footprint
        now = time.perf_counter() - t0
        for r in group:
            queue.remove(r)
            r.admitted = now
        images = [r.image for r in group]
        for r in group:
            r.image = None  # the batch takes the images over
        try:
            with self._timed("batch_setup_s"):
                b = self._batch(images, [r.lam for r in group])
        except Exception as e:  # noqa: BLE001 - the batch's requests fail, the loop goes on
            for r in group:
                self._fail(r, f"batch set-up failed: {e}", t0)
            return _FAILED
        for m, r in enumerate(group):
            r.dt, r.n_steps = b.dt(m), b.n_steps(m)
        self.events.append((time.perf_counter() - t0, "admit", [r.rid for r in group], [r.n_steps for r in group]))
        return _Slot(b, group, need, large, list(range(len(group))))

    def _batch_head(self, queue: deque, now: float, upstream: bool) -> Optional[Request]:
        """The oldest request of the first key (by its oldest request) whose
        group is ready to run: full, or waited batch_wait seconds, or with no
        more requests to come; None when every group is still gathering."""
        seen = {}
        for r in queue:
            if r.key not in seen:
                seen[r.key] = [r, 0]
            seen[r.key][1] += 1
        for head, n in seen.values():
            if n >= self.max_batch or not upstream or now - head.ready >= self.batch_wait:
                return head
        return None

    def _coherence(self, r: Request, step: float) -> None:
        q = r.sim.receivers[0].quantities[0]
        recs = [rec for rec in r.sim.receivers if (rec.name, q) in r.traces]
        x = np.array([r.traces[(rec.name, q)][:, 0] for rec in recs])
        z = np.array([rec.position[1] for rec in recs])
        r.stc = post.stc(x, z, step, **self.stc)

    def _deliver(self, slot: _Slot, m: int, pool: ThreadPoolExecutor, t0: float) -> Optional[Future]:
        """Returns model m's traces (its steps are done and the slot idle);
        the coherence of a borehole array goes to the workers."""
        b, r = slot.batch, slot.requests[m]
        if r.error is not None:
            return None
        r.traces = b.traces_done(m)
        if not all(np.isfinite(v).all() for v in r.traces.values()):
            r.error = "non-finite traces: the model diverged"
            r.done = time.perf_counter() - t0
            return None
        if self.stc is not None and b.models[m].axisymmetric and len(r.sim.receivers) > 1:
            step = b.models[m].time.record_every * r.dt

            def job(r=r, step=step):
                self._coherence(r, step)
                r.done = time.perf_counter() - t0

            return pool.submit(job)
        r.done = time.perf_counter() - t0
        return None

    def _fail(self, r: Request, msg: str, t0: float) -> None:
        if r.error is None:
            r.error = msg
            r.done = time.perf_counter() - t0

    def serve(self, arrivals: Sequence[Tuple[float, Simulation]]) -> List[Request]:
        """Runs the stream (arrival times in s from now) to completion; returns
        the requests in arrival order with their timings and results."""
        reqs = [Request(i, sim, float(t)) for i, (t, sim) in enumerate(sorted(arrivals, key=lambda a: a[0]))]
        self.stats = {"iterations": 0, "power_start_s": 0.0, "power_result_s": 0.0, "batch_setup_s": 0.0,
                      "advance_s": 0.0, "deliver_s": 0.0, "close_s": 0.0}
        self.events = []  # (s since start, "admit", request ids, their step counts)
        pending = deque(reqs)
        building: deque = deque()   # (request, future of its setup) in arrival order
        timing: deque = deque()     # requests waiting for a power-iteration slot
        power: List[Tuple[Request, object]] = []
        queues = {False: deque(), True: deque()}   # waiting for admission, by slot class
        slots: List[_Slot] = []
        jobs: List[Future] = []
        used = 0
        t0 = time.perf_counter()

        def expired(r: Request, now: float) -> bool:
            return self.timeout is not None and r.error is None and now - r.arrival > self.timeout

        with ThreadPoolExecutor(max_workers=max(1, self.workers)) as pool:
            while pending or building or timing or power or queues[False] or queues[True] or slots or jobs:
                now = time.perf_counter() - t0
                busy = False
                while pending and pending[0].arrival <= now:
                    r = pending.popleft()
                    building.append((r, pool.submit(self._prepare, r)))
                # 1. setups, in arrival order
                while building and building[0][1].done():
                    r, f = building.popleft()
                    try:
                        f.result()
                    except Exception as e:  # noqa: BLE001 - the request fails, the loop goes on
                        self._fail(r, f"setup failed: {e}", t0)
                        continue
                    if r.footprint > self.hbm_budget:
                        self._fail(r, f"refused: needs about {r.footprint} device bytes, more than the budget of "
                                      f"{self.hbm_budget}", t0)
                        r.image = None
                        continue
                    timing.append(r)
                # 2. power iterations: started within the budget, polled
                # without waiting
                while timing and len(power) < self.max_power_jobs and used + timing[0].footprint <= self.hbm_budget:
                    r = timing.popleft()
                    try:
                        power.append([r, self._start_power(pool, r), None])
                        used += r.footprint
                    except Exception as e:  # noqa: BLE001 - the request fails, the loop goes on
                        self._fail(r, f"power iteration failed: {e}", t0)
                        r.image = None
                    busy = True
                for entry in list(power):
                    r, start, job = entry
                    if job is None:
                        try:
                            job = start.poll()
                        except Exception as e:  # noqa: BLE001 - the request fails, the loop goes on
                            power.remove(entry)
                            used -= r.footprint
                            self._fail(r, f"power iteration failed: {e}", t0)
                            r.image = None
                            continue
                        if job is None:
                            continue
                        entry[2] = job
                        busy = True
                    if not job.ready():
                        continue
                    power.remove(entry)
                    if r.large:
                        # the job's context becomes the request's solver: its
                        # bytes stay counted until the slot closes
                        r.job = job
                        r.ready = now
                        queues[True].append(r)
                        busy = True
                        continue
                    used -= r.footprint
                    try:
                        with self._timed("power_result_s"):
                            r.lam = job.lambda_max()
                        r.dt, r.n_steps = r.image.time_step(r.lam)
                    except Exception as e:  # noqa: BLE001
                        self._fail(r, f"power iteration failed: {e}", t0)
                        r.image = None
                        continue
                    finally:
                        job.close()
                    r.ready = now
                    queues[False].append(r)
                    busy = True
                # requests past their deadline before admission
                for q in queues.values():
                    for r in [r for r in q if expired(r, now)]:
                        q.remove(r)
                        r.image = None
                        if r.job is not None:
                            r.job.close()
                            r.job = None
                            used -= r.footprint
                        self._fail(r, f"timed out after {self.timeout:g} s waiting for a slot", t0)
                # 3. admission, per slot class
                upstream = bool(pending or building or timing or power)
                for large, cap in ((True, self.max_large_slots), (False, self.max_slots)):
                    q = queues[large]
                    while q and sum(s.large == large for s in slots) < cap:
                        head = q[0] if large else self._batch_head(q, now, upstream)
                        if head is None:
                            break  # every key's group is still gathering requests
                        slot = self._admit(q, large, t0, used, head)
                        if slot is _FAILED:
                            continue
                        if slot is None:
                            break
                        slots.append(slot)
                        if not slot.held:
                            used += slot.footprint
                        busy = True
                # 4. stepping and results
                for slot in list(slots):
                    b = slot.batch
                    done = b.device_step()
                    for m in [m for m in slot.open if slot.requests[m].n_steps <= done]:
                        slot.open.remove(m)
                        with self._timed("deliver_s"):
                            j = self._deliver(slot, m, pool, t0)
                        if j is not None:
                            jobs.append(j)
                    for m in [m for m in slot.open if expired(slot.requests[m], now)]:
                        slot.open.remove(m)
                        self._fail(slot.requests[m], f"timed out after {self.timeout:g} s", t0)
                    if not slot.open:
                        with self._timed("close_s"):
                            b.close()  # waits for steps still queued (of finished or expired models)
                        slots.remove(slot)
                        used -= slot.footprint
                        continue
                    last = max(slot.requests[m].n_steps for m in slot.open)
                    while b.step < last and b.step - done < self.lookahead * self.chunk:
                        with self._timed("advance_s"):
                            b.advance(min(self.chunk, last - b.step))
                        busy = True
                for j in [j for j in jobs if j.done()]:
                    j.result()
                    jobs.remove(j)
                self.stats["iterations"] += 1
                if not busy:
                    idle = not (building or timing or power or queues[False] or queues[True] or slots or jobs)
                    if idle and pending:
                        time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                    else:
                        time.sleep(2e-4)
        return reqs


def summary(reqs: Sequence[Request], wall: float) -> dict:
    """Solves/s and latency percentiles (ms) of a served stream's finished
    requests, and the count of failed ones."""
    ok = [r for r in reqs if r.ok]
    lat = np.array([r.latency for r in ok] or [np.nan]) * 1e3
    wait = np.array([r.admitted - r.arrival for r in ok] or [np.nan]) * 1e3
    return {"solves": len(ok), "failed": len(reqs) - len(ok), "solves_per_s": len(ok) / wall,
            "latency_p50_ms": float(np.percentile(lat, 50)), "latency_p99_ms": float(np.percentile(lat, 99)),
            "queue_wait_p50_ms": float(np.percentile(wait, 50)), "queue_wait_p99_ms": float(np.percentile(wait, 99))}
