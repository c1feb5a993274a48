"""The request loop's host logic on the CPU: stages, slot classes, memory
budget, early return of finished models and failure handling, with the
device-facing steps (setup, power iteration, batch) replaced by stand-ins
that count what the loop asks of them.  The device path itself is covered
by the GPU tests of the served mix."""

from __future__ import annotations

import threading
import time
from types import SimpleNamespace

import numpy as np
import pytest

from insem_amd import serve


class FakeImage:
    """What the loop reads of a setup image."""

    def __init__(self, spec):
        self.spec = spec
        self.ngll = spec.get("order", 4) + 1
        self.axisymmetric = spec.get("axisym", True)
        self.azimuthal_order = 1
        self.domains = [dict(medium="elastic", ncomp=3, n_global=spec["points"], n_elements=spec["points"] // 16)]
        self.time = SimpleNamespace(record_every=1)
        self.n_dofs = 3 * spec["points"]

    def time_step(self, lam):
        return 2.0 / np.sqrt(lam), self.spec["steps"]


def solo_traces(spec):
    """The traces a model gives alone: a function of its own spec only."""
    n = spec["steps"] + 1
    v = np.sin(np.arange(n) * 0.01 * (1 + spec["seed"]))[:, None]
    if spec.get("diverge"):
        v = v.copy()
        v[-1] = np.nan
    return {("r0", "displacement"): v}


class FakeJob:
    def __init__(self, img, log):
        self.img, self.polls, self.log = img, 0, log

    def ready(self):
        self.polls += 1
        return self.polls > self.img.spec.get("power_polls", 2)

    def lambda_max(self):
        if self.img.spec.get("power_fails"):
            raise RuntimeError("device error")
        return 4.0 / self.img.spec["dt"] ** 2

    def close(self):
        self.log.append(("power_closed", self.img.spec["seed"]))


class FakeBatch:
    """Steps advance at once; each model's traces are its solo traces."""

    def __init__(self, images, lambdas, log):
        self.models = images
        self.lambdas = lambdas
        self.step = 0
        self.max_steps = max(im.spec["steps"] for im in images)
        self.log = log
        self.closed = False
        log.append(("batch", [im.spec["seed"] for im in images]))

    def dt(self, m):
        return 2.0 / np.sqrt(self.lambdas[m])

    def n_steps(self, m):
        return self.models[m].spec["steps"]

    def idle(self):
        return True

    def device_step(self):
        return self.step

    def traces_done(self, m):
        return self.traces(m)

    def advance(self, n):
        assert not self.closed
        time.sleep(1e-4)  # a device call releases the interpreter lock, as this does
        self.step = min(self.max_steps, self.step + n)
        self.log.append(("advance", [im.spec["seed"] for im in self.models], self.step))

    def traces(self, m):
        assert self.step >= self.n_steps(m)
        return solo_traces(self.models[m].spec)

    def close(self):
        self.closed = True
        self.log.append(("closed", [im.spec["seed"] for im in self.models]))


class Started:
    def __init__(self, job):
        self.job = job

    def poll(self):
        return self.job


class FakeServer(serve.Server):
    def __init__(self, **kw):
        kw.setdefault("hbm_budget", 1 << 40)
        super().__init__(**kw)
        self.log = []

    def _setup(self, r):
        if r.sim.get("setup_fails"):
            raise ValueError("bad mesh")
        return FakeImage(r.sim)

    def _power_job(self, img):
        self.log.append(("power", img.spec["seed"]))
        return FakeJob(img, self.log)

    def _stage_size(self, img):
        return 0

    def _start_power(self, pool, r):
        # started at once on the loop thread, so that the polls alone order
        # the requests
        return Started(self._power_job(r.image))


    def _batch(self, images, lambdas):
        if any(im.spec.get("batch_fails") for im in images):
            raise RuntimeError("out of memory")
        return FakeBatch(images, lambdas, self.log)

    def _solo(self, job, image):
        # a large request runs alone on its power job's context
        return FakeBatch([image], [job.lambda_max()], self.log)


def spec(seed, steps=100, points=1000, dt=1e-6, **kw):
    return dict(seed=seed, steps=steps, points=points, dt=dt, **kw)


def check_solo(r):
    ref = solo_traces(r.sim)
    assert r.ok, r.error
    assert r.traces.keys() == ref.keys()
    for k in ref:
        assert np.array_equal(r.traces[k], ref[k])


def test_every_request_gets_its_own_results():
    specs = [spec(i, steps=50 + 37 * i, order=4 + i % 2) for i in range(12)]
    srv = FakeServer(max_slots=2, max_batch=3, chunk=16)
    reqs = srv.serve([(0.001 * i, s) for i, s in enumerate(specs)])
    assert [r.rid for r in reqs] == list(range(12))
    for r in reqs:
        check_solo(r)
        assert r.dt == pytest.approx(r.sim["dt"]) and r.n_steps == r.sim["steps"]
        assert r.arrival <= r.admitted <= r.done
    batches = [e[1] for e in srv.log if e[0] == "batch"]
    assert sorted(s for b in batches for s in b) == list(range(12))
    for b in batches:  # batched only with requests of the same order, at most max_batch
        assert len(b) <= 3 and len({specs[s]["order"] for s in b}) == 1
    s = serve.summary(reqs, max(r.done for r in reqs))
    assert s["solves"] == 12 and s["failed"] == 0 and s["latency_p99_ms"] >= s["latency_p50_ms"]


def test_finished_models_return_before_their_batch_ends():
    # request 9 holds the only slot while 0 and 1 get their time steps, so
    # that those two are admitted together when it ends
    srv = FakeServer(max_slots=1, max_batch=4, chunk=10)
    reqs = srv.serve([(0.0, spec(9, steps=300, power_polls=0)), (0.0, spec(0, steps=20)),
                      (0.0, spec(1, steps=400))])
    assert [e[1] for e in srv.log if e[0] == "batch"] == [[9], [0, 1]]
    for r in reqs:
        check_solo(r)
    assert reqs[1].done < reqs[2].done
    # the short model was returned at the first idle point past its 20
    # steps, long before its batch closed
    start = srv.log.index(("batch", [0, 1]))
    advances = [e for e in srv.log[start:] if e[0] == "advance"]
    assert advances[1][2] == 20 and advances[-1][2] == 400
    assert reqs[1].done < reqs[2].done


def test_large_requests_have_their_own_slots():
    specs = [spec(0, steps=3000, points=10**6, axisym=False)] + [spec(i, steps=60) for i in range(1, 9)] + \
            [spec(9, steps=3000, points=10**6, axisym=False)]
    srv = FakeServer(max_slots=1, max_batch=8, chunk=20, max_large_slots=1, large_dofs=2_000_000)
    reqs = srv.serve([(0.0, s) for s in specs])
    for r in reqs:
        check_solo(r)
    assert reqs[0].large and reqs[9].large and not any(r.large for r in reqs[1:9])
    batches = [e[1] for e in srv.log if e[0] == "batch"]
    assert [0] in batches and [9] in batches  # each large request alone
    # the small requests ran (and finished) while the first large one was on
    # the device, and the two large ones never ran at once
    first_large_closed = srv.log.index(("closed", [0]))
    small_closed = [i for i, e in enumerate(srv.log) if e[0] == "closed" and e[1][0] in range(1, 9)]
    assert small_closed and max(small_closed) < first_large_closed
    assert srv.log.index(("batch", [9])) > first_large_closed
    # with no large slots, large requests share the ordinary ones
    srv = FakeServer(max_slots=2, max_batch=8, chunk=20, max_large_slots=0)
    reqs = srv.serve([(0.0, s) for s in specs])
    assert not any(r.large for r in reqs)


def test_power_iterations_do_not_stall_running_batches():
    # the second request's power iteration takes many polls; the first
    # request's batch keeps receiving steps meanwhile
    srv = FakeServer(max_slots=2, max_batch=1, chunk=5)
    reqs = srv.serve([(0.0, spec(0, steps=200, power_polls=0)), (0.0, spec(1, steps=10, power_polls=30, order=5))])
    for r in reqs:
        check_solo(r)
    start1 = srv.log.index(("power", 1))
    batch1 = srv.log.index(("batch", [1]))
    advances_between = [e for e in srv.log[start1:batch1] if e[0] == "advance" and e[1] == [0]]
    assert len(advances_between) >= 10


def test_memory_budget_and_refusal():
    fp = serve.footprint_estimate(FakeImage(spec(0)))
    srv = FakeServer(max_slots=4, max_batch=4, chunk=50, hbm_budget=2 * fp + fp // 2, max_power_jobs=4)
    big = spec(3, points=100_000)
    reqs = srv.serve([(0.0, spec(0)), (0.0, spec(1)), (0.0, spec(2)), (0.0, big), (0.0, spec(4))])
    assert not reqs[3].ok and reqs[3].error.startswith("refused") and not reqs[3].traces
    for r in reqs[:3] + reqs[4:]:
        check_solo(r)
    # power contexts and batches together never exceed the budget: at most
    # two models of this size on the device at any time
    live, peak = set(), 0
    for e in srv.log:
        if e[0] == "power":
            live.add(("p", e[1]))
        elif e[0] == "power_closed":
            live.discard(("p", e[1]))
        elif e[0] == "batch":
            live |= {("b", s) for s in e[1]}
        elif e[0] == "closed":
            live -= {("b", s) for s in e[1]}
        peak = max(peak, len(live))
    assert peak <= 2


def test_failures_end_one_request_only():
    specs = [spec(0), spec(1, diverge=True), spec(2, setup_fails=True), spec(3, power_fails=True, order=6),
             spec(4, order=7, batch_fails=True), spec(5)]
    srv = FakeServer(max_slots=2, max_batch=1, chunk=25)
    reqs = srv.serve([(0.0, s) for s in specs])
    check_solo(reqs[0]), check_solo(reqs[5])
    assert reqs[1].error.startswith("non-finite traces")
    assert reqs[2].error.startswith("setup failed: bad mesh")
    assert reqs[3].error.startswith("power iteration failed")
    assert reqs[4].error.startswith("batch set-up failed")
    for r in reqs[1:5]:
        assert r.done == r.done  # a time is recorded
    s = serve.summary(reqs, 1.0)
    assert s["solves"] == 2 and s["failed"] == 4


def test_timeout_stops_a_request_and_its_batch():
    class Slow(FakeBatch):
        def device_step(self):
            time.sleep(2e-3)
            return self.step

    class SlowServer(FakeServer):
        def _batch(self, images, lambdas):
            return Slow(images, lambdas, self.log)

    srv = SlowServer(max_slots=1, max_batch=1, chunk=1, timeout=0.15)
    reqs = srv.serve([(0.0, spec(0, steps=10)), (0.0, spec(1, steps=100_000))])
    check_solo(reqs[0])
    assert reqs[1].error.startswith("timed out")
    assert reqs[1].done - reqs[1].arrival < 1.0
    assert ("closed", [1]) in srv.log  # its batch was stopped


def test_batches_group_similar_step_counts():
    # request 9 holds the only slot while the others get their time steps;
    # then the oldest waiting request (0, 100 steps) is batched with the two
    # whose step counts are closest to its own
    srv = FakeServer(max_slots=1, max_batch=3, chunk=50)
    specs = [spec(9, steps=400, power_polls=0), spec(0, steps=100), spec(1, steps=1000), spec(2, steps=120),
             spec(3, steps=5000), spec(4, steps=90)]
    reqs = srv.serve([(0.0, s) for s in specs])
    for r in reqs:
        check_solo(r)
    batches = [e[1] for e in srv.log if e[0] == "batch"]
    assert batches[0] == [9] and sorted(batches[1]) == [0, 2, 4] and sorted(batches[2]) == [1, 3]


def test_batch_wait_gathers_requests_that_share_a_batch():
    # six requests of one key arriving 10 ms apart: admitted as they come
    # (batch_wait = 0) the first runs alone; with a window they share one
    # batch, and each still gets its solo results
    arrivals = [(0.01 * i, spec(i, steps=40, power_polls=0)) for i in range(6)]
    srv = FakeServer(max_slots=4, max_batch=8, chunk=10)
    srv.serve(arrivals)
    first = [e[1] for e in srv.log if e[0] == "batch"][0]
    assert first == [0]
    srv = FakeServer(max_slots=4, max_batch=8, chunk=10, batch_wait=0.5)
    reqs = srv.serve(arrivals)
    for r in reqs:
        check_solo(r)
    batches = [e[1] for e in srv.log if e[0] == "batch"]
    assert batches == [[0, 1, 2, 3, 4, 5]]
    # a full group does not wait
    srv = FakeServer(max_slots=4, max_batch=3, chunk=10, batch_wait=30.0)
    reqs = srv.serve(arrivals)
    for r in reqs:
        check_solo(r)
    assert sorted(len(b) for b in (e[1] for e in srv.log if e[0] == "batch")) == [3, 3]



def test_power_iterations_start_on_the_loop_thread():
    # every device call of the loop, a power iteration's start included, is
    # made on the loop's own thread; a start that fails ends its request only

    class Starts(FakeServer):
        threads = set()

        def _power_job(self, img):
            Starts.threads.add(threading.get_ident())
            if img.spec.get("start_fails"):
                raise RuntimeError("no device memory")
            return super()._power_job(img)

    srv = Starts(max_slots=2, max_batch=2, chunk=50)
    reqs = srv.serve([(0.0, spec(0)), (0.0, spec(1, start_fails=True)), (0.0, spec(2)), (0.0, spec(3))])
    assert Starts.threads == {threading.get_ident()}
    for r in (reqs[0], reqs[2], reqs[3]):
        check_solo(r)
    assert reqs[1].error.startswith("power iteration failed")


def test_staged_start_splits_host_and_device_work(monkeypatch):
    # the server's own start: the block is reserved and the staged context
    # started on the loop thread, the host half staged on a worker; a
    # staging that fails gives the block back and fails its request only
    loop = threading.get_ident()
    calls = []

    def reserve(n):
        calls.append(("reserve", threading.get_ident()))
        return 4096 + len(calls)

    def unreserve(block):
        calls.append(("unreserve", threading.get_ident(), block))

    class FakeStaged:
        def __init__(self, image, block):
            calls.append(("stage", threading.get_ident()))
            if image.spec.get("stage_fails"):
                raise RuntimeError("staging failed")
            self.image = image

    class FakePowerJob:
        @staticmethod
        def from_staged(staged):
            calls.append(("start", threading.get_ident()))
            return FakeJob(staged.image, [])

    monkeypatch.setattr(serve, "reserve_block", reserve)
    monkeypatch.setattr(serve, "unreserve_block", unreserve)
    monkeypatch.setattr(serve, "StagedPower", FakeStaged)
    monkeypatch.setattr(serve, "PowerJob", FakePowerJob)

    class Real(FakeServer):
        def _start_power(self, pool, r):
            return serve.Server._start_power(self, pool, r)

    srv = Real(max_slots=2, max_batch=2, chunk=50)
    reqs = srv.serve([(0.0, spec(0)), (0.0, spec(1, stage_fails=True)), (0.0, spec(2))])
    for r in (reqs[0], reqs[2]):
        check_solo(r)
    assert reqs[1].error.startswith("power iteration failed")
    assert all(c[1] == loop for c in calls if c[0] in ("reserve", "unreserve", "start"))
    assert all(c[1] != loop for c in calls if c[0] == "stage")
    assert sum(c[0] == "unreserve" for c in calls) == 1
