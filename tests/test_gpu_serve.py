"""The request loop: each served request's results equal the same request
run alone, whatever it was batched with and whenever it arrived."""

from __future__ import annotations

import numpy as np
import pytest

from insem_amd import serve
from insem_amd.configs import cased_hole, layered_coupled, open_hole

pytestmark = pytest.mark.gpu


def make(kind, seed):
    rng = np.random.default_rng(seed)
    small = dict(duration=2.5e-4, z=(-0.4, 0.8), formation_radius=0.6, dz=0.08, order=4, f_max=8e3,
                 n_receivers=3, receiver_z0=0.2)
    if kind == "cased":
        return cased_hole(rng, **small)
    if kind == "open":
        return open_hole(rng, **small)
    return layered_coupled(nx=36, nz_fluid=3, nz_solid=(5, 6), duration=0.01 + 0.01 * (seed % 3), n_receivers=5,
                           width=60.0)


STREAM = [(0.00, "cased", 21), (0.01, "open", 31), (0.02, "cased", 22), (0.03, "layered", 41),
          (0.04, "cased", 23), (0.05, "open", 32), (0.06, "layered", 42)]


def layered_class_threshold():
    """A large_dofs that puts the toy layered requests (and only them) in
    the large slot class."""
    dofs = {kind: [make(kind, seed).build_setup().n_dofs for _, k, seed in STREAM if k == kind]
            for kind in ("cased", "open", "layered")}
    lo = max(dofs["cased"] + dofs["open"])
    assert min(dofs["layered"]) > lo
    return lo + 1


@pytest.mark.parametrize("large_slots", [0, 1])
def test_served_results_equal_solo_runs(large_slots):
    large_dofs = layered_class_threshold() if large_slots else 0
    srv = serve.Server(max_slots=2, max_batch=2, chunk=64, max_large_slots=large_slots, large_dofs=large_dofs,
                       stc=dict(slowness=np.linspace(100e-6, 1500e-6, 57), window=5e-5,
                                times=np.linspace(0.0, 1.9e-4, 39)))
    reqs = srv.serve([(t, make(kind, seed)) for t, kind, seed in STREAM])
    assert [r.rid for r in reqs] == list(range(len(STREAM)))
    for r, (t, kind, seed) in zip(reqs, STREAM):
        assert r.ok, r.error
        assert r.large == (large_slots == 1 and kind == "layered")
        assert r.arrival == t
        assert r.admitted >= r.arrival and r.done > r.admitted
        alone = make(kind, seed).solver
        assert alone.dt == r.dt and alone.n_steps == r.n_steps
        alone.run()
        ref = alone.traces
        assert ref.keys() == r.traces.keys()
        for k in ref:
            assert np.array_equal(ref[k], r.traces[k]), (r.rid, k)
        if kind != "layered":
            assert r.stc is not None and r.stc.coherence.shape[0] == 57
            assert 0.0 <= r.stc.coherence.max() <= 1.0 + 1e-12
    s = serve.summary(reqs, max(r.done for r in reqs))
    assert s["solves"] == len(STREAM) and s["latency_p99_ms"] >= s["latency_p50_ms"] > 0


def test_admission_respects_the_memory_budget():
    # a budget that holds one model at a time: requests wait for memory, and
    # a request larger than the budget is refused
    fps = sorted(serve.footprint_estimate(make("cased", s).build_setup()) for s in (21, 22, 23))
    budget = fps[0] + fps[1] - 1  # every model fits, no two together
    assert budget >= fps[2]
    one = fps[0]
    srv = serve.Server(max_slots=4, max_batch=4, chunk=64, hbm_budget=budget)
    reqs = srv.serve([(0.0, make("cased", 21)), (0.0, make("cased", 22)), (0.0, make("cased", 23))])
    spans = sorted((r.admitted, r.done) for r in reqs)
    for (a0, d0), (a1, _) in zip(spans, spans[1:]):
        assert a1 >= d0 - 1e-9  # never two on the device at once
    small = serve.Server(hbm_budget=one // 4)
    refused, = small.serve([(0.0, make("cased", 21))])
    assert not refused.ok and refused.error.startswith("refused") and not refused.traces
