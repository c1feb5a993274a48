"""Batched models on the device: each model's results in a batch equal its
results alone, bit for bit (same time step, same step count, same traces),
and match the CPU oracle within the trace bound."""

from __future__ import annotations

import numpy as np
import pytest

import insem_amd as insem
from insem_amd.configs import cased_hole, layered_coupled
from oracle_driver import OracleRun

pytestmark = pytest.mark.gpu

TRACE_RTOL = 1e-5  # relative L2 of each trace against the oracle


def small_cased(seed, order=4):
    """A small cased-hole dipole model (m = 1): the cased-hole draws on a
    0.6 m formation radius and a short window, so that it runs in seconds."""
    rng = np.random.default_rng(seed)
    return cased_hole(rng, duration=2.5e-4, z=(-0.4, 0.8), formation_radius=0.6, dz=0.08, order=order, f_max=8e3,
                      n_receivers=3, receiver_z0=0.2)


def assert_same_traces(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("order", [4, 5])
def test_cased_hole_batch_equals_alone(order):
    seeds = [11, 12, 13]
    batch = insem.BatchSolver([small_cased(s, order) for s in seeds])
    assert batch.n_models == 3
    steps = [batch.n_steps(m) for m in range(3)]
    assert len(set(steps)) > 1, steps  # uneven step counts: models finish at different union steps
    batch.run()
    for m, s in enumerate(seeds):
        alone = small_cased(s, order).solver
        assert alone.dt == batch.dt(m)
        assert alone.n_steps == batch.n_steps(m)
        alone.run()
        assert_same_traces(batch.traces(m), alone.traces)


def test_batch_results_do_not_depend_on_model_order():
    # the batch steps its models longest first whatever order they come in;
    # each model keeps its place in the caller's list
    seeds = [11, 12, 13, 14]
    fwd = insem.BatchSolver([small_cased(s) for s in seeds])
    rev = insem.BatchSolver([small_cased(s) for s in reversed(seeds)])
    steps = [fwd.n_steps(m) for m in range(4)]
    assert len(set(steps)) > 1, steps  # then at least one of the two orders is not longest first
    fwd.run()
    rev.run()
    for m in range(4):
        assert fwd.n_steps(m) == rev.n_steps(3 - m) and fwd.dt(m) == rev.dt(3 - m)
        assert_same_traces(fwd.traces(m), rev.traces(3 - m))


def test_coupled_cartesian_batch_equals_alone_and_oracle():
    def make(k):
        return layered_coupled(nx=24 + 8 * k, nz_fluid=3, nz_solid=(5 + k, 6), duration=0.02 + 0.01 * k,
                               n_receivers=5, width=2000.0 * (24 + 8 * k) / 800.0)
    batch = insem.BatchSolver([make(k) for k in range(3)])
    # advance in uneven calls: chunk boundaries and model ends fall anywhere
    while batch.step < batch.max_steps:
        batch.advance(13)
    for k in range(3):
        sim = make(k)
        alone = sim.solver
        alone.run()
        tb = batch.traces(k)
        assert_same_traces(tb, alone.traces)
        o = OracleRun(sim.build_setup(), dt=alone.dt)
        o.run()
        to = o.traces()
        for key, ref in to.items():
            if np.linalg.norm(ref) > 0:
                err = np.linalg.norm(tb[key] - ref) / np.linalg.norm(ref)
                assert err <= TRACE_RTOL, (k, key, err)


def test_solve_many_equals_solo_runs():
    # mixed orders and kinds (open-hole monopole, cased-hole dipole) in
    # batches of two, all side by side; every model's traces equal its solo
    # run and come back in the order given
    from insem_amd.configs import open_hole

    def small_open(seed, order):
        rng = np.random.default_rng(seed)
        return open_hole(rng, duration=2.5e-4, z=(-0.4, 0.8), formation_radius=0.6, dz=0.08, order=order, f_max=8e3,
                         n_receivers=3, receiver_z0=0.2)

    specs = [(small_cased, 21, 4), (small_open, 31, 4), (small_cased, 22, 5), (small_cased, 23, 4),
             (small_open, 32, 4), (small_cased, 24, 4), (small_cased, 25, 5)]
    seen = []
    res = insem.solve_many([f(s, o) for f, s, o in specs], batch_size=2,
                           on_done=lambda i, t, n: seen.append((i, n)))
    assert sorted(i for i, _ in seen) == list(range(len(specs)))
    for (f, s, o), r, (i, n) in zip(specs, res, sorted(seen)):
        alone = f(s, o).solver
        assert alone.n_steps == n
        alone.run()
        assert_same_traces(alone.traces, r)
        alone.close()
