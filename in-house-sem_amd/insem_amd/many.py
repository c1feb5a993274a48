"""Solving many independent models on one GPU (no counterpart in the
reference, which runs one model per Solver).

`solve_many` groups the models that can share a batch (same polynomial
order, coordinate system, azimuthal order, media and record cadence),
sorts each group by step count and cuts it into batches of up to
`batch_size` models, and runs all batches side by side, each on its own
stream: a batch runs until its longest model ends, so batches of models of
similar length waste few steps, and many small batches side by side keep
the device busy while the longer ones finish.  The models' host setups
(which make no device calls) are built on worker threads; then every
model's power iteration runs (a window of them at a time, overlapping on
the device) to fix its time step and step count; then the batches are
planned, and each starts stepping as soon as it is created, while the
next ones are set up.  Every device call is made from the calling thread.
Each model's traces are bitwise those of running it alone.
"""

from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .core import BatchSolver, PowerJob, Simulation
from .serve import batch_key


def plan_batches(keys: Sequence, sizes: Sequence[int], batch_size: int) -> List[List[int]]:
    """Model indices in batches: models of equal key (those that can share a
    batch), sorted by size (any per-model figure: the step count in
    solve_many), cut into ceil(n / batch_size) batches of nearly equal count
    (at most batch_size each); groups in order of their first model, every
    model in exactly one batch."""
    if batch_size < 1:
        raise ValueError("batch_size must be at least 1")
    groups: Dict = {}
    for i, k in enumerate(keys):
        groups.setdefault(k, []).append(i)
    out = []
    for idx in groups.values():
        idx = sorted(idx, key=lambda i: (sizes[i], i))
        n = -(-len(idx) // batch_size)
        out += [idx[p * len(idx) // n:(p + 1) * len(idx) // n] for p in range(n)]
    return out


def power_lambdas(images: Sequence, window: int = 16) -> List[float]:
    """Every model's lambda_max (its power iteration, as its Solver runs
    it), a window of power jobs queued on the device at a time so that they
    overlap."""
    out: List[float] = []
    for w0 in range(0, len(images), window):
        jobs = [PowerJob(im) for im in images[w0:w0 + window]]
        for j in jobs:
            out.append(j.lambda_max())
            j.close()
    return out


def solve_many(sims: Sequence[Simulation], batch_size: int = 5, chunk: int = 512, workers: int = 8,
               on_done: Optional[Callable[[int, float, int], None]] = None) -> List[Dict[Tuple[str, str], np.ndarray]]:
    """Traces of every model, in the order given.  on_done(i, seconds,
    n_steps) is called when model i's traces are ready (seconds since the
    call; the model's own step count)."""
    t0 = time.perf_counter()
    out: List[Optional[Dict[Tuple[str, str], np.ndarray]]] = [None] * len(sims)
    live = []
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        images = [pool.submit(sim.build_setup) for sim in sims]
        imgs = [f.result() for f in images]
    lams = power_lambdas(imgs)
    steps = [im.time_step(lam)[1] for im, lam in zip(imgs, lams)]
    for sub in plan_batches([batch_key(im) for im in imgs], steps, batch_size):
        b = BatchSolver([imgs[i] for i in sub], [lams[i] for i in sub])
        b.advance(4 * chunk)
        live.append((sub, b))
    while live:
        for _, b in live:
            if b.step < b.max_steps:
                b.advance(chunk)
        for sub, b in [g for g in live if g[1].step >= g[1].max_steps and g[1].idle()]:
            for m, i in enumerate(sub):
                out[i] = b.traces(m)
                if on_done:
                    on_done(i, time.perf_counter() - t0, b.n_steps(m))
            b.close()
        live = [g for g in live if g[1]._handle is not None]
        if live and all(g[1].step >= g[1].max_steps for g in live):
            time.sleep(5e-4)
    return out
