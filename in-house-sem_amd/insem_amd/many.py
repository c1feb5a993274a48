"""Solving many independent models on one GPU (no counterpart in the
reference, which runs one model per Solver).

`solve_many` groups the models that can share a batch (same polynomial
order, coordinate system, azimuthal order, media and record cadence),
sorts each group by size and cuts it into batches of up to `batch_size`
models, and runs all batches side by side, each on its own stream: a
batch runs until its longest model ends, so batches of similar models
waste few steps, and many small batches side by side keep the device busy
while the longer ones finish.  The models' host setups (which make no
device calls) are built on worker threads while the batches already
created run; every device call is made from the calling thread.  Each
model's traces are bitwise those of running it alone.
"""

from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .core import BatchSolver, Simulation
from .serve import batch_key


def solve_many(sims: Sequence[Simulation], batch_size: int = 5, chunk: int = 512, workers: int = 8,
               on_done: Optional[Callable[[int, float, int], None]] = None) -> List[Dict[Tuple[str, str], np.ndarray]]:
    """Traces of every model, in the order given.  on_done(i, seconds,
    n_steps) is called when model i's traces are ready (seconds since the
    call; the model's own step count)."""
    if batch_size < 1:
        raise ValueError("batch_size must be at least 1")
    t0 = time.perf_counter()
    out: List[Optional[Dict[Tuple[str, str], np.ndarray]]] = [None] * len(sims)
    live = []
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        images = [pool.submit(sim.build_setup) for sim in sims]
        groups: Dict[Tuple, List[int]] = {}
        for i in range(len(sims)):
            groups.setdefault(batch_key(images[i].result()), []).append(i)
        for idx in groups.values():
            idx.sort(key=lambda i: sum(d["n_elements"] for d in images[i].result().domains))
            n = -(-len(idx) // batch_size)
            for part in range(n):
                sub = idx[part * len(idx) // n:(part + 1) * len(idx) // n]
                b = BatchSolver([images[i].result() for i in sub])
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
