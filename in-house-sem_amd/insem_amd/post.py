"""Array processing of receiver traces (reference: python/insem/post.py).

Only the slowness-time coherence used by the served borehole responses is
here (post.py:347-384 of the reference): the semblance of Kimball &
Marzetta (1984) over linearly shifted windows of an array of traces.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np


@dataclass
class STC:
    """Slowness-time coherence of an array: coherence[slowness, time]
    (reference: post.py:329-345)."""

    slowness: np.ndarray  # s/m
    time: np.ndarray      # window start times at the first receiver (s)
    coherence: np.ndarray

    def peak(self, t_min: float = -np.inf, t_max: float = np.inf, s_min: float = 0.0,
             s_max: float = np.inf) -> Tuple[float, float, float]:
        """(slowness, time, coherence) of the largest coherence in a region."""
        rows = (self.slowness >= s_min) & (self.slowness <= s_max)
        cols = (self.time >= t_min) & (self.time <= t_max)
        masked = np.full(self.coherence.shape, -np.inf)
        masked[np.ix_(rows, cols)] = self.coherence[np.ix_(rows, cols)]
        a, b = np.unravel_index(int(np.argmax(masked)), masked.shape)
        return float(self.slowness[a]), float(self.time[b]), float(self.coherence[a, b])


def _sample(x: np.ndarray, q: np.ndarray) -> np.ndarray:
    """x[i] at fractional sample positions q[i, ...] by linear interpolation,
    0 outside [0, n - 1]."""
    n = x.shape[1]
    inside = (q >= 0.0) & (q <= n - 1)
    qc = np.clip(q, 0.0, n - 1)
    i0 = np.floor(qc).astype(np.int64)
    i1 = np.minimum(i0 + 1, n - 1)
    w = qc - i0
    rows = np.arange(x.shape[0]).reshape((-1,) + (1,) * (q.ndim - 1))
    v = (1.0 - w) * x[rows, i0] + w * x[rows, i1]
    return np.where(inside, v, 0.0)


def stc(traces: np.ndarray, offsets: Sequence[float], dt: float, slowness: Sequence[float], window: float,
        times: Optional[Sequence[float]] = None, t0: float = 0.0, min_energy: float = 1e-6) -> STC:
    """Slowness-time coherence of an array (reference: post.py:347-384).

    traces (n_receivers, n_samples) sampled at t0 + i dt; offsets (m) along
    the array; slowness trial values (s/m); window length (s); times: window
    starts at the first receiver (default every sample that leaves a whole
    window).  For a trial slowness s and window start T, with x_i shifted by
    s (z_i - z_0):

        coherence = sum_t (sum_i x_i)^2 / (N sum_t sum_i x_i^2),  t in [T, T + window],

    and 0 where the window's energy is below min_energy times the largest
    one (silent windows are otherwise spuriously coherent)."""
    x = np.asarray(traces, dtype=float)
    z = np.asarray(offsets, dtype=float)
    z = z - z[0]
    n_rec, n = x.shape
    t = t0 + dt * np.arange(n)
    s_grid = np.asarray(slowness, dtype=float)
    starts = t[t <= t[-1] - window] if times is None else np.asarray(times, dtype=float)
    nw = max(2, int(round(window / dt)) + 1)
    tau = dt * np.arange(nw)
    base = (starts[:, None] + tau[None, :] - t0) / dt  # fractional sample of the first receiver
    coh = np.zeros((len(s_grid), len(starts)))
    energy = np.zeros_like(coh)
    for a, s in enumerate(s_grid):
        q = base[None, :, :] + (s * z / dt)[:, None, None]  # [receiver, start, lag]
        y = _sample(x, q)
        stack = (y.sum(axis=0) ** 2).sum(axis=1)
        power = n_rec * (y * y).sum(axis=(0, 2))
        coh[a] = np.divide(stack, power, out=np.zeros_like(stack), where=power > 0)
        energy[a] = power
    if energy.size and energy.max() > 0:
        coh[energy < min_energy * energy.max()] = 0.0
    return STC(s_grid, starts, coh)
