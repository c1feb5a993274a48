"""Slowness-time coherence (reference: python/insem/post.py:347-384; its
plane-wave test is tests/python/test_borehole.py:154-162)."""

from __future__ import annotations

import numpy as np
import pytest

from insem_amd import post


def test_stc_recovers_plane_wave_slowness():
    # the reference's known answer: a Gaussian pulse moving at 312 us/m
    dt, s_true = 1e-5, 312e-6
    t = np.arange(400) * dt
    offsets = np.arange(6) * 0.15
    x = np.array([np.exp(-((t - 1e-3 - s_true * z) / 1e-4) ** 2) for z in offsets])
    r = post.stc(x, offsets, dt, np.linspace(100e-6, 600e-6, 501), window=4e-4)
    s, T, c = r.peak()
    assert s == pytest.approx(s_true, abs=1.5e-6)
    assert c > 0.99
    # the picked window overlaps the pulse at the first receiver
    assert T <= 1.1e-3 and T + 4e-4 >= 0.9e-3


def test_stc_is_a_normalised_semblance():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((5, 300))
    r = post.stc(x, np.arange(5) * 0.2, 1e-5, np.linspace(0, 1e-3, 41), window=2e-4)
    assert r.coherence.min() >= 0.0 and r.coherence.max() <= 1.0 + 1e-12
    # incoherent noise: semblance about 1/N
    assert r.coherence.mean() < 0.4
    # identical traces at zero offset: coherence 1 wherever there is energy
    same = np.tile(rng.standard_normal(200), (4, 1))
    r1 = post.stc(same, np.zeros(4), 1e-5, [0.0, 1e-4], window=1e-4)
    assert np.allclose(r1.coherence[r1.coherence > 0], 1.0)


def test_stc_matches_brute_force_interpolation():
    # every window sum written out with np.interp (zero outside the record)
    rng = np.random.default_rng(9)
    dt, t0 = 2e-5, 1e-4
    x = rng.standard_normal((3, 40))
    offsets = np.array([0.0, 0.13, 0.29])
    grid = [0.0, 1.7e-4, 3.9e-4]
    times = np.array([1.2e-4, 3.3e-4, 6.0e-4])
    window = 1e-4
    r = post.stc(x, offsets, dt, grid, window, times=times, t0=t0, min_energy=0.0)
    t = t0 + dt * np.arange(x.shape[1])
    nw = int(round(window / dt)) + 1
    for a, s in enumerate(grid):
        for b, T in enumerate(times):
            y = np.array([[np.interp(T + k * dt + s * (z - offsets[0]), t, xi, left=0.0, right=0.0)
                           for k in range(nw)] for xi, z in zip(x, offsets)])
            num = (y.sum(axis=0) ** 2).sum()
            den = 3 * (y ** 2).sum()
            assert r.coherence[a, b] == pytest.approx(num / den, rel=1e-12, abs=1e-15)
