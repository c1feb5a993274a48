"""Analytic 2D solutions used to check end-to-end accuracy.

Fields follow numpy's FFT convention (e^{+i w t}), so outgoing waves use the
Hankel function H^(2).  The scalar Green's function of (lap + k^2) g = -delta
in 2D is g = -(i/4) H0^(2)(k r) (a textbook result), and a monopole source
(1/kappa) p_tt = div(grad p / rho) + s(t) delta radiates P(w) = rho g S(w).
"""

from __future__ import annotations

import numpy as np
from scipy.special import hankel2


def spectrum(t, s, pad=8):
    n = len(t)
    buf = np.zeros(pad * n)
    buf[:n] = s
    w = 2 * np.pi * np.fft.rfftfreq(pad * n, t[1] - t[0])
    return np.fft.rfft(buf), w, pad * n


def monopole_pressure(t, stf_values, x, z, c, rho):
    S, w, n = spectrum(t, stf_values)
    r = np.hypot(x, z)
    G = np.zeros_like(S)
    G[1:] = -0.25j * hankel2(0, w[1:] * r / c)
    return np.fft.irfft(rho * G * S, n)[: len(t)]


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
