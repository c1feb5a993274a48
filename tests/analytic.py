"""Closed-form wave fields of homogeneous media, used to check end-to-end
accuracy of the oracle (the reference checks the same quantities in
tests/python/test_verification.py, test_axisym.py and test_dipole.py).

Conventions.  Spectra follow numpy's FFT (time dependence e^{+i w t}), so
outgoing waves are H^(2)(k R) in 2D and e^{-i k R} / R in 3D.  The scalar
Green's function g of (lap + k^2) g = -delta is
    2D: g = -(i/4) H0^(2)(k R)          3D: g = e^{-i k R} / (4 pi R).
With G_ij = (1/(rho w^2)) [k_s^2 g_s delta_ij + d_i d_j (g_s - g_p)] the
displacement of a point force F is u_i = G_ij F_j, and that of a moment
tensor M (body force -div(M delta)) is u_i = -M_nj d_j G_in.  Because
d_i d_j g(R) = g'' g_i g_j + (g'/R)(delta_ij - g_i g_j) (g_i = x_i / R), the
tensor is G = A I + B (gamma gamma^T) with
    A = [k_s^2 g_s + (g_s' - g_p')/R] / (rho w^2)
    B = [(g_s'' - g_p'') - (g_s' - g_p')/R] / (rho w^2).
Acoustic media: (1/kappa) p_tt = div(grad p / rho) + s.  A monopole
s = S(t) delta radiates P = rho S g; a dipole s = -div(f S delta / rho)
radiates P = -S f . grad g (no density dependence); the particle
velocity is V = -grad P / (i w rho).
"""

from __future__ import annotations

import numpy as np
from scipy.special import hankel2


# ----------------------------------------------------------------- spectra

def _spec(t, s, pad=8):
    """One-sided spectrum of the sampled signal s(t), zero-padded pad times."""
    n = pad * len(t)
    w = 2.0 * np.pi * np.fft.rfftfreq(n, t[1] - t[0])
    return np.fft.rfft(s, n), w, n


def _synth(H, S, n, m):
    """Time series of the spectrum H * S (the static w = 0 term dropped),
    first m samples."""
    X = H * S
    X = np.where(np.isfinite(X), X, 0.0)
    X[0] = 0.0
    return np.fft.irfft(X, n)[:m]


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ------------------------------------------------- scalar Green's functions

def scalar_green(dim, k, R):
    """(g, g', g'') of the outgoing scalar Green's function at distance R for
    wavenumbers k (array, k > 0)."""
    if dim == 2:
        x = k * R
        h0, h1 = hankel2(0, x), hankel2(1, x)
        g = -0.25j * h0
        gp = 0.25j * k * h1  # d/dR H0 = -k H1
        gpp = 0.25j * k * k * (h0 - h1 / x)  # d/dx H1 = H0 - H1/x
        return g, gp, gpp
    g = np.exp(-1j * k * R) / (4.0 * np.pi * R)
    q = -1j * k - 1.0 / R  # g' = q g
    return g, q * g, (q * q + 1.0 / (R * R)) * g


def tensor_coefficients(dim, w, R, vp, vs, rho):
    """A(w), B(w) of G = A I + B gamma gamma^T at distance R."""
    ks, kp = w / vs, w / vp
    gs, gs1, gs2 = scalar_green(dim, ks, R)
    _, gp1, gp2 = scalar_green(dim, kp, R)
    rw2 = rho * w * w
    A = (ks * ks * gs + (gs1 - gp1) / R) / rw2
    B = ((gs2 - gp2) - (gs1 - gp1) / R) / rw2
    return A, B


def green_tensor(dim, w, x, vp, vs, rho):
    """G_ij(w) at receiver offset x (length dim: (x, z) or (x, y, z)):
    array [dim][dim][n_w]."""
    x = np.asarray(x, dtype=np.float64)
    R = float(np.linalg.norm(x))
    A, B = tensor_coefficients(dim, w, R, vp, vs, rho)
    gam = x / R
    return np.eye(dim)[:, :, None] * A[None, None, :] + np.outer(gam, gam)[:, :, None] * B[None, None, :]


def _positive(w):
    """Angular frequencies with the w = 0 entry replaced by a harmless value
    (its spectral line is dropped by _synth)."""
    W = w.copy()
    W[0] = W[1]
    return W


# --------------------------------------------------------------- 2D fields

def acoustic_monopole_2d(t, stf, x, z, c, rho):
    S, w, n = _spec(t, stf)
    g, _, _ = scalar_green(2, _positive(w) / c, np.hypot(x, z))
    return _synth(rho * g, S, n, len(t))


def acoustic_dipole_2d(t, stf, x, z, c, fx, fz):
    S, w, n = _spec(t, stf)
    R = np.hypot(x, z)
    _, g1, _ = scalar_green(2, _positive(w) / c, R)
    return _synth(-(fx * x + fz * z) / R * g1, S, n, len(t))


def acoustic_particle_velocity_2d(t, stf, x, z, c, rho):
    """(v_x, v_z) of a monopole: V = -grad P / (i w rho) = -S g' gamma / (i w)."""
    S, w, n = _spec(t, stf)
    W = _positive(w)
    R = np.hypot(x, z)
    _, g1, _ = scalar_green(2, W / c, R)
    v = -g1 / (1j * W)
    return np.stack([_synth(v * x / R, S, n, len(t)), _synth(v * z / R, S, n, len(t))], axis=1)


def elastic_point_source_2d(t, stf, x, z, vp, vs, rho, force=None, moment=None, h=1e-4):
    """(u_x, u_z) at (x, z) of a line force force = (fx, fz), or a moment
    tensor moment = [[mxx, mxz], [mxz, mzz]], at the origin."""
    S, w, n = _spec(t, stf)
    W = _positive(w)
    p = np.array([x, z], dtype=np.float64)
    if force is not None:
        U = np.einsum("ijw,j->iw", green_tensor(2, W, p, vp, vs, rho), np.asarray(force, dtype=np.float64))
    else:
        M = np.asarray(moment, dtype=np.float64)
        U = np.zeros((2, W.size), dtype=complex)
        for j in range(2):  # u_i = -M_nj d_j G_in, central differences in the receiver position
            e = np.zeros(2)
            e[j] = h
            dG = (green_tensor(2, W, p + e, vp, vs, rho) - green_tensor(2, W, p - e, vp, vs, rho)) / (2 * h)
            U -= np.einsum("inw,n->iw", dG, M[:, j])
    return np.stack([_synth(U[i], S, n, len(t)) for i in range(2)], axis=1)


# --------------------------------------------------------------- 3D fields

def acoustic_monopole_3d(t, stf, r, z, c, rho):
    S, w, n = _spec(t, stf)
    g, _, _ = scalar_green(3, _positive(w) / c, np.hypot(r, z))
    return _synth(rho * g, S, n, len(t))


def acoustic_dipole_3d(t, stf, r, z, c, fx=1.0):
    """Amplitude P of p = P cos(theta) for a horizontal dipole fx e_x."""
    S, w, n = _spec(t, stf)
    R = np.hypot(r, z)
    _, g1, _ = scalar_green(3, _positive(w) / c, R)
    return _synth(-fx * (r / R) * g1, S, n, len(t))


def elastic_vertical_force_3d(t, stf, r, z, vp, vs, rho, fz=1.0):
    """(u_r, u_z) of a vertical point force fz e_z at the origin."""
    S, w, n = _spec(t, stf)
    G = green_tensor(3, _positive(w), (r, 0.0, z), vp, vs, rho)
    return np.stack([_synth(fz * G[0, 2], S, n, len(t)), _synth(fz * G[2, 2], S, n, len(t))], axis=1)


def elastic_explosion_3d(t, stf, r, z, vp, vs, rho, m0=1.0, h=1e-4):
    """(u_r, u_z) of an isotropic moment tensor m0 I at the origin."""
    S, w, n = _spec(t, stf)
    W = _positive(w)
    p = np.array([r, 0.0, z])
    U = np.zeros((3, W.size), dtype=complex)
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        dG = (green_tensor(3, W, p + e, vp, vs, rho) - green_tensor(3, W, p - e, vp, vs, rho)) / (2 * h)
        U -= m0 * dG[:, j, :]  # -M_nj d_j G_in with M = m0 I
    return np.stack([_synth(U[0], S, n, len(t)), _synth(U[2], S, n, len(t))], axis=1)


def elastic_horizontal_force_3d(t, stf, r, z, vp, vs, rho, fx=1.0):
    """Mode amplitudes (U_r, U_theta, U_z), u = (U_r cos, U_theta sin, U_z cos)
    of theta, of a horizontal point force fx e_x at the origin.  With
    G = A I + B gamma gamma^T at (r cos, r sin, z): u_r = cos(theta) (A + B r^2/R^2),
    u_theta = -sin(theta) A, u_z = cos(theta) B r z / R^2."""
    S, w, n = _spec(t, stf)
    R = np.hypot(r, z)
    A, B = tensor_coefficients(3, _positive(w), R, vp, vs, rho)
    comps = (A + B * r * r / (R * R), -A, B * r * z / (R * R))
    return np.stack([_synth(fx * c, S, n, len(t)) for c in comps], axis=1)


# ----------------------------------------------------- interface formulas

def rayleigh_speed(vp, vs):
    """Root of the Rayleigh function (2 - x^2)^2 = 4 sqrt(1 - x^2 vs^2/vp^2) sqrt(1 - x^2), x = c / vs."""
    from scipy.optimize import brentq

    k2 = (vs / vp) ** 2

    def f(x):
        return (2 - x * x) ** 2 - 4 * np.sqrt(1 - k2 * x * x) * np.sqrt(1 - x * x)

    return vs * brentq(f, 0.5, 0.99999)
