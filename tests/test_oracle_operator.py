"""The oracle's element operators against closed forms (restating
tests/cpp/test_kernels.cpp:29-154, test_axisym.cpp:85-122,
test_dipole.cpp:85-143 and test_refinement.cpp:156-192 of the reference).

GLL quadrature is exact for quadratic fields on affine elements, so the
assembled internal force at interior points equals mass * div(sigma) (or
mass * div(grad p / rho)) to round-off.  A dropped term, a wrong sign, a
transposed derivative matrix or a swapped metric entry breaks one of them.
"""

from __future__ import annotations

import numpy as np
import pytest

import insem_amd as insem
from models import grid, quarter_annulus
from oracle_driver import OracleRun

RHO, LAM, MU = 2.0, 3.0, 1.5  # elastic: vp^2 = (lam + 2 mu) / rho, vs^2 = mu / rho


def elastic_material():
    return insem.Material.elastic(np.sqrt((LAM + 2 * MU) / RHO), np.sqrt(MU / RHO), RHO)


def run_operator(sim, field):
    img = sim.build_setup()
    o = OracleRun(img, dt=1e-9)
    d = img.domains[0]
    x, z = img.coordinates(0)
    u = np.array(field(x, z), dtype=np.float64).reshape(d["ncomp"], -1).astype(np.float32)
    f, _ = o.internal_forces(0, u)
    return img, f, x, z


def interior(x, z, x0, x1, z0, z1, shear=0.0, eps=1e-9):
    xs = x - shear * z
    return (np.abs(z - z0) > eps) & (np.abs(z - z1) > eps) & (np.abs(xs - x0) > eps) & (np.abs(xs - x1) > eps)


def check_cartesian(shear, ngll, elastic, field, expected):
    m = grid(3, 2, 0.0, 1.5, 0.0, 1.0, shear)
    sim = insem.Simulation(m, insem.Discretization(order=ngll - 1, uniform_level=1))
    sim.set_material(0, elastic_material() if elastic else insem.Material.acoustic(1.0, RHO))
    sim.set_time(n_steps=1, dt=1e-9)
    img, f, x, z = run_operator(sim, field)
    # unit-density lumped mass: elastic mass = w J rho, acoustic = w J / (rho vp^2)
    unit = img.mass(0) / RHO if elastic else img.mass(0) * RHO * 1.0
    _, fs, _, _ = run_operator(sim, lambda x, z: [x * x + z * z, x * z] if elastic else [x * x + z * z])
    scale = float(np.max(np.abs(fs)))
    sel = interior(x, z, 0.0, 1.5, 0.0, 1.0, shear)
    assert sel.sum() > 10
    for c in range(f.shape[0]):
        np.testing.assert_allclose(f[c, sel], expected[c] * unit[sel], atol=1e-4 * scale)  # test_kernels.cpp:159


@pytest.mark.parametrize("shear", [0.0, 0.3])
@pytest.mark.parametrize("ngll", [3, 5, 9])
def test_elastic_operator_exact_for_quadratics(shear, ngll):
    cases = [
        (lambda x, z: [x * x, 0 * x], (2 * (LAM + 2 * MU), 0.0)),
        (lambda x, z: [z * z, 0 * x], (2 * MU, 0.0)),
        (lambda x, z: [x * z, 0 * x], (0.0, LAM + MU)),
        (lambda x, z: [0 * x, z * z + x], (0.0, 2 * (LAM + 2 * MU))),
    ]
    for field, expected in cases:
        check_cartesian(shear, ngll, True, field, expected)


@pytest.mark.parametrize("shear", [0.0, 0.3])
@pytest.mark.parametrize("ngll", [3, 5, 9])
def test_acoustic_operator_exact_for_quadratics(shear, ngll):
    # p = x^2 + 3 z^2 + x z: div(grad p / rho) = 8 / rho
    check_cartesian(shear, ngll, False, lambda x, z: [x * x + 3 * z * z + x * z], (8.0 / RHO,))


def test_rigid_motions_and_constants_are_annihilated():
    check_cartesian(0.2, 5, True, lambda x, z: [-z + 1.0, x - 2.0], (0.0, 0.0))
    check_cartesian(0.2, 5, False, lambda x, z: [4.0 + 0 * x], (0.0,))


def test_curved_blocks_annihilate_rigid_motions():
    # Quad9 quarter annulus: every element is Curved; a rigid rotation has
    # zero strain, which the per-point metric path must reproduce.
    sim = insem.Simulation(quarter_annulus(2, 3), insem.Discretization(order=4, uniform_level=1))
    sim.set_material(0, elastic_material())
    sim.set_time(n_steps=1, dt=1e-9)
    assert sim.geometry_kind_counts()["curved"] == sim.n_leaves
    _, f, _, _ = run_operator(sim, lambda x, z: [-z + 0.5, x - 0.2])
    _, fs, _, _ = run_operator(sim, lambda x, z: [x * x, z * x])
    assert np.max(np.abs(f)) < 1e-4 * np.max(np.abs(fs))


def test_curved_path_matches_affine_path():
    # Quad9 cells of a sheared grid with mid-nodes nudged by 1e-9 relative:
    # classified Curved, geometrically affine to 1e-9.
    def sheared_quad9(eps):
        nx, nz, shear = 2, 2, 0.4
        mx, mz = 2 * nx + 1, 2 * nz + 1
        nodes = []
        for j in range(mz):
            for i in range(mx):
                z = j / (mz - 1)
                x = i / (mx - 1) + shear * z
                if i % 2 == 1 and j % 2 == 1:
                    x += eps
                nodes.append((x, z))
        node = lambda i, j: i + mx * j  # noqa: E731
        cells = []
        for j in range(nz):
            for i in range(nx):
                a, b = 2 * i, 2 * j
                cells.append((node(a, b), node(a + 2, b), node(a + 2, b + 2), node(a, b + 2), node(a + 1, b),
                              node(a + 2, b + 1), node(a + 1, b + 2), node(a, b + 1), node(a + 1, b + 1)))
        return insem.MacroMesh("quad9", "cartesian", np.array(nodes), np.array(cells), np.zeros(len(cells)))

    out = {}
    for eps in (0.0, 1e-9):
        sim = insem.Simulation(sheared_quad9(eps), insem.Discretization(order=4, uniform_level=1))
        sim.set_material(0, elastic_material())
        sim.set_time(n_steps=1, dt=1e-9)
        kinds = sim.geometry_kind_counts()
        _, f, _, _ = run_operator(sim, lambda x, z: [np.sin(3 * x + z), np.cos(2 * z - x)])
        out[eps] = (kinds, f)
    assert out[0.0][0]["affine"] == 16 and out[1e-9][0]["curved"] > 0
    np.testing.assert_allclose(out[1e-9][1], out[0.0][1], atol=1e-3 * np.max(np.abs(out[0.0][1])))


def test_elastic_operator_exact_on_refined_affine_mesh():
    regions = ((0.6, 0.8, 0.3, 0.5, 3),)
    m = grid(3, 2, 0, 1.5, 0, 1, 0.25)
    sim = insem.Simulation(m, insem.Discretization(order=4, uniform_level=1, regions=regions))
    sim.set_material(0, elastic_material())
    sim.set_time(n_steps=1, dt=1e-9)
    img, f, x, z = run_operator(sim, lambda x, z: [x * z, 0 * x])
    assert img.domains[0]["constrained"]
    unit = img.mass(0) / RHO
    _, fs, _, _ = run_operator(sim, lambda x, z: [x * x + z * z, x * z])
    sel = interior(x, z, 0.0, 1.5, 0.0, 1.0, 0.25)
    scale = float(np.max(np.abs(fs)))
    np.testing.assert_allclose(f[0, sel], 0.0, atol=2e-5 * scale)
    np.testing.assert_allclose(f[1, sel], (LAM + MU) * unit[sel], atol=2e-5 * scale)


# ---------------------------------------------------------- axisymmetric

def axisym_acceleration(order, level, material, field, azimuthal_order=0):
    """a = M^-1 f_int(u) at t = 0 (initialisation computes it), as the
    reference's initial_acceleration helper (test_axisym.cpp:24-45)."""
    m = grid(3, 4, 0.0, 1.0, -1.0, 1.0, coords="axisymmetric")
    sim = insem.Simulation(m, insem.Discretization(order=order, uniform_level=level, azimuthal_order=azimuthal_order))
    sim.set_material(0, material)
    sim.set_time(n_steps=1, dt=1e-9)
    img = sim.build_setup()
    o = OracleRun(img, dt=1e-9)
    d = img.domains[0]
    r, z = img.coordinates(0)
    u = np.array(field(r, z), dtype=np.float64).reshape(d["ncomp"], -1).astype(np.float32)
    o.set_state(0, u, np.zeros_like(u))
    o.advance(0)
    return o.field(0, 2).astype(np.float64), r, z


def check_axisym_interior(a, r, z, expected, skip_axis_comps=()):
    sel = (r < 1.0 - 1e-9) & (z > -1.0 + 1e-9) & (z < 1.0 - 1e-9)
    assert np.sum(sel & (r == 0.0)) > 0
    scale = max([float(np.max(np.abs(a)))] + [abs(e) for e in expected if not callable(e)])
    for c in range(a.shape[0]):
        s = sel & ~((r == 0.0) & (c in skip_axis_comps))
        exp = expected[c](r[s], z[s]) if callable(expected[c]) else expected[c]
        np.testing.assert_allclose(a[c, s], exp, atol=1e-4 * scale)


@pytest.mark.parametrize("order", [3, 4, 6])
@pytest.mark.parametrize("level", [0, 1])
def test_axisym_acoustic_exact_for_quadratic_pressure(order, level):
    vp, rho = 1.7, 1.3
    # p = r^2 + 3 z^2: lap p = 10, p_tt = vp^2 lap p
    a, r, z = axisym_acceleration(order, level, insem.Material.acoustic(vp, rho), lambda r, z: [r * r + 3 * z * z])
    check_axisym_interior(a, r, z, (vp * vp * 10.0,))


@pytest.mark.parametrize("order", [3, 4, 6])
@pytest.mark.parametrize("level", [0, 1])
def test_axisym_elastic_exact_with_hoop_terms(order, level):
    vp, vs, rho = 2.0, 1.1, 1.4
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    cases = [
        (lambda r, z: [r, 0 * r], (0.0, 0.0)),
        (lambda r, z: [r * z, 0 * r], (0.0, 2 * (lam + mu))),
        (lambda r, z: [0 * r, r * r], (0.0, 4 * mu)),
        (lambda r, z: [0 * r, z * z], (0.0, 2 * (lam + 2 * mu))),
    ]
    for field, div_sigma in cases:
        a, r, z = axisym_acceleration(order, level, insem.Material.elastic(vp, vs, rho), field)
        check_axisym_interior(a, r, z, (div_sigma[0] / rho, div_sigma[1] / rho), skip_axis_comps=(0,))


@pytest.mark.parametrize("order", [3, 4, 6])
@pytest.mark.parametrize("level", [0, 1])
def test_azimuthal1_acoustic_exact_for_cubic_pressure(order, level):
    vp, rho = 1.7, 1.3
    # p = P cos(theta): P = r z^2 is x z^2 (lap = 2 x); P = r^3 is x (x^2 + y^2) (lap = 8 x)
    for field, lap in [(lambda r, z: [r * z * z], 2.0), (lambda r, z: [r ** 3], 8.0)]:
        a, r, z = axisym_acceleration(order, level, insem.Material.acoustic(vp, rho), field, azimuthal_order=1)
        check_axisym_interior(a, r, z, (lambda r, z, lap=lap: vp * vp * lap * r,), skip_axis_comps=(0,))


@pytest.mark.parametrize("order", [3, 4, 6])
@pytest.mark.parametrize("level", [0, 1])
def test_azimuthal1_elastic_exact_with_1_over_r_terms(order, level):
    vp, vs, rho = 2.0, 1.1, 1.4
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    s2 = np.sqrt(2.0)
    cases = [
        (lambda r, z: [0 * r, s2 * z * z, 0 * r], (0.0, s2 * 2 * mu, 0.0)),
        (lambda r, z: [0 * r, 0 * r, r * z], (0.0, s2 * (lam + mu), 0.0)),
        (lambda r, z: [0 * r, s2 * r * r, 0 * r], (0.0, s2 * (2 * lam + 6 * mu), 0.0)),
        (lambda r, z: [r * r / s2, r * r / s2, 0 * r], (0.0, s2 * (3 * lam + 5 * mu), 0.0)),
        (lambda r, z: [0 * r, 0 * r, r ** 3], (0.0, 0.0, lambda r, z: 8 * mu * r)),
    ]
    for field, ds in cases:
        a, r, z = axisym_acceleration(order, level, insem.Material.elastic(vp, vs, rho), field, azimuthal_order=1)
        exp = tuple((lambda r, z, f=e: f(r, z) / rho) if callable(e) else e / rho for e in ds)
        check_axisym_interior(a, r, z, exp, skip_axis_comps=(0, 2))
