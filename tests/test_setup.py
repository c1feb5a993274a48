"""Host setup that feeds the boundary: quadrature, Lagrange basis, DofMap
numbering, gather/assemble and the lumped mass, checked against closed
forms and invariants (restating tests/cpp/test_quadrature.cpp:14-84,
test_lagrange.cpp:29-91, test_dof_map.cpp:13-79 and
test_refinement.cpp:111-154 of the reference)."""

from __future__ import annotations

import numpy as np
import pytest

import insem_amd as insem
from models import grid, quarter_annulus
from oracle_driver import OracleRun


def acoustic_sim(m, order, level=0, regions=()):
    sim = insem.Simulation(m, insem.Discretization(order=order, uniform_level=level, regions=tuple(regions)))
    sim.set_material(0, insem.Material.acoustic(1.0, 1.0))
    sim.set_time(n_steps=1, dt=1e-9)
    return sim


# ------------------------------------------------------------- quadrature

def test_gll5_closed_form():
    # GLL(5): 0, +-sqrt(3/7), +-1 with weights 32/45, 49/90, 1/10
    p, w = insem.gll(5)
    np.testing.assert_allclose(p, [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], atol=1e-15)
    np.testing.assert_allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], atol=1e-15)


@pytest.mark.parametrize("n", range(2, 11))
def test_gll_points_are_legendre_derivative_roots(n):
    p, w = insem.gll(n)
    interior = np.sort(np.polynomial.legendre.Legendre.basis(n - 1).deriv().roots().real)
    np.testing.assert_allclose(p[1:-1], interior, atol=1e-13)
    # exact for degree <= 2n - 3
    for k in range(2 * n - 2):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert np.dot(w, p ** k) == pytest.approx(exact, abs=1e-13)


@pytest.mark.parametrize("n", range(3, 11))
def test_glj01_points_and_exactness(n):
    from scipy.special import roots_jacobi

    p, w = insem.glj01(n)
    np.testing.assert_allclose(p[1:-1], np.sort(roots_jacobi(n - 2, 1, 2)[0]), atol=1e-13)
    # int (1 + x) x^k dx exactly for k <= 2n - 3
    for k in range(2 * n - 2):
        exact = (1 - (-1) ** (k + 1)) / (k + 1) + (1 - (-1) ** (k + 2)) / (k + 2)
        assert np.dot(w, p ** k) == pytest.approx(exact, abs=1e-12)


@pytest.mark.parametrize("n", [2, 3, 5, 9])
def test_derivative_matrix_is_exact_for_polynomials(n):
    p, _ = insem.gll(n)
    D = insem.derivative_matrix(p)
    for k in range(n):
        np.testing.assert_allclose(D @ p ** k, k * p ** max(k - 1, 0) if k else 0 * p, atol=1e-11)


# ----------------------------------------------------------------- DofMap

@pytest.mark.parametrize("ngll", [2, 3, 5, 8])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_global_point_count(ngll, level):
    img = acoustic_sim(grid(3, 2), ngll - 1, level).build_setup()
    s, n = 1 << level, ngll - 1
    assert img.domains[0]["n_elements"] == 3 * 2 * s * s
    assert img.domains[0]["n_global"] == (3 * s * n + 1) * (2 * s * n + 1)


def test_multiplicity_histogram_and_assemble_gather_one():
    img = acoustic_sim(grid(2, 2), 3).build_setup()
    o = OracleRun(img, dt=1e-9)
    ng = img.domains[0]["n_global"]
    mult = o.assemble(0, o.gather(0, np.ones(ng, dtype=np.float32)))[0]
    assert np.all(mult == np.round(mult))
    count = np.bincount(mult.astype(int), minlength=5)
    assert count[4] == 1  # centre vertex
    assert count[2] == 4 * 2 + 4  # interior edges x 2 points + boundary vertices
    assert count[1] == ng - count[4] - count[2]


def test_assemble_is_transpose_of_gather_on_curved_mesh():
    m = quarter_annulus(2, 2)
    sim = insem.Simulation(m, insem.Discretization(order=4, uniform_level=1))
    sim.set_material(0, insem.Material.elastic(2.0, 1.0, 1.0))
    sim.set_time(n_steps=1, dt=1e-9)
    img = sim.build_setup()
    o = OracleRun(img, dt=1e-9)
    d = img.domains[0]
    rng = np.random.default_rng(42)
    g = rng.uniform(-1, 1, size=(2, d["n_global"])).astype(np.float32)
    loc = rng.uniform(-1, 1, size=d["n_elements"] * 2 * 25).astype(np.float32)
    lhs = np.dot(o.assemble(0, loc).ravel().astype(np.float64), g.ravel().astype(np.float64))
    rhs = np.dot(loc.astype(np.float64), o.gather(0, g).astype(np.float64))
    assert lhs == pytest.approx(rhs, rel=1e-5)


def test_assemble_is_bitwise_deterministic_across_threads():
    img = acoustic_sim(grid(4, 4), 4, 2).build_setup()
    rng = np.random.default_rng(7)
    loc = rng.uniform(-1, 1, size=img.domains[0]["n_elements"] * 25).astype(np.float32)
    outs = []
    for threads in (1, 3, 8):
        o = OracleRun(img, dt=1e-9, threads=threads)
        outs.append(o.assemble(0, loc))
    for a in outs[1:]:
        assert np.array_equal(a, outs[0])


def _refined_grid_sim(order, shear=0.3):
    # Local refinement next to two corners gives hanging faces at two levels.
    regions = ((0.75, 1.5, 0.45, 1.0, 3), (0.0, 0.3, 0.0, 0.3, 2))
    return acoustic_sim(grid(3, 2, 0, 1.5, 0, 1, shear), order, 1, regions)


@pytest.mark.parametrize("ngll", [4, 5, 7])
def test_constrained_gather_reproduces_polynomials(ngll):
    sim = _refined_grid_sim(ngll - 1)
    img = sim.build_setup()
    assert img.domains[0]["constrained"]
    o = OracleRun(img, dt=1e-9)
    gx, gz = img.coordinates(0)
    n = ngll - 1
    poly = lambda x, z: 0.3 + x ** n - 0.7 * z ** n + 0.4 * x * z ** (n - 1)  # noqa: E731
    loc = o.gather(0, poly(gx, gz).astype(np.float32))
    lx, lz = sim.leaf_points()
    exact = poly(lx, lz)
    assert np.max(np.abs(loc - exact)) < 2e-5 * max(1.0, np.max(np.abs(exact)))


def test_lumped_mass_keeps_area_with_constraints():
    sim = _refined_grid_sim(4)
    img = sim.build_setup()
    mass = img.mass(0)  # acoustic, rho = vp = 1: the quadrature weights
    assert np.all(mass > 0)
    assert mass.sum() == pytest.approx(1.5, rel=1e-6)


def test_setup_counts_of_layered_coupled_model():
    from models import layered_coupled

    sim = layered_coupled(nx=80, nz_fluid=8, nz_solid=(16, 16), duration=0.01)
    img = sim.build_setup()
    f, s = img.domains
    assert (f["medium"], s["medium"]) == ("acoustic", "elastic")
    # (nx N + 1) x (nz N + 1) points per domain, shared interface counted once per domain
    assert f["n_global"] == (80 * 4 + 1) * (8 * 4 + 1)
    assert s["n_global"] == (80 * 4 + 1) * (32 * 4 + 1)
    assert img.n_coupling == 80 * 4 + 1
    assert f["n_dirichlet"] == 80 * 4 + 1  # free top of the fluid
    assert sim.geometry_kind_counts() == {"rect": 80 * 8, "affine": 0, "curved": 80 * 32}


def test_closed_or_taken_over_handles_raise_instead_of_crashing():
    # a batch takes its images over (their handles become None); any later
    # use of such an image, or of a closed one, is an InsemError
    from insem_amd.configs import layered_coupled

    img = layered_coupled(nx=8, nz_fluid=2, nz_solid=(2, 2), duration=1e-3).build_setup()
    assert img.time_step(4.0e8)[1] > 0
    img._owned, img._h = False, None  # what BatchSolver does to the images it takes
    for call in (lambda: img.time_step(4.0e8), img.start_vectors, lambda: img.mass(0), lambda: img.coordinates(0),
                 lambda: img.amplitudes(1e-6, 3), lambda: insem.SetupImage(None)):
        with pytest.raises(insem.InsemError, match="null handle"):
            call()


def test_stage_size_is_host_only_and_refuses_unreserved_blocks():
    # the staged set-up's measuring pass makes no device call (this runs
    # without a GPU) and is deterministic; staging needs a reserved block
    import ctypes

    from insem_amd import _lib
    from insem_amd.configs import cased_hole, layered_coupled

    for sim in (layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=2e-3),
                cased_hole(np.random.default_rng(3))):
        img = sim.build_setup()
        n = img.stage_size()
        assert n == img.stage_size()
        assert n > 4 * img.n_dofs * 4  # at least u, v, a and f of every DOF
        ctx = ctypes.c_void_p()
        rc = _lib.hip().insem_hip_stage(ctypes.c_void_p(img.view), ctypes.c_void_p(1 << 20), ctypes.byref(ctx))
        assert rc != 0 and not ctx.value
        assert b"not reserved" in _lib.hip().insem_hip_last_error()
