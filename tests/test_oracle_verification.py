"""The oracle's functions pinned by closed-form physics, independently of its
own formulas: every source kind and coefficient, the fluid-solid coupling,
Stacey absorbing boundaries (Cartesian elastic and acoustic, axisymmetric
m = 0 and m = 1 with its T C T^T transform), the receivers' particle
velocity and m = 1 rotation, and the coupled power-iteration Δt.

Restates the reference's verification tests (tests/python/test_verification.py:
28-100, 143-191, 223-254; test_axisym.py:26-65; test_dipole.py:23-77;
tests/cpp/test_dipole.cpp:174-200) with this repository's own analytic
solutions (tests/analytic.py).  The error bounds are the reference's; the
figures in comments are the oracle's (float build).
"""

from __future__ import annotations

import numpy as np
import pytest

import analytic
import insem_amd as insem
from insem_amd import mesh
from models import grid
from oracle_driver import OracleRun

STF = insem.ricker(1.0)


def open_box(half, n, order, material, coords="cartesian", r1=None, azimuthal_order=0):
    """Square [-half, half]^2 (Cartesian) or [0, r1] x [-half, half]
    (axisymmetric, the left side the axis), absorbing elsewhere."""
    if coords == "cartesian":
        m = mesh.rectangle((-half, half), (-half, half), nx=n, nz=n)
        tags = (mesh.BOTTOM, mesh.RIGHT, mesh.TOP, mesh.LEFT)
    else:
        m = mesh.rectangle((0.0, r1), (-half, half), nx=n // 2, nz=n, coords="axisymmetric")
        tags = (mesh.BOTTOM, mesh.RIGHT, mesh.TOP)
    sim = insem.Simulation(m, insem.Discretization(order=order, azimuthal_order=azimuthal_order))
    sim.set_material(0, material)
    for tag in tags:
        sim.set_boundary(tag, "absorbing")
    return sim


def run(sim, state=None):
    img = sim.build_setup()
    o = OracleRun(img)
    if state is not None:
        for medium, fn in state.items():
            di = img.domain_index(medium)
            d = img.domains[di]
            x, z = img.coordinates(di)
            u, v = fn(x, z, d["ncomp"])
            o.set_state(di, u, v)
    o.run()
    tr = o.traces()
    n = next(iter(tr.values())).shape[0] if tr else 0
    return o, tr, np.arange(n) * o.dt * img.time.record_every


# ------------------------------------------------------------ point sources

@pytest.mark.parametrize("kind,force,moment", [
    ("force", (0.3, 1.0), None),
    ("explosion", None, ((1.0, 0.0), (0.0, 1.0))),
    ("double-couple", None, ((0.0, 1.0), (1.0, 0.0))),
])
def test_elastic_point_sources_match_green_function(kind, force, moment):
    # Force: f phi; moment tensor: M . grad phi (solver.cpp:474-488).  0.14-0.20 %.
    vp, vs, rho, xr, zr = 2.0, 1.15, 1.0, 1.9, 1.1
    sim = open_box(5.0, 48, 5, insem.Material.elastic(vp, vs, rho))
    if force is not None:
        sim.add_source(insem.Source.force(0.0, 0.0, force[0], force[1], STF))
    else:
        sim.add_source(insem.Source.moment_tensor(0.0, 0.0, moment[0][0], moment[1][1], moment[0][1], STF))
    sim.add_receiver("r", xr, zr, ["displacement"])
    sim.set_time(duration=4.0)
    _, tr, t = run(sim)
    exact = analytic.elastic_point_source_2d(t, STF(t), xr, zr, vp, vs, rho, force=force, moment=moment)
    assert analytic.rel_l2(tr[("r", "displacement")], exact) < 0.01


def test_acoustic_dipole_matches_green_function():
    # Dipole: f . grad phi / rho (solver.cpp:497-502); rho != 1 so a wrong
    # density factor shows.  0.45 %.
    c, rho, fx, fz = 1.0, 1.3, 0.6, -0.8
    sim = open_box(4.0, 32, 6, insem.Material.acoustic(c, rho))
    sim.add_source(insem.Source.dipole(0.0, 0.0, fx, fz, STF))
    receivers = {"a": (1.5, 0.7), "b": (-0.9, 1.2)}
    for name, (x, z) in receivers.items():
        sim.add_receiver(name, x, z, ["pressure"])
    sim.set_time(duration=4.5)
    _, tr, t = run(sim)
    for name, (x, z) in receivers.items():
        exact = analytic.acoustic_dipole_2d(t, STF(t), x, z, c, fx, fz)
        assert analytic.rel_l2(tr[(name, "pressure")][:, 0], exact) < 0.01


def test_acoustic_particle_velocity_matches_green_function():
    # -grad p / rho at the receiver (d_dx, d_dz weights) integrated by the
    # trapezoid (solver.cpp:883-893, 945-958).  0.36 %.
    c, rho, xr, zr = 1.0, 1.0, 1.2, -0.9
    sim = open_box(4.0, 40, 5, insem.Material.acoustic(c, rho))
    sim.add_source(insem.Source.monopole(0.0, 0.0, STF))
    sim.add_receiver("r", xr, zr, ["particle_velocity"])
    sim.set_time(duration=4.0)
    _, tr, t = run(sim)
    v = tr[("r", "particle_velocity")]
    assert v.shape[1] == 2
    assert analytic.rel_l2(v, analytic.acoustic_particle_velocity_2d(t, STF(t), xr, zr, c, rho)) < 0.01


AXIS_RECEIVERS = {"axis": (0.0, 1.6), "off": (1.2, -0.9), "near": (0.3, 0.4)}


def test_axisymmetric_monopole_matches_3d_point_source():
    # 1/(2 pi) scaling, GLJ axial elements (solver.cpp:441-446).
    # Order 4: 0.20-0.68 %; order 6: 0.05-0.17 %.
    c, rho = 1.0, 1.0
    err = {}
    for order in (4, 6):
        sim = open_box(4.0, 48, order, insem.Material.acoustic(c, rho), "axisymmetric", 4.0)
        sim.add_source(insem.Source.monopole(0.0, 0.0, STF))
        for name, (r, z) in AXIS_RECEIVERS.items():
            sim.add_receiver(name, r, z, ["pressure"])
        sim.set_time(duration=4.5)
        _, tr, t = run(sim)
        for name, (r, z) in AXIS_RECEIVERS.items():
            exact = analytic.acoustic_monopole_3d(t, STF(t), r, z, c, rho)
            err[order, name] = analytic.rel_l2(tr[(name, "pressure")][:, 0], exact)
    for name in AXIS_RECEIVERS:
        assert err[4, name] < 0.015
        assert err[6, name] < 0.5 * err[4, name]


@pytest.mark.parametrize("kind", ["force", "explosion"])
def test_axisymmetric_elastic_sources_match_3d_point_sources(kind):
    # Vertical force on the axis; isotropic moment tensor with the hoop term
    # M_tt = M_rr (solver.cpp:484-488).  0.10-0.64 %.
    vp, vs, rho = 2.0, 1.15, 1.0
    sim = open_box(5.0, 48, 5, insem.Material.elastic(vp, vs, rho), "axisymmetric", 5.0)
    if kind == "force":
        sim.add_source(insem.Source.force(0.0, 0.0, 0.0, 1.0, STF))
    else:
        sim.add_source(insem.Source.moment_tensor(0.0, 0.0, 1.0, 1.0, 0.0, STF))
    receivers = {"axis": (0.0, 1.8), "off": (1.5, 1.0), "off2": (2.0, -1.2)}
    for name, (r, z) in receivers.items():
        sim.add_receiver(name, r, z, ["displacement"])
    sim.set_time(duration=4.0)
    _, tr, t = run(sim)
    for name, (r, z) in receivers.items():
        if kind == "force":
            exact = analytic.elastic_vertical_force_3d(t, STF(t), r, z, vp, vs, rho)
        else:
            exact = analytic.elastic_explosion_3d(t, STF(t), r, z, vp, vs, rho)
        assert analytic.rel_l2(tr[(name, "displacement")], exact) < 0.012


def test_azimuthal_order_1_fluid_dipole_matches_3d():
    # m = 1 pressure, 1/pi scaling, dipole coefficient with rho != 1.  0.2-1.1 %.
    c, rho = 1.0, 1.3
    sim = open_box(4.0, 48, 4, insem.Material.acoustic(c, rho), "axisymmetric", 4.0, azimuthal_order=1)
    sim.add_source(insem.Source.dipole(0.0, 0.0, 1.0, 0.0, STF))
    receivers = {"near": (0.3, 0.4), "off": (1.2, -0.9), "up": (0.8, 1.6)}
    for name, (r, z) in receivers.items():
        sim.add_receiver(name, r, z, ["pressure"])
    sim.set_time(duration=4.5)
    _, tr, t = run(sim)
    for name, (r, z) in receivers.items():
        exact = analytic.acoustic_dipole_3d(t, STF(t), r, z, c)
        assert analytic.rel_l2(tr[(name, "pressure")][:, 0], exact) < 0.015


def test_azimuthal_order_1_horizontal_force_matches_3d():
    # m = 1 force (f_x split over U+ and U-), and the receivers' rotation
    # (U+, U-) -> (U_r, U_theta) (solver.cpp:477-480, 800-806).  0.15-0.20 %.
    vp, vs, rho = 2.0, 1.15, 1.0
    sim = open_box(5.0, 48, 5, insem.Material.elastic(vp, vs, rho), "axisymmetric", 5.0, azimuthal_order=1)
    sim.add_source(insem.Source.force(0.0, 0.0, 1.0, 0.0, STF))
    receivers = {"axis": (0.0, 1.8), "off": (1.5, 1.0), "off2": (2.0, -1.2)}
    for name, (r, z) in receivers.items():
        sim.add_receiver(name, r, z, ["displacement"])
    sim.set_time(duration=4.0)
    _, tr, t = run(sim)
    for name, (r, z) in receivers.items():
        u = tr[(name, "displacement")]
        assert u.shape[1] == 3
        exact = analytic.elastic_horizontal_force_3d(t, STF(t), r, z, vp, vs, rho)
        assert analytic.rel_l2(u, exact) < 0.012
        if r > 0:  # U_theta carries a third of the signal off the axis
            assert analytic.rel_l2(u[:, 1], exact[:, 1]) < 0.02


# ----------------------------------------------------- Stacey boundaries

def _plane_pulse(z0, sigma, speed, comp):
    """Downgoing plane pulse f(z + speed t) in component comp: u = f, v = speed f'."""
    def state(x, z, ncomp):
        f = np.exp(-((z - z0) ** 2) / (2 * sigma ** 2))
        u = np.zeros((ncomp, x.size))
        v = np.zeros_like(u)
        u[comp] = f
        v[comp] = speed * (-(z - z0) / sigma ** 2) * f
        return u, v
    return state


def test_stacey_absorbs_acoustic_wave_at_normal_incidence():
    # Rigid-walled channel, absorbing bottom (test_verification.py:167-191).
    c = 1.0
    sim = insem.Simulation(mesh.rectangle((0.0, 0.5), (-6.0, 0.0), nx=2, nz=48), insem.Discretization(order=4))
    sim.set_material(0, insem.Material.acoustic(c, 1.0))
    sim.set_boundary(mesh.BOTTOM, "absorbing")
    sim.add_receiver("r", 0.25, -2.0, ["pressure"])
    sim.set_time(duration=10.0)
    _, tr, t = run(sim, {"acoustic": _plane_pulse(-2.0, 0.3, c, 0)})
    p = tr[("r", "pressure")][:, 0]
    incident = np.abs(p[t < 1.5]).max()
    assert np.abs(p[t > 6.5]).max() < 0.01 * incident  # the bottom echo returns after t = 8


def test_stacey_absorbs_elastic_p_wave_at_normal_incidence():
    # With lambda = 0 (vp = sqrt 2 vs) a plane P wave is an exact solution
    # with traction-free side walls, so a narrow channel carries it to the
    # absorbing bottom at normal incidence: the rho vp n n^T part of C.
    vs = 1.0
    vp = np.sqrt(2.0) * vs
    sim = insem.Simulation(mesh.rectangle((0.0, 0.5), (-6.0, 0.0), nx=2, nz=48), insem.Discretization(order=4))
    sim.set_material(0, insem.Material.elastic(vp, vs, 1.3))
    sim.set_boundary(mesh.BOTTOM, "absorbing")
    sim.add_receiver("r", 0.25, -2.0, ["displacement"])
    sim.set_time(duration=8.0 / vp + 2.5)
    _, tr, t = run(sim, {"elastic": _plane_pulse(-2.0, 0.3, vp, 1)})
    uz = tr[("r", "displacement")][:, 1]
    incident = np.abs(uz[t < 1.0]).max()
    echo = (t > 8.0 / vp - 1.0) & (t < 8.0 / vp + 1.5)
    assert incident > 0.5
    assert np.abs(uz[echo]).max() < 0.01 * incident
    # without the absorbing bottom the echo is total
    sim.set_boundary(mesh.BOTTOM, "free")
    _, tr, t = run(sim, {"elastic": _plane_pulse(-2.0, 0.3, vp, 1)})
    assert np.abs(tr[("r", "displacement")][echo, 1]).max() > 0.8 * incident


def test_stacey_absorbs_elastic_s_wave_at_normal_incidence():
    # Plane S wave in a domain wide enough that nothing from its side edges
    # reaches the centre before the bottom echo would: the rho vs (I - n n^T)
    # part of C.
    vp, vs = 2.0, 1.0
    sim = insem.Simulation(mesh.rectangle((-16.0, 16.0), (-4.0, 0.0), nx=128, nz=16), insem.Discretization(order=4))
    sim.set_material(0, insem.Material.elastic(vp, vs, 1.1))
    for tag in (mesh.BOTTOM, mesh.LEFT, mesh.RIGHT):
        sim.set_boundary(tag, "absorbing")
    sim.add_receiver("r", 0.0, -1.5, ["displacement"])
    sim.set_time(duration=6.5)
    _, tr, t = run(sim, {"elastic": _plane_pulse(-1.5, 0.25, vs, 0)})
    ux = tr[("r", "displacement")][:, 0]
    incident = np.abs(ux[t < 0.8]).max()
    echo = (t > 4.0) & (t < 6.5)  # 5 units of travel at vs: t = 5
    assert incident > 0.5
    assert np.abs(ux[echo]).max() < 0.02 * incident


@pytest.mark.parametrize("medium", ["acoustic", "elastic"])
@pytest.mark.parametrize("azimuthal_order", [0, 1])
def test_axisymmetric_stacey_boundaries_absorb(medium, azimuthal_order):
    # Energy never grows and falls below 5 % (tests/cpp/test_dipole.cpp:174-200;
    # for m = 1 elastic the damping matrix is T C T^T, solver.cpp:313-319).
    m = grid(4, 6, 0.0, 1.0, -1.0, 1.0, coords="axisymmetric")
    sim = insem.Simulation(m, insem.Discretization(order=4, azimuthal_order=azimuthal_order))
    sim.set_material(0, insem.Material.elastic(2.0, 1.1, 1.3) if medium == "elastic"
                     else insem.Material.acoustic(1.5, 1.0))
    for tag in (mesh.BOTTOM, mesh.RIGHT, mesh.TOP):
        sim.set_boundary(tag, "absorbing")
    sim.set_time(duration=3.0, energy_every=20)

    def state(r, z, ncomp):
        # component 1 g, the others r g (zero on the axis), as the reference
        g = np.exp(-(r * r + (z - 0.1) ** 2) / (2 * 0.15 ** 2))
        u = np.array([(1.0 if c == 1 else r) * g for c in range(ncomp)])
        return u, np.zeros_like(u)

    o, _, _ = run(sim, {medium: state})
    e = o.energy()
    total = e[:, 1] + e[:, 2]
    assert total[0] > 0
    assert np.all(total[1:] <= total[:-1] * (1 + 1e-6))
    assert total[-1] < 0.05 * total[0]


# ------------------------------------------------------------- coupling

def test_fluid_solid_reflection_and_transmission_at_normal_incidence():
    # Plane pressure pulse from a fluid onto a solid (test_verification.py:
    # 223-254): the coupling terms cx p and -(cx a_r + cz a_z) with their
    # face normals and weights (solver.cpp:357-418, 693-727).
    vf, rf, vp, vs, rs = 1.5, 1.0, 3.0, 1.6, 2.2
    m = mesh.layered((-4.0, 4.0), [-4.0, 0.0, 4.0], nz=[24, 24], regions=[1, 2], nx=24)
    sim = insem.Simulation(m, insem.Discretization(order=4))
    sim.set_material(1, insem.Material.elastic(vp, vs, rs))
    sim.set_material(2, insem.Material.acoustic(vf, rf))
    sim.set_boundary(mesh.TOP, "absorbing")
    sim.set_boundary(mesh.BOTTOM, "absorbing")
    sim.add_receiver("fluid", 0.0, 1.0, ["pressure"])
    sim.add_receiver("solid", 0.0, -1.0, ["velocity"])
    sim.set_time(duration=2.4)
    _, tr, t = run(sim, {"acoustic": _plane_pulse(2.0, 0.15, vf, 0)})
    p = tr[("fluid", "pressure")][:, 0]
    v = tr[("solid", "velocity")][:, 1]

    def peak(x):
        return x[np.argmax(np.abs(x))]

    incident = peak(p[t < 1.2])
    zf, zs = rf * vf, rs * vp
    assert peak(p[(t > 1.7) & (t < 2.3)]) / incident == pytest.approx((zs - zf) / (zs + zf), rel=0.01)
    assert peak(v[(t > 1.4) & (t < 2.0)]) / incident == pytest.approx(-2.0 / (zs + zf), rel=0.01)


def _coupled_box(azimuthal_order):
    """Fluid beside a solid (Cartesian) or a fluid column r < 0.25 inside a
    solid (axisymmetric), absorbing outer sides."""
    if azimuthal_order == 0:
        m = grid(4, 4, material=[0, 0, 1, 1] * 4)
        tags = (mesh.BOTTOM, mesh.RIGHT, mesh.TOP, mesh.LEFT)
    else:
        m = grid(4, 8, 0.0, 1.0, -1.0, 1.0, coords="axisymmetric")
        mat = np.where(m.nodes[m.cells].mean(axis=1)[:, 0] < 0.25, 0, 1)
        m = insem.MacroMesh(m.cell_type, m.coords, m.nodes, m.cells, mat, m.boundary)
        tags = (mesh.BOTTOM, mesh.RIGHT, mesh.TOP)
    sim = insem.Simulation(m, insem.Discretization(order=4, azimuthal_order=azimuthal_order))
    sim.set_material(0, insem.Material.acoustic(1.5, 1.3))
    sim.set_material(1, insem.Material.elastic(3.0, 1.6, 2.2))
    for tag in tags:
        sim.set_boundary(tag, "absorbing")
    sim.set_time(n_steps=600)
    return sim


def _one_trace(azimuthal_order, source, at, quantity, medium, dt=None):
    sim = _coupled_box(azimuthal_order)
    sim.add_source(source)
    sim.add_receiver("r", at[0], at[1], [quantity], medium)
    img = sim.build_setup()
    o = OracleRun(img, dt=dt)
    o.run()
    return o.dt, o.traces()[("r", quantity)]


def _trapezoid_rates(v, dt):
    """The sequence a with v[n] = v[n-1] + dt/2 (a[n-1] + a[n]), a[0] = 0
    (the particle-velocity trapezoid inverted; zero initial state)."""
    a = np.zeros(v.size)
    for n in range(1, v.size):
        a[n] = 2.0 * (v[n] - v[n - 1]) / dt - a[n - 1]
    return a


def test_fluid_solid_reciprocity_cartesian():
    # In the variables (u, p / s^2) the coupled system matrix is symmetric
    # (solid: f += C p; fluid: f -= C^T a, solver.cpp:693-727), so the
    # pressure at A from a force at B equals minus the solid acceleration at
    # B from a monopole at A, sample for sample (discrete time keeps it),
    # to round-off.  A wrong sign or factor on one side of the coupling
    # breaks it.
    stf = insem.ricker(6.0)
    A, B = (0.3, 0.55), (0.7, 0.4)
    dt, aB = _one_trace(0, insem.Source.monopole(*A, stf), B, "acceleration", "elastic")
    for comp, (fx, fz) in ((0, (1.0, 0.0)), (1, (0.0, 1.0))):
        _, pA = _one_trace(0, insem.Source.force(*B, fx, fz, stf), A, "pressure", "acoustic", dt)
        assert np.abs(pA).max() > 0.1
        assert analytic.rel_l2(pA[:, 0], -aB[:, comp]) < 1e-4  # 3-4e-6


def test_fluid_solid_reciprocity_azimuthal_order_1():
    # m = 1: the solid takes sqrt(1/2) cx p on U+ and U-, the fluid
    # a_r = sqrt(1/2)(a+ + a-) (solver.cpp:704-725).  The fluid's -dp/dr / rho
    # at A (particle-velocity weights, its trapezoid inverted) from a force at
    # B equals the solid's acceleration at B from a dipole at A, for U_r and
    # U_z.  Dropping either sqrt(1/2) gives a 29 % difference.
    stf = insem.ricker(6.0)
    A, B = (0.1, 0.2), (0.6, -0.1)
    dt, aB = _one_trace(1, insem.Source.dipole(*A, 1.0, 0.0, stf), B, "acceleration", "elastic")
    for comp, (fx, fz) in ((0, (1.0, 0.0)), (2, (0.0, 1.0))):
        _, vA = _one_trace(1, insem.Source.force(*B, fx, fz, stf), A, "particle_velocity", "acoustic", dt)
        ar = _trapezoid_rates(vA[:, 0], dt)
        assert np.abs(ar).max() > 1.0
        assert analytic.rel_l2(ar, aB[:, comp]) < 1e-4  # 1-2e-6


def test_coupled_power_iteration_gives_the_stability_limit():
    # The coupled lambda_max (solver.cpp:621-667) must be the stability limit
    # of the coupled central-difference step: a run at 0.97 of 2/sqrt(lambda)
    # stays bounded, one at 1.04 of it blows up from the same state.
    m = grid(3, 2, material=[0, 0, 0, 1, 1, 1])
    sim = insem.Simulation(m, insem.Discretization(order=4))
    sim.set_material(0, insem.Material.elastic(3.0, 1.7, 2.0))
    sim.set_material(1, insem.Material.acoustic(1.5, 1.0))
    sim.set_time(n_steps=1)
    img = sim.build_setup()
    assert img.n_coupling > 0
    o = OracleRun(img)
    lam = o.lambda_max()
    # each domain alone (dense M^-1 K, no coupling) is softer: the coupled
    # iteration decides lambda here
    alone = []
    for di, d in enumerate(img.domains):
        n = d["ncomp"] * d["n_global"]
        K = np.zeros((n, n))
        for k in range(n):
            e = np.zeros(n, dtype=np.float32)
            e[k] = 1.0
            K[:, k] = -o.internal_forces(di, e.reshape(d["ncomp"], -1))[0].ravel()
        alone.append(np.linalg.eigvals(K / np.tile(img.mass(di), d["ncomp"])[:, None]).real.max())
    assert lam > 1.01 * max(alone)
    rng = np.random.default_rng(3)
    growth = {}
    for f in (0.97, 1.04):
        sim.set_time(n_steps=3000)
        img = sim.build_setup()
        o = OracleRun(img, dt=f * 2.0 / np.sqrt(lam))
        for di, d in enumerate(img.domains):
            o.set_state(di, rng.uniform(-1, 1, (d["ncomp"], d["n_global"])), np.zeros((d["ncomp"], d["n_global"])))
        o.run()
        growth[f] = max(np.abs(o.field(di, 0)).max() for di in range(len(img.domains)))
    # a rough random start reaches a few hundred and stays there (it is the
    # same after 1000 and 3000 steps); past the limit it grows without bound
    assert growth[0.97] < 1e3
    assert not np.isfinite(growth[1.04]) or growth[1.04] > 1e6


def test_rayleigh_wave_speed_on_free_surface():
    # Traction-free top (test_verification.py:143-164).  0.2 %.
    vp, vs = 2.0, 1.0
    sim = insem.Simulation(mesh.rectangle((-8.0, 8.0), (-6.0, 0.0), nx=64, nz=24), insem.Discretization(order=4))
    sim.set_material(0, insem.Material.elastic(vp, vs, 1.0))
    for tag in (mesh.BOTTOM, mesh.LEFT, mesh.RIGHT):
        sim.set_boundary(tag, "absorbing")
    sim.add_source(insem.Source.force(0.0, 0.0, 0.0, -1.0, insem.ricker(1.5)))
    sim.add_receiver("near", 2.0, 0.0, ["velocity"])
    sim.add_receiver("far", 5.0, 0.0, ["velocity"])
    sim.set_time(duration=7.5)
    _, tr, t = run(sim)
    a, b = tr[("near", "velocity")][:, 1], tr[("far", "velocity")][:, 1]
    cc = np.correlate(b, a, "full")
    i = int(np.argmax(cc))
    frac = 0.5 * (cc[i - 1] - cc[i + 1]) / (cc[i - 1] - 2 * cc[i] + cc[i + 1])
    lag = (i + frac - (len(a) - 1)) * (t[1] - t[0])
    assert 3.0 / lag == pytest.approx(analytic.rayleigh_speed(vp, vs), rel=0.01)


# ------------------------------------------------------ the analytic module

def test_analytic_explosion_is_a_pure_p_wave():
    # -grad . G for M = I equals -grad g_p / (rho vp^2): the finite-difference
    # moment-tensor field against the closed form.
    vp, vs, rho, r, z = 2.0, 1.15, 1.3, 1.4, -0.7
    t = np.arange(1200) * 0.005
    s = STF(t)
    u = analytic.elastic_explosion_3d(t, s, r, z, vp, vs, rho)
    S, w, n = analytic._spec(t, s)
    W = analytic._positive(w)
    R = np.hypot(r, z)
    _, g1, _ = analytic.scalar_green(3, W / vp, R)
    exact = np.stack([analytic._synth(-g1 * r / R / (rho * vp * vp), S, n, t.size),
                      analytic._synth(-g1 * z / R / (rho * vp * vp), S, n, t.size)], axis=1)
    assert analytic.rel_l2(u, exact) < 1e-6


def test_analytic_2d_tensor_is_reciprocal_and_far_field_transverse():
    # G_ij = G_ji, and far away the S part is transverse: gamma^T G gamma
    # tends to the P wave alone (1/(rho vp^2)) g_p-like amplitude.
    w = np.linspace(0.5, 20.0, 40)
    G = analytic.green_tensor(2, w, (3.0, 4.0), 2.0, 1.0, 1.0)
    assert np.allclose(G[0, 1], G[1, 0])
    gam = np.array([0.6, 0.8])
    far = analytic.green_tensor(2, w, 2000.0 * gam, 2.0, 1.0, 1.0)
    longi = np.einsum("i,ijw,j->w", gam, far, gam)
    gp, _, _ = analytic.scalar_green(2, w / 2.0, 2000.0)
    assert np.allclose(longi, gp / (1.0 * 4.0), rtol=1e-2)
