"""The oracle's double build (liboracle64.so, the reference's INSEM_REAL=double
build, CMakeLists.txt:8) against its float build, and pinned on its own.

FP32_VS_FP64 is the stated error of FP32 arithmetic against the FP64
oracle: the relative L2 of every trace, same setup and dt.  The figures in
comments are the float oracle's; the GPU's float traces equal the float
oracle's to 1e-5 (tests/test_gpu_parity.py) and are checked against the
double build with the same bound in tests/test_gpu_configs.py.
"""

from __future__ import annotations

import numpy as np
import pytest

import analytic
import insem_amd as insem
from models import acoustic_box, grid, layered_coupled
from oracle_driver import OracleRun

FP32_VS_FP64 = 1e-4


def small_borehole(azimuthal_order):
    bh = insem.mesh.borehole((-0.6, 0.6), [("mud", 0.1)], 0.5, dz=0.1)
    disc = insem.Discretization(order=4, f_max=1.2e4, points_per_wavelength=5.0, azimuthal_order=azimuthal_order)
    sim = insem.Simulation(bh.mesh, disc)
    sim.set_material(bh.regions["mud"], insem.Material.acoustic(1500.0, 1000.0))
    sim.set_material(bh.regions["formation"], insem.Material.elastic(4000.0, 2300.0, 2500.0))
    for name in ("bottom", "outer", "top"):
        sim.set_boundary(bh.boundaries[name], "absorbing")
    if azimuthal_order == 0:
        sim.add_source(insem.Source.monopole(0.0, 0.0, insem.ricker(8000.0)))
    else:
        sim.add_source(insem.Source.dipole(0.0, 0.0, 1.0, 0.0, insem.ricker(4000.0)))
    for k, z in enumerate((0.1, 0.25, 0.4)):
        sim.add_receiver(f"p{k}", 0.05, z, ["pressure", "particle_velocity"])
        sim.add_receiver(f"s{k}", 0.3, z, ["displacement", "velocity"])
    sim.set_time(duration=4e-4)
    return sim


def worst_trace_error(sim):
    img = sim.build_setup()
    f32 = OracleRun(img)
    f32.run()
    f64 = OracleRun(img, dt=f32.dt, double=True)
    f64.run()
    a, b = f32.traces(), f64.traces()
    assert a.keys() == b.keys()
    return max(analytic.rel_l2(a[k], b[k]) for k in a)


@pytest.mark.parametrize("name", ["acoustic_box", "layered", "borehole_m0", "borehole_m1"])
def test_float_build_within_stated_error_of_double_build(name):
    # acoustic_box 4.6e-7; layered 40x20 1.8e-5; boreholes 2.0e-6 (m = 0), 1.2e-5 (m = 1)
    sim = {"acoustic_box": acoustic_box, "layered": lambda: layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.3,
                                                              n_receivers=5),
           "borehole_m0": lambda: small_borehole(0), "borehole_m1": lambda: small_borehole(1)}[name]()
    err = worst_trace_error(sim)
    assert 0.0 < err < FP32_VS_FP64


def test_double_build_conserves_energy_to_double_round_off():
    # The reference's double build conserves energy to 1e-10
    # (tests/cpp/test_solver.cpp:71-72), its float build to 1e-4.  The
    # double oracle reads the float setup, whose mass and inverse mass are
    # rounded to float separately (the update conserves the energy of
    # 1/inv_mass, the monitor weighs by mass: ~6e-8 apart), so its floor is
    # 1e-7 rather than 1e-10 (measured 1.5e-8), still far below the float
    # build's drift.
    sim = insem.Simulation(grid(6, 6), insem.Discretization(order=4))
    sim.set_material(0, insem.Material.elastic(2.0, 1.1, 1.3))
    sim.set_time(n_steps=400, energy_every=20)
    img = sim.build_setup()
    x, z = img.coordinates(0)
    u = np.zeros((2, x.size))
    u[0] = np.exp(-((x - 0.45) ** 2 + (z - 0.55) ** 2) / (2 * 0.08 ** 2))
    drift = {}
    for double in (False, True):
        o = OracleRun(img, double=double)
        o.set_state(0, u, np.zeros_like(u))
        o.run()
        e = o.energy()
        total = e[:, 1] + e[:, 2]
        drift[double] = np.abs(total / total[0] - 1.0).max()
    assert drift[True] < 1e-7
    assert drift[False] > 20 * drift[True]  # float 1.1e-6


def test_double_build_has_the_same_discretisation_error():
    # FP64 arithmetic changes the analytic error only at the round-off level:
    # the 2D Green's function error (discretisation-bound, ~1.3 %) agrees.
    c, rho, xr, zr = 1.0, 1.0, 1.5, 0.7
    m = insem.mesh.rectangle((-4.0, 4.0), (-4.0, 4.0), nx=32, nz=32)
    sim = insem.Simulation(m, insem.Discretization(order=4))
    sim.set_material(0, insem.Material.acoustic(c, rho))
    for tag in (1, 2, 3, 4):
        sim.set_boundary(tag, "absorbing")
    stf = insem.ricker(1.0)
    sim.add_source(insem.Source.monopole(0.0, 0.0, stf))
    sim.add_receiver("r", xr, zr, ["pressure"])
    sim.set_time(duration=4.5)
    img = sim.build_setup()
    err = {}
    for double in (False, True):
        o = OracleRun(img, double=double)
        o.run()
        p = o.traces()[("r", "pressure")][:, 0]
        t = np.arange(p.size) * o.dt
        err[double] = analytic.rel_l2(p, analytic.acoustic_monopole_2d(t, stf(t), xr, zr, c, rho))
    assert err[True] < 0.03
    assert abs(err[True] - err[False]) < 1e-4
