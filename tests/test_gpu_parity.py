"""Parity of the HIP backend (through the C-ABI) against the CPU oracle.

Element kernels, gather and assembly follow the reference's operation order
exactly and are compiled without FMA contraction on both sides, so the
operator results must be bitwise equal.  Whole runs may differ in the last
bits of the double Stacey solve and receiver sums (libm and summation
details), so traces are compared with the relative L2 bound of 1e-5.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import insem_amd as insem
from insem_amd import _lib
from models import acoustic_box, grid, layered_coupled
from oracle_driver import OracleRun

pytestmark = pytest.mark.gpu

TRACE_RTOL = 1e-5  # relative L2 of each trace against the oracle


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def hip_internal_forces(solver, di, u):
    d = solver.image.domains[di]
    u = np.ascontiguousarray(u, dtype=np.float32)
    f = np.zeros((d["ncomp"], d["n_global"]), dtype=np.float32)
    dot = ctypes.c_double()
    _lib.check_hip(_lib.hip().insem_hip_internal_forces(
        ctypes.c_void_p(solver.context), di, u.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(dot)))
    return f, dot.value


def hip_block_forces(solver, di, bi, ul, fl):
    ul = np.ascontiguousarray(ul, dtype=np.float32)
    fl = np.ascontiguousarray(fl, dtype=np.float32).copy()
    _lib.check_hip(_lib.hip().insem_hip_block_forces(
        ctypes.c_void_p(solver.context), di, bi, ul.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        fl.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return fl


def compare_runs(sim, steps=None, state=None):
    """Runs the model on the GPU and the oracle with the same dt and returns
    the worst relative L2 over all traces and the two runs."""
    s = sim.solver
    img_o = sim.build_setup()
    o = OracleRun(img_o, dt=s.dt)
    assert o.n_steps == s.n_steps
    if state is not None:
        for medium, (u, v) in state.items():
            s.set_state(medium, u, v)
            o.set_state(img_o.domain_index(medium), u, v)
    if steps is None:
        s.run()
        o.run()
    else:
        s.advance(steps)
        o.advance(steps)
    tg, to = s.traces, o.traces()
    assert tg.keys() == to.keys()
    errors = {}
    for k in to:
        assert tg[k].shape == to[k].shape, k
        if np.linalg.norm(to[k]) > 0:
            errors[k] = rel_l2(tg[k], to[k])
    for k, e in sorted(errors.items(), key=lambda kv: -kv[1])[:5]:
        print(f"trace {k}: relative L2 {e:.3e}, peak {np.abs(to[k]).max():.3e}")
    worst = max(errors.values(), default=0.0)
    return worst, s, o, img_o


def test_device_visible():
    hip = _lib.hip()
    assert hip.insem_hip_device_count() >= 1
    dev = ctypes.c_int32(-1)
    assert hip.insem_hip_get_device(ctypes.byref(dev)) == 0
    assert 0 <= dev.value < hip.insem_hip_device_count()
    assert hip.insem_hip_set_device(dev.value) == 0


def test_acoustic_box_stable_dt_matches_oracle():
    sim = acoustic_box()
    s = sim.solver
    img = sim.build_setup()
    o = OracleRun(img, dt=1e-6)
    lam = o.lambda_max()
    assert s.stable_dt == pytest.approx(2.0 / np.sqrt(lam), rel=1e-9)


def test_acoustic_box_internal_forces_bitwise():
    sim = acoustic_box()
    s = sim.solver
    img = sim.build_setup()
    o = OracleRun(img, dt=s.dt)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, size=(1, img.domains[0]["n_global"])).astype(np.float32)
    fg, dg = hip_internal_forces(s, 0, u)
    fo, do = o.internal_forces(0, u)
    assert np.array_equal(fg, fo)
    assert dg == pytest.approx(do, rel=1e-12)


def test_acoustic_box_traces_match_oracle():
    worst, s, o, _ = compare_runs(acoustic_box())
    assert s.step == s.n_steps
    assert worst <= TRACE_RTOL, worst


@pytest.mark.parametrize("shear", [0.0, 0.3])
def test_elastic_block_forces_bitwise(shear):
    m = grid(6, 4, 0.0, 1.5, 0.0, 1.0, shear)
    sim = insem.Simulation(m, insem.Discretization(order=4, uniform_level=1))
    sim.set_material(0, insem.Material.elastic(3.0, 1.5, 2.0))
    sim.set_time(n_steps=4)
    s = sim.solver
    img = sim.build_setup()
    o = OracleRun(img, dt=s.dt)
    d = img.domains[0]
    npe = 25
    rng = np.random.default_rng(2)
    ul = rng.uniform(-1, 1, size=d["n_elements"] * 2 * npe).astype(np.float32)
    fl0 = rng.uniform(-1, 1, size=ul.size).astype(np.float32)
    for bi in range(d["n_blocks"]):
        assert np.array_equal(hip_block_forces(s, 0, bi, ul, fl0), o.block_forces(0, bi, ul, fl0))


def small_borehole(azimuthal_order, order=4):
    """Coupled axisymmetric borehole (mud column + elastic formation) with
    wavelength refinement, so hanging faces and GLJ axial elements appear."""
    bh = insem.mesh.borehole((-0.6, 0.6), [("mud", 0.1)], 0.5, dz=0.1)
    disc = insem.Discretization(order=order, f_max=1.2e4, points_per_wavelength=5.0,
                                azimuthal_order=azimuthal_order)
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
    sim.set_time(duration=4e-4, energy_every=25)
    return sim


@pytest.mark.parametrize("azimuthal_order", [0, 1])
def test_borehole_axisym_parity(azimuthal_order):
    sim = small_borehole(azimuthal_order)
    worst, s, o, img = compare_runs(sim)
    assert any(d["constrained"] for d in img.domains)
    assert img.n_coupling > 0
    assert worst <= TRACE_RTOL, worst
    rng = np.random.default_rng(5)
    for di, d in enumerate(img.domains):
        u = rng.uniform(-1, 1, size=(d["ncomp"], d["n_global"])).astype(np.float32)
        assert np.array_equal(hip_internal_forces(s, di, u)[0], o.internal_forces(di, u)[0])
    eg, eo = s.energy, o.energy()
    assert eg.shape == eo.shape
    np.testing.assert_allclose(eg[:, 1:], eo[:, 1:], atol=1e-5 * np.abs(eo[:, 1:]).max())


def test_borehole_stable_dt_matches_oracle():
    sim = small_borehole(1)
    s = sim.solver
    o = OracleRun(sim.build_setup(), dt=1e-9)
    assert s.stable_dt == pytest.approx(2.0 / np.sqrt(o.lambda_max()), rel=1e-9)


def test_mixed_kind_blocks_parity():
    # One domain with Rect, Affine and Curved blocks (a Quad9 mesh whose
    # right column is sheared and whose top row is curved).
    nx, nz = 3, 3
    mx, mz = 2 * nx + 1, 2 * nz + 1
    nodes = []
    for j in range(mz):
        for i in range(mx):
            x, z = i / (mx - 1), j / (mz - 1)
            if i >= mx - 3:
                x += 0.15 * z
            if j == mz - 1:
                z += 0.05 * np.sin(np.pi * x)
            nodes.append((x, z))
    node = lambda i, j: i + mx * j  # noqa: E731
    cells, bnd = [], []
    for j in range(nz):
        for i in range(nx):
            a, b = 2 * i, 2 * j
            c = len(cells)
            cells.append((node(a, b), node(a + 2, b), node(a + 2, b + 2), node(a, b + 2), node(a + 1, b),
                          node(a + 2, b + 1), node(a + 1, b + 2), node(a, b + 1), node(a + 1, b + 1)))
            if j == 0:
                bnd.append((c, 0, 1))
            if i == nx - 1:
                bnd.append((c, 1, 2))
    m = insem.MacroMesh("quad9", "cartesian", np.array(nodes), np.array(cells), np.zeros(len(cells)), np.array(bnd))
    sim = insem.Simulation(m, insem.Discretization(order=5, uniform_level=2))
    sim.set_material(0, insem.Material.elastic(2.0, 1.0, 1.2))
    sim.set_boundary(1, "absorbing")
    sim.set_boundary(2, "absorbing")
    sim.add_source(insem.Source.force(0.4, 0.5, 1.0, -0.5, insem.ricker(6.0)))
    sim.add_receiver("a", 0.7, 0.3, ["displacement", "acceleration"])
    sim.set_time(n_steps=300, energy_every=50)
    kinds = sim.geometry_kind_counts()
    assert min(kinds.values()) > 0, kinds
    worst, s, o, img = compare_runs(sim)
    assert worst <= TRACE_RTOL, worst
    assert img.domains[0]["n_blocks"] == 3


def test_acoustic_box_particle_velocity_and_fields():
    sim = acoustic_box(duration=2e-3)
    sim.add_receiver("pv", -0.7, 1.1, ["pressure", "particle_velocity"])
    worst, s, o, img = compare_runs(sim)
    assert worst <= TRACE_RTOL, worst
    for which in range(3):
        g = s.state("acoustic")[which]
        c = o.field(0, which).astype(np.float64)
        assert rel_l2(g, c) <= 1e-5


def test_runs_are_bitwise_reproducible():
    out = []
    for _ in range(2):
        sim = layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.03, n_receivers=7)
        s = sim.solver
        s.run()
        out.append(s.traces)
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k])


def test_advance_in_chunks_equals_one_call():
    # The predictor is fused into the previous step's tail inside one call;
    # splitting a run into calls must not change a bit.
    ref = layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.03, n_receivers=7)
    chunked = layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.03, n_receivers=7)
    a, b = ref.solver, chunked.solver
    a.run()
    while b.step < b.n_steps:
        b.advance(7)
    for k, v in a.traces.items():
        assert np.array_equal(v, b.traces[k])
    for medium in ("acoustic", "elastic"):
        for x, y in zip(a.state(medium), b.state(medium)):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("model,switch", [("layered", "INSEM_HIP_FUSE"), ("borehole_m1", "INSEM_HIP_FUSE"),
                                          ("layered", "INSEM_HIP_PAIR"), ("borehole_m0", "INSEM_HIP_PAIR"),
                                          ("borehole_m1", "INSEM_HIP_PAIR"), ("layered", "INSEM_HIP_LOCAL"),
                                          ("borehole_m1", "INSEM_HIP_LOCAL"), ("layered", "INSEM_HIP_BATCH_ORDER"),
                                          ("borehole_m0", "INSEM_HIP_BATCH_ORDER")])
def test_step_schedules_are_bitwise_equal(model, switch, monkeypatch):
    # Points owned by one element slot are updated inside the force launch;
    # the same run with every point updated by the tail kernels
    # (INSEM_HIP_FUSE=0) must give the same bits in traces and fields.  The
    # elastic and acoustic blocks share one force launch (INSEM_HIP_PAIR=0:
    # one after the other), points whose slots lie in one work-group are
    # summed there (INSEM_HIP_LOCAL=0: in the tail), and blocks launch their
    # elements in compact batches (INSEM_HIP_BATCH_ORDER=0: block order): every
    # variant gives the same bits.
    def make():
        if model == "layered":
            return layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.03, n_receivers=7,
                                   energy_every=10)
        return small_borehole(0 if model == "borehole_m0" else 1)

    fused = make()
    a = fused.solver
    monkeypatch.setenv(switch, "0")
    plain = make()
    b = plain.solver
    monkeypatch.delenv(switch)
    owned = ctypes.c_int64()
    read = ctypes.c_int64()
    _lib.check_hip(_lib.hip().insem_hip_block_points(ctypes.c_void_p(a.context), 0, 0, ctypes.byref(owned),
                                                     ctypes.byref(read)))
    assert owned.value > 0
    a.advance(40)
    b.advance(40)
    for k, v in a.traces.items():
        assert np.array_equal(v, b.traces[k]), k
    for medium in a.media:
        for x, y in zip(a.state(medium), b.state(medium)):
            assert np.array_equal(x, y)
    np.testing.assert_allclose(a.energy, b.energy, rtol=1e-9, atol=1e-9 * np.abs(b.energy).max())


def test_layered_coupled_small_parity():
    sim = layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.05, n_receivers=11, energy_every=10)
    worst, s, o, img = compare_runs(sim)
    assert worst <= TRACE_RTOL, worst
    # the operator of each domain is bitwise the oracle's
    rng = np.random.default_rng(3)
    for di, d in enumerate(img.domains):
        u = rng.uniform(-1, 1, size=(d["ncomp"], d["n_global"])).astype(np.float32)
        assert np.array_equal(hip_internal_forces(s, di, u)[0], o.internal_forces(di, u)[0])
    eg, eo = s.energy, o.energy()
    assert eg.shape == eo.shape
    np.testing.assert_allclose(eg[:, 0], eo[:, 0], rtol=1e-12)
    scale = np.abs(eo[:, 1:]).max()
    np.testing.assert_allclose(eg[:, 1:], eo[:, 1:], atol=1e-5 * scale)


# Every polynomial order the kernels are built for (NG = 2 .. 9), on each
# physics: the element forces bitwise against the oracle, and a short run of
# the fused time step at two more orders.

@pytest.mark.parametrize("order", [1, 2, 3, 5, 6, 7, 8])
@pytest.mark.parametrize("azimuthal_order", [0, 1])
def test_borehole_forces_bitwise_every_order(order, azimuthal_order):
    sim = small_borehole(azimuthal_order, order)
    s = sim.solver
    img = sim.build_setup()
    assert img.ngll == order + 1
    o = OracleRun(img, dt=s.dt)
    rng = np.random.default_rng(order)
    for di, d in enumerate(img.domains):
        u = rng.uniform(-1, 1, size=(d["ncomp"], d["n_global"])).astype(np.float32)
        assert np.array_equal(hip_internal_forces(s, di, u)[0], o.internal_forces(di, u)[0]), (order, di)


@pytest.mark.parametrize("order", [1, 2, 3, 5, 6, 7, 8])
def test_cartesian_forces_bitwise_every_order(order):
    # sheared (Curved) cells: the general metric path of both media
    for medium in ("elastic", "acoustic"):
        sim = insem.Simulation(grid(3, 2, 0.0, 1.5, 0.0, 1.0, 0.3), insem.Discretization(order=order))
        sim.set_material(0, insem.Material.elastic(3.0, 1.5, 2.0) if medium == "elastic"
                         else insem.Material.acoustic(1.5, 1.0))
        sim.set_time(n_steps=4)
        s = sim.solver
        img = sim.build_setup()
        o = OracleRun(img, dt=s.dt)
        d = img.domains[0]
        rng = np.random.default_rng(10 + order)
        u = rng.uniform(-1, 1, size=(d["ncomp"], d["n_global"])).astype(np.float32)
        assert np.array_equal(hip_internal_forces(s, 0, u)[0], o.internal_forces(0, u)[0]), (order, medium)
        s.close()


@pytest.mark.parametrize("order", [2, 7])
def test_borehole_step_parity_other_orders(order):
    sim = small_borehole(1, order)
    sim.set_time(n_steps=120, energy_every=40)
    worst, s, o, img = compare_runs(sim)
    assert s.step == s.n_steps == 120
    assert worst <= TRACE_RTOL, worst
