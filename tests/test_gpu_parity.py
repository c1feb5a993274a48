"""Parity of the HIP backend (through the C-ABI) against the CPU oracle.

Element kernels, gather and assembly follow the reference's operation order
exactly and are compiled without FMA contraction on both sides, so the
operator results must be bitwise equal.  Whole runs differ only in the FP32
Stacey solve and receiver sums (double in the reference), so traces are
compared with the relative L2 bound of 1e-5.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import insem_amd as insem
from insem_amd import _lib
from models import cfg0, grid, layered_coupled
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
    worst = 0.0
    for k in to:
        assert tg[k].shape == to[k].shape, k
        if np.linalg.norm(to[k]) > 0:
            worst = max(worst, rel_l2(tg[k], to[k]))
    return worst, s, o, img_o


def test_device_visible():
    assert _lib.hip().insem_hip_device_count() >= 1


def test_cfg0_stable_dt_matches_oracle():
    sim = cfg0()
    s = sim.solver
    img = sim.build_setup()
    o = OracleRun(img, dt=1e-6)
    lam = o.lambda_max()
    assert s.stable_dt == pytest.approx(2.0 / np.sqrt(lam), rel=1e-9)


def test_cfg0_internal_forces_bitwise():
    sim = cfg0()
    s = sim.solver
    img = sim.build_setup()
    o = OracleRun(img, dt=s.dt)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, size=(1, img.domains[0]["n_global"])).astype(np.float32)
    fg, dg = hip_internal_forces(s, 0, u)
    fo, do = o.internal_forces(0, u)
    assert np.array_equal(fg, fo)
    assert dg == pytest.approx(do, rel=1e-12)


def test_cfg0_traces_match_oracle():
    worst, s, o, _ = compare_runs(cfg0())
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
