"""Staged contexts (insem_hip_stage / insem_hip_realize): the host half of a
context's set-up made without device calls, then copied to a reserved
block, gives the same context as insem_hip_create, bit for bit."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import insem_amd as insem
from insem_amd import _lib
from insem_amd.configs import cased_hole, layered_coupled
from insem_amd.core import PowerJob, StagedPower, reserve_block, unreserve_block

pytestmark = pytest.mark.gpu


def models():
    return [layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.01, n_receivers=5),
            cased_hole(np.random.default_rng(21), duration=2.5e-4, z=(-0.4, 0.8), formation_radius=0.6, dz=0.08,
                       order=4, f_max=8e3, n_receivers=3, receiver_z0=0.2)]


@pytest.mark.parametrize("k", [0, 1])
def test_staged_power_job_and_solver_equal_direct_ones(k):
    direct = PowerJob(models()[k].build_setup())
    lam = direct.lambda_max()
    direct.close()
    img = models()[k].build_setup()
    block = reserve_block(img.stage_size() + 65536)
    staged = StagedPower(img, block)
    job = PowerJob.from_staged(staged)
    solver = job.into_solver(img)
    job.close()
    ref = models()[k].solver
    assert solver.dt == ref.dt and solver.n_steps == ref.n_steps
    assert lam > 0
    solver.run()
    ref.run()
    a, b = solver.traces, ref.traces
    assert a.keys() == b.keys()
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    solver.close()


def test_staged_context_refuses_calls_and_gives_its_block_back():
    hip = _lib.hip()
    img = models()[1].build_setup()
    n = img.stage_size()
    small = reserve_block(n // 2)
    ctx = ctypes.c_void_p()
    assert hip.insem_hip_stage(ctypes.c_void_p(img.view), ctypes.c_void_p(small), ctypes.byref(ctx)) != 0
    assert b"reserved for it" in hip.insem_hip_last_error()
    unreserve_block(small)
    block = reserve_block(n)
    _lib.check_hip(hip.insem_hip_stage(ctypes.c_void_p(img.view), ctypes.c_void_p(block), ctypes.byref(ctx)))
    assert hip.insem_hip_advance(ctx, 1) != 0
    assert b"staged" in hip.insem_hip_last_error()
    # one staged context per block; a claimed block is not given back by hand
    ctx2 = ctypes.c_void_p()
    assert hip.insem_hip_stage(ctypes.c_void_p(img.view), ctypes.c_void_p(block), ctypes.byref(ctx2)) != 0
    assert b"in use" in hip.insem_hip_last_error()
    with pytest.raises(insem.InsemError, match="staged context"):
        unreserve_block(block)
    _lib.check_hip(hip.insem_hip_destroy(ctx))  # never realized: the block goes back
    with pytest.raises(insem.InsemError, match="not reserved"):
        unreserve_block(block)
