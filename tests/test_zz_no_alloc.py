"""insem_hip_block_forces_noalloc against insem_hip_block_forces, through the
C-ABI: bitwise equal f_local on two calls in a row, the second on a smaller
model, and the size limit refused with INSEM_HIP_EINVAL."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import insem_amd as insem
from insem_amd import _lib
from models import grid

pytestmark = pytest.mark.gpu

F32P = ctypes.POINTER(ctypes.c_float)


def forces(fn, solver, di, bi, ul, fl):
    ul = np.ascontiguousarray(ul, dtype=np.float32)
    out = np.ascontiguousarray(fl, dtype=np.float32).copy()
    rc = fn(ctypes.c_void_p(solver.context), di, bi, ul.ctypes.data_as(F32P), out.ctypes.data_as(F32P))
    return rc, out


def both(solver, di, bi, ul, fl):
    hip = _lib.hip()
    rc0, ref = forces(hip.insem_hip_block_forces, solver, di, bi, ul, fl)
    rc1, got = forces(hip.insem_hip_block_forces_noalloc, solver, di, bi, ul, fl)
    assert rc0 == 0 and rc1 == 0, _lib.hip().insem_hip_last_error()
    return ref, got


def model(nx, nz, shear, elastic):
    sim = insem.Simulation(grid(nx, nz, 0.0, 1.5, 0.0, 1.0, shear), insem.Discretization(order=4))
    sim.set_material(0, insem.Material.elastic(3.0, 1.5, 2.0) if elastic else insem.Material.acoustic(2.0, 1.2))
    sim.set_time(n_steps=4)
    return sim


def test_noalloc_equals_allocating_entry_point_on_two_calls():
    rng = np.random.default_rng(11)
    # first call: a sheared elastic grid (2 components); second: a smaller
    # acoustic grid (1 component), so the second input is smaller than the
    # first and any leftover of the first call would show in its result
    for nx, nz, shear, elastic in ((12, 9, 0.3, True), (5, 4, 0.0, False)):
        sim = model(nx, nz, shear, elastic)
        s = sim.solver
        d = s.image.domains[0]
        n = d["n_elements"] * d["ncomp"] * 25
        ul = rng.uniform(-1, 1, size=n).astype(np.float32)
        fl = rng.uniform(-1, 1, size=n).astype(np.float32)
        for bi in range(d["n_blocks"]):
            ref, got = both(s, 0, bi, ul, fl)
            assert np.array_equal(ref, got)
            assert not np.array_equal(got, fl)  # the block's elements were written
        s.close()


def test_noalloc_refuses_fields_beyond_its_limit():
    # 410 x 410 elastic elements of order 4: 168,100 * 2 * 25 = 8,405,000
    # element-local values, above the 8,388,608 the fixed array holds
    sim = model(410, 410, 0.0, True)
    s = sim.solver
    d = s.image.domains[0]
    n = d["n_elements"] * d["ncomp"] * 25
    assert n > 8 * 2**20
    ul = np.zeros(n, dtype=np.float32)
    rc, _ = forces(_lib.hip().insem_hip_block_forces_noalloc, s, 0, 0, ul, ul)
    assert rc == -1
    assert b"no-allocation limit" in _lib.hip().insem_hip_last_error()
    s.close()
