"""The autotuner's variants of the paired force launch give the same bits.

Each variant (register budget for the default, at most 6 or at most 4 waves
per SIMD with as many, 6 or 4 work-groups per CU; LLVM's max-ilp or default
scheduling strategy) runs the same arithmetic in the same order, so a run
with any variant forced equals the run with the tuned table's choice bit for
bit in every trace, field and energy row: the layered (Cartesian curved
elastic + rectangular acoustic) and the borehole (axisymmetric m = 0 and
m = 1, GLJ axial elements, hanging faces) models, at orders 4, 5 and 6,
where the variants exist.  Also: the tuned table's choice is what the
context reports, and out-of-range variants are refused.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from insem_amd import _lib
from models import layered_coupled
from test_gpu_parity import small_borehole

pytestmark = pytest.mark.gpu

VARIANTS = 6


def variant_of(solver):
    v, p = ctypes.c_int32(), ctypes.c_int32()
    _lib.check_hip(_lib.hip().insem_hip_force_variant(ctypes.c_void_p(solver.context), ctypes.byref(v),
                                                       ctypes.byref(p)))
    return v.value, p.value


def run(make, variant, steps=30):
    sim = make()
    s = sim.solver
    if variant is not None:
        _lib.check_hip(_lib.hip().insem_hip_set_force_variant(ctypes.c_void_p(s.context), variant))
    s.advance(steps)
    out = {"traces": s.traces, "state": {m: s.state(m) for m in s.media}, "energy": s.energy,
           "variant": variant_of(s)}
    s.close()
    return out


def models():
    for order in (4, 5, 6):
        yield f"layered-o{order}", lambda o=order: layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), order=o,
                                                                   duration=0.03, n_receivers=7, energy_every=10)
        for m in (0, 1):
            yield f"borehole-m{m}-o{order}", lambda m=m, o=order: small_borehole(m, order=o)


@pytest.mark.parametrize("name,make", list(models()), ids=[n for n, _ in models()])
def test_every_force_variant_gives_the_same_bits(name, make):
    base = run(make, None)
    assert base["variant"][1] == 1, "the model must have a paired force launch"
    for v in range(VARIANTS):
        r = run(make, v)
        assert r["variant"] == (v, 1)
        for k, t in base["traces"].items():
            assert np.array_equal(t, r["traces"][k]), (name, v, k)
        for m, fields in base["state"].items():
            for x, y in zip(fields, r["state"][m]):
                assert np.array_equal(x, y), (name, v, m)
        assert np.array_equal(base["energy"], r["energy"]), (name, v)


def test_out_of_range_variant_is_refused():
    sim = small_borehole(0)
    s = sim.solver
    for bad in (-2, VARIANTS):
        rc = _lib.hip().insem_hip_set_force_variant(ctypes.c_void_p(s.context), bad)
        assert rc != 0
        assert b"force variant" in _lib.hip().insem_hip_last_error()
    s.close()


@pytest.mark.parametrize("name,make", [("layered-o4", lambda: layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8),
                                                                              duration=0.03, n_receivers=7,
                                                                              energy_every=10)),
                                       ("borehole-m1-o5", lambda: small_borehole(1, order=5))])
def test_chunk_length_changes_no_bit(name, make):
    def go(chunk):
        sim = make()
        s = sim.solver
        ctx = ctypes.c_void_p(s.context)
        if chunk:
            _lib.check_hip(_lib.hip().insem_hip_set_chunk(ctx, chunk))
        k = ctypes.c_int32()
        _lib.check_hip(_lib.hip().insem_hip_chunk(ctx, ctypes.byref(k)))
        assert chunk == 0 or k.value == chunk
        s.advance(45)
        s.advance(s.n_steps)
        out = (s.traces, {m: s.state(m) for m in s.media}, s.energy)
        s.close()
        return out
    base = go(0)
    for chunk in (1, 7, 128):
        t, st, e = go(chunk)
        for k, v in base[0].items():
            assert np.array_equal(v, t[k]), (name, chunk, k)
        for m, fields in base[1].items():
            for x, y in zip(fields, st[m]):
                assert np.array_equal(x, y), (name, chunk, m)
        assert np.array_equal(base[2], e), (name, chunk)
