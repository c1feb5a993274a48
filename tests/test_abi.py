"""The C-ABI library loads on a host without a GPU and exports every entry
point include/insem_hip.h declares; host-side validation and the host
setup's error messages (reference: the INSEM_CHECK sites of
cpp/src/solver/solver.cpp:98-104, 281-283, 431-436, 533-535 and
description.cpp:286-295) behave without touching a device."""

from __future__ import annotations

import ctypes
import os
import re

import pytest

import insem_amd as insem
from insem_amd import _lib
from models import grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "insem_hip.h")).read()
    return sorted(set(re.findall(r"\b(insem_hip_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.HIP_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_device_count_never_fails():
    assert _lib.hip().insem_hip_device_count() >= 0


def test_device_selection_rejects_missing_devices():
    hip = _lib.hip()
    n = hip.insem_hip_device_count()
    assert hip.insem_hip_set_device(-1) == -1
    assert hip.insem_hip_set_device(n) == -1
    assert b"no device" in hip.insem_hip_last_error()
    assert hip.insem_hip_get_device(None) == -1


def test_create_rejects_unsupported_order_before_touching_a_device():
    class Setup(ctypes.Structure):
        _fields_ = [("ngll", ctypes.c_int32)]

    s = Setup(12)
    ctx = ctypes.c_void_p()
    rc = _lib.hip().insem_hip_create(ctypes.byref(s), ctypes.byref(ctx))
    assert rc == -1
    assert b"GLL points" in _lib.hip().insem_hip_last_error()
    assert _lib.hip().insem_hip_create(None, ctypes.byref(ctx)) == -1


def test_null_context_is_an_error():
    assert _lib.hip().insem_hip_advance(None, 1) == -1
    assert b"null context" in _lib.hip().insem_hip_last_error()


def test_invalid_materials_are_rejected():
    sim = insem.Simulation(grid(2, 2), insem.Discretization(order=3))
    with pytest.raises(insem.InsemError, match="must be below the P velocity"):
        sim.set_material(0, insem.Material.elastic(1.0, 2.0, 1.0))
    with pytest.raises(insem.InsemError, match="positive density"):
        sim.set_material(0, insem.Material.acoustic(1.0, -1.0))


def test_sources_and_receivers_outside_their_medium_fail_clearly():
    sim = insem.Simulation(grid(2, 2), insem.Discretization(order=3))
    sim.set_material(0, insem.Material.elastic(2.0, 1.0, 1.0))
    sim.set_time(n_steps=1)
    sim.add_source(insem.Source.monopole(0.5, 0.5, insem.ricker(5.0)))
    with pytest.raises(insem.InsemError, match="acoustic"):
        sim.build_setup()
    sim.clear_sources()
    sim.add_receiver("r1", 3.0, 0.5, ["velocity"])
    with pytest.raises(insem.InsemError, match="outside"):
        sim.build_setup()
    sim.clear_receivers()
    sim.add_receiver("r1", 0.5, 0.5, ["pressure"])
    with pytest.raises(insem.InsemError, match="is not recorded in elastic media"):
        sim.build_setup()


def test_axis_takes_no_boundary_condition():
    sim = insem.Simulation(grid(2, 2, 0.0, 1.0, -1.0, 1.0, coords="axisymmetric"), insem.Discretization(order=3))
    sim.set_material(0, insem.Material.acoustic(1.5, 1.0))
    sim.set_boundary(4, "absorbing")  # left face = the axis
    with pytest.raises(insem.InsemError, match="symmetry axis"):
        sim.build_setup()


def test_non_axisymmetric_sources_on_the_axis_are_rejected():
    m = grid(2, 2, 0.0, 1.0, -1.0, 1.0, coords="axisymmetric")
    sim = insem.Simulation(m, insem.Discretization(order=3))
    sim.set_material(0, insem.Material.acoustic(1.5, 1.0))
    sim.add_source(insem.Source.dipole(0.0, 0.0, 1.0, 0.0, insem.ricker(5.0)))
    with pytest.raises(insem.InsemError, match="azimuthal order 1, not 0"):
        sim.build_setup()
    sim1 = insem.Simulation(m, insem.Discretization(order=3, azimuthal_order=1))
    sim1.set_material(0, insem.Material.acoustic(1.5, 1.0))
    sim1.add_source(insem.Source.monopole(0.0, 0.0, insem.ricker(5.0)))
    with pytest.raises(insem.InsemError, match="not a force or a dipole"):
        sim1.build_setup()


def test_azimuthal_order_needs_axisymmetric_mesh():
    with pytest.raises(insem.InsemError, match="axisymmetric mesh"):
        insem.Simulation(grid(2, 2), insem.Discretization(order=3, azimuthal_order=1))


def test_unsupported_polynomial_order_is_rejected():
    sim = insem.Simulation(grid(1, 1), insem.Discretization(order=9))
    sim.set_material(0, insem.Material.acoustic(1.5, 1.0))
    with pytest.raises(insem.InsemError, match="polynomial orders 1 to 8"):
        sim.build_setup()
