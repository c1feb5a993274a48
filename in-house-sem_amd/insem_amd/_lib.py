"""Loads the native libraries of the package (built in-tree by `make`).

libinsem_host.so holds the host setup and the Solver facade; it links
libinsem_hip.so, the gfx950 kernels behind the C-ABI of include/insem_hip.h.
There is no fallback: a missing library is an error.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_PATH = os.path.join(HERE, "libinsem_host.so")
HIP_PATH = os.path.join(HERE, "libinsem_hip.so")

c_int = ctypes.c_int32
c_i64 = ctypes.c_int64
c_double = ctypes.c_double
c_void = ctypes.c_void_p
c_char = ctypes.c_char_p
P = ctypes.POINTER


class InsemError(RuntimeError):
    """Error raised by the simulator (reference: insem.InsemError)."""


_host = None
_hip = None


def _declare(lib, name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


def hip():
    """The C-ABI library (include/insem_hip.h)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_PATH):
            raise ImportError(f"{HIP_PATH} is missing: build the package with `make -C in-house-sem_amd`")
        lib = ctypes.CDLL(HIP_PATH)
        _declare(lib, "insem_hip_last_error", c_char)
        _declare(lib, "insem_hip_device_count", c_int)
        _declare(lib, "insem_hip_get_device", c_int, P(c_int))
        _declare(lib, "insem_hip_set_device", c_int, c_int)
        _declare(lib, "insem_hip_create", c_int, c_void, P(c_void))
        _declare(lib, "insem_hip_destroy", c_int, c_void)
        _declare(lib, "insem_hip_stage_size", c_int, c_void, P(c_i64))
        _declare(lib, "insem_hip_reserve", c_int, c_i64, P(c_void))
        _declare(lib, "insem_hip_unreserve", c_int, c_void)
        _declare(lib, "insem_hip_stage", c_int, c_void, c_void, P(c_void))
        _declare(lib, "insem_hip_realize", c_int, c_void)
        _declare(lib, "insem_hip_set_time", c_int, c_void, c_double, c_int, P(ctypes.c_float))
        _declare(lib, "insem_hip_lambda_max", c_int, c_void, P(ctypes.c_float), P(c_double))
        _declare(lib, "insem_hip_lambda_start", c_int, c_void, P(ctypes.c_float))
        _declare(lib, "insem_hip_lambda_result", c_int, c_void, P(c_double))
        _declare(lib, "insem_hip_set_state", c_int, c_void, c_int, P(ctypes.c_float), P(ctypes.c_float))
        _declare(lib, "insem_hip_get_field", c_int, c_void, c_int, c_int, P(ctypes.c_float))
        _declare(lib, "insem_hip_advance", c_int, c_void, c_int)
        _declare(lib, "insem_hip_prepare", c_int, c_void, c_int)
        _declare(lib, "insem_hip_set_force_variant", c_int, c_void, c_int)
        _declare(lib, "insem_hip_force_variant", c_int, c_void, P(c_int), P(c_int))
        _declare(lib, "insem_hip_set_chunk", c_int, c_void, c_int)
        _declare(lib, "insem_hip_progress", c_int, c_void, P(c_int))
        _declare(lib, "insem_hip_chunk", c_int, c_void, P(c_int))
        _declare(lib, "insem_hip_step", c_int, c_void, P(c_int))
        _declare(lib, "insem_hip_n_samples", c_int, c_void, P(c_int))
        _declare(lib, "insem_hip_read_traces", c_int, c_void, P(c_double))
        _declare(lib, "insem_hip_n_energy", c_int, c_void, P(c_int))
        _declare(lib, "insem_hip_read_energy", c_int, c_void, P(c_double))
        _declare(lib, "insem_hip_internal_forces", c_int, c_void, c_int, P(ctypes.c_float), P(ctypes.c_float), P(c_double))
        _declare(lib, "insem_hip_block_forces", c_int, c_void, c_int, c_int, P(ctypes.c_float), P(ctypes.c_float))
        _declare(lib, "insem_hip_block_forces_noalloc", c_int, c_void, c_int, c_int, P(ctypes.c_float),
                 P(ctypes.c_float))
        _declare(lib, "insem_hip_synchronize", c_int, c_void)
        _declare(lib, "insem_hip_query", c_int, c_void, P(c_int))
        _declare(lib, "insem_hip_device_memory", c_int, P(c_i64), P(c_i64))
        _declare(lib, "insem_hip_release_cache", c_int, P(c_i64))
        _declare(lib, "insem_hip_set_timing", c_int, c_void, c_int)
        _declare(lib, "insem_hip_kernel_time", c_int, c_void, P(c_double), P(c_int), P(c_double), P(c_int))
        _declare(lib, "insem_hip_block_time", c_int, c_void, c_int, c_int, P(c_double), P(c_int))
        _declare(lib, "insem_hip_block_points", c_int, c_void, c_int, c_int, P(c_i64), P(c_i64))
        _declare(lib, "insem_hip_block_local_slots", c_int, c_void, c_int, c_int, P(c_i64))
        _declare(lib, "insem_hip_footprint", c_int, c_void, P(c_i64))
        _declare(lib, "insem_hip_create_batch", c_int, c_void, c_void, P(c_void))
        _declare(lib, "insem_hip_set_time_batch", c_int, c_void, P(c_double), P(c_int), P(ctypes.c_float))
        _hip = lib
    return _hip


def host():
    """The host library (setup and Solver facade)."""
    global _host
    if _host is None:
        hip()
        if not os.path.exists(HOST_PATH):
            raise ImportError(f"{HOST_PATH} is missing: build the package with `make -C in-house-sem_amd`")
        lib = ctypes.CDLL(HOST_PATH)
        d = lambda n, r, *a: _declare(lib, n, r, *a)  # noqa: E731
        d("insem_host_last_error", c_char)
        d("insem_basis_rule", c_int, c_int, c_int, c_int, P(c_double), P(c_double))
        d("insem_basis_derivative_matrix", c_int, c_int, P(c_double), P(c_double))
        d("insem_mesh_new", c_void, c_int, c_int, P(c_double), c_int, P(c_int), c_int, P(c_int), P(c_int), c_int)
        d("insem_mesh_free", None, c_void)
        d("insem_mesh_info", c_int, c_void, P(c_int))
        d("insem_mesh_arrays", c_int, c_void, P(c_int), P(c_int))
        d("insem_mesh_map", c_int, c_void, c_int, c_double, c_double, P(c_double))
        d("insem_mesh_is_axis_face", c_int, c_void, c_int, c_int)
        d("insem_model_new", c_void, c_void, c_int, c_int, c_double, c_double, c_int, P(c_double), c_int, c_int)
        d("insem_model_free", None, c_void)
        d("insem_model_set_material", c_int, c_void, c_int, c_int, c_double, c_double, c_double)
        d("insem_model_set_boundary", c_int, c_void, c_int, c_int)
        d("insem_model_add_source", c_int, c_void, c_int, c_double, c_double, c_double, c_double, c_double, c_int,
          c_double, c_double, c_double, P(c_double), c_int, c_double)
        d("insem_model_clear_sources", c_int, c_void)
        d("insem_model_add_receiver", c_int, c_void, c_char, c_double, c_double, c_char, c_char)
        d("insem_model_clear_receivers", c_int, c_void)
        d("insem_model_set_time", c_int, c_void, c_double, c_double, c_int, c_double, c_int, c_int)
        d("insem_model_setup", c_int, c_void)
        d("insem_model_info", c_int, c_void, P(c_int))
        d("insem_model_leaf_points", c_int, c_void, P(c_double), P(c_double))
        d("insem_model_build", c_void, c_void)
        d("insem_image_free", None, c_void)
        d("insem_image_view", c_void, c_void)
        d("insem_image_domain_info", c_int, c_void, c_int, P(c_int))
        d("insem_image_n_domains", c_int, c_void)
        d("insem_image_block_info", c_int, c_void, c_int, c_int, P(c_int))
        d("insem_image_coords", c_int, c_void, c_int, P(c_double))
        d("insem_image_mass", c_int, c_void, c_int, P(c_double))
        d("insem_image_start_len", c_i64, c_void)
        d("insem_image_start_vectors", c_int, c_void, P(ctypes.c_float))
        d("insem_image_amplitudes", c_int, c_void, c_double, c_int, P(ctypes.c_float))
        d("insem_image_time", c_int, c_void, P(c_double), P(c_int))
        d("insem_image_model_time", c_int, c_void, c_double, P(c_double), P(c_int))
        d("insem_image_n_traces", c_int, c_void)
        d("insem_image_trace_info", c_int, c_void, c_int, c_char, c_int, c_char, c_int, P(c_int))
        d("insem_image_counts", c_int, c_void, P(c_i64))
        d("insem_solver_new", c_void, c_void)
        d("insem_solver_free", None, c_void)
        d("insem_solver_image", c_void, c_void)
        d("insem_solver_context", c_void, c_void)
        d("insem_solver_info", c_int, c_void, P(c_double), P(c_int))
        d("insem_solver_advance", c_int, c_void, c_int)
        d("insem_solver_set_state", c_int, c_void, c_int, P(c_double), P(c_double), c_i64)
        d("insem_solver_field", c_int, c_void, c_int, c_int, P(c_double))
        d("insem_solver_n_samples", c_int, c_void)
        d("insem_solver_traces", c_int, c_void, P(c_double))
        d("insem_solver_n_energy", c_int, c_void)
        d("insem_solver_energy", c_int, c_void, P(c_double))
        d("insem_batch_new", c_void, P(c_void), c_int)
        d("insem_batch_new_lambdas", c_void, P(c_void), P(c_double), c_int)
        d("insem_batch_free", None, c_void)
        d("insem_power_start", c_void, c_void)
        d("insem_image_stage_size", c_int, c_void, P(c_i64))
        d("insem_power_stage", c_void, c_void, c_void)
        d("insem_power_start_staged", c_void, c_void)
        d("insem_power_staged_free", None, c_void)
        d("insem_power_ready", c_int, c_void, P(c_int))
        d("insem_power_lambda", c_int, c_void, P(c_double))
        d("insem_power_free", None, c_void)
        d("insem_power_solver", c_void, c_void, c_void)
        d("insem_batch_context", c_void, c_void)
        d("insem_batch_model_image", c_void, c_void, c_int)
        d("insem_batch_info", c_int, c_void, c_int, P(c_double), P(c_int))
        d("insem_batch_advance", c_int, c_void, c_int)
        d("insem_batch_traces", c_int, c_void, c_int, P(c_double))
        d("insem_batch_traces_done", c_int, c_void, c_int, P(c_double))
        d("insem_batch_device_step", c_int, c_void, P(c_int))
        d("insem_batch_footprint", c_i64, c_void)
        _host = lib
    return _host


def check(rc: int) -> None:
    if rc != 0:
        raise InsemError(host().insem_host_last_error().decode())


def check_count(n: int) -> int:
    """A count or flag from the host library (negative on failure)."""
    if n < 0:
        raise InsemError(host().insem_host_last_error().decode())
    return n


def check_ptr(p):
    if not p:
        raise InsemError(host().insem_host_last_error().decode())
    return p


def check_hip(rc: int) -> None:
    if rc != 0:
        raise InsemError(hip().insem_hip_last_error().decode())
