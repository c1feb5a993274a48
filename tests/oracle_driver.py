"""Test infrastructure: drives the CPU oracle (oracle/liboracle.so) on a setup
image exactly as the product Solver drives the HIP backend -- same start
vectors, same dt / n_steps logic (reference: cpp/src/solver/solver.cpp:131-144),
same source amplitude tables -- so the two runs can be compared call for call.
"""

from __future__ import annotations

import ctypes
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "in-house-sem_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
# the same restatement computing in double (the reference's INSEM_REAL=double build)
ORACLE64_PATH = os.path.join(ROOT, "oracle", "liboracle64.so")

_libs = {}
P = ctypes.POINTER
F32 = P(ctypes.c_float)
F64 = P(ctypes.c_double)
I32 = P(ctypes.c_int32)


def lib(double=False):
    path = ORACLE64_PATH if double else ORACLE_PATH
    if path not in _libs:
        L = ctypes.CDLL(path)
        sig = {
            "oracle_last_error": (ctypes.c_char_p, []),
            "oracle_create": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_void_p)]),
            "oracle_destroy": (ctypes.c_int, [ctypes.c_void_p]),
            "oracle_set_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, F32]),
            "oracle_lambda_max": (ctypes.c_int, [ctypes.c_void_p, F32, F64]),
            "oracle_set_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, F32, F32]),
            "oracle_get_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, F32]),
            "oracle_advance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
            "oracle_step": (ctypes.c_int, [ctypes.c_void_p, I32]),
            "oracle_n_samples": (ctypes.c_int, [ctypes.c_void_p, I32]),
            "oracle_read_traces": (ctypes.c_int, [ctypes.c_void_p, F64]),
            "oracle_n_energy": (ctypes.c_int, [ctypes.c_void_p, I32]),
            "oracle_read_energy": (ctypes.c_int, [ctypes.c_void_p, F64]),
            "oracle_internal_forces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, F32, F32, F64]),
            "oracle_block_forces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, F32, F32]),
            "oracle_set_threads": (ctypes.c_int, [ctypes.c_int32]),
            "oracle_gather": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, F32, F32]),
            "oracle_assemble": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, F32, F32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[path] = L
    return _libs[path]


def _ok(rc, L=None):
    if rc != 0:
        raise RuntimeError((L or lib()).oracle_last_error().decode())


def f32(a):
    return a.ctypes.data_as(F32)


def f64(a):
    return a.ctypes.data_as(F64)


class OracleRun:
    """The oracle's time loop on one setup image (the image must outlive it);
    double=True runs the double-arithmetic build."""

    def __init__(self, image, dt=None, lambda_max=None, threads=0, double=False):
        self.image = image
        self.L = L = lib(double)
        if threads:
            L.oracle_set_threads(threads)
        h = ctypes.c_void_p()
        _ok(L.oracle_create(ctypes.c_void_p(image.view), ctypes.byref(h)), L)
        self.h = h
        t = image.time
        if dt is None:
            if lambda_max is None:
                lambda_max = self.lambda_max()
            self.stable_dt = 2.0 / math.sqrt(lambda_max)
            dt = t.dt if t.dt > 0 else t.safety * self.stable_dt
        self.dt = float(dt)
        if t.n_steps > 0:
            self.n_steps = t.n_steps
        elif t.duration > 0:
            self.n_steps = int(math.ceil(t.duration / self.dt - 1e-9))
        else:
            self.n_steps = 0
        amps = image.amplitudes(self.dt, self.n_steps)
        _ok(self.L.oracle_set_time(self.h, self.dt, self.n_steps, f32(amps)), self.L)

    def lambda_max(self):
        start = self.image.start_vectors()
        out = ctypes.c_double()
        _ok(self.L.oracle_lambda_max(self.h, f32(start), ctypes.byref(out)), self.L)
        return out.value

    def set_state(self, di, u, v):
        u = np.ascontiguousarray(u, dtype=np.float32)
        v = np.ascontiguousarray(v, dtype=np.float32)
        _ok(self.L.oracle_set_state(self.h, di, f32(u), f32(v)), self.L)

    def field(self, di, which):
        d = self.image.domains[di]
        out = np.zeros((d["ncomp"], d["n_global"]), dtype=np.float32)
        _ok(self.L.oracle_get_field(self.h, di, which, f32(out)), self.L)
        return out

    def advance(self, n):
        _ok(self.L.oracle_advance(self.h, int(n)), self.L)

    def run(self):
        s = ctypes.c_int32()
        _ok(self.L.oracle_step(self.h, ctypes.byref(s)), self.L)
        self.advance(self.n_steps - s.value)

    def traces(self):
        ns = ctypes.c_int32()
        _ok(self.L.oracle_n_samples(self.h, ctypes.byref(ns)), self.L)
        n = ns.value
        total = sum(nc for _, _, nc in self.image.traces) * n
        buf = np.zeros(max(total, 1))
        _ok(self.L.oracle_read_traces(self.h, f64(buf)), self.L)
        out, o = {}, 0
        for name, q, nc in self.image.traces:
            out[(name, q)] = buf[o:o + nc * n].reshape(n, nc).copy()
            o += nc * n
        return out

    def energy(self):
        ne = ctypes.c_int32()
        _ok(self.L.oracle_n_energy(self.h, ctypes.byref(ne)), self.L)
        buf = np.zeros((max(ne.value, 1), 3))
        _ok(self.L.oracle_read_energy(self.h, f64(buf)), self.L)
        return buf[:ne.value]

    def internal_forces(self, di, u, with_dot=True):
        d = self.image.domains[di]
        u = np.ascontiguousarray(u, dtype=np.float32)
        f = np.zeros((d["ncomp"], d["n_global"]), dtype=np.float32)
        dot = ctypes.c_double()
        _ok(self.L.oracle_internal_forces(self.h, di, f32(u), f32(f), ctypes.byref(dot) if with_dot else None), self.L)
        return f, dot.value

    def block_forces(self, di, bi, ul, fl):
        ul = np.ascontiguousarray(ul, dtype=np.float32)
        fl = np.ascontiguousarray(fl, dtype=np.float32).copy()
        _ok(self.L.oracle_block_forces(self.h, di, bi, f32(ul), f32(fl)), self.L)
        return fl

    def gather(self, di, g):
        d = self.image.domains[di]
        g = np.ascontiguousarray(g, dtype=np.float32)
        out = np.zeros(d["n_elements"] * d["ncomp"] * self.image.ngll ** 2, dtype=np.float32)
        _ok(self.L.oracle_gather(self.h, di, f32(g), f32(out)), self.L)
        return out

    def assemble(self, di, loc):
        d = self.image.domains[di]
        loc = np.ascontiguousarray(loc, dtype=np.float32)
        out = np.zeros((d["ncomp"], d["n_global"]), dtype=np.float32)
        _ok(self.L.oracle_assemble(self.h, di, f32(loc), f32(out)), self.L)
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
