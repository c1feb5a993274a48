"""Reads every array of an insem_hip_setup (include/insem_hip.h) through
ctypes, for the golden fixtures of the host setup: a flat dict of numpy
arrays named by their place in the struct."""

from __future__ import annotations

import ctypes

import numpy as np

P = ctypes.POINTER
i32, f32, f64, u8 = ctypes.c_int32, ctypes.c_float, ctypes.c_double, ctypes.c_uint8


class Block(ctypes.Structure):
    _fields_ = [("kind", i32), ("azimuthal_order", i32), ("n_elements", i32), ("elements", P(i32)),
                ("metric", P(f32)), ("coef0", P(f32)), ("coef1", P(f32)), ("wr", P(f32)), ("hoop", P(f32)),
                ("r", P(f32)), ("axial", P(u8))]


class Domain(ctypes.Structure):
    _fields_ = [("medium", i32), ("ncomp", i32), ("azimuthal_order", i32), ("n_elements", i32), ("n_global", i32),
                ("l2g", P(i32)), ("n_rows", i32), ("row_offsets", P(i32)), ("row_dofs", P(i32)),
                ("row_weights", P(f64)), ("t_offsets", P(i32)), ("t_slots", P(i32)), ("t_weights", P(f64)),
                ("mass", P(f32)), ("inv_mass", P(f32)), ("n_absorbing", i32), ("absorbing_points", P(i32)),
                ("absorbing_coef", P(f32)), ("n_dirichlet", i32), ("dirichlet_points", P(i32)), ("n_axis", i32),
                ("axis_points", P(i32)), ("n_axis_components", i32), ("axis_components", P(i32)),
                ("n_blocks", i32), ("blocks", P(Block))]


class Coupling(ctypes.Structure):
    _fields_ = [("n_points", i32), ("fluid_points", P(i32)), ("solid_points", P(i32)), ("cx", P(f32)),
                ("cz", P(f32))]


class Source(ctypes.Structure):
    _fields_ = [("domain", i32), ("n_points", i32), ("points", P(i32)), ("coef", P(f32))]


class Receiver(ctypes.Structure):
    _fields_ = [("domain", i32), ("n_points", i32), ("points", P(i32)), ("value", P(f32)), ("d_dx", P(f32)),
                ("d_dz", P(f32)), ("rho", f64), ("n_quantities", i32), ("quantities", P(i32))]


class Setup(ctypes.Structure):
    _fields_ = [("ngll", i32), ("D", P(f32)), ("Dj", P(f32)), ("w2", P(f32)), ("n_domains", i32),
                ("domains", P(Domain)), ("coupling", Coupling), ("n_sources", i32), ("sources", P(Source)),
                ("n_receivers", i32), ("receivers", P(Receiver)), ("record_every", i32), ("energy_every", i32)]


def _arr(ptr, n, dtype):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def setup_arrays(view) -> dict:
    """Every array of the setup at address `view` (SetupImage.view)."""
    s = ctypes.cast(ctypes.c_void_p(view), P(Setup)).contents
    ng2 = s.ngll * s.ngll
    out = {"ngll": np.array([s.ngll, s.record_every, s.energy_every], dtype=np.int32),
           "D": _arr(s.D, ng2, np.float32), "Dj": _arr(s.Dj, ng2, np.float32), "w2": _arr(s.w2, ng2, np.float32)}
    for di in range(s.n_domains):
        d = s.domains[di]
        p = f"d{di}."
        nslots = d.n_elements * ng2
        nrows = d.n_rows
        out[p + "info"] = np.array([d.medium, d.ncomp, d.azimuthal_order, d.n_elements, d.n_global, nrows,
                                    d.n_absorbing, d.n_dirichlet, d.n_axis, d.n_blocks], dtype=np.int32)
        out[p + "l2g"] = _arr(d.l2g, nslots, np.int32)
        out[p + "row_offsets"] = _arr(d.row_offsets, nrows + 1 if nrows else 0, np.int32)
        nq = int(out[p + "row_offsets"][-1]) if nrows else 0
        out[p + "row_dofs"] = _arr(d.row_dofs, nq, np.int32)
        out[p + "row_weights"] = _arr(d.row_weights, nq, np.float64)
        out[p + "t_offsets"] = _arr(d.t_offsets, d.n_global + 1, np.int32)
        nt = int(out[p + "t_offsets"][-1])
        out[p + "t_slots"] = _arr(d.t_slots, nt, np.int32)
        out[p + "t_weights"] = _arr(d.t_weights, nt, np.float64)
        out[p + "mass"] = _arr(d.mass, d.n_global, np.float32)
        out[p + "inv_mass"] = _arr(d.inv_mass, d.n_global, np.float32)
        out[p + "absorbing_points"] = _arr(d.absorbing_points, d.n_absorbing, np.int32)
        out[p + "absorbing_coef"] = _arr(d.absorbing_coef, d.n_absorbing * d.ncomp * d.ncomp, np.float32)
        out[p + "dirichlet_points"] = _arr(d.dirichlet_points, d.n_dirichlet, np.int32)
        out[p + "axis_points"] = _arr(d.axis_points, d.n_axis, np.int32)
        out[p + "axis_components"] = _arr(d.axis_components, d.n_axis_components, np.int32)
        for bi in range(d.n_blocks):
            b = d.blocks[bi]
            q = f"{p}b{bi}."
            ne = b.n_elements
            out[q + "info"] = np.array([b.kind, b.azimuthal_order, ne], dtype=np.int32)
            out[q + "elements"] = _arr(b.elements, ne, np.int32)
            nm = {0: 3 * ne, 1: 5 * ne}.get(b.kind, 5 * ne * ng2)
            out[q + "metric"] = _arr(b.metric, nm, np.float32)
            out[q + "coef0"] = _arr(b.coef0, ne * ng2, np.float32)
            out[q + "coef1"] = _arr(b.coef1, ne * ng2, np.float32)
            if b.kind == 3:
                out[q + "wr"] = _arr(b.wr, ne * ng2, np.float32)
                out[q + "hoop"] = _arr(b.hoop, ne * ng2, np.float32)
                out[q + "r"] = _arr(b.r, ne * ng2, np.float32)
                out[q + "axial"] = _arr(b.axial, ne, np.uint8)
    c = s.coupling
    out["coupling.fluid_points"] = _arr(c.fluid_points, c.n_points, np.int32)
    out["coupling.solid_points"] = _arr(c.solid_points, c.n_points, np.int32)
    out["coupling.cx"] = _arr(c.cx, c.n_points, np.float32)
    out["coupling.cz"] = _arr(c.cz, c.n_points, np.float32)
    for si in range(s.n_sources):
        src = s.sources[si]
        nc = s.domains[src.domain].ncomp
        out[f"src{si}.points"] = _arr(src.points, src.n_points, np.int32)
        out[f"src{si}.coef"] = _arr(src.coef, nc * src.n_points, np.float32)
        out[f"src{si}.domain"] = np.array([src.domain], dtype=np.int32)
    for ri in range(s.n_receivers):
        r = s.receivers[ri]
        p = f"rec{ri}."
        out[p + "points"] = _arr(r.points, r.n_points, np.int32)
        out[p + "value"] = _arr(r.value, r.n_points, np.float32)
        out[p + "d_dx"] = _arr(r.d_dx, r.n_points, np.float32)
        out[p + "d_dz"] = _arr(r.d_dz, r.n_points, np.float32)
        out[p + "rho"] = np.array([r.rho])
        out[p + "quantities"] = _arr(r.quantities, r.n_quantities, np.int32)
        out[p + "domain"] = np.array([r.domain], dtype=np.int32)
    return out
