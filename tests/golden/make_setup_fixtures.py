"""Writes tests/golden/setup_digests.json: SHA-256 digests (with shape, dtype,
the first values and, for floating arrays, a position-sensitive
fingerprint) of every array the host setup produces for a set of
small models covering each path of the setup -- mesh templates, forest
refinement and 2:1 balance, DofMap numbering and hanging-node rows, the
transpose CSR, geometry kinds and block metrics, mass, Stacey, Dirichlet and
axis points, fluid-solid coupling, sources of every kind, receivers -- plus
the quadrature rules and derivative matrices.

tests/test_setup_golden.py recomputes them: the host setup must reproduce
them bit for bit.  Run from the repository root after building:
    python tests/golden/make_setup_fixtures.py
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "in-house-sem_amd"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import insem_amd as insem  # noqa: E402
from insem_amd import configs, mesh  # noqa: E402
from setup_arrays import setup_arrays  # noqa: E402

OUT = os.path.join(HERE, "setup_digests.json")
FIXED_DT = 1.234e-6  # source amplitude tables are digested at this dt, 40 steps
N_SPOTS = 32


def digest(a):
    """SHA-256 of the bytes, and for floating arrays a position-sensitive
    fingerprint (sums, a cosine-weighted sum, extremes and N_SPOTS evenly
    spaced values) that a comparison within a few ulps can use."""
    a = np.ascontiguousarray(a)
    out = {"sha256": hashlib.sha256(a.tobytes()).hexdigest(), "shape": list(a.shape), "dtype": str(a.dtype),
           "head": [float(x) for x in a.ravel()[:6]]}
    if a.dtype.kind == "f" and a.size:
        x = a.ravel().astype(np.float64)
        c = np.cos(0.1 * np.arange(x.size))
        idx = np.unique(np.linspace(0, x.size - 1, N_SPOTS).astype(np.int64))
        out["fp"] = {"sum": float(x.sum()), "sum2": float((x * x).sum()), "wsum": float((c * x).sum()),
                     "absmax": float(np.abs(x).max()), "min": float(x.min()), "max": float(x.max()),
                     "spot_idx": [int(i) for i in idx], "spots": [float(x[i]) for i in idx]}
    return out


def _borehole(azimuthal_order, order=4):
    bh = mesh.borehole((-0.6, 0.6), [("mud", 0.1)], 0.5, dz=0.1)
    disc = insem.Discretization(order=order, f_max=1.2e4, points_per_wavelength=5.0, azimuthal_order=azimuthal_order)
    sim = insem.Simulation(bh.mesh, disc)
    sim.set_material(bh.regions["mud"], insem.Material.acoustic(1500.0, 1000.0))
    sim.set_material(bh.regions["formation"], insem.Material.elastic(4000.0, 2300.0, 2500.0))
    for name in ("bottom", "outer", "top"):
        sim.set_boundary(bh.boundaries[name], "absorbing")
    if azimuthal_order == 0:
        sim.add_source(insem.Source.monopole(0.0, 0.0, insem.ricker(8000.0)))
        sim.add_source(insem.Source.moment_tensor(0.23, 0.1, 1.0, 2.0, 0.5, insem.ricker(7000.0)))
        sim.add_source(insem.Source.force(0.25, -0.2, 0.3, 1.0, insem.ricker(6000.0)))
    else:
        sim.add_source(insem.Source.dipole(0.0, 0.0, 1.0, 0.0, insem.ricker(4000.0)))
        sim.add_source(insem.Source.force(0.3, 0.05, 0.4, -0.7, insem.ricker(5000.0)))
    for k, z in enumerate((0.1, 0.25, 0.4)):
        sim.add_receiver(f"p{k}", 0.05, z, ["pressure", "particle_velocity"])
        sim.add_receiver(f"s{k}", 0.3, z, ["displacement", "velocity", "acceleration"])
    sim.add_receiver("axis", 0.0, 0.33, ["pressure"])
    sim.set_time(duration=4e-4, energy_every=25)
    return sim


def _refined_grid():
    disc = insem.Discretization(order=5, uniform_level=1, regions=((0.3, 0.6, 0.4, 0.7, 3),))
    sim = insem.Simulation(configs.grid(3, 3, shear=0.2), disc)
    sim.set_material(0, insem.Material.elastic(2.0, 1.1, 1.3))
    for tag in (1, 2):
        sim.set_boundary(tag, "absorbing")
    sim.set_boundary(3, "rigid")
    sim.add_source(insem.Source.moment_tensor(0.47, 0.52, 1.0, 0.5, 0.2, insem.ricker(8.0)))
    sim.add_receiver("a", 0.55, 0.61, ["displacement"])
    sim.set_time(n_steps=10)
    return sim


def _annulus():
    sim = insem.Simulation(configs.quarter_annulus(3, 4), insem.Discretization(order=6))
    sim.set_material(0, insem.Material.acoustic(1.2, 0.9))
    for tag in (1, 2, 3, 4):
        sim.set_boundary(tag, "absorbing")
    sim.add_source(insem.Source.dipole(0.9, 0.9, 0.3, 0.4, insem.ricker(3.0)))
    sim.add_receiver("r", 1.1, 0.7, ["pressure", "particle_velocity"])
    sim.set_time(n_steps=10)
    return sim


def _small_cased():
    return configs.cased_hole(np.random.default_rng(11), duration=2.5e-4, z=(-0.4, 0.8), formation_radius=0.6,
                              dz=0.08, order=5, f_max=8e3, n_receivers=3, receiver_z0=0.2)


def _small_open():
    return configs.open_hole(np.random.default_rng(31), duration=2.5e-4, z=(-0.4, 0.8), formation_radius=0.6,
                             dz=0.08, order=4, f_max=8e3, n_receivers=3, receiver_z0=0.2)


MODELS = {
    "acoustic_box": lambda: configs.acoustic_box(),
    "layered_40x20": lambda: configs.layered_coupled(nx=40, nz_fluid=4, nz_solid=(8, 8), duration=0.05,
                                                     n_receivers=7, width=100.0),
    "borehole_m0": lambda: _borehole(0),
    "borehole_m1": lambda: _borehole(1),
    "borehole_m1_order6": lambda: _borehole(1, order=6),
    "refined_sheared_grid": _refined_grid,
    "quarter_annulus": _annulus,
    "cased_hole_small": _small_cased,
    "open_hole_small": _small_open,
}


def model_digests(sim):
    img = sim.build_setup()
    out = {k: digest(v) for k, v in setup_arrays(img.view).items()}
    for di in range(img.n_domains):
        out[f"d{di}.coords"] = digest(img.coordinates(di))
    out["start_vectors"] = digest(img.start_vectors())
    out["amplitudes"] = digest(img.amplitudes(FIXED_DT, 40))
    out["traces"] = {"names": [list(t) for t in img.traces]}
    return out


def mesh_digests():
    """The mesh templates' macro meshes."""
    zs = lambda x: -600.0 + 20.0 * np.sin(2.0 * np.pi * x / 500.0)  # noqa: E731
    cases = {
        "rectangle": mesh.rectangle((-4.0, 4.0), (-3.0, 2.0), nx=5, nz=4),
        "rectangle_axi": mesh.rectangle((0.0, 2.0), (-1.0, 1.0), nx=3, nz=4, coords="axisymmetric"),
        "layered": mesh.layered((0.0, 2000.0), [-1000.0, zs, -200.0, 0.0], [6, 5, 3], [2, 1, 0], nx=9),
        "layered_quadratic": mesh.layered((-5.0, 5.0), [-5.0, lambda x: 0.6 + 0.4 * np.sin(1.3 * x), 5.0], nz=6,
                                          regions=[0, 1], nx=8, quadratic=True),
        "borehole_open": mesh.borehole((-1.5, 4.5), [("mud", 0.1)], 4.0, beds=[(-1.5, "formation")], dz=0.08).mesh,
        "borehole_cased": mesh.borehole((-0.4, 0.8), [("mud", 0.1), ("casing", 0.11), ("cement", 0.14)], 0.6,
                                        beds=[(-0.4, "bed0"), (0.3, "bed1")], dz=0.08).mesh,
    }
    out = {}
    for name, m in cases.items():
        for field in ("nodes", "cells", "material", "boundary"):
            out[f"{name}.{field}"] = digest(np.asarray(getattr(m, field)))
        out[f"{name}.kind"] = {"cell_type": m.cell_type, "coords": m.coords}
    return out


def basis_digests():
    out = {}
    for n in range(2, 11):
        for a, b in ((0, 0), (0, 1)):
            x, w = insem.gauss_lobatto_jacobi(n, a, b)
            out[f"rule{n}.{a}{b}.x"] = digest(x)
            out[f"rule{n}.{a}{b}.w"] = digest(w)
    return out


def compute():
    return {"models": {k: model_digests(f()) for k, f in MODELS.items()}, "meshes": mesh_digests(),
            "basis": basis_digests()}


if __name__ == "__main__":
    data = compute()
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=0, sort_keys=True)
    print(f"wrote {OUT}: {sum(len(v) for v in data['models'].values())} model arrays, "
          f"{len(data['meshes'])} mesh arrays, {len(data['basis'])} basis arrays")
