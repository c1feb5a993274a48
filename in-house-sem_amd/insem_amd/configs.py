"""Model builders: the reference's test fixtures (tests/cpp/test_helpers.hpp:16-76)
and the benchmark workloads."""

from __future__ import annotations

import numpy as np

from . import core as insem
from . import mesh


def grid(nx, nz, x0=0.0, x1=1.0, z0=0.0, z1=1.0, shear=0.0, coords="cartesian", material=None):
    """nx x nz Quad4 cells on [x0, x1] x [z0, z1], tags 1..4 bottom/right/top/left,
    x += shear * z (test_helpers.hpp:16-46)."""
    nodes = []
    for j in range(nz + 1):
        for i in range(nx + 1):
            z = z0 + (z1 - z0) * j / nz
            nodes.append((x0 + (x1 - x0) * i / nx + shear * z, z))
    cells, mat, bnd = [], [], []
    node = lambda i, j: i + (nx + 1) * j  # noqa: E731
    for j in range(nz):
        for i in range(nx):
            c = len(cells)
            cells.append((node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)))
            mat.append(0)
            if j == 0:
                bnd.append((c, 0, 1))
            if i == nx - 1:
                bnd.append((c, 1, 2))
            if j == nz - 1:
                bnd.append((c, 2, 3))
            if i == 0:
                bnd.append((c, 3, 4))
    if material is not None:
        mat = list(material)
    return insem.MacroMesh("quad4", coords, np.array(nodes), np.array(cells), np.array(mat), np.array(bnd))


def quarter_annulus(nr, nt, r0=1.0, r1=2.0):
    """Quad9 quarter annulus with nodes on circles (test_helpers.hpp:48-76)."""
    mr, mt = 2 * nr, 2 * nt
    nodes = []
    for j in range(mt + 1):
        for i in range(mr + 1):
            r = r0 + (r1 - r0) * i / mr
            t = 0.5 * np.pi * j / mt
            nodes.append((r * np.cos(t), r * np.sin(t)))
    node = lambda i, j: i + (mr + 1) * j  # noqa: E731
    cells = []
    for j in range(nt):
        for i in range(nr):
            a, b = 2 * i, 2 * j
            cells.append((node(a, b), node(a + 2, b), node(a + 2, b + 2), node(a, b + 2), node(a + 1, b),
                          node(a + 2, b + 1), node(a + 1, b + 2), node(a, b + 1), node(a + 1, b + 1)))
    return insem.MacroMesh("quad9", "cartesian", np.array(nodes), np.array(cells), np.zeros(len(cells)))


def acoustic_box(duration=4e-3):
    """2D acoustic box, 32x32 quads, order 4, Stacey on all sides, monopole
    Ricker 400 Hz at the centre, pressure receiver at (1.5, 0) (the reference's README.md:85-95)."""
    m = mesh.rectangle((-4.0, 4.0), (-4.0, 4.0), nx=32, nz=32)
    sim = insem.Simulation(m, insem.Discretization(order=4))
    sim.set_material(0, insem.Material.acoustic(vp=1500.0, rho=1000.0))
    for tag in (mesh.BOTTOM, mesh.RIGHT, mesh.TOP, mesh.LEFT):
        sim.set_boundary(tag, "absorbing")
    sim.add_source(insem.Source.monopole(0.0, 0.0, insem.ricker(400.0)))
    sim.add_receiver("h1", 1.5, 0.0, ["pressure"])
    sim.set_time(duration=duration)
    return sim


def layered_coupled(nx=800, nz_fluid=80, nz_solid=(160, 160), order=4, duration=1.0, n_receivers=101,
                    n_steps=0, energy_every=0, width=2000.0):
    """Coupled acoustic-elastic layered model: 2000 x 1000 m, fluid on top
    (200 m), two elastic layers separated by a sinusoidal interface
    (amplitude 20 m, wavelength 500 m); free top, absorbing elsewhere;
    explosion 50 m below the fluid base, 25 Hz Ricker; receivers recording
    pressure in the fluid and displacement in the solid."""
    zs = lambda x: -600.0 + 20.0 * np.sin(2.0 * np.pi * x / 500.0)  # noqa: E731
    m = mesh.layered((0.0, width), [-1000.0, zs, -200.0, 0.0], [nz_solid[0], nz_solid[1], nz_fluid], [2, 1, 0],
                     nx=nx)
    sim = insem.Simulation(m, insem.Discretization(order=order))
    sim.set_material(0, insem.Material.acoustic(1500.0, 1000.0))
    sim.set_material(1, insem.Material.elastic(2500.0, 1200.0, 2200.0))
    sim.set_material(2, insem.Material.elastic(4000.0, 2300.0, 2500.0))
    sim.set_boundary(mesh.TOP, "free")
    for tag in (mesh.BOTTOM, mesh.LEFT, mesh.RIGHT):
        sim.set_boundary(tag, "absorbing")
    x0 = 0.5 * width
    sim.add_source(insem.Source.moment_tensor(x0, -250.0, 1e6, 1e6, 0.0, insem.ricker(25.0)))
    xs = np.linspace(0.05 * width, 0.95 * width, n_receivers)
    for i, x in enumerate(xs):
        if i % 2 == 0:
            sim.add_receiver(f"p{i}", x, -100.0, ["pressure"])
        else:
            sim.add_receiver(f"s{i}", x, -400.0, ["displacement"])
    sim.set_time(duration=duration, n_steps=n_steps, energy_every=energy_every)
    return sim


def cased_hole(rng, duration=5e-3, z=(-1.5, 4.5), formation_radius=4.0, dz=0.08, order=None, f_max=None,
               n_receivers=8, receiver_z0=1.0):
    """Axisymmetric cased-hole dipole model (azimuthal order 1) drawn from
    `rng` (numpy Generator): borehole radius U[0.08, 0.12] m, steel casing
    8-12 mm (vp 5900, vs 3200, rho 7850), cement 20-40 mm (vp 2800, vs 1500,
    rho 1900), 1-4 horizontal formation beds with vp U[2000, 5000],
    vs = vp / U[1.6, 2.2], rho U[2000, 2700]; order in {4, 5, 6}; refinement
    to 6 points per wavelength at f_max U[5, 10] kHz; a dipole point force in
    the fluid on the axis with a U[2, 5] kHz Ricker; pressure receivers at
    r = 0.05 m from z = receiver_z0 up in 0.15 m steps."""
    rb = rng.uniform(0.08, 0.12)
    t_steel = rng.uniform(0.008, 0.012)
    t_cem = rng.uniform(0.020, 0.040)
    n_beds = int(rng.integers(1, 5))
    bed_z = np.sort(rng.uniform(z[0] + 0.5, z[1] - 0.5, size=n_beds - 1))
    beds = [(z[0], "bed0")] + [(float(b), f"bed{i + 1}") for i, b in enumerate(bed_z)]
    form = []
    for _ in range(n_beds):
        vp = rng.uniform(2000.0, 5000.0)
        form.append((vp, vp / rng.uniform(1.6, 2.2), rng.uniform(2000.0, 2700.0)))
    order = int(rng.choice([4, 5, 6])) if order is None else order
    f_max = rng.uniform(5e3, 1e4) if f_max is None else f_max
    f0 = rng.uniform(2e3, 5e3)
    bh = mesh.borehole(z, [("mud", rb), ("casing", rb + t_steel), ("cement", rb + t_steel + t_cem)],
                       formation_radius, beds=beds, dz=dz)
    disc = insem.Discretization(order=order, f_max=f_max, points_per_wavelength=6.0, azimuthal_order=1)
    sim = insem.Simulation(bh.mesh, disc)
    sim.set_material(bh.regions["mud"], insem.Material.acoustic(1500.0, 1000.0))
    sim.set_material(bh.regions["casing"], insem.Material.elastic(5900.0, 3200.0, 7850.0))
    sim.set_material(bh.regions["cement"], insem.Material.elastic(2800.0, 1500.0, 1900.0))
    for (_, name), (vp, vs, rho) in zip(beds, form):
        sim.set_material(bh.regions[name], insem.Material.elastic(vp, vs, rho))
    for name in ("bottom", "outer", "top"):
        sim.set_boundary(bh.boundaries[name], "absorbing")
    sim.add_source(insem.Source.dipole(0.0, 0.0, 1.0, 0.0, insem.ricker(f0)))
    for k in range(n_receivers):
        sim.add_receiver(f"r{k}", 0.05, receiver_z0 + 0.15 * k, ["pressure"])
    sim.set_time(duration=duration)
    return sim


def open_hole_monopole(duration=3e-3):
    """The axisymmetric open-hole monopole workload (azimuthal order 0): mud
    column of radius 0.1 m (vp 1500, rho 1000) in a fast formation (vp 4000,
    vs 2300, rho 2500) out to r = 4 m, z in [-1.5, 4.5] m with dz = 0.08 m,
    order 4 with GLJ elements on the axis, refined to 6 points per
    wavelength at 20 kHz (2:1 balanced); absorbing outer, top and bottom
    edges; an on-axis pressure monopole with an 8 kHz Ricker; 13 pressure
    receivers at r = 0.05 m, z = 1.0 ... 2.8 m in 0.15 m steps."""
    bh = mesh.borehole((-1.5, 4.5), [("mud", 0.1)], 4.0, beds=[(-1.5, "formation")], dz=0.08)
    sim = insem.Simulation(bh.mesh, insem.Discretization(order=4, f_max=20e3, points_per_wavelength=6.0))
    sim.set_material(bh.regions["mud"], insem.Material.acoustic(1500.0, 1000.0))
    sim.set_material(bh.regions["formation"], insem.Material.elastic(4000.0, 2300.0, 2500.0))
    for name in ("bottom", "outer", "top"):
        sim.set_boundary(bh.boundaries[name], "absorbing")
    sim.add_source(insem.Source.monopole(0.0, 0.0, insem.ricker(8e3)))
    for k in range(13):
        sim.add_receiver(f"r{k}", 0.05, 1.0 + 0.15 * k, ["pressure"])
    sim.set_time(duration=duration)
    return sim


def open_hole(rng, duration=3e-3, z=(-1.5, 4.5), formation_radius=4.0, dz=0.08, order=None, f_max=None,
              n_receivers=8, receiver_z0=1.0):
    """Axisymmetric open-hole monopole model (azimuthal order 0) drawn from
    `rng`: mud column of radius U[0.08, 0.12] m (vp 1500, rho 1000), 1-4
    formation beds with vp U[2000, 5000], vs = vp / U[1.6, 2.2], rho
    U[2000, 2700]; order in {4, 5, 6}; refinement to 6 points per
    wavelength at f_max U[5, 10] kHz; an on-axis pressure monopole with a
    Ricker of f_max / 2.5; pressure receivers at r = 0.05 m from z =
    receiver_z0 up in 0.15 m steps."""
    rb = rng.uniform(0.08, 0.12)
    n_beds = int(rng.integers(1, 5))
    bed_z = np.sort(rng.uniform(z[0] + 0.5, z[1] - 0.5, size=n_beds - 1))
    beds = [(z[0], "bed0")] + [(float(b), f"bed{i + 1}") for i, b in enumerate(bed_z)]
    form = []
    for _ in range(n_beds):
        vp = rng.uniform(2000.0, 5000.0)
        form.append((vp, vp / rng.uniform(1.6, 2.2), rng.uniform(2000.0, 2700.0)))
    order = int(rng.choice([4, 5, 6])) if order is None else order
    f_max = rng.uniform(5e3, 1e4) if f_max is None else f_max
    bh = mesh.borehole(z, [("mud", rb)], formation_radius, beds=beds, dz=dz)
    sim = insem.Simulation(bh.mesh, insem.Discretization(order=order, f_max=f_max, points_per_wavelength=6.0))
    sim.set_material(bh.regions["mud"], insem.Material.acoustic(1500.0, 1000.0))
    for (_, name), (vp, vs, rho) in zip(beds, form):
        sim.set_material(bh.regions[name], insem.Material.elastic(vp, vs, rho))
    for name in ("bottom", "outer", "top"):
        sim.set_boundary(bh.boundaries[name], "absorbing")
    sim.add_source(insem.Source.monopole(0.0, 0.0, insem.ricker(f_max / 2.5)))
    for k in range(n_receivers):
        sim.add_receiver(f"r{k}", 0.05, receiver_z0 + 0.15 * k, ["pressure"])
    sim.set_time(duration=duration)
    return sim


def served_request(kind, seed):
    """One request of the served workload: "cased" (cased-hole dipole, 3-8 ms),
    "open" (open-hole monopole, 3-8 ms) or "layered" (2D coupled layered
    model of 1-4 M nodes at order 4, 0.5-1 s, 101 receivers), drawn from
    numpy.random.default_rng(seed)."""
    r = np.random.default_rng(int(seed))
    if kind == "cased":
        return cased_hole(r, duration=r.uniform(3e-3, 8e-3))
    if kind == "open":
        return open_hole(r, duration=r.uniform(3e-3, 8e-3))
    nx = int(r.integers(250, 501)) * 2
    nzf, nzs = nx // 10, (nx * 2) // 10
    return layered_coupled(nx=nx, nz_fluid=nzf, nz_solid=(nzs, nzs), duration=r.uniform(0.5, 1.0), n_receivers=101)


def served_stream(n_requests, seed=7):
    """The served workload's Poisson stream at unit rate: (arrival times,
    kinds, seeds), 60 % cased-hole, 30 % open-hole and 10 % layered
    requests; scale the arrival times by 1 / rate for another rate."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0, n_requests))
    kinds = rng.choice(["cased", "open", "layered"], p=[0.6, 0.3, 0.1], size=n_requests)
    seeds = rng.integers(0, 2**31, size=n_requests)
    return arrivals, kinds, seeds
