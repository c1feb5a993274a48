"""insem_amd._core: the drop-in for the reference's compiled module, and the
case/results files (CPU only; the GPU solver is exercised in test_gpu_io.py).

Checks:
* every name the reference's package imports from its core
  (python/insem/__init__.py:7-29, basis.py:3-10) exists;
* mesh, forest, point map and geometry objects agree with the package's
  own model (same leaves and kinds) and with mathematics: gather and
  assemble are adjoint, gather reproduces linear fields at every element
  point (hanging faces included), global coordinates have no mismatch;
* the basis functions against closed forms and scipy;
* a case file written by save() loads back to the same description, is
  rewritten byte-equivalently (h5dump), carries the reference's format
  attributes and layout, and loads with the reference's optional groups
  missing; errors name the problem;
* insem-run's usage and failure exits.
"""

from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.special

import insem_amd as pkg
from insem_amd import _core as core
from insem_amd import mesh as templates

RUNNER = os.path.join(os.path.dirname(pkg.__file__), "insem-run")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

# python/insem/__init__.py:7-29 and basis.py:3-10
REFERENCE_NAMES = [
    "Discretization", "DofMap", "Forest", "Geometry", "InsemError", "MacroMesh", "Material", "Receiver",
    "Simulation", "Solver", "Source", "SourceTimeFunction", "TimeConfig", "__version__", "gaussian",
    "gaussian_derivative", "global_coordinates", "real_dtype", "ricker", "sampled_stf",
    "derivative_matrix", "gauss_lobatto_jacobi", "gll", "glj01", "interpolation_matrix", "jacobi",
]


def core_mesh(m):
    """A _core.MacroMesh from one of the package's mesh templates."""
    return core.MacroMesh(m.cell_type, m.coords, m.nodes, m.cells, m.material, m.boundary)


def small_borehole():
    return templates.borehole((-1.0, 1.0), [("mud", 0.1)], 1.0, dz=0.25)


def refined_borehole(order=4):
    """Axisymmetric borehole (fluid column, solid) with a refined patch, so
    that the forest has hanging faces."""
    bh = small_borehole()
    sim = core.Simulation(core_mesh(bh.mesh), core.Discretization(order=order, regions=[(0.0, 0.3, -0.3, 0.3, 1)]))
    sim.set_material(bh.regions["mud"], core.Material.acoustic(1500.0, 1000.0))
    sim.set_material(bh.regions["formation"], core.Material.elastic(4000.0, 2300.0, 2500.0))
    return sim


def test_every_reference_name_is_exported():
    missing = [n for n in REFERENCE_NAMES if not hasattr(core, n)]
    assert not missing
    assert core.real_dtype == "float32"
    assert core.__version__ == "0.1.0"
    assert issubclass(core.InsemError, RuntimeError)


def test_macro_mesh_round_trips_its_arrays():
    m = templates.rectangle((-4.0, 4.0), (-2.0, 2.0), nx=8, nz=4)
    cm = core_mesh(m)
    assert (cm.n_cells, cm.n_nodes, cm.cell_type, cm.coords) == (32, 45, "quad4", "cartesian")
    np.testing.assert_array_equal(cm.cells, m.cells)
    np.testing.assert_array_equal(cm.boundary, m.boundary)
    np.testing.assert_array_equal(cm.nodes, m.nodes)
    x, z = cm.map(0, 1.0, 1.0)
    assert (x, z) == pytest.approx((-3.0, -1.0))
    with pytest.raises(core.InsemError, match="unknown cell type"):
        core.MacroMesh("tri3", "cartesian", m.nodes, m.cells, m.material, m.boundary)
    with pytest.raises(core.InsemError, match="shape"):
        core.MacroMesh("quad4", "cartesian", m.nodes[:, :1], m.cells, m.material, m.boundary)


def test_setup_matches_the_package_model():
    sim = refined_borehole()
    sim.setup()
    bh = small_borehole()
    own = pkg.Simulation(bh.mesh, pkg.Discretization(order=4, regions=((0.0, 0.3, -0.3, 0.3, 1),)))
    own.set_material(bh.regions["mud"], pkg.Material.acoustic(1500.0, 1000.0))
    own.set_material(bh.regions["formation"], pkg.Material.elastic(4000.0, 2300.0, 2500.0))
    own.setup()
    assert sim.forest.n_leaves == own.n_leaves
    assert sim.forest.max_level == own.max_level == 1
    assert sim.geometry_kind_counts() == own.geometry_kind_counts()
    assert sim.forest.leaves.shape == (sim.forest.n_leaves, 4)
    g = sim.geometry
    assert g.x.shape == (sim.forest.n_leaves, 5, 5)
    x_own, z_own = own.leaf_points()
    np.testing.assert_array_equal(g.x.ravel(), np.asarray(x_own).ravel())
    np.testing.assert_array_equal(g.z.ravel(), np.asarray(z_own).ravel())


def test_point_map_gather_and_assemble():
    sim = refined_borehole(order=5)
    sim.setup()
    d, g = sim.dofs, sim.geometry
    assert d.ngll == 6 and d.n_elements == sim.forest.n_leaves
    assert (d.local_to_global < 0).any(), "the refined patch must leave hanging faces"
    xz, mismatch = core.global_coordinates(d, g)
    assert xz.shape == (2, d.n_global)
    assert mismatch < 1e-12
    # linear fields are reproduced at every element point, constrained or not
    lin = 2.0 * xz[0] - 3.0 * xz[1] + 0.5
    np.testing.assert_allclose(d.gather(lin), 2.0 * g.x - 3.0 * g.z + 0.5, rtol=0, atol=1e-12)
    # gather and assemble are adjoint, for one and for two components
    rng = np.random.default_rng(3)
    for shape in ((d.n_global,), (2, d.n_global)):
        gl = rng.standard_normal(shape)
        loc = rng.standard_normal(d.gather(gl).shape)
        assert np.vdot(d.assemble(loc), gl) == pytest.approx(np.vdot(loc, d.gather(gl)), rel=1e-12)
    # assembling ones counts each point's unconstrained copies plus its constraint weights
    ones = d.assemble(np.ones((d.n_elements, 6, 6)))
    assert ones.sum() == pytest.approx(d.n_elements * 36)
    assert (d.multiplicity >= 1).all()
    with pytest.raises(core.InsemError, match="global field"):
        d.gather(np.zeros(d.n_global + 1))


def test_forest_neighbours_and_corners():
    sim = refined_borehole()
    sim.setup()
    f = sim.forest
    for leaf in range(f.n_leaves):
        c = f.leaf_corners(leaf)
        assert c.shape == (4, 2)
        for face in range(4):
            nb, nface, dl = f.face_neighbor(leaf, face)
            if nb >= 0:
                assert abs(dl) <= 1 and 0 <= nface < 4
    box = f.leaf_box(0)
    assert len(box) == 4


def test_basis_functions():
    for n in (2, 5, 9):
        x, w = core.gll(n)
        assert w.sum() == pytest.approx(2.0)
        xj, wj = core.glj01(n)
        assert wj.sum() == pytest.approx(2.0)  # integral of (1 + x) over [-1, 1]
        xa, wa = core.gauss_lobatto_jacobi(n, 0, 0)
        np.testing.assert_array_equal(xa, x)
        D = core.derivative_matrix(x)
        k = min(n - 1, 3)
        np.testing.assert_allclose(D @ x ** k, k * x ** (k - 1), atol=1e-11)
        t = np.linspace(-1, 1, 7)
        I = core.interpolation_matrix(x, t)
        np.testing.assert_allclose(I @ (x ** (n - 1)), t ** (n - 1), atol=1e-12)
        np.testing.assert_allclose(core.interpolation_matrix(x, x), np.eye(n), atol=1e-14)
    for n, a, b in ((0, 0, 0), (3, 0, 0), (4, 1, 2), (7, 0.5, 1.5), (9, 2, 0)):
        for xv in (-1.0, -0.3, 0.2, 0.9, 1.0):
            assert core.jacobi(n, a, b, xv) == pytest.approx(scipy.special.eval_jacobi(n, a, b, xv), rel=1e-12, abs=1e-12)


def test_source_time_functions_and_descriptions():
    r = core.ricker(25.0)
    assert r.kind == "ricker" and r.t0 == pytest.approx(1.2 / 25.0)
    t = np.linspace(0, 0.1, 11)
    a = (np.pi * 25.0 * (t - r.t0)) ** 2
    np.testing.assert_allclose(r(t), (1 - 2 * a) * np.exp(-a), rtol=1e-14, atol=1e-15)
    assert r(r.t0) == pytest.approx(1.0)
    g = core.gaussian(10.0, t0=0.05, amplitude=2.0)
    assert g(0.05) == pytest.approx(2.0) and g.kind == "gaussian"
    gd = core.gaussian_derivative(10.0, t0=0.05)
    assert gd(0.05) == pytest.approx(0.0, abs=1e-15)
    s = core.sampled_stf([0.0, 1.0, 0.0], 0.5)
    assert s(0.25) == pytest.approx(0.5) and s(2.0) == 0.0 and s.kind == "samples"
    with pytest.raises(core.InsemError):
        core.ricker(-1.0)
    src = core.Source.moment_tensor(1.0, 2.0, 1.0, 1.0, 0.0, r)
    assert src.kind == "moment_tensor" and src.position == [1.0, 2.0] and src.components == [1.0, 1.0, 0.0]
    assert core.Source.monopole(0, 0, r).components == [1.0, 0.0, 0.0]
    mat = core.Material.elastic(4000.0, 2300.0, 2500.0)
    assert (mat.medium, mat.vp, mat.vs, mat.rho) == ("elastic", 4000.0, 2300.0, 2500.0)
    assert repr(mat) == "<Material elastic vp=4000 vs=2300 rho=2500>"
    with pytest.raises(core.InsemError):
        core.Material.elastic(2000.0, 2500.0, 2000.0)
    d = core.Discretization(order=5, regions=[(0, 1, 0, 1, 2)], azimuthal_order=1)
    assert (d.order, d.regions, d.azimuthal_order, d.max_level) == (5, [(0.0, 1.0, 0.0, 1.0, 2)], 1, 12)


def full_case():
    """A layered coupled model with every kind of description entry."""
    m = templates.layered((0.0, 200.0), [-200.0, -100.0, 0.0], [2, 2], [2, 1], nx=8)
    sim = core.Simulation(core_mesh(m), core.Discretization(order=3, f_max=5.0, points_per_wavelength=4.0,
                                                            max_level=2, regions=[(50, 60, -20, 0, 1)]))
    sim.set_material(1, core.Material.acoustic(1500.0, 1000.0))
    sim.set_material(2, core.Material.elastic(4000.0, 2300.0, 2500.0))
    sim.set_boundary(templates.TOP, "free")
    for tag in (templates.BOTTOM, templates.LEFT, templates.RIGHT):
        sim.set_boundary(tag, "absorbing")
    sim.add_source(core.Source.moment_tensor(100.0, -150.0, 1.0, 1.0, 0.0, core.ricker(25.0)))
    sim.add_source(core.Source.force(80.0, -120.0, 0.0, 1.0, core.gaussian(20.0, t0=0.06, amplitude=3.0)))
    sim.add_source(core.Source.monopole(50.0, -20.0, core.sampled_stf(np.sin(np.arange(40) * 0.3), 1e-3, 2.0)))
    sim.add_source(core.Source.dipole(60.0, -30.0, 1.0, 0.0, core.gaussian_derivative(15.0)))
    sim.add_receiver("r0", 120.0, -40.0, ["pressure", "particle_velocity"])
    sim.add_receiver("r1", 120.0, -150.0, ["displacement", "velocity"], medium="elastic")
    sim.set_time(duration=0.1, safety=0.7, record_every=2, energy_every=5, snapshot_every=10,
                 snapshot_quantities=["velocity", "pressure"])
    return sim


def describe(sim):
    """Every field of a simulation's description, in plain Python values."""
    d = sim.discretization
    t = sim.time
    return {
        "mesh": (sim.mesh.cell_type, sim.mesh.coords, sim.mesh.nodes.tolist(), sim.mesh.cells.tolist(),
                 sim.mesh.material.tolist(), sim.mesh.boundary.tolist()),
        "disc": (d.order, d.uniform_level, d.f_max, d.points_per_wavelength, d.max_level, d.regions, d.azimuthal_order),
        "materials": {k: (v.medium, v.vp, v.vs, v.rho) for k, v in sim.materials.items()},
        "boundaries": dict(sim.boundaries),
        "sources": [(s.kind, s.position, s.components, s.stf.kind, s.stf.f0, s.stf.t0, s.stf.amplitude,
                     s.stf(np.linspace(0, 0.05, 9)).tolist()) for s in sim.sources],
        "receivers": [(r.name, r.position, r.quantities, r.medium) for r in sim.receivers],
        "time": (t.dt, t.safety, t.n_steps, t.duration, t.record_every, t.energy_every, t.snapshot_every,
                 t.snapshot_quantities),
    }


def test_case_file_round_trip(tmp_path):
    sim = full_case()
    a, b = str(tmp_path / "a.h5"), str(tmp_path / "b.h5")
    sim.save(a)
    back = core.Simulation.load(a)
    assert describe(back) == describe(sim)
    back.save(b)
    if os.path.exists(H5DUMP):
        da = subprocess.run([H5DUMP, a], capture_output=True, text=True, check=True).stdout.replace(a, "F")
        db = subprocess.run([H5DUMP, b], capture_output=True, text=True, check=True).stdout.replace(b, "F")
        assert da == db
    # the reference's layout and attribute names (simulation.cpp:233-322)
    assert core._read_attribute(a, "/", "format") == "insem-case"
    assert core._read_attribute(a, "/", "version") == 1
    assert core._read_attribute(a, "/mesh", "cell_type") == "quad4"
    assert core._read_attribute(a, "/sources/2", "stf") == "samples"
    assert core._read_attribute(a, "/sources", "count") == 4
    assert core._read_attribute(a, "/receivers/1", "quantities") == "displacement,velocity"
    assert core._read_attribute(a, "/boundary_conditions", str(templates.TOP)) == "free"
    assert core._read_attribute(a, "/time", "snapshot_quantities") == "velocity,pressure"
    assert core._read_dataset(a, "/mesh/nodes").shape == (sim.mesh.n_nodes, 2)
    assert core._read_dataset(a, "/discretization/regions").shape == (1, 5)
    np.testing.assert_array_equal(core._read_dataset(a, "/sources/2/stf_samples"), np.sin(np.arange(40) * 0.3))
    # the set-up of the loaded case is the same model
    sim.setup()
    back.setup()
    assert back.forest.n_leaves == sim.forest.n_leaves
    np.testing.assert_array_equal(back.dofs.local_to_global, sim.dofs.local_to_global)


def test_mesh_only_case_and_errors(tmp_path):
    m = templates.rectangle((-1.0, 1.0), (-1.0, 1.0), nx=2, nz=2)
    sim = core.Simulation(core_mesh(m))
    p = str(tmp_path / "mesh_only.h5")
    sim.save(p)
    back = core.Simulation.load(p)
    assert back.sources == [] and back.receivers == [] and back.materials == {}
    with pytest.raises(core.InsemError, match="cannot open"):
        core.Simulation.load(str(tmp_path / "missing.h5"))
    with pytest.raises(core.InsemError, match="unknown boundary type|is none of"):
        sim.set_boundary(1, "sticky")
    with pytest.raises(core.InsemError, match="receiver"):
        sim.add_receiver("a/b", 0.0, 0.0, ["pressure"])


def test_insem_run_usage_and_failures(tmp_path):
    assert os.path.exists(RUNNER), "insem-run is not built"
    r = subprocess.run([RUNNER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage: insem-run" in r.stderr
    r = subprocess.run([RUNNER, "--help"], capture_output=True, text=True)
    assert r.returncode == 0
    r = subprocess.run([RUNNER, str(tmp_path / "nope.h5")], capture_output=True, text=True)
    assert r.returncode == 1 and "insem-run: cannot open HDF5 file" in r.stderr
