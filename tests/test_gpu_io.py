"""The drop-in front end on the GPU: insem_amd._core's Simulation and
Solver, snapshots, results files and the insem-run command.

* _core runs the same model as the package's own front end to the same
  bits (both drive the one backend) and matches the oracle (traces within
  1e-5 relative L2);
* snapshots are the device fields at the snapshot steps, gathered to the
  element points: equal to the oracle's fields at those steps on a model
  without hanging faces (where the gather copies values), and splitting the
  run at the snapshot steps changes no trace;
* insem-run loads a saved case, runs it and writes a results file whose
  traces, times, energy and snapshots equal the Solver's.
"""

from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import insem_amd as pkg
from insem_amd import _core as core
from insem_amd import mesh as templates
from oracle_driver import OracleRun

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(pkg.__file__), "insem-run")


def box_case(lib, snapshot_every=0):
    """The 2D acoustic box (8 x 8 m, order 4) at 16 x 16 quads (absorbing
    sides, monopole, pressure receiver), in either front end."""
    m = templates.rectangle((-4.0, 4.0), (-4.0, 4.0), nx=16, nz=16)
    mm = core.MacroMesh(m.cell_type, m.coords, m.nodes, m.cells, m.material, m.boundary) if lib is core else m
    sim = lib.Simulation(mm, lib.Discretization(order=4))
    sim.set_material(0, lib.Material.acoustic(vp=1500.0, rho=1000.0))
    for tag in (templates.BOTTOM, templates.RIGHT, templates.TOP, templates.LEFT):
        sim.set_boundary(tag, "absorbing")
    sim.add_source(lib.Source.monopole(0.0, 0.0, lib.ricker(400.0)))
    sim.add_receiver("h1", 1.5, 0.0, ["pressure", "particle_velocity"])
    if lib is core:
        sim.set_time(duration=2e-3, energy_every=10, snapshot_every=snapshot_every,
                     snapshot_quantities=["pressure"])
    else:
        sim.set_time(duration=2e-3, energy_every=10)
    return sim


def test_core_front_end_equals_package_and_oracle():
    a = box_case(core)
    sa = a.solver
    sa.run()
    b = box_case(pkg)
    sb = b.solver
    sb.run()
    assert sa.dt == sb.dt and sa.n_steps == sb.n_steps and sa.step == sa.n_steps
    for q in ("pressure", "particle_velocity"):
        np.testing.assert_array_equal(sa.trace("h1", q), sb.trace("h1", q))
    np.testing.assert_array_equal(sa.energy, sb.energy)
    np.testing.assert_array_equal(sa.times, sb.times)
    o = OracleRun(b.build_setup(), dt=sb.dt)
    o.run()
    ref = o.traces()[("h1", "pressure")]
    got = sa.trace("h1", "pressure")
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-5
    u, v, acc = sa.state("acoustic")
    assert u.shape == (1, sa.coordinates("acoustic").shape[1])


def test_snapshots_are_the_fields_at_their_steps():
    every = 20
    sim = box_case(core, snapshot_every=every)
    s = sim.solver
    s.run()
    times = s.snapshot_times
    steps = np.rint(times / s.dt).astype(int)
    np.testing.assert_array_equal(steps, np.arange(0, s.n_steps + 1, every))
    snap = s.snapshot("acoustic", "pressure")
    npe = 25
    assert snap.shape == (len(steps), 1, sim.forest.n_leaves * npe)
    pts = s.snapshot_points("acoustic")
    cells = s.snapshot_cells("acoustic")
    assert pts.shape == (2, sim.forest.n_leaves * npe) and cells.shape == (sim.forest.n_leaves * 16, 4)
    # the same model through the package front end and the oracle
    b = box_case(pkg)
    o = OracleRun(b.build_setup(), dt=s.dt)
    l2g = sim.dofs.local_to_global.reshape(-1)
    assert (l2g >= 0).all()
    done = 0
    for k, st in enumerate(steps):
        o.advance(st - done)
        done = st
        field = o.field(0, 0)[0]
        np.testing.assert_array_equal(snap[k, 0], field[l2g].astype(np.float32))
    # stopping at the snapshot steps changed no trace
    plain = box_case(core).solver
    plain.run()
    np.testing.assert_array_equal(plain.trace("h1", "pressure"), s.trace("h1", "pressure"))


def test_insem_run_writes_the_solver_results(tmp_path):
    case = str(tmp_path / "box.h5")
    sim = box_case(core, snapshot_every=25)
    sim.save(case)
    r = subprocess.run([RUNNER, case], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "box_results.h5")
    assert "results:         " + out in r.stdout
    s = sim.solver
    s.run()
    assert core._read_attribute(out, "/", "format") == "insem-results"
    assert core._read_attribute(out, "/", "steps_done") == s.n_steps
    assert core._read_attribute(out, "/", "dt") == s.dt
    np.testing.assert_array_equal(core._read_dataset(out, "/time"), s.times)
    for q in ("pressure", "particle_velocity"):
        np.testing.assert_array_equal(core._read_dataset(out, f"/receivers/h1/{q}"), s.trace("h1", q))
    assert core._read_attribute(out, "/receivers/h1", "medium") == "acoustic"
    np.testing.assert_array_equal(core._read_dataset(out, "/energy"), s.energy)
    assert core._read_attribute(out, "/energy", "columns") == "time,kinetic,potential"
    np.testing.assert_array_equal(core._read_dataset(out, "/snapshots/acoustic/pressure"),
                                  s.snapshot("acoustic", "pressure"))
    np.testing.assert_array_equal(core._read_dataset(out, "/snapshots/time"), s.snapshot_times)
    assert core._read_dataset(out, "/snapshots/acoustic/points").shape == (s.snapshot_points("acoustic").shape[1], 2)
    # Solver.write_results writes the same file content
    own = str(tmp_path / "own.h5")
    s.write_results(own)
    np.testing.assert_array_equal(core._read_dataset(own, "/receivers/h1/pressure"),
                                  core._read_dataset(out, "/receivers/h1/pressure"))
