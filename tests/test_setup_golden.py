"""The host setup reproduces what it produced before it was rewritten
(tests/golden/, written by tests/golden/make_setup_fixtures.py): every
integer and float32 array bit for bit -- numbering, constraint rows,
transpose CSR, block metrics and coefficients, mass, boundary, coupling,
source and receiver arrays, start vectors, source tables, mesh templates --
and the float64 arrays (coordinates, constraint weights, quadrature rules)
bit for bit or within 1e-13 of their scale (a few ulps: an independently
written quadrature may round a node's last bit differently)."""

from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_setup_fixtures as fx  # noqa: E402

with open(fx.OUT) as fh:
    GOLDEN = json.load(fh)
FULL64 = dict(np.load(fx.OUT64))
ULP_TOL = 1e-13


def _close64(key):
    """A float64 array whose digest changed: within ULP_TOL of its scale."""
    if key not in FULL64 or key not in fx.FULL64:
        return False
    a, b = fx.FULL64[key], FULL64[key]
    if a.shape != b.shape:
        return False
    scale = max(float(np.abs(b).max()) if b.size else 0.0, 1e-300)
    return bool(np.all(np.abs(a - b) <= ULP_TOL * scale))


def _compare(got, want, what):
    assert got.keys() == want.keys(), (what, sorted(set(got) ^ set(want)))
    prefix = {"meshes": "meshes/", "basis": "basis/"}.get(what, what + "/")
    bad = [k for k in want if got[k] != want[k] and not _close64(prefix + k)]
    for k in bad[:8]:
        print(what, k, "got", {q: got[k].get(q) for q in ("shape", "head")} if isinstance(got[k], dict) else got[k],
              "want", {q: want[k].get(q) for q in ("shape", "head")} if isinstance(want[k], dict) else want[k])
    assert not bad, (what, bad[:20])


@pytest.mark.parametrize("name", sorted(fx.MODELS))
def test_model_setup_bitwise(name):
    _compare(fx.model_digests(fx.MODELS[name]()), GOLDEN["models"][name], name)


def test_mesh_templates_bitwise():
    _compare(fx.mesh_digests(), GOLDEN["meshes"], "meshes")


def test_quadrature_rules_bitwise():
    _compare(fx.basis_digests(), GOLDEN["basis"], "basis")


def test_fixtures_cover_hanging_faces_and_every_kind():
    models = GOLDEN["models"]
    # constrained maps (hanging faces) in the refined and borehole models
    for name in ("borehole_m0", "borehole_m1", "refined_sheared_grid"):
        assert any(k.endswith("row_dofs") and v["shape"][0] > 0 for k, v in models[name].items()), name
    kinds = {int(v["head"][0]) for m in models.values() for k, v in m.items() if k.endswith(".info") and ".b" in k}
    assert kinds == {0, 1, 2, 3}
