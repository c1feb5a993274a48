"""The host setup reproduces the committed digests of every array it
produced before it was rewritten (tests/golden/setup_digests.json, written by
tests/golden/make_setup_fixtures.py): numbering, constraint rows, transpose
CSR, block metrics, mass, boundary, coupling, source and receiver arrays,
coordinates, power-iteration start vectors, source tables, mesh templates
and quadrature rules, bit for bit."""

from __future__ import annotations

import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_setup_fixtures as fx  # noqa: E402

with open(fx.OUT) as fh:
    GOLDEN = json.load(fh)


def _compare(got, want, what):
    assert got.keys() == want.keys(), (what, sorted(set(got) ^ set(want)))
    bad = [k for k in want if got[k] != want[k]]
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
