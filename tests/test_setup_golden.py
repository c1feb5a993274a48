"""The host setup against what it produced before it was rewritten
(tests/golden/setup_digests.json, written by tests/golden/make_setup_fixtures.py).

Integer arrays -- element numbering, l2g, constraint-row structure, the
transpose CSR, Stacey/Dirichlet/axis/coupling/source/receiver point lists,
mesh connectivity -- must be bitwise equal.  Floating arrays are bitwise
equal or within a few ulps: the rewritten quadrature rounds every node and
weight correctly (the earlier one was off by up to ~1e2 ulps in some
weights), so values derived from them may differ in their last bits.  Their
fingerprints (sums, a cosine-weighted sum, extremes and 32 spot values at
fixed positions) must agree to FLOAT_RTOL of the array's scale, which any
wrong index, sign, factor or order would break.
"""

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

FLOAT_RTOL = {"float32": 4 * 2.0 ** -23, "float64": 1e-12}
STATS = {"bitwise": 0, "ulps": 0}


def _same(got, want):
    if got == want:
        STATS["bitwise"] += 1
        return True
    if not isinstance(want, dict) or "sha256" not in want:
        return False
    if got["shape"] != want["shape"] or got["dtype"] != want["dtype"] or "fp" not in want or "fp" not in got:
        return False
    g, w = got["fp"], want["fp"]
    rtol = FLOAT_RTOL[want["dtype"]]
    n = max(1, int(np.prod(want["shape"])))
    scale = max(w["absmax"], 1e-300)
    if g["spot_idx"] != w["spot_idx"]:
        return False
    if not np.allclose(g["spots"], w["spots"], rtol=rtol, atol=rtol * scale):
        return False
    for k in ("min", "max"):
        if abs(g[k] - w[k]) > rtol * scale:
            return False
    for k, tol in (("sum", rtol * scale * n), ("wsum", rtol * scale * n), ("sum2", 2 * rtol * scale * scale * n)):
        if abs(g[k] - w[k]) > tol:
            return False
    STATS["ulps"] += 1
    return True


def _compare(got, want, what):
    assert got.keys() == want.keys(), (what, sorted(set(got) ^ set(want)))
    bad = [k for k in want if not _same(got[k], want[k])]
    for k in bad[:8]:
        print(what, k, "got", {q: got[k].get(q) for q in ("shape", "head")} if isinstance(got[k], dict) else got[k],
              "want", {q: want[k].get(q) for q in ("shape", "head")} if isinstance(want[k], dict) else want[k])
    assert not bad, (what, bad[:20])


@pytest.mark.parametrize("name", sorted(fx.MODELS))
def test_model_setup_matches(name):
    _compare(fx.model_digests(fx.MODELS[name]()), GOLDEN["models"][name], name)


def test_mesh_templates_match():
    _compare(fx.mesh_digests(), GOLDEN["meshes"], "meshes")


def test_quadrature_rules_match():
    _compare(fx.basis_digests(), GOLDEN["basis"], "basis")


def test_integer_arrays_are_compared_bitwise():
    # every integer array of the fixture carries no fingerprint: only its
    # digest can match
    for m in GOLDEN["models"].values():
        for v in m.values():
            if isinstance(v, dict) and v.get("dtype", "").startswith(("int", "uint")):
                assert "fp" not in v


def test_fixtures_cover_hanging_faces_and_every_kind():
    models = GOLDEN["models"]
    # constrained maps (hanging faces) in the refined and borehole models
    for name in ("borehole_m0", "borehole_m1", "refined_sheared_grid"):
        assert any(k.endswith("row_dofs") and v["shape"][0] > 0 for k, v in models[name].items()), name
    kinds = {int(v["head"][0]) for m in models.values() for k, v in m.items() if k.endswith(".info") and ".b" in k}
    assert kinds == {0, 1, 2, 3}
