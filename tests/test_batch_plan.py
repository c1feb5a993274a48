"""The batching policy of insem_amd.solve_many (host only): models that
can share a batch, sorted by size, in batches of at most batch_size, every
model exactly once."""

from __future__ import annotations

import numpy as np
import pytest

from insem_amd import plan_batches


def test_every_model_once_and_batches_are_homogeneous():
    rng = np.random.default_rng(3)
    keys = [("order", int(k)) for k in rng.integers(4, 7, size=128)]
    sizes = [int(s) for s in rng.integers(1000, 60000, size=128)]
    plan = plan_batches(keys, sizes, 5)
    flat = sorted(i for b in plan for i in b)
    assert flat == list(range(128))
    for b in plan:
        assert 1 <= len(b) <= 5
        assert len({keys[i] for i in b}) == 1
        assert [sizes[i] for i in b] == sorted(sizes[i] for i in b)
    # per key: ceil(n / 5) batches of nearly equal count, in size order
    for k in set(keys):
        bs = [b for b in plan if keys[b[0]] == k]
        n = sum(1 for kk in keys if kk == k)
        assert len(bs) == -(-n // 5)
        assert max(map(len, bs)) - min(map(len, bs)) <= 1
        assert all(sizes[bs[j][-1]] <= sizes[bs[j + 1][0]] for j in range(len(bs) - 1))


def test_small_inputs():
    assert plan_batches([], [], 3) == []
    assert plan_batches(["a"], [7], 3) == [[0]]
    assert plan_batches(["a", "b", "a"], [5, 1, 2], 1) == [[2], [0], [1]]
    with pytest.raises(ValueError):
        plan_batches(["a"], [1], 0)
