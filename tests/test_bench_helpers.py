"""Host-side helpers of bench.py: the grouping of the multi-mesh batch's FP64 solves."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_consecutive_groups_cover_in_order_within_budget():
    rng = np.random.default_rng(4)
    for _ in range(50):
        sizes = rng.uniform(0.0, 10.0, rng.integers(1, 40))
        budget = float(rng.uniform(5.0, 40.0))
        groups = bench.consecutive_groups(sizes, budget)
        assert [i for g in groups for i in g] == list(range(len(sizes)))
        for g in groups:
            assert g == list(range(g[0], g[-1] + 1))
            assert len(g) == 1 or sizes[g].sum() <= budget
        # maximal runs: the next index did not fit
        for a, b in zip(groups, groups[1:]):
            assert sizes[a].sum() + sizes[b[0]] > budget


def test_consecutive_groups_oversize_and_empty():
    assert bench.consecutive_groups([5.0, 50.0, 5.0], 10.0) == [[0], [1], [2]]
    assert bench.consecutive_groups([], 10.0) == []
