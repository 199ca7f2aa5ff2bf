"""How a mosaic's tiles are cut into concurrent batched solves
(imcom_amd.batch.split_tiles): every tile exactly once, largest first, runs
of about equal n^3 work."""
import numpy as np
import pytest

from imcom_amd.batch import split_tiles


def test_split_covers_every_tile_once_largest_first():
    rng = np.random.default_rng(3)
    ns = rng.integers(1024, 6145, size=64)
    groups = split_tiles(ns, 3)
    flat = [i for g in groups for i in g]
    assert sorted(flat) == list(range(64))
    sizes = ns[flat]
    assert np.all(np.diff(sizes) <= 0)           # runs are consecutive in decreasing size
    work = np.array([np.sum(ns[g].astype(float) ** 3) for g in groups])
    total = work.sum()
    # each run but the last reaches its third of the work with its last tile
    for j, w in enumerate(np.cumsum(work)[:-1], start=1):
        assert w >= total * j / 3
        assert w - float(ns[groups[j - 1][-1]]) ** 3 < total * j / 3


def test_split_edge_cases():
    assert split_tiles([100], 3) == [[0]]
    assert sorted(i for g in split_tiles([100, 200], 3) for i in g) == [0, 1]
    one = split_tiles([5, 5, 5, 5], 1)
    assert len(one) == 1 and sorted(one[0]) == [0, 1, 2, 3]
    with pytest.raises(ValueError):
        split_tiles([], 3)
    with pytest.raises(ValueError):
        split_tiles([10], 0)
