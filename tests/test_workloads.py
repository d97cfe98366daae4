"""Host-side workload helpers: the tile orders of primary rays."""
import numpy as np

from shinobi_amd import workloads as W


def test_tile_order_is_a_permutation_in_tile_order():
    o = W.tile_order(10, 7, (2, 4))
    assert sorted(o.tolist()) == list(range(70))
    # first tile: columns 0-1 of rows 0-3, row-major inside the tile
    np.testing.assert_array_equal(o[:8], [0, 1, 10, 11, 20, 21, 30, 31])
    # the partial tiles of the bottom edge (rows 4-6) come last
    assert set(o[-6:].tolist()) == {48, 49, 58, 59, 68, 69}


def test_square_tile_and_identity():
    np.testing.assert_array_equal(W.tile_order(16, 8, 4), W.tile_order(16, 8, (4, 4)))
    np.testing.assert_array_equal(W.tile_order(5, 3, 1), np.arange(15))
