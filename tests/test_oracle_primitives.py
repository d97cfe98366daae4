"""Pins for the oracle's scan, stable sort and range extraction.

* scan: the reference's own known answer (ScanTest, src/Test/ScanTest.h:37-125)
  and numpy's cumsum with uint32 wrap-around;
* sort: numpy's stable argsort (the semantics of chag::pp::sort as driven by
  src/Utils/Sort.cu:28-56);
* ranges: searchsorted left/right (the host re-derivation sketched in
  src/Control/Algorithm/FastSimpleGridBuilder.h:214-246)."""
import numpy as np
import pytest


def test_scantest_known_answer(oracle):
    n = 31 * 1024 * 1024 - 1
    ones = np.ones(n, np.uint32)
    inc = oracle.inclusive_scan(ones)
    exc = oracle.exclusive_scan(ones)
    ref = np.arange(n, dtype=np.uint32)
    assert np.array_equal(exc, ref)
    assert np.array_equal(inc, ref + 1)


@pytest.mark.parametrize("n", [0, 1, 5, 1000, 100_001])
def test_scan_wraps_like_uint32(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    inc = np.cumsum(a.astype(np.uint64)).astype(np.uint64) % (1 << 32)
    assert np.array_equal(oracle.inclusive_scan(a), inc.astype(np.uint32))
    exc = np.concatenate([[0], inc[:-1]]).astype(np.uint32) if n else inc.astype(np.uint32)
    assert np.array_equal(oracle.exclusive_scan(a), exc)


@pytest.mark.parametrize("n,hi,threads", [(10, 4, 1), (5000, 2**32 - 1, 1), (300_000, 17, 4),
                                          (300_000, 2**24, 8), (70_000, 2**31, 3)])
def test_sort_is_numpy_stable_sort(oracle, n, hi, threads):
    rng = np.random.default_rng(n + hi)
    keys = rng.integers(0, hi + 1, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = oracle.sort_pairs(keys, vals, nthreads=threads)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


def test_ranges_are_searchsorted(oracle):
    rng = np.random.default_rng(3)
    C = 500
    keys = np.sort(rng.integers(0, C + 1, 20000).astype(np.uint32))   # C is the dummy key
    cells = oracle.ranges(keys, C).reshape(C, 2)
    lo = np.searchsorted(keys, np.arange(C), "left")
    hi = np.searchsorted(keys, np.arange(C), "right")
    nonempty = hi > lo
    assert np.array_equal(cells[nonempty, 0], lo[nonempty])
    assert np.array_equal(cells[nonempty, 1], hi[nonempty])
    assert (cells[~nonempty] == 0).all()
    # a run longer than one 256-thread tile at the very end, no dummy tail
    # (defect D2 of the reference misses its end)
    keys2 = np.concatenate([np.zeros(10, np.uint32), np.full(40000, 7, np.uint32)])
    cells2 = oracle.ranges(keys2, 8).reshape(8, 2)
    assert tuple(cells2[7]) == (10, 40010) and tuple(cells2[0]) == (0, 10)
