"""GPU scan and radix sort against their textbook definitions and the
reference's ScanTest known answer (src/Test/ScanTest.h:37-125)."""
import numpy as np
import pytest

import shinobi_amd as sa

pytestmark = pytest.mark.gpu


def test_scantest_known_answer(ctx):
    """ScanTest: n = 31*2^20 - 1 ones; inclusive[i] = i+1, exclusive[i] = i"""
    n = 31 * 1024 * 1024 - 1
    ones = np.ones(n, np.uint32)
    inc = sa.scan_u32(ctx, ones, inclusive=True)
    exc = sa.scan_u32(ctx, ones, inclusive=False)
    ref = np.arange(n, dtype=np.uint32)
    np.testing.assert_array_equal(exc, ref)
    np.testing.assert_array_equal(inc, ref + 1)


@pytest.mark.parametrize("n", [0, 1, 7, 2047, 2048, 2049, 100_003, 3_000_001])
def test_scan_random_matches_oracle(ctx, oracle, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(sa.scan_u32(ctx, a, False), oracle.exclusive_scan(a))
    np.testing.assert_array_equal(sa.scan_u32(ctx, a, True), oracle.inclusive_scan(a))


@pytest.mark.parametrize("radix_bits", [4, 6, 8])
@pytest.mark.parametrize("n,key_bits,dup", [(2, 32, False), (4095, 32, False), (4096, 17, True),
                                            (4097, 20, False), (1_000_003, 32, False), (2_000_000, 12, True)])
def test_sort_is_stable_and_matches_numpy(ctx, radix_bits, n, key_bits, dup):
    rng = np.random.default_rng(n + key_bits)
    hi = 64 if dup else 2**key_bits
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = sa.sort_pairs(ctx, keys, vals, key_bits=key_bits, radix_bits=radix_bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])
