"""The traversal divides with rcp_rn / div_rn (sg_device.hpp) instead of the
compiler's IEEE division sequence.  Their results must be the IEEE ones bit
for bit, or closest-hit distances would drift from the oracle: checked on the
device over every binary32 input (reciprocal), and over every divisor with
pseudo-random and signed-zero dividends (division)."""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "_build", "libsg_arith_check.so")


@pytest.fixture(scope="module")
def check(ctx):
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    lib.sg_arith_check.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32)]
    lib.sg_arith_check.restype = C.c_int

    def run(mode, seed=0):
        n, ex = C.c_ulonglong(0), C.c_uint32(0)
        rc = lib.sg_arith_check(0, mode, seed, C.byref(n), C.byref(ex))
        assert rc == 0, f"sg_arith_check({mode}) failed with {rc}"
        return n.value, ex.value

    return run


def test_rcp_rn_equals_ieee_for_every_input(check):
    n, ex = check(0)
    assert n == 0, f"{n} inputs differ, e.g. bits {ex:08x}"


@pytest.mark.parametrize("seed", [1, 2])
def test_div_rn_equals_ieee_for_every_divisor(check, seed):
    n, ex = check(1, seed)
    assert n == 0, f"{n} divisors differ, e.g. bits {ex:08x}"


@pytest.mark.parametrize("mode", [2, 3], ids=["plus_zero", "minus_zero"])
def test_div_rn_zero_dividend(check, mode):
    n, ex = check(mode)
    assert n == 0, f"{n} divisors differ, e.g. bits {ex:08x}"
