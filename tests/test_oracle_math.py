"""Pins for the oracle's scalar definitions: the deterministic pow behind the
grid resolutions, the quarter-turn sin/cos of the bounce sampler, and the
resolution closed forms of GridBuilder::computeResolution
(src/Control/Algorithm/GridBuilder.h:41-62) and of the two-level top grid
(src/Control/Algorithm/FastTwoLevelGridBuilder.h:87-102)."""
import math

import numpy as np
import pytest


def test_det_pow_matches_libm(oracle):
    rng = np.random.default_rng(0)
    xs = np.exp(rng.uniform(math.log(1e-6), math.log(1e12), 20000))
    y = float(np.float32(0.3333333))
    for x in xs:
        a = oracle.det_pow(float(x), y)
        b = math.pow(float(x), y)
        assert abs(a - b) <= 4e-15 * b


def test_cbrtpow_is_rounded_pow(oracle):
    """cbrtpow(x) = (float)pow((double)x, (double)0.3333333f); the
    deterministic evaluation may differ only when the double result sits on
    a float rounding boundary (never in this sample)."""
    rng = np.random.default_rng(1)
    xs = np.float32(np.exp(rng.uniform(math.log(1e-3), math.log(1e11), 20000)))
    y = float(np.float32(0.3333333))
    exact = np.array([np.float32(math.pow(float(x), y)) for x in xs])
    got = np.array([oracle.cbrtpow(float(x)) for x in xs], np.float32)
    assert np.array_equal(got, exact)


@pytest.mark.parametrize("x,y,expected", [(8.0, 1.0 / 3.0, 2.0), (2.0, 10.0, 1024.0), (1.0, 0.3, 1.0),
                                          (0.25, 0.5, 0.5), (1e300, 0.5, 1e150)])
def test_det_pow_closed_forms(oracle, x, y, expected):
    assert oracle.det_pow(x, y) == pytest.approx(expected, rel=2e-15)


def test_det_pow_special_values(oracle):
    assert oracle.det_pow(0.0, 0.33) == 0.0
    assert math.isnan(oracle.det_pow(-1.0, 0.33))
    assert math.isinf(oracle.det_pow(float("inf"), 0.33))


def test_sincos_2pi(oracle):
    us = np.float32(np.linspace(0.0, 1.0, 4097, endpoint=False))
    for u in us:
        s, c = oracle.sincos_2pi(float(u))
        assert abs(s - math.sin(2 * math.pi * float(u))) < 1.2e-7
        assert abs(c - math.cos(2 * math.pi * float(u))) < 1.2e-7
    # exact quarter turns
    assert oracle.sincos_2pi(0.0) == (0.0, 1.0)
    assert oracle.sincos_2pi(0.25) == (1.0, 0.0) or oracle.sincos_2pi(0.25)[0] == 1.0
    assert oracle.sincos_2pi(0.5)[1] == -1.0


def test_uniform_resolution_closed_form(oracle):
    # bounds [0,2]x[0,1]x[0,1], N=1000, lambda 5: V=2, 5*1000/2=2500,
    # 2500^(1/3) = 13.572 -> res = (int)(2*13.572), (int)13.572, (int)13.572
    st, res = oracle.resolution([0, 0, 0], [2, 1, 1], 1000, 5.0, 1)
    assert st == 0 and res == (27, 13, 13)
    # unit cube, N = 10000: (5e4)^(1/3) = 36.84 -> 36^3 (the 10k-triangle workload)
    st, res = oracle.resolution([0, 0, 0], [1, 1, 1], 10000, 5.0, 1)
    assert st == 0 and res == (36, 36, 36)


def test_twolevel_top_resolution_closed_form(oracle):
    # unit cube, N = 2e6: (1e7)^(1/3) = 215.44 -> 215 / 6 = 35 per axis
    st, res = oracle.resolution([0, 0, 0], [1, 1, 1], 2_000_000, 5.0, 6)
    assert st == 0 and res == (35, 35, 35)
    # thin slab: the clamp to >= 2 (FastTwoLevelGridBuilder.h:100-102)
    st, res = oracle.resolution([0, 0, 0], [1, 1, 0.001], 1000, 5.0, 6)
    assert st == 0 and res[2] == 2


def test_invalid_scenes(oracle):
    assert oracle.resolution([0, 0, 0], [1, 1, 0], 100, 5.0, 1)[0] == 1      # zero extent
    assert oracle.resolution([0, 0, 0], [1, 1, 1], 0, 5.0, 1)[0] == 1        # empty
    # resolution < 1 on an axis (the reference does not clamp the uniform grid)
    assert oracle.resolution([0, 0, 0], [1, 1, 0.0001], 1, 5.0, 1)[0] == 1
