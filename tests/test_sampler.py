"""Owen-scrambled Sobol sampler: the oracle against independent pins.

Pins: scipy's unscrambled Sobol (same Joe-Kuo data; reference dim d is scipy dim d+1),
the stratification that nested (Owen) scrambling preserves, the reference's own range
tests (SamplerTest.cpp:7-31, math_random_generatorsTest.cpp:6-12) and the committed
golden vectors.
"""
import json
import os

import numpy as np
import pytest
from scipy.stats import qmc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sobol_golden.json")


def unscrambled(table, idx, dim):
    g = np.asarray(idx, dtype=np.uint64) ^ (np.asarray(idx, dtype=np.uint64) >> np.uint64(1))
    acc = np.zeros(len(g), dtype=np.uint64)
    for b in range(32):
        acc ^= np.where((g >> np.uint64(b)) & np.uint64(1), np.uint64(table[dim, b]), np.uint64(0))
    return acc


def test_direction_numbers_match_scipy(oracle):
    """All 32 bits of every direction number: reference dim d = scipy dim d + 1."""
    table = oracle.direction_table()
    sv = qmc.Sobol(256, scramble=False, bits=32)._sv
    np.testing.assert_array_equal(table[:255], sv[1:256])


@pytest.mark.parametrize("start", [0, 1 << 16])
def test_unscrambled_points_match_scipy(oracle, start):
    table = oracle.direction_table()
    n = 64
    s = qmc.Sobol(256, scramble=False, bits=32)
    if start:
        s.fast_forward(start)
    pts = np.round(s.random(n) * 2.0 ** 32).astype(np.uint64)
    idx = np.arange(start, start + n)
    for d in range(255):
        np.testing.assert_array_equal(unscrambled(table, idx, d), pts[:, d + 1], err_msg=f"dim {d}")


def test_last_direction_row_is_zero(oracle):
    table = oracle.direction_table()
    assert not table[255].any() and table[:255].all(axis=1).all()


@pytest.mark.parametrize("m", [4, 8, 12])
def test_owen_scrambled_points_stay_stratified(oracle, m):
    """For 2^m aligned indices, every dimension hits each [k/2^m,(k+1)/2^m) exactly once."""
    base = 7 << m
    idx = np.arange(base, base + (1 << m), dtype=np.uint32)
    for d in list(range(0, 40)) + [100, 254]:
        bits = oracle.sobol_bits(idx, d)
        strata = np.sort(bits >> np.uint32(32 - m))
        np.testing.assert_array_equal(strata, np.arange(1 << m), err_msg=f"dim {d}")


def test_float_conversion_is_exact_2_pow_minus_32(oracle):
    idx = np.arange(0, 4096, dtype=np.uint32)
    for d in (0, 1, 7, 200):
        bits = oracle.sobol_bits(idx, d)
        f = oracle.sobol_float(idx, d)
        np.testing.assert_array_equal(f, (bits.astype(np.float32) * np.float32(2.0 ** -32)))


def test_reference_range_tests(oracle):
    # SobolSamplerTest.sampler3D: indices 0..99, dims 0..2 in [-1, 1), not all zero
    idx = np.repeat(np.arange(100, dtype=np.uint32), 3)
    dims = np.tile(np.arange(3, dtype=np.uint32), 100)
    v = oracle.sobol_float(idx, dims)
    assert (v >= -1).all() and (v < 1).all() and (v != 0).any()
    # math_random_generatorsTest.generate: dim 0, indices 0..9 in [0, 1)
    v = oracle.sobol_float(np.arange(10, dtype=np.uint32), 0)
    assert (v >= 0).all() and (v < 1).all()


def test_golden_vectors(oracle):
    g = json.load(open(GOLDEN))
    idx = np.array(g["index"], dtype=np.uint32)
    dims = np.array(g["dim"], dtype=np.uint32)
    np.testing.assert_array_equal(oracle.sobol_bits(idx, dims), np.array(g["bits"], dtype=np.uint32))


def test_acos_table_closed_form(oracle):
    """generate_luts_trigonometrics.py:11-20: (float)acos((i/P)*2-1) with P = 500000."""
    import math
    lut = oracle.acos_lut()
    i = np.concatenate([np.arange(0, 500000, 997), [0, 1, 249999, 250000, 499999]])
    expect = np.array([math.acos((k / 500000) * 2 - 1) for k in i], dtype=np.float32)
    np.testing.assert_array_equal(lut[i], expect)


def test_atan2_approx_error_bound_and_quadrants(oracle):
    rng = np.random.default_rng(3)
    y = rng.uniform(-5, 5, 20000).astype(np.float32)
    x = rng.uniform(-5, 5, 20000).astype(np.float32)
    a = oracle.atan2_approx(y, x)
    # the reference formula (math_utils_approx.h:24-40) maps x<0,y<0 to atan2 - 0 and y<0,x>=0 stays negative
    assert np.max(np.abs(a - np.arctan2(y, x))) < 2e-5


def test_sample_index_policies(oracle):
    """CPU policy (Integrator.h:38-42,80): idx = 4*pixel, pixel row flipped by vertical_invert,
    times the frame plus renderer_max_samples; GPU raygen policy (integrator_optix.cu:43):
    (x*H + y)*sample_max + frame. Pinned by properties a wrong term or index would break:
    under the GPU policy the indices of all pixels and frames 0..S-1 are exactly
    0..W*H*S-1 (W != H, so swapping the row and column strides fails), and a pixel's S
    frames are consecutive; under the CPU policy frame 0 gives every pixel the offset, and
    pixel indices are multiples of 4 whose rows swap under vertical_invert; u32 wraparound."""
    W, H, S = 7, 5, 6
    idx = np.array([[[oracle.sample_index(W, H, x, y, f, offset=S, policy=1) for f in range(S)] for y in range(H)] for x in range(W)])
    assert sorted(idx.ravel().tolist()) == list(range(W * H * S))
    assert np.all(np.diff(idx, axis=2) == 1)
    assert all(oracle.sample_index(W, H, x, y, 0, offset=17) == 17 for x in range(W) for y in range(H))
    one = np.array([[oracle.sample_index(W, H, x, y, 1) for x in range(W)] for y in range(H)])
    assert np.all(one % 4 == 0) and sorted((one // 4).ravel().tolist()) == list(range(W * H))
    flip = np.array([[oracle.sample_index(W, H, x, y, 1, vertical_invert=True) for x in range(W)] for y in range(H)])
    np.testing.assert_array_equal(flip, one[::-1])
    assert oracle.sample_index(1920, 1080, 1919, 1079, 600) == (4 * 2073599 * 600) % (1 << 32)
    assert oracle.sample_index(1920, 1080, 1919, 1079, 600, offset=1 << 20, policy=1) == ((1919 * 1080 + 1079) * (1 << 20) + 600) % (1 << 32)
