"""Owen-scrambled Sobol sampler: the oracle against independent pins.

Pins: vectors written by the reference's own Python generator
(scripts/generate_luts_sobol.py: its direction-number recurrence, gray code, LK hash and
Owen scramble; golden/make_sobol_ref_golden.py), scipy's unscrambled Sobol (same Joe-Kuo
data; reference dim d is scipy dim d+1), the stratification that nested (Owen)
scrambling preserves and the reference's own range tests (SamplerTest.cpp:7-31,
math_random_generatorsTest.cpp:6-12).
"""
import os

import numpy as np
import pytest
from scipy.stats import qmc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sobol_ref_golden.npz")


def reference_vectors():
    """The fixture written from scripts/generate_luts_sobol.py (loaded without pickle)."""
    with np.load(GOLDEN) as g:
        return {k: g[k] for k in g.files}


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


def test_direction_table_equals_reference_generator(oracle):
    """The 256 x 32 table equals the header generate_luts_sobol.py writes with -d 256 -s 32
    (math/CMakeLists.txt:71-85), zero row 255 included."""
    np.testing.assert_array_equal(oracle.direction_table(), reference_vectors()["table"])


def test_scrambled_bits_equal_reference_generator(oracle):
    """Owen-scrambled integers of all 256 dimensions at index 0, the workloads' maximum
    indices, edge and random u32 indices against the reference script's gray_code, hash,
    reverse_bits and owen_scramble with seed 0xDEADBEEF (generate_luts_sobol.py:100-125).
    A changed hash constant, seed, shift or bit-reversal step fails here."""
    g = reference_vectors()
    I, D = np.meshgrid(g["index"], np.arange(256, dtype=np.uint32), indexing="ij")
    np.testing.assert_array_equal(oracle.sobol_bits(I.ravel(), D.ravel()).reshape(I.shape), g["bits"])


def test_float_samples_equal_reference_generator(oracle):
    """float32 of the script's sobol_sample (accum / 2^32 in double, :128-138) equals the
    oracle's float(accum) * 2^-32 at every nonzero index of dims 0..254."""
    g = reference_vectors()
    keep = g["index"] != 0
    I, D = np.meshgrid(g["index"][keep], np.arange(255, dtype=np.uint32), indexing="ij")
    got = oracle.sobol_float(I.ravel(), D.ravel()).reshape(I.shape)
    np.testing.assert_array_equal(got, g["value"][keep][:, :255])


def test_reference_vectors_catch_a_wrong_hash_constant():
    """The fixture separates the reference hash from a one-bit change of any constant or of
    the seed (a restatement with a typo would not match the vectors)."""
    def owen(x, seed, c1=0x3d20adea, c2=0x05526c56, c3=0x53a22864):
        def rev(v):
            return int(f"{v & 0xFFFFFFFF:032b}"[::-1], 2)
        x = rev(x)
        x = (x ^ (x * c1)) & 0xFFFFFFFF
        x = (x + seed) & 0xFFFFFFFF
        x = (x * ((seed >> 16) | 1)) & 0xFFFFFFFF
        x = (x ^ (x * c2)) & 0xFFFFFFFF
        x = (x ^ (x * c3)) & 0xFFFFFFFF
        return rev(x)
    g = reference_vectors()
    table = g["table"].astype(np.uint64)
    n = 8

    def bits_for(seed=0xDEADBEEF, **consts):
        out = np.zeros((n, 256), dtype=np.uint64)
        for r, i in enumerate(g["index"][:n].tolist()):
            gc = i ^ (i >> 1)
            for d in range(256):
                acc = 0
                for b in range(32):
                    if (gc >> b) & 1:
                        acc ^= int(table[d, b])
                out[r, d] = owen(acc, seed, **consts)
        return out
    np.testing.assert_array_equal(bits_for(), g["bits"][:n])
    for kw in ({"seed": 0xDEADBEEE}, {"c1": 0x3d20adeb}, {"c2": 0x05526c57}, {"c3": 0x53a22865}):
        assert (bits_for(**kw) != g["bits"][:n]).any(), kw


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
