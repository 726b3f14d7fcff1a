"""Pins the oracle's DDF storage codecs (src/lbm/domain.rs:773-784,
sim_kernels.cl:79-90) against independent references: numpy's IEEE binary16
conversion, the known answers written next to the reference codecs, and a
brute-force nearest-code search."""
import numpy as np

import oracle_lib as O

ALL = np.arange(65536, dtype=np.uint16)


def test_half_encode_matches_numpy_rne():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(3_000_000) * np.exp2(rng.integers(-30, 18, 3_000_000))).astype(np.float32)
    edges = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, -65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                      2.0 ** -14, 2.0 ** -14 * (1 - 2 ** -11), np.inf, -np.inf], np.float32)
    x = np.concatenate([x, edges])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(O.half_encode_rte(x), ref)


def test_half_decode_matches_numpy_all_codes():
    got = O.half_decode(ALL)
    ref = ALL.view(np.float16).astype(np.float32)
    finite = ~np.isnan(ref)
    assert np.array_equal(got[finite].view(np.uint32), ref[finite].view(np.uint32))
    assert np.isnan(got[~finite]).all()


def test_half_nan_is_canonical():
    x = np.array([np.nan, -np.nan], np.float32)
    assert list(O.half_encode_rte(x)) == [0x7E00, 0x7E00]


def test_fp16s_scale_is_exact_power_of_two():
    # the FP16S load factor 3.0517578E-5f of domain.rs:775 is 2^-15 exactly
    assert np.float32(3.0517578e-5) == np.float32(2.0 ** -15)
    # hence store(load(h)) == h for every non-NaN code
    vals = O.fp16s_load(ALL)
    ok = ~np.isnan(vals)
    assert np.array_equal(O.fp16s_store(vals[ok]), ALL[ok])


def test_fp16s_store_is_the_ieee_half_of_x_times_2_to_15():
    """vstore_half_rte(x * 32768) (domain.rs:773-784) against numpy's IEEE
    round-to-nearest-even conversion of the same exact product, on values
    across the range and on the exact midpoints between adjacent codes, where
    the rounding direction (and any change of the scale) shows."""
    rng = np.random.default_rng(11)
    x = (rng.standard_normal(1 << 18) * np.exp2(rng.integers(-40, 2, 1 << 18))).astype(np.float32)
    codes = rng.integers(0, 0x7BFF, 1 << 16).astype(np.uint16)
    lo = codes.view(np.float16).astype(np.float64)
    hi = (codes + 1).astype(np.uint16).view(np.float16).astype(np.float64)
    mid = ((lo + hi) / 2 / 32768.0).astype(np.float32)  # exact: 12 significant bits
    assert np.array_equal(mid.astype(np.float64) * 32768.0, (lo + hi) / 2)
    for v in (x, mid, -mid):
        with np.errstate(over="ignore"):  # beyond the half range both give the infinity code
            ref = (v * np.float32(32768.0)).astype(np.float16).view(np.uint16)
        assert np.array_equal(O.fp16s_store(v), ref)
    # ties go to the even code
    assert np.all(O.fp16s_store(mid) % 2 == 0)


def test_fp16s_known_answers():
    # domain.rs:773: "+-1.99902344, +-1.86446416E-9, +-1.81898936E-12" (max, smallest
    # normal, smallest denormal). The middle value of the comment is one ulp
    # above the smallest normal 2^-29 = 1.8626451E-9: it is code 0x0401.
    v = O.fp16s_load(np.array([0x7BFF, 0x0401, 0x0001], np.uint16))
    np.testing.assert_allclose(v, [1.99902344, 1.86446416e-9, 1.81898936e-12], rtol=5e-8)  # comment digits are truncated
    assert O.fp16s_load(np.array([0x0400], np.uint16))[0] == np.float32(2.0 ** -29)


def test_fp16c_known_answers():
    # sim_kernels.cl:79: "+-1.99951168, +-6.10351562E-5, +-2.98023224E-8"
    v = O.fp16c_decode(np.array([0x7FFF, 0x0800, 0x0001, 0xFFFF, 0x8800, 0x8001], np.uint16))
    np.testing.assert_allclose(v, [1.99951168, 6.10351562e-5, 2.98023224e-8, -1.99951168, -6.10351562e-5,
                                   -2.98023224e-8], rtol=5e-8)  # comment digits are truncated


def test_fp16c_decode_encode_identity_all_codes():
    assert np.array_equal(O.fp16c_encode(O.fp16c_decode(ALL)), ALL)


def test_fp16c_decode_monotonic():
    pos = O.fp16c_decode(ALL[:0x8000])
    assert np.all(np.diff(pos.astype(np.float64)) > 0)


def test_fp16c_encode_is_nearest_with_ties_away_from_zero():
    """Brute force: every encoded value is the nearest FP16C value; exact
    midpoints round away from zero (the +0x800 rounding of sim_kernels.cl:80)."""
    table = O.fp16c_decode(ALL[:0x8000]).astype(np.float64)  # ascending, 0 .. 1.9995
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1.9995, 400_000) * np.exp2(-rng.integers(0, 27, 400_000))
    mids = (table[1:] + table[:-1]) / 2
    x = np.concatenate([x, mids[::7]]).astype(np.float32).astype(np.float64)
    k = np.searchsorted(table, x)
    k = np.clip(k, 1, len(table) - 1)
    lo, hi = table[k - 1], table[k]
    nearest = np.where(x - lo < hi - x, k - 1, k)  # ties -> the larger magnitude
    got = O.fp16c_encode(x.astype(np.float32))
    below = x < table[1] / 2  # below half the smallest denormal: flushes to 0
    nearest = np.where(below, 0, nearest)
    assert np.array_equal(got.astype(np.int64), nearest)
    # odd symmetry
    assert np.array_equal(O.fp16c_encode(-x.astype(np.float32)), got | 0x8000)


def test_fp16c_nan_encodes_as_positive_nan_pattern():
    assert list(O.fp16c_encode(np.array([np.nan, -np.nan], np.float32))) == [0x7C00, 0x7C00]
