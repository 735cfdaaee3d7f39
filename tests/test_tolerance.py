"""The float parity bounds of helpers.assert_close can fail (CPU only).

A numpy model of the 3-term BF16 split product (hi = RNE_bf16(x), lo = RNE_bf16(x - hi),
C = ah*bh + ah*bl + al*bh summed exactly and rounded once to Float32) passes the bf16x3 bound
on the shape and seed of the GPU test of the 256 x 256-tile kernel, while the same model with
one correction product dropped fails it, and the model itself fails the fp32 bound: the bound
separates the arithmetic the device is meant to run from a plausible mistake in it."""
import numpy as np
import pytest

from helpers import assert_close, norm_bound, rand, rel_err, rng


def bf16_rne(x):
    """Round Float32 values to the nearest BF16 (ties to even), returned as Float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split(x):
    hi = bf16_rne(x)
    lo = bf16_rne(x - hi)
    return hi.astype(np.float64), lo.astype(np.float64)


@pytest.fixture(scope="module")
def operands():
    # the GPU test test_packed_bf16x3_256_tiles: ij,jk->ik at i = k = 2900, j = 520, seed 1415
    r = rng(1415)
    a = rand(r, (2900, 520), np.float32)
    b = rand(r, (520, 2900), np.float32)
    ref = a.astype(np.float64) @ b.astype(np.float64)
    return a, b, ref


def test_split_model_is_exact_where_it_should_be():
    # hi + lo reproduces a value with at most 16 significant bits exactly, and BF16 rounding
    # is to nearest with ties to even
    x = np.float32(1.0 + 2.0 ** -7 + 2.0 ** -15)
    hi, lo = split(np.array([x]))
    assert hi[0] + lo[0] == float(x)
    assert bf16_rne(np.array([1.0 + 2.0 ** -8], np.float32))[0] == 1.0            # tie -> even
    assert bf16_rne(np.array([1.0 + 3 * 2.0 ** -8], np.float32))[0] == 1.0 + 2.0 ** -6


def test_three_term_split_passes_the_bf16x3_bound(operands):
    a, b, ref = operands
    (ah, al), (bh, bl) = split(a), split(b)
    c = (ah @ bh + ah @ bl + al @ bh).astype(np.float32)
    e = rel_err(c, ref)
    assert e <= norm_bound("bf16x3", 520)
    assert_close(c, ref, np.float32, 16 * 520, "3-term split", precision="bf16x3", depth=520)
    # ... and is not FP32 arithmetic: the fp32 bound rejects it
    assert e > norm_bound("fp32", 520)


@pytest.mark.parametrize("dropped", ["ah*bl", "al*bh"])
def test_two_term_split_fails_the_bf16x3_bound(operands, dropped):
    a, b, ref = operands
    (ah, al), (bh, bl) = split(a), split(b)
    c = ah @ bh + (al @ bh if dropped == "ah*bl" else ah @ bl)
    c = c.astype(np.float32)
    assert rel_err(c, ref) > 10 * norm_bound("bf16x3", 520)
    with pytest.raises(AssertionError, match="norm-wise"):
        assert_close(c, ref, np.float32, 16 * 520, "2-term split", precision="bf16x3", depth=520)


def test_fp32_rounding_passes_the_fp32_bound(operands):
    # FP32 operands, products summed in FP32 in order over K: within the fp32 bound
    a, b, ref = operands
    rows = slice(0, 64)
    acc = np.zeros((64, 2900), np.float32)
    for k in range(520):
        acc += np.outer(a[rows, k], b[k]).astype(np.float32)
    assert_close(acc, ref[rows], np.float32, 520, "fp32 sequential", precision="fp32", depth=520)
    # a dropped K-slice is caught by the norm-wise check even where the elementwise one passes
    bad = acc - np.outer(a[rows, 519], b[519]).astype(np.float32) * np.float32(1e-3)
    with pytest.raises(AssertionError, match="norm-wise"):
        assert_close(bad, ref[rows], np.float32, 16 * 520, "perturbed", precision="fp32", depth=520)
