"""The float parity bounds of helpers.assert_close can fail (CPU only).

A numpy model of the 3-term BF16 split product (hi = RNE_bf16(x), lo = RNE_bf16(x - hi),
C = ah*bh + ah*bl + al*bh summed exactly and rounded once to Float32) passes the bf16x3 bound
on the shape and seed of the GPU test of the 256 x 256-tile kernel, while the same model with
one correction product dropped fails it, and the model itself fails the fp32 bound: the bound
separates the arithmetic the device is meant to run from a plausible mistake in it.

The same for the 3-plane split (hi, mid, lo; the six products a_p*b_q with p + q <= 2): summed
in FP32 over K at depth 256 it passes the bf16x6 bound, and each model that drops one of the
three 2^-16 products (a_lo*b_hi, a_mid*b_mid, a_hi*b_lo) fails it."""
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


def split3(x):
    x = np.asarray(x, np.float32)
    hi = bf16_rne(x)
    r1 = (x - hi).astype(np.float32)
    mid = bf16_rne(r1)
    lo = bf16_rne((r1 - mid).astype(np.float32))
    return [v.astype(np.float64) for v in (hi, mid, lo)]


SIX = [(2, 0), (1, 1), (0, 2), (1, 0), (0, 1), (0, 0)]


@pytest.fixture(scope="module")
def operands256():
    # depth 256: where the GPU tests of the six-product kernel pin it
    r = rng(1416)
    a = rand(r, (256, 256), np.float32)
    b = rand(r, (256, 1200), np.float32)
    return a, b, a.astype(np.float64) @ b.astype(np.float64)


def six_product_fp32(a, b, products):
    """C accumulated in FP32 over K in order, the listed products of each k added in turn."""
    A, B = split3(a), split3(b)
    acc = np.zeros((a.shape[0], b.shape[1]), np.float32)
    for k in range(a.shape[1]):
        for p, q in products:
            acc += np.outer(A[p][:, k], B[q][k]).astype(np.float32)
    return acc


def test_three_plane_split_reconstructs_to_fp32():
    r = rng(1417)
    x = rand(r, (4096,), np.float32)
    hi, mid, lo = split3(x)
    # hi + mid + lo is within 2^-24 |x| of x (Float32 has 24 bits)
    assert np.all(np.abs(hi + mid + lo - x.astype(np.float64)) <= 2.0 ** -24 * np.abs(x))
    assert np.all(np.abs(mid) <= 2.0 ** -8 * np.abs(x)) and np.all(np.abs(lo) <= 2.0 ** -16 * np.abs(x))


def test_six_product_split_passes_the_bf16x6_bound(operands256):
    a, b, ref = operands256
    c = six_product_fp32(a, b, SIX)
    assert_close(c, ref, np.float32, 256, "6-product split", precision="bf16x6", depth=256)


@pytest.mark.parametrize("dropped", [(2, 0), (1, 1), (0, 2)])
def test_five_product_split_fails_the_bf16x6_bound(operands256, dropped):
    a, b, ref = operands256
    c = six_product_fp32(a, b, [pq for pq in SIX if pq != dropped])
    assert rel_err(c, ref) > 1.5 * norm_bound("bf16x6", 256)
    with pytest.raises(AssertionError, match="norm-wise"):
        assert_close(c, ref, np.float32, 256, "5-product split", precision="bf16x6", depth=256)
