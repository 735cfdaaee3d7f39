"""Shared helpers for the parity tests: seeded inputs, device calls, oracle comparison."""
import json
import os

import numpy as np

F_TOL = {np.float32: 2e-5, np.complex64: 2e-5, np.float64: 1e-12, np.complex128: 1e-12}


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def rand(r, shape, dtype):
    shape = tuple(int(s) for s in shape)
    dtype = np.dtype(dtype)
    if dtype == np.int64:
        return np.asarray(r.integers(-8, 9, size=shape), dtype=np.int64)
    if dtype.kind == "c":
        re = r.uniform(-1, 1, size=shape)
        im = r.uniform(-1, 1, size=shape)
        return np.asarray(re + 1j * im, dtype=dtype)
    return np.asarray(r.uniform(-1, 1, size=shape), dtype=dtype)


def parse(spec):
    l, r_ = spec.replace(" ", "").split("->")
    return [list(w) for w in l.split(",")], list(r_)


def wide(dtype):
    return {np.dtype(np.float32): np.float64, np.dtype(np.complex64): np.complex128}.get(
        np.dtype(dtype), dtype)


# Norm-wise bound ||got - ref||_F / ||ref||_F for Float32 / ComplexF32 results, by the
# arithmetic the product ran in (the ``precision`` argument of assert_close):
#   fp32    FP32 operands and FP32 accumulation: the accumulation error of a depth-d sum of
#           random-sign products is about u*sqrt(d/6) relative (u = 2^-24): 1.6e-6 at d = 4096.
#           Bound 2e-6, growing as sqrt(d) beyond d = 4096.
#   3m_*    Gauss's 3-product complex form: the imaginary part is (ar+ai)(br+bi) - ar*br - ai*bi,
#           whose terms are up to twice the result's size: 3x the fp32 bound.
#   bf16x3  3-term BF16 split: each operand keeps ~16 bits (x - hi - lo ~ 2^-17 |x|), so every
#           product errs by ~2^-16.5 relative whatever the depth; 1e-5 up to d = 65536.  A split
#           that drops one correction product errs by ~1e-3 and fails (tests/test_tolerance.py).
#   bf16x6  3-plane BF16 split, six products: every operand keeps ~24 bits and the dropped
#           products are below 2^-23 relative, so what remains is the FP32 accumulation
#           (u*sqrt(d/6), 4e-7 at d = 256).  Bound 1e-6 up to d = 256, growing as sqrt(d):
#           a kernel that drops one of the three 2^-16 products errs by ~2.2e-6 whatever the
#           depth and fails it at small depth (tests/test_tolerance.py).
#   fp64    FP64 products rounded once to Float32: the fp32 bound.
# d counts the real products summed into one output component: K for a real product, 2K for a
# complex one (the realified GEMM's depth).
NORM_TOL = {"fp32": 2e-6, "fp64": 2e-6, "3m_fp32": 6e-6, "3m_fp64": 6e-6, "bf16x3": 1e-5,
            "3m_bf16x3": 3e-5, "bf16x6": 1e-6}


def norm_bound(precision, depth, complex_=False):
    d = max(1.0, float(depth)) * (2.0 if complex_ else 1.0)
    if precision.endswith("bf16x3"):
        return NORM_TOL[precision] * max(1.0, np.sqrt(d / 65536.0))
    if precision == "bf16x6":
        return NORM_TOL[precision] * max(1.0, np.sqrt(d / 256.0))
    return NORM_TOL[precision] * max(1.0, np.sqrt(d / 4096.0))


def _log_error(what, dtype, precision, depth, err, bound):
    path = os.environ.get("OME_TEST_ERRLOG")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"test": os.environ.get("PYTEST_CURRENT_TEST", "?"), "what": str(what),
                                "dtype": np.dtype(dtype).name, "precision": precision,
                                "depth": float(depth), "rel_err": err, "bound": bound}) + "\n")


def assert_close(got, ref64, dtype, k_terms=1, what="", precision="fp32", depth=None, norm=True):
    """Floating point: |got - ref| <= tol * sqrt(k) * (max|ref| + 1) elementwise, with the
    tolerance written per type in F_TOL; for Float32 / ComplexF32 also the norm-wise bound
    ||got - ref||_F / ||ref||_F <= norm_bound(precision, depth) (depth defaults to k_terms;
    norm=False leaves it out for a result whose size is set by cancellation, such as a sum of
    random-sign terms to one scalar); integer/tropical: bit-exact."""
    got = np.asarray(got)
    ref64 = np.asarray(ref64)
    assert got.shape == ref64.shape, (what, got.shape, ref64.shape)
    dt = np.dtype(dtype)
    if dt.kind in "iu":
        assert np.array_equal(got, ref64), what
        return
    tol = F_TOL[dt.type] * max(1.0, np.sqrt(k_terms))
    scale = np.max(np.abs(ref64)) if ref64.size else 0.0
    err = np.max(np.abs(got.astype(wide(dt)) - ref64)) if ref64.size else 0.0
    assert err <= tol * (scale + 1.0), "%s: max err %.3e > %.3e (scale %.3e)" % (what, err, tol * (scale + 1), scale)
    if norm and dt.type in (np.float32, np.complex64) and ref64.size and np.any(ref64 != 0):
        d = k_terms if depth is None else depth
        e = rel_err(got, ref64)
        bound = norm_bound(precision, d, dt.kind == "c")
        _log_error(what, dt, precision, d, e, bound)
        assert e <= bound, "%s: norm-wise error %.3e > %.3e (%s, depth %g)" % (what, e, bound, precision, d)


def rel_err(got, ref64):
    got = np.asarray(got).astype(np.complex128)
    ref64 = np.asarray(ref64).astype(np.complex128)
    n = np.linalg.norm(ref64.ravel())
    return float(np.linalg.norm((got - ref64).ravel()) / (n if n > 0 else 1.0))


# ---- double-word (Double64) inputs and tolerance ---------------------------------------

def rand_dd(r, shape, spread=0):
    """Double-words hi + i*lo (complex128 storage): hi uniform in +-[0.5, 1) * 2^e with e in
    [-spread, spread], lo a random fraction of half an ulp of hi."""
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape)) if shape else 1
    hi = r.uniform(0.5, 1.0, n) * np.where(r.random(n) < 0.5, -1.0, 1.0)
    hi = hi * 2.0 ** r.integers(-spread, spread + 1, n)
    lo = hi * r.uniform(-1.0, 1.0, n) * 2.0 ** -54
    s = hi + lo
    lo = lo - (s - hi)
    return (s + 1j * lo).reshape(shape)


def dd_value(z):
    """hi + lo of complex128 double-word storage, as exact Python Fractions (object array)."""
    from fractions import Fraction
    z = np.asarray(z)
    out = np.empty(z.shape, dtype=object)
    for idx in np.ndindex(z.shape):
        out[idx] = Fraction(float(z[idx].real)) + Fraction(float(z[idx].imag))
    return out


def assert_dd_close(got, ref, absmag, n_ops, what=""):
    """Double-word results: |got - ref| <= n_ops * 2^-100 * (sum of |terms|) per element,
    where absmag is that sum (float64) and n_ops bounds the operations in one output's chain
    (each double-word operation errs by at most 4u^2 = 2^-104 of its magnitude)."""
    from fractions import Fraction
    got, ref, absmag = np.asarray(got), np.asarray(ref), np.asarray(absmag, dtype=np.float64)
    assert got.shape == ref.shape == absmag.shape, (what, got.shape, ref.shape, absmag.shape)
    g, f = dd_value(got), dd_value(ref)
    tol = Fraction(n_ops, 2 ** 100)
    for idx in np.ndindex(got.shape):
        bound = tol * Fraction(float(absmag[idx])) * (1 + Fraction(1, 2 ** 40))
        assert abs(g[idx] - f[idx]) <= bound, "%s at %s: |err| %.3e > %.3e" % (
            what, idx, float(abs(g[idx] - f[idx])), float(bound))
