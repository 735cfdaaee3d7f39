"""Shared helpers for the parity tests: seeded inputs, device calls, oracle comparison."""
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


def assert_close(got, ref64, dtype, k_terms=1, what=""):
    """Floating point: |got - ref| <= tol * sqrt(k) * (max|ref| + 1) elementwise, with the
    tolerance written per type in F_TOL; integer/tropical: bit-exact."""
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
