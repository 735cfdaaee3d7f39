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
