"""Pins the oracle's double-word (Double64) arithmetic on the CPU.

DoubleFloats.jl, which provides the Double64 of the reference's GPU test
(test/roceinsum.jl:56-72), is a test dependency that is not part of the reference tree, so
the oracle restates its + and * as the published double-word algorithms of Joldes, Muller
and Popescu (ACM TOMS 44(2), 2017): Algorithm 6 (AccurateDWPlusDW, relative error at most
3u^2 + 13u^3) and Algorithm 12 (DWTimesDW3, at most 4u^2), u = 2^-53.  These tests check the
oracle against exact rational arithmetic (fractions.Fraction): the error bounds, results
that are exact double-words, normalisation of every result, and whole contractions against
an exact brute-force sum.  No GPU.
"""
import itertools
from fractions import Fraction as Fr

import numpy as np

from oracle import loop

U = Fr(1, 2 ** 53)


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def normalize(hi, lo):
    """Fast2Sum: the double-word nearest to hi + lo with |lo| <= ulp(hi) / 2."""
    s = hi + lo
    return s, lo - (s - hi)


def rand_dd(r, n, spread=20):
    """n double-words hi + i*lo: random sign and exponent of hi, lo a random fraction of
    half an ulp of hi."""
    hi = r.uniform(0.5, 1.0, n) * np.where(r.random(n) < 0.5, -1.0, 1.0)
    hi = hi * 2.0 ** r.integers(-spread, spread + 1, n)
    lo = hi * r.uniform(-1.0, 1.0, n) * 2.0 ** -54
    h, l = normalize(hi, lo)
    return h + 1j * l


def exact(z):
    return Fr(float(z.real)) + Fr(float(z.imag))


def normalized(z):
    return float(z.real) + float(z.imag) == float(z.real)


def check_bound(op, x, y, z, bound):
    for a, b, c in zip(x, y, z):
        e = exact(a) + exact(b) if op == "+" else exact(a) * exact(b)
        assert abs(exact(c) - e) <= bound * abs(e), (op, a, b, c)
        assert normalized(c), (op, a, b, c)


def test_dw_add_error_bound():
    r = rng(7001)
    x = rand_dd(r, 4000)
    y = rand_dd(r, 4000)
    # near-cancellation: y close to -x
    y[:1000] = -x[:1000] * (1 + r.uniform(-1, 1, 1000) * 2.0 ** -40)
    y[:1000] = np.array([complex(*normalize(v.real, v.imag)) for v in y[:1000]])
    check_bound("+", x, y, loop.dw_op("+", x, y), 3 * U ** 2 + 13 * U ** 3)


def test_dw_mul_error_bound():
    r = rng(7002)
    x = rand_dd(r, 4000)
    y = rand_dd(r, 4000)
    check_bound("*", x, y, loop.dw_op("*", x, y), 4 * U ** 2)


def test_dw_exact_cases():
    t = 2.0 ** -52
    cases = [
        # (x, y, op, exact double-word result)
        ((1 + t, 0.0), (1 + t, 0.0), "*", (1 + 2 * t, 2.0 ** -104)),   # (1 + 2^-52)^2
        ((3.0, 0.0), (5.0, 0.0), "*", (15.0, 0.0)),
        ((1.0, 2.0 ** -60), (1.0, 0.0), "*", (1.0, 2.0 ** -60)),          # one
        ((0.75, 2.0 ** -57), (0.0, 0.0), "+", (0.75, 2.0 ** -57)),        # zero
        ((0.75, 2.0 ** -57), (-0.75, -2.0 ** -57), "+", (0.0, 0.0)),      # cancellation
        ((1.0, 0.0), (2.0 ** -60, 0.0), "+", (1.0, 2.0 ** -60)),          # tail kept in lo
        # 1 + 2^-53 + 2^-60 is above the midpoint: hi rounds up, lo is negative
        ((1.0, 2.0 ** -60), (2.0 ** -53, 0.0), "+", (1 + t, -2.0 ** -53 + 2.0 ** -60)),
        ((-2.0, 0.0), (1 + t, 2.0 ** -70), "*", (-2 - 2 * t, -2.0 ** -69)),
    ]
    for x, y, op, want in cases:
        got = loop.dw_op(op, np.array([complex(*x)]), np.array([complex(*y)]))[0]
        assert (got.real, got.imag) == want, (x, op, y, got)
        assert exact(np.complex128(complex(*want))) == (
            (Fr(x[0]) + Fr(x[1])) + (Fr(y[0]) + Fr(y[1])) if op == "+" else
            (Fr(x[0]) + Fr(x[1])) * (Fr(y[0]) + Fr(y[1])))


def naive_dw_add(x, y):
    """A plausible mistake: componentwise sums, one renormalisation, no error terms."""
    h, l = normalize(x.real + y.real, x.imag + y.imag)
    return np.complex128(h + 1j * l)


def test_bound_check_catches_a_naive_add():
    r = rng(7003)
    x = rand_dd(r, 200, spread=2)
    y = rand_dd(r, 200, spread=2)
    bad = np.array([naive_dw_add(a, b) for a, b in zip(x, y)])
    bound = 3 * U ** 2 + 13 * U ** 3
    assert any(abs(exact(c) - (exact(a) + exact(b))) > bound * abs(exact(a) + exact(b))
               for a, b, c in zip(x, y, bad))


def exact_einsum(ixs, iy, xs, sizes):
    """Brute-force rational sum over every index assignment (column-major offsets)."""
    labels = []
    for ix in list(ixs) + [iy]:
        for l in ix:
            if l not in labels:
                labels.append(l)
    out = {}
    for vals in itertools.product(*[range(sizes[l]) for l in labels]):
        a = dict(zip(labels, vals))
        p = Fr(1)
        mag = Fr(1)
        for ix, x in zip(ixs, xs):
            v = exact(x[tuple(a[l] for l in ix)])
            p *= v
            mag *= abs(v)
        key = tuple(a[l] for l in iy)
        s, m = out.get(key, (Fr(0), Fr(0)))
        out[key] = (s + p, m + mag)
    return out


def check_contraction(ixs, iy, xs, sizes, terms_per_output):
    got = loop.loop_einsum(ixs, iy, xs, sizes, dd=True)
    ref = exact_einsum(ixs, iy, xs, sizes)
    nfac = len(ixs)
    # each product carries (nfac - 1) products of error <= 4u^2, each sum step <= 3u^2 + 13u^3
    # of the running sum; bound everything by the sum of |terms|
    c = (4 * (nfac - 1) + 4 * terms_per_output) * U ** 2
    for key, (s, m) in ref.items():
        v = got[key] if key else got[()]
        assert abs(exact(v) - s) <= c * m, (key, v, float(s))
        assert normalized(v)


def test_loop_einsum_dd_matmul_exact():
    r = rng(7004)
    sizes = {"i": 3, "j": 6, "k": 4}
    a = rand_dd(r, 18, spread=3).reshape(3, 6)
    b = rand_dd(r, 24, spread=3).reshape(6, 4)
    check_contraction([["i", "j"], ["j", "k"]], ["i", "k"], [a, b], sizes, 6)


def test_loop_einsum_dd_reference_network():
    # the first code of the doublefloats testset (test/roceinsum.jl:56-66), D = 2
    r = rng(7005)
    D = 2
    code = "abewcd,bfixgh,ajeycd,jfizgh->wxyz"
    lhs, iy = code.split("->")
    ixs = [list(w) for w in lhs.split(",")]
    xs = [rand_dd(r, D ** 6, spread=0).reshape((D,) * 6) for _ in range(4)]
    sizes = {l: D for ix in ixs for l in ix}
    inner = len(set(l for ix in ixs for l in ix) - set(iy))
    check_contraction(ixs, list(iy), xs, sizes, D ** inner)
