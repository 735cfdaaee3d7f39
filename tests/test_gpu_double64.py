"""GPU parity of Double64 (DoubleFloats' double-word numbers, element type OME_DD) through
the C ABI against the oracle's double-word restatement.

Covers the reference's own GPU test for this type, test/roceinsum.jl:56-72 (two D = 2
networks, the second nested, against loop_einsum of the flat code), plus the GEMM-shaped
binary rule, the index walk, the unary rules of test/roceinsum.jl:74-102, scaling, permutes,
device tree plans, slicing and gradients.  The oracle's double-word arithmetic is pinned
against exact rational arithmetic in tests/test_oracle_double64.py.  Results agree to the
double-word error bound (assert_dd_close), far inside one Float64 ulp, and each comparison
also checks that plain Float64 arithmetic would have missed that bound.
"""
import numpy as np
import pytest

from helpers import assert_dd_close, parse, rand_dd, rng
from oracle import loop

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield
    import omeinsum_mi355x as om
    om.set_einsum_backend(om.DefaultBackend())


def om():
    import omeinsum_mi355x
    return omeinsum_mi355x


def dev(z):
    return om().Double64(torch.from_numpy(np.ascontiguousarray(z)).cuda())


def backends():
    return [om().DefaultBackend(), om().MI355XBackend()]


def oracle(ixs, iy, xs, sd, y0=None, sx=True, sy=False):
    y = None if y0 is None else np.array(y0, dtype=np.complex128)
    return loop.loop_einsum(ixs, iy, xs, sd, y=y, sx=sx, sy=sy, dd=True)


def magnitude(ixs, iy, xs, sd):
    """Sum over each output's terms of the product of |hi| (float64), the scale of the
    double-word error bound."""
    return np.abs(loop.loop_einsum(ixs, iy, [np.abs(x.real) for x in xs], sd))


def check(got, ref, mag, nops, what):
    assert_dd_close(got, ref, mag, nops, what)
    # the bound has teeth: Float64 arithmetic alone lands outside it somewhere
    plain = np.asarray(got).real.astype(np.complex128)
    if np.any(np.abs(np.asarray(ref).imag) > 0):
        with pytest.raises(AssertionError):
            assert_dd_close(plain, ref, mag, nops, what + " (float64 only)")


def run(code, xs, sd, backend):
    o = om()
    o.set_einsum_backend(backend)
    try:
        y = o.einsum(code, [dev(x) for x in xs], sd)
        torch.cuda.synchronize()
        return o.to_numpy(y)
    finally:
        o.set_einsum_backend(o.DefaultBackend())


def test_reference_doublefloats_networks():
    # test/roceinsum.jl:56-72: T = rand(Double64, D^6), U = rand(Double64, D^3), D = 2
    r = rng(7101)
    D = 2
    T = rand_dd(r, (D,) * 6)
    U = rand_dd(r, (D,) * 3)
    sd = {l: D for l in "abcdefghijuvwxyz"}
    for be in backends():
        ixs, iy = parse("abewcd,bfixgh,ajeycd,jfizgh->wxyz")
        xs = [T, T, T, T]
        got = run(om().ein("abewcd,bfixgh,ajeycd,jfizgh->wxyz"), xs, sd, be)
        check(got, oracle(ixs, iy, xs, sd), magnitude(ixs, iy, xs, sd), 4 * 2 ** 10 + 16,
              "network 1 %s" % type(be).__name__)
        nested = om().ein("(ubcdef,fjz),dhx,(bvghij,eiy),cgw->uvwxyz")
        ixs, iy = parse("ubcdef,fjz,dhx,bvghij,eiy,cgw->uvwxyz")
        xs = [T, U, U, T, U, U]
        got = run(nested, xs, sd, be)
        check(got, oracle(ixs, iy, xs, sd), magnitude(ixs, iy, xs, sd), 6 * 2 ** 8 + 24,
              "network 2 %s" % type(be).__name__)


BINARY = [("ij,jk->ik", {"i": 40, "j": 50, "k": 30}),          # tile GEMM
          ("ijb,jkb->ikb", {"i": 20, "j": 17, "k": 24, "b": 3}),  # batched
          ("ij,j->i", {"i": 33, "j": 70}),
          ("ij,ij->", {"i": 64, "j": 64}),                      # few outputs, long K
          ("ij,jk->ki", {"i": 5, "j": 3, "k": 7})]              # small, direct


@pytest.mark.parametrize("spec,sd", BINARY)
def test_binary_rules(spec, sd):
    ixs, iy = parse(spec)
    r = rng(7102)
    xs = [rand_dd(r, [sd[l] for l in ix]) for ix in ixs]
    k = int(np.prod([sd[l] for l in set(ixs[0]) & set(ixs[1]) - set(iy)] or [1]))
    ref = oracle(ixs, iy, xs, sd)
    mag = magnitude(ixs, iy, xs, sd)
    for be in backends():
        got = run(om().EinCode(ixs, iy), xs, sd, be)
        check(got, ref, mag, 4 * k + 8, "%s %s" % (spec, type(be).__name__))


def test_scaling_and_loop_einsum():
    o = om()
    ixs, iy = parse("ij,jk->ik")
    sd = {"i": 24, "j": 20, "k": 18}
    r = rng(7103)
    xs = [rand_dd(r, [sd[l] for l in ix]) for ix in ixs]
    y0 = rand_dd(r, [sd[l] for l in iy])
    mag = magnitude(ixs, iy, xs, sd) * 2.0 + 3.0 * np.abs(y0.real)
    for be in backends():
        o.set_einsum_backend(be)
        try:
            y = dev(y0)
            o.einsum_inplace(o.EinCode(ixs, iy), [dev(x) for x in xs], y, 2.0, 3.0, sd)
            torch.cuda.synchronize()
            got = o.to_numpy(y)
        finally:
            o.set_einsum_backend(o.DefaultBackend())
        check(got, oracle(ixs, iy, xs, sd, y0=y0, sx=2.0, sy=3.0), mag, 4 * 20 + 16,
              "scaled %s" % type(be).__name__)
    # the index walk itself (loop_einsum, src/loop_einsum.jl:8-13)
    got = o.to_numpy(o.loop_einsum(o.EinCode(ixs, iy), [dev(x) for x in xs], sd))
    check(got, oracle(ixs, iy, xs, sd), magnitude(ixs, iy, xs, sd), 4 * 20 + 8, "loop_einsum")


UNARY = ["ij->", "ij->j", "ii->", "ii->i", "ijk->kij", "a->aaaa", "ac->acc", "i->ik",
         "illljkk->kijjcc"]


@pytest.mark.parametrize("spec", UNARY)
def test_unary_rules(spec):
    # test/roceinsum.jl:74-102 with D = 2 (output-only labels sized 2)
    ixs, iy = parse(spec)
    sd = {l: 2 for ix in ixs + [iy] for l in ix}
    r = rng(7104)
    xs = [rand_dd(r, [sd[l] for l in ix]) for ix in ixs]
    ref = oracle(ixs, iy, xs, sd)
    mag = magnitude(ixs, iy, xs, sd)
    for be in backends():
        o = om()
        o.set_einsum_backend(be)
        try:
            got = o.to_numpy(o.einsum(o.EinCode(ixs, iy), [dev(x) for x in xs], sd))
        finally:
            o.set_einsum_backend(o.DefaultBackend())
        assert_dd_close(got, ref, mag, 64, "%s %s" % (spec, type(be).__name__))


def test_permute_is_exact():
    o = om()
    r = rng(7105)
    x = rand_dd(r, (5, 6, 7, 3))
    y = o.Double64(torch.empty((3, 6, 5, 7), dtype=torch.complex128, device="cuda"))
    o.tensorpermute(y, dev(x), [3, 1, 0, 2])
    torch.cuda.synchronize()
    assert np.array_equal(o.to_numpy(y), np.transpose(x, (3, 1, 0, 2)))


def test_tree_plan_sliced_and_gradient():
    o = om()
    code = o.ein("(ij,jk),(kl,lm)->im")
    sd = {"i": 9, "j": 12, "k": 10, "l": 11, "m": 8}
    ixs, iy = parse("ij,jk,kl,lm->im")
    r = rng(7106)
    xs = [rand_dd(r, [sd[l] for l in ix]) for ix in ixs]
    ref = oracle(ixs, iy, xs, sd)
    mag = magnitude(ixs, iy, xs, sd)
    sliced = o.SlicedEinsum(["j", "l"], code)
    for be in backends():
        for c in (code, sliced):
            got = run(c, xs, sd, be)
            check(got, ref, mag, 4 * 12 * 10 * 11 + 64, "%s %s" % (type(c).__name__, type(be).__name__))
    # gradient of the scalar sum(A B C D) (src/bp.jl:122-130) against the oracle's einsum_grad
    # restatement: grad_i = einsum(ixs with ixs[i] -> iy, xs with xs[i] -> ybar) -> ixs[i]
    scode = o.ein("(ij,jk),(kl,li)->")
    sxs_ix, _ = parse("ij,jk,kl,li->")
    sd2 = {"i": 6, "j": 7, "k": 5, "l": 4}
    sxs = [rand_dd(r, [sd2[l] for l in ix]) for ix in sxs_ix]
    for be in backends():
        o.set_einsum_backend(be)
        try:
            y, grads = o.cost_and_gradient(scode, [dev(x) for x in sxs])
            torch.cuda.synchronize()
            grads = [o.to_numpy(g) for g in grads]
        finally:
            o.set_einsum_backend(o.DefaultBackend())
        one = np.array(1.0 + 0j)
        for i, ix in enumerate(sxs_ix):
            gixs = [list(v) for v in sxs_ix]
            gixs[i] = []
            gxs = list(sxs)
            gxs[i] = one
            ref_g = oracle(gixs, ix, gxs, sd2)
            mag_g = np.abs(loop.loop_einsum(gixs, ix, [np.abs(x.real) for x in gxs], sd2))
            assert_dd_close(grads[i], ref_g, mag_g, 4 * 7 * 6 * 5 + 64,
                            "grad %d %s" % (i, type(be).__name__))
