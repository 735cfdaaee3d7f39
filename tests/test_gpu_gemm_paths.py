"""GPU parity of the GEMM routes the planner picks by shape and precision, through the C ABI
against the oracle: complex products as real matrix-core GEMMs (4M realified, 3M Gauss),
Float32 at FP64 precision (widened operands, one rounding), and the few-outputs / long-K
inner products."""
import numpy as np
import pytest

from helpers import assert_close, parse, rand, rng, wide
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


def dev(x, tropical=False):
    t = torch.from_numpy(np.array(x, order="C")).cuda()
    return om().Tropical(t) if tropical else t


def run(ixs, iy, xs, sd, backend, y0=None, sx=True, sy=False, tropical=False):
    o = om()
    o.set_einsum_backend(backend)
    try:
        code = o.EinCode(ixs, iy)
        dxs = [dev(x, tropical) for x in xs]
        if y0 is None:
            y = o.einsum(code, dxs, sd)
        else:
            y = dev(y0)
            o.einsum_inplace(code, dxs, y, sx, sy, sd)
        torch.cuda.synchronize()
        return o.to_numpy(y)
    finally:
        o.set_einsum_backend(o.DefaultBackend())


def oracle(ixs, iy, xs, sd, y0=None, sx=True, sy=False, tropical=False):
    T = np.result_type(*[x.dtype for x in xs])
    W = T if tropical else wide(T)
    y = None if y0 is None else np.asarray(y0, dtype=W)
    return loop.loop_einsum(ixs, iy, [np.asarray(x, dtype=W) for x in xs], sd, y=y, sx=sx,
                            sy=sy, tropical=tropical)


SPECS = ["ij,jk->ik", "ji,kj->ik", "ji,jk->ik", "ij,kj->ik", "ijb,jkb->ikb", "ajb,bkj->bka",
         "ij,jk->ki"]
SIZES = {"i": 70, "j": 90, "k": 65, "a": 33, "b": 3}


@pytest.mark.parametrize("spec", SPECS)
@pytest.mark.parametrize("algo", ["4m", "3m"])
@pytest.mark.parametrize("dtype,prec", [(np.complex64, "fp32"), (np.complex64, "bf16x3"),
                                        (np.complex64, "fp64"), (np.complex128, "fp32")])
def test_complex_as_real_gemm(spec, algo, dtype, prec):
    # M*N*K >= 2^16 with M, 2N >= 32: the planner takes the real-GEMM route
    ixs, iy = parse(spec)
    r = rng(2101)
    xs = [rand(r, [SIZES[l] for l in ix], dtype) for ix in ixs]
    y0 = rand(r, [SIZES[l] for l in iy], dtype)
    be = om().MI355XBackend(precision=prec, complex_algo=algo)
    got = run(ixs, iy, xs, SIZES, be)
    ref = oracle(ixs, iy, xs, SIZES)
    k = SIZES["j"]
    # 3M forms (ar+ai)(br+bi) - ar*br - ai*bi: operands up to twice as large, three terms
    nprec = ("3m_" if algo == "3m" else "") + prec
    assert_close(got, ref, dtype, 4 * k * (3 if algo == "3m" else 1), spec, precision=nprec, depth=k)
    sx, sy = 1.5 - 0.5j, 0.25 + 1j
    got = run(ixs, iy, xs, SIZES, be, y0=y0, sx=sx, sy=sy)
    ref = oracle(ixs, iy, xs, SIZES, y0=y0, sx=sx, sy=sy)
    assert_close(got, ref, dtype, 16 * k * (3 if algo == "3m" else 1), spec + " scaled",
                 precision=nprec, depth=k)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_complex_split_k_and_edges(prec):
    # one 128 x 64 output tile and a long K: the ComplexF32 matrix-core kernel splits K and
    # sums the partials in a second pass; odd sizes leave partial tiles and a partial K-step
    be = om().MI355XBackend(precision=prec)
    r = rng(2102)
    for spec, sd in [("ij,jk->ik", {"i": 40, "j": 3001, "k": 24}),
                     ("ji,kj->ki", {"i": 129, "j": 77, "k": 65}),
                     ("ijb,jkb->ikb", {"i": 33, "j": 250, "k": 17, "b": 5})]:
        ixs, iy = parse(spec)
        xs = [rand(r, [sd[l] for l in ix], np.complex64) for ix in ixs]
        got = run(ixs, iy, xs, sd, be)
        assert_close(got, oracle(ixs, iy, xs, sd), np.complex64, 4 * sd["j"], spec,
                     precision=prec, depth=sd["j"])
        y0 = rand(r, [sd[l] for l in iy], np.complex64)
        got = run(ixs, iy, xs, sd, be, y0=y0, sx=0.5 + 2j, sy=-1 + 0.5j)
        ref = oracle(ixs, iy, xs, sd, y0=y0, sx=0.5 + 2j, sy=-1 + 0.5j)
        assert_close(got, ref, np.complex64, 4 * sd["j"] + 8, spec + " scaled",
                     precision=prec, depth=sd["j"])


def scattered(r, m_sizes, k_sizes, n_sizes):
    """A binary code whose M, K and N labels are interleaved at random in both operands and in
    the output (no operand is a GEMM-ready run of groups)."""
    ml = list(range(len(m_sizes)))
    kl = [100 + q for q in range(len(k_sizes))]
    nl = [200 + q for q in range(len(n_sizes))]
    sd = dict(zip(ml, m_sizes)) | dict(zip(kl, k_sizes)) | dict(zip(nl, n_sizes))
    a = [int(x) for x in r.permutation(ml + kl)]
    b = [int(x) for x in r.permutation(kl + nl)]
    y = [int(x) for x in r.permutation(ml + nl)]
    return [a, b], y, sd


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_complex_scattered_labels(prec):
    # ComplexF32 products whose labels are scattered over both operands and the output (the
    # tensor-network case): network-like size-2 labels, mixed sizes, a long K (split), a short
    # K (other routes), and scaled accumulation into y
    be = om().MI355XBackend(precision=prec)
    r = rng(2103)
    cases = [([2] * 6, [2] * 5, [2] * 5),          # M 64, K 32, N 32
             ([8, 9], [3, 4, 5], [5, 6]),           # M 72, K 60, N 30
             ([2] * 5, [2] * 12, [2] * 4),          # M 32, K 4096, N 16: split K
             ([3, 5, 7], [2, 3], [4, 4, 2])]        # M 105, K 6
    for ms, ks, ns in cases:
        ixs, iy, sd = scattered(r, ms, ks, ns)
        xs = [rand(r, [sd[l] for l in ix], np.complex64) for ix in ixs]
        k = int(np.prod(ks))
        got = run(ixs, iy, xs, sd, be)
        assert_close(got, oracle(ixs, iy, xs, sd), np.complex64, 4 * k, str((ms, ks, ns)),
                     precision=prec, depth=k)
        y0 = rand(r, [sd[l] for l in iy], np.complex64)
        got = run(ixs, iy, xs, sd, be, y0=y0, sx=0.5 - 1j, sy=2 + 0.25j)
        ref = oracle(ixs, iy, xs, sd, y0=y0, sx=0.5 - 1j, sy=2 + 0.25j)
        assert_close(got, ref, np.complex64, 4 * k + 8, str((ms, ks, ns)) + " scaled",
                     precision=prec, depth=k)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
def test_fp64_precision_rounds_once(dtype):
    # precision="fp64": Float32 / ComplexF32 operands are widened and multiplied on the FP64
    # matrix cores, so every output component is the FP64 result rounded once to Float32:
    # |got - exact| <= 2^-24 |exact| (plus the FP64 reference's own error).
    r = rng(2102)
    ixs, iy = parse("ij,jk->ik")
    sd = {"i": 200, "j": 300, "k": 150}
    xs = [rand(r, [sd[l] for l in ix], dtype) for ix in ixs]
    got = run(ixs, iy, xs, sd, om().MI355XBackend(precision="fp64"))
    ref = oracle(ixs, iy, xs, sd)
    scale = np.max(np.abs(ref))
    parts = [(np.real, np.real)] + ([(np.imag, np.imag)] if dtype == np.complex64 else [])
    for f, g in parts:
        e = np.abs(f(got).astype(np.float64) - g(ref))
        assert np.all(e <= 2.0 ** -24 * np.abs(g(ref)) + 1e-13 * scale), e.max()
    # the FP32 matrix-core product misses that bound somewhere (the check has teeth)
    got32 = run(ixs, iy, xs, sd, om().MI355XBackend(precision="fp32"))
    e32 = np.abs(np.real(got32).astype(np.float64) - np.real(ref))
    assert np.any(e32 > 2.0 ** -24 * np.abs(np.real(ref)) + 1e-13 * scale)


LONG_K = [("ij,ij->", {"i": 256, "j": 300}),            # one output, K = 76800
          ("ijk,jkl->il", {"i": 3, "j": 100, "k": 100, "l": 5}),  # 15 outputs, K = 10^4
          ("abc,abd->cd", {"a": 64, "b": 64, "c": 2, "d": 3}),    # 6 outputs, K = 4096
          # the one-launch panel kernel: 1024 < outputs <= 16384, K >= 512
          ("ijk,jkl->il", {"i": 64, "j": 64, "k": 64, "l": 64}),  # 64^3 inputs, 4096 outputs
          ("ijk,jkl->il", {"i": 30, "j": 20, "k": 30, "l": 37}),  # ragged 4x4 tiles
          ("ij,jk->ik", {"i": 2, "j": 1000, "k": 700}),            # M < 4: 1 x 16 tiles
          ("ij,jk->ik", {"i": 900, "j": 700, "k": 3}),             # N < 4: 16 x 1 tiles
          ("ijb,jkb->ikb", {"i": 9, "j": 600, "k": 10, "b": 13}),  # batched
          ("aib,abk->ik", {"a": 16, "b": 40, "i": 37, "k": 33})]   # strided operands


@pytest.mark.parametrize("spec,sd", LONG_K)
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.int64, "tropical"])
def test_few_outputs_long_k(spec, sd, dtype):
    trop = dtype == "tropical"
    dt = np.float64 if trop else dtype
    ixs, iy = parse(spec)
    r = rng(2103)
    xs = [rand(r, [sd[l] for l in ix], dt) for ix in ixs]
    if trop:
        xs = [np.round(x * 8) for x in xs]  # exact sums: bit-exact max-plus
    for be in (om().DefaultBackend(), om().MI355XBackend()):
        got = run(ixs, iy, xs, sd, be, tropical=trop)
        ref = oracle(ixs, iy, xs, sd, tropical=trop)
        if trop or dt == np.int64:
            assert np.array_equal(got, ref), spec
        else:
            k = int(np.prod([sd[l] for l in set(ixs[0]) & set(ixs[1]) - set(iy)]))
            assert_close(got, ref, dt, k, spec)
        y0 = rand(r, [sd[l] for l in iy], dt)
        if not trop:
            sx, sy = (3, -2) if dt == np.int64 else (0.5, 2.0)
            got = run(ixs, iy, xs, sd, be, y0=y0, sx=sx, sy=sy)
            ref = oracle(ixs, iy, xs, sd, y0=y0, sx=sx, sy=sy)
            if dt == np.int64:
                assert np.array_equal(got, ref), spec
            else:
                assert_close(got, ref, dt, 4 * k, spec + " scaled", depth=k)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
def test_c4_shaped_backward_rule(prec, dtype):
    # forward + the ChainRules rule (src/autodiff.jl:24-33) of abcd,cdef->abef at size 16
    # (M = N = K = 256: the matrix-core paths) against the oracle's einsum_grad in FP64
    from oracle import default_backend as db
    o = om()
    ixs, iy = parse("abcd,cdef->abef")
    sd = {l: 16 for l in "abcdef"}
    r = rng(2104)
    xs = [rand(r, [16] * 4, dtype) for _ in range(2)]
    dy = rand(r, [16] * 4, dtype)
    o.set_einsum_backend(o.MI355XBackend(precision=prec))
    try:
        code = o.EinCode(ixs, iy)
        dxs = [dev(x) for x in xs]
        y = o.einsum(code, dxs, sd)
        grads = o.einsum_backward_rule(code, dxs, y, sd, dev(dy))
        torch.cuda.synchronize()
        got = [o.to_numpy(g) for g in grads]
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    W = wide(dtype)
    for i in range(2):
        ref = db.einsum_grad(ixs, [x.astype(W) for x in xs], iy, sd, dy.astype(W), i)
        # each gradient contracts two labels of size 16: depth 256
        assert_close(got[i], ref, dtype, 16 * 4096, "grad %d" % i, precision=prec, depth=256)


def test_c4_shaped_gradient_finite_differences():
    # directional central differences of L(A, B) = sum(dy * (A . B)) in FP64 on the device
    o = om()
    ixs, iy = parse("abcd,cdef->abef")
    sd = {l: 6 for l in "abcdef"}
    r = rng(2105)
    a, b, dy, va, vb = (rand(r, [6] * 4, np.float64) for _ in range(5))
    o.set_einsum_backend(o.MI355XBackend(precision="fp64"))
    try:
        code = o.EinCode(ixs, iy)

        def loss(x, z):
            yy = o.to_numpy(o.einsum(code, (dev(x), dev(z)), sd))
            return float(np.sum(dy * yy))

        y = o.einsum(code, (dev(a), dev(b)), sd)
        ga, gb = (o.to_numpy(g) for g in o.einsum_backward_rule(code, (dev(a), dev(b)), y, sd, dev(dy)))
        h = 1e-4
        fa = (loss(a + h * va, b) - loss(a - h * va, b)) / (2 * h)
        fb = (loss(a, b + h * vb) - loss(a, b - h * vb)) / (2 * h)
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    assert abs(fa - np.sum(ga * va)) <= 1e-7 * max(1.0, abs(fa))
    assert abs(fb - np.sum(gb * vb)) <= 1e-7 * max(1.0, abs(fb))


def test_one_launch_few_outputs_is_deterministic_across_calls():
    """Few outputs / long K run in one launch, each workgroup summing its tile over all of K
    and reducing its threads' sums in a fixed tree: repeated calls, interleaved with calls of
    other shapes, give bit-equal results, within the FP32 bound of the FP64 oracle."""
    o = om()
    cases = [("ijk,jkl->il", {"i": 64, "j": 64, "k": 64, "l": 64}),
             ("ijk,jkl->il", {"i": 30, "j": 20, "k": 30, "l": 37}),
             ("aib,abk->ik", {"a": 16, "b": 40, "i": 37, "k": 33})]
    r = rng(2107)
    data = []
    for spec, sd in cases:
        ixs, iy = parse(spec)
        xs = [rand(r, [sd[l] for l in ix], np.float32) for ix in ixs]
        data.append((ixs, iy, xs, sd, [dev(x) for x in xs]))
    first = {}
    for rep in range(4):
        for c, (ixs, iy, xs, sd, dxs) in enumerate(data):
            y = o.einsum(o.EinCode(ixs, iy), dxs, sd)
            torch.cuda.synchronize()
            got = o.to_numpy(y)
            if rep == 0:
                first[c] = got
                k = int(np.prod([sd[l] for l in set(ixs[0]) & set(ixs[1]) - set(iy)]))
                assert_close(got, oracle(ixs, iy, xs, sd), np.float32, k, str(c))
            else:
                assert np.array_equal(got, first[c]), (c, rep)
