"""GPU parity of flat einsum (unary, binary, n-ary), permute and batched GEMM through the C ABI
against the oracle on identical seeded inputs.  Integer and tropical results are bit-exact;
floating point is checked against the FP64 oracle with the tolerances in helpers.F_TOL."""
import itertools

import numpy as np
import pytest

from helpers import assert_close, parse, rand, rng, wide
from oracle import default_backend as db
from oracle import loop

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import omeinsum_mi355x as om
    om.set_einsum_backend(om.DefaultBackend())
    yield


def om():
    import omeinsum_mi355x
    return omeinsum_mi355x


def dev(x, tropical=False):
    t = torch.from_numpy(np.array(x, order="C")).cuda()
    return om().Tropical(t) if tropical else t


def host(y):
    return om().to_numpy(y)


def run(ixs, iy, xs, size_dict=None, tropical=False, y0=None, sx=True, sy=False):
    o = om()
    code = o.EinCode(ixs, iy)
    dxs = [dev(x, tropical) for x in xs]
    sd = o.get_size_dict(code.ixs, dxs, size_dict)
    if y0 is None:
        y = o.einsum(code, dxs, sd)
    else:
        y = dev(y0, tropical)
        o.einsum_inplace(code, dxs, y, sx, sy, sd)
    torch.cuda.synchronize()
    return host(y)


def reference(ixs, iy, xs, size_dict=None, tropical=False, y0=None, sx=True, sy=False):
    T = np.result_type(*[x.dtype for x in xs])
    # tropical numbers are checked bit-exactly in their own type (the oracle rounds every
    # max/plus to it); everything else against the FP64 oracle
    W = T if tropical else wide(T)
    xs64 = [np.asarray(x, dtype=W) for x in xs]
    y = None if y0 is None else np.asarray(y0, dtype=W)
    if tropical:
        sx = sx if isinstance(sx, bool) else float(sx)
    return loop.loop_einsum(ixs, iy, xs64, size_dict, y=y, sx=sx, sy=sy, tropical=tropical)


DTYPES = [np.float64, np.float32, np.complex128, np.complex64, np.int64]
UNARY = ["ij->", "ij->j", "ii->", "ii->i", "ijk->kij", "a->aaaa", "ac->acc", "->ii", "i->ik",
         "->ik", "illljkk->kijjcc", "ijkl->ljki", "iijj->ij", "ijk->ik"]


@pytest.mark.parametrize("spec", UNARY)
@pytest.mark.parametrize("dtype", DTYPES + ["tropical"])
def test_unary_codes(spec, dtype):
    # test/roceinsum.jl:74-99 plus permutation / partial-trace forms
    trop = dtype == "tropical"
    dt = np.float64 if trop else dtype
    ixs, iy = parse(spec)
    r = rng(101)
    sd = {l: 3 for l in "abcdefijkl"}
    xs = [rand(r, [3] * len(ix), dt) for ix in ixs]
    got = run(ixs, iy, xs, sd, trop)
    ref = reference(ixs, iy, xs, sd, trop)
    assert_close(got, ref, np.int64 if trop else dt, k_terms=27, what=spec)


def binary_sweep():
    out = []
    for has_batch in (True, False):
        for i1 in [(), ("i",), ("j",), ("i", "j"), ("j", "i")]:
            for i2 in [(), ("k",), ("j",), ("k", "j"), ("j", "k")]:
                for i3 in [(), ("i",), ("k",), ("i", "k"), ("k", "i")]:
                    b = ("l",) if has_batch else ()
                    out.append((list(i1 + b), list(i2 + b), list(i3 + b)))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.int64, np.complex64, "tropical"])
def test_binary_sweep(dtype):
    # all 5x5x5x2 label patterns of test/binaryrules.jl:18-45
    trop = dtype == "tropical"
    dt = np.float32 if trop else dtype
    r = rng(202)
    sd = {l: 5 for l in "ijkl"}
    for ix1, ix2, iy in binary_sweep():
        a = rand(r, [5] * len(ix1), dt)
        b = rand(r, [5] * len(ix2), dt)
        got = run([ix1, ix2], iy, [a, b], sd, trop)
        ref = reference([ix1, ix2], iy, [a, b], sd, trop)
        assert_close(got, ref, np.int64 if trop else dt, k_terms=25, what=str((ix1, ix2, iy)))


COMPOSITE = ["ijl,jl->il", "ab,bc->ac", "abb,bc->ac", "ab,bc->acc", "ab,bce->ac",
             "ab,bc->ace", "bal,bcl->lcae", "bal,bcl->ca", "bal,bc->lca", "ddebal,bcf->lcac",
             "c,c->cc", "c,cc->c", "cc,c->cc", "cc,cc->cc", "cc,cb->bc", "cb,bc->cc",
             "ac,cc->ac", "mk,ijk->ijm", "ijkl,klmn->ijmn", "abcd,cdef->abef", "abef,cdef->abcd",
             "abcd,abef->cdef", "ijk,jkl->il", "ijb,jkb->ikb", "kjb,bij->bki"]


@pytest.mark.parametrize("spec", COMPOSITE)
@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.int64, "tropical"])
def test_composite_codes(spec, dtype):
    # test/einsum.jl:57-75, test/roceinsum.jl:101-120, test/einsum.jl:311-319
    trop = dtype == "tropical"
    dt = np.float64 if trop else dtype
    ixs, iy = parse(spec)
    r = rng(303)
    sd = {l: 3 for l in "abcdefijklmn"}
    xs = [rand(r, [3] * len(ix), dt) for ix in ixs]
    got = run(ixs, iy, xs, sd, trop)
    ref = reference(ixs, iy, xs, sd, trop)
    assert_close(got, ref, np.int64 if trop else dt, k_terms=81, what=spec)


def test_regression_code():
    # test/binaryrules.jl:77-81
    r = rng(404)
    x, y = rand(r, [2] * 8, np.float64), rand(r, [2] * 5, np.float64)
    ixs, iy = [[10, 5, 4, 2, 6, 6, 7, 9], [1, 9, 8, 2, 10]], [4, 10, 9, 7, 10]
    assert_close(run(ixs, iy, [x, y]), reference(ixs, iy, [x, y]), np.float64, 64)


@pytest.mark.parametrize("sx,sy", [(True, False), (5.0, 3.0), (5.0, 1.0), (2.0, 0.0), (0.0, 2.5)])
def test_scaling_semantics(sx, sy):
    # test/einsum.jl:55-66 (einsum! with sx, sy) and the NaN-safe overwrite when sy == 0
    r = rng(505)
    sd = {1: 3, 2: 3, 3: 3, 4: 4, 5: 5}
    ix, iy, iz = [1, 2, 3, 3, 4], [3, 5, 1, 1, 2, 5], [1, 2, 3, 4, 5, 5]
    x = rand(r, (3, 3, 3, 3, 4), np.float64)
    y = rand(r, (3, 5, 3, 3, 3, 5), np.float64)
    z = rand(r, (3, 3, 3, 4, 5, 5), np.float64)
    if sy is False or sy == 0.0:
        z[...] = np.nan
    got = run([ix, iy], iz, [x, y], sd, y0=z, sx=sx, sy=sy)
    ref = db.einsum_inplace([ix, iy], iz, [x, y], np.asfortranarray(z.copy()),
                            1.0 if sx is True else sx, 0.0 if sy is False else sy, sd)
    assert_close(got, ref, np.float64, 81)


def test_loop_einsum_scaling_with_duplicate_output():
    # test/roceinsum.jl:10-21
    o = om()
    r = rng(606)
    a = rand(r, (3, 3), np.float64)
    xs = (dev(a), dev(a))
    sd = {1: 3, 2: 3, 3: 3}
    y = dev(np.ones((3, 3)))
    o.loop_einsum_inplace([[1, 2], [2, 3]], [1, 3], xs, y, 4.0, 2.0, sd)
    assert np.allclose(host(y), 4 * a @ a + 2)
    y3 = dev(np.ones((3, 3, 3)))
    o.loop_einsum_inplace([[1, 2], [2, 3]], [1, 3, 3], xs, y3, 4.0, 2.0, sd)
    ref = 2 * np.ones((3, 3, 3))
    res = 4 * a @ a
    for k in range(3):
        ref[:, k, k] += res[:, k]
    assert np.allclose(host(y3), ref)


NARY = ["ij,ik,il->jkl", "ij,ij,il->l", "ij,ij,ij->ij", "ij,jk,kl->il", "npu,por,dom,lmn->urdl",
        "abewcd,bfixgh,ajeycd,jfizgh->wxyz"]


@pytest.mark.parametrize("spec", NARY)
@pytest.mark.parametrize("dtype", [np.float64, np.complex64, np.int64])
def test_nary_codes(spec, dtype):
    # star / star-and-contract (test/einsum.jl:159-171), fallback (test/roceinsum.jl:23-27,56-72)
    ixs, iy = parse(spec)
    r = rng(707)
    sd = {l: 2 for l in "abcdefghijklmnopruwxyz"}
    xs = [rand(r, [2] * len(ix), dtype) for ix in ixs]
    assert_close(run(ixs, iy, xs, sd), reference(ixs, iy, xs, sd), dtype, 64, spec)


def _s_tensor():
    s = np.zeros((3, 3, 3), dtype=np.int64)
    for i, j, k in itertools.product(range(3), repeat=3):
        s[i, j, k] = int(len({i, j, k}) == 3)
    return s


def test_coloring_known_answers():
    # docs/src/applications.md:20-34: 6 colourings of two vertices, none for Petersen
    s = _s_tensor()
    assert run([list("ijk"), list("ijk")], [], [s, s])[()] == 6
    ixs, iy = parse("afl,bhn,cjf,dlh,enj,ago,big,cki,dmk,eom->")
    assert run(ixs, iy, [s] * 10)[()] == 0


def test_trivial_and_broadcast():
    o = om()
    # ein"->ii"(asarray(1); size_info i=5) == I(5) (test/einsum.jl:191-194)
    y = run([[]], ["i", "i"], [np.array(1, dtype=np.int64)], {"i": 5})
    assert np.array_equal(y, np.eye(5, dtype=np.int64))
    # ein",->"(3, 4) === 12 (test/einsum.jl:198-200)
    y = run([[], []], [], [np.array(3, dtype=np.int64), np.array(4, dtype=np.int64)])
    assert y.dtype == np.int64 and y[()] == 12
    # issue 136 (test/einsum.jl:306-309)
    y = run([[1, 2, 3], [2]], [1, 3], [np.ones((2, 2, 1)), np.ones(2)])
    assert np.array_equal(y, np.array([[2.0], [2.0]]))
    y = run([[1, 2, 3], [2]], [1, 3], [np.ones((2, 2, 0)), np.ones(2)])
    assert y.shape == (2, 0)
    # ein"ij,jk->ik" with an empty inner label writes sx*0 + sy*y
    y = run([[1, 2], [2, 3]], [1, 3], [np.ones((3, 0)), np.ones((0, 4))])
    assert np.array_equal(y, np.zeros((3, 4)))
    assert o.asscalar(o.asarray(3)) == 3


def test_errors():
    o = om()
    a = dev(np.ones((3, 3)))
    with pytest.raises(o.ArgumentError):
        o.einsum(o.ein("ij,jk->ik"), (a, a, a))
    with pytest.raises(o.ArgumentError):
        o.einsum(o.ein("ij,ijk->ik"), (a, a))
    with pytest.raises(o.DimensionMismatch):
        o.einsum(o.ein("ij,jk->ik"), (dev(np.ones((2, 3))), dev(np.ones((2, 2)))))
    # allow_loops(false) makes the index walk an error (test/einsum.jl:323-330)
    o.allow_loops(False)
    try:
        from omeinsum_mi355x._lib import LoopsForbidden
        with pytest.raises(LoopsForbidden):
            o.loop_einsum(o.EinCode([[1, 2, 3, 4], [2, 3]], [1, 4]),
                          (dev(np.ones((5, 5, 5, 5))), dev(np.ones((5, 5)))))
        # the binary path does not need loops
        t = rand(rng(1), (5, 5, 5, 5), np.float64)
        m = rand(rng(2), (5, 5), np.float64)
        got = run([[1, 2, 3, 4], [2, 3]], [1, 4], [t, m])
        assert_close(got, reference([[1, 2, 3, 4], [2, 3]], [1, 4], [t, m]), np.float64, 25)
    finally:
        o.allow_loops(True)


def test_permute_high_rank():
    o = om()
    r = rng(808)
    # 18-D permute (test/roceinsum.jl:158-161)
    shape = [4] + [int(v) for v in r.integers(1, 4, size=17)]
    c = rand(r, shape, np.float64)
    perm = list(range(17, -1, -1))
    y = torch.empty([shape[p] for p in perm], dtype=torch.float64, device="cuda")
    o.tensorpermute(y, dev(c), perm)
    assert np.array_equal(host(y), np.transpose(c, perm))
    # 28-D dimension-2 permute (benchmark/large.jl:5-8)
    c = rand(r, [2] * 28, np.float32)
    perm = [int(p) for p in r.permutation(28)]
    y = torch.empty([2] * 28, dtype=torch.float32, device="cuda")
    o.tensorpermute(y, dev(c), perm)
    assert np.array_equal(host(y), np.transpose(c, perm))
    # scaled permute with accumulation
    a = rand(r, (33, 65, 17), np.complex64)
    y0 = rand(r, (17, 33, 65), np.complex64)
    y = dev(y0)
    o.tensorpermute(y, dev(a), [2, 0, 1], 2.0, 0.5)
    assert np.allclose(host(y), 2.0 * np.transpose(a, (2, 0, 1)) + 0.5 * y0, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int64])
def test_config1_ijk_jkl_il(dtype):
    # ein"ijk,jkl->il" on two 64x64x64 tensors (Int64 entries in [-8, 8], bit-exact)
    r = rng(1000)
    a, b = rand(r, (64, 64, 64), dtype), rand(r, (64, 64, 64), dtype)
    ixs, iy = parse("ijk,jkl->il")
    got = run(ixs, iy, [a, b])
    if dtype == np.int64:
        ref = loop.loop_einsum(ixs, iy, [a, b])
    else:
        ref = db.einsum(ixs, iy, [a.astype(np.float64), b.astype(np.float64)])
    assert_close(got, ref, dtype, 4096, "ijk,jkl->il")


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("spec", ["ij,jk->ik", "ji,jk->ik", "ij,kj->ik", "ji,kj->ik", "ij,jk->ki",
                                  "ijb,jkb->ikb", "ajb,bkj->bka"])
def test_mfma_f32_layouts(spec, prec):
    o = om()
    ixs, iy = parse(spec)
    sd = {"i": 129, "j": 300, "k": 77, "a": 65, "b": 3}
    r = rng(909)
    xs = [rand(r, [sd[l] for l in ix], np.float32) for ix in ixs]
    from omeinsum_mi355x import _lib
    ctx = _lib.context()
    o.set_einsum_backend(o.MI355XBackend(precision=prec))
    # below the packed path's size BF16x3 runs as FP32 unless the MFMA kernel is forced:
    # force it so that the in-loop split kernel is what runs here
    ctx.set_option(_lib.OPT_GEMM_KERNEL, 2 if prec == "bf16x3" else -1)
    try:
        got = run(ixs, iy, xs, sd)
    finally:
        ctx.set_option(_lib.OPT_GEMM_KERNEL, -1)
        o.set_einsum_backend(o.DefaultBackend())
    ref = db.einsum(ixs, iy, [x.astype(np.float64) for x in xs], sd)
    tol_k = 300
    if prec == "bf16x3":
        # 3-term split: per-product error <= ~3*2^-18 relative; elementwise at 4x the FP32
        # bound, norm-wise at the BF16x3 bound
        assert_close(got, ref, np.float32, 16 * tol_k, spec, precision="bf16x3", depth=tol_k)
    else:
        assert_close(got, ref, np.float32, tol_k, spec)


@pytest.mark.parametrize("spec", ["ij,jk->ik", "ji,kj->ik", "ijb,jkb->ikb"])
def test_mfma_f64(spec):
    ixs, iy = parse(spec)
    sd = {"i": 200, "j": 129, "k": 150, "b": 2}
    r = rng(919)
    xs = [rand(r, [sd[l] for l in ix], np.float64) for ix in ixs]
    got = run(ixs, iy, xs, sd)
    ref = db.einsum(ixs, iy, xs, sd)
    assert_close(got, ref, np.float64, 129, spec)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex128, np.int64])
def test_batched_gemm(ta, tb, dtype):
    # _batched_gemm! with every transpose combination (test/utils.jl:35-45)
    o = om()
    r = rng(1111)
    bsz, m, n, k = 3, 40, 33, 50
    A = rand(r, (bsz, m, k) if ta == "N" else (bsz, k, m), dtype)
    B = rand(r, (bsz, k, n) if tb == "N" else (bsz, n, k), dtype)
    C0 = rand(r, (bsz, m, n), dtype)
    C = dev(C0)
    o.batched_gemm(ta, tb, 2, dev(A), dev(B), 3, C)
    opA = A if ta == "N" else np.transpose(A, (0, 2, 1))
    opB = B if tb == "N" else np.transpose(B, (0, 2, 1))
    ref = 2 * np.matmul(opA.astype(wide(dtype)), opB.astype(wide(dtype))) + 3 * C0.astype(wide(dtype))
    assert_close(host(C), ref, dtype, k, "batched_gemm %s%s" % (ta, tb))


def test_int64_wraps_like_julia():
    # Int64 arithmetic wraps (two's complement) exactly as Julia's does
    r = rng(1212)
    a = np.asarray(r.integers(-2 ** 62, 2 ** 62, size=(17, 23)), dtype=np.int64)
    b = np.asarray(r.integers(-2 ** 62, 2 ** 62, size=(23, 19)), dtype=np.int64)
    got = run([["i", "j"], ["j", "k"]], ["i", "k"], [a, b])
    ref = loop.loop_einsum([["i", "j"], ["j", "k"]], ["i", "k"], [a, b])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tropical_gemm_exact(dtype):
    r = rng(1313)
    a = np.asarray(r.integers(-50, 50, size=(70, 90)), dtype=dtype)
    b = np.asarray(r.integers(-50, 50, size=(90, 65)), dtype=dtype)
    got = run([["i", "j"], ["j", "k"]], ["i", "k"], [a, b], tropical=True)
    ref = np.max(a[:, :, None] + b[None, :, :], axis=1)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
@pytest.mark.parametrize("spec", ["ij,jk->ik", "ji,kj->ki"])
def test_packed_bf16x3_256_tiles(dtype, spec):
    # products with >= 128 tiles of 256 x 256 take the 8-wave 256 x 256 kernel; ragged edges
    o = om()
    ixs, iy = parse(spec)
    sd = {"i": 2900, "j": 520, "k": 2900}
    r = rng(1415)
    xs = [rand(r, [sd[l] for l in ix], dtype) for ix in ixs]
    o.set_einsum_backend(o.MI355XBackend(precision="bf16x3"))
    try:
        got = run(ixs, iy, xs, sd)
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    W = wide(dtype)
    a, b = (x.astype(W) for x in xs)
    ref = (a @ b) if spec == "ij,jk->ik" else (b @ a)
    assert_close(got, ref, dtype, 16 * 520, spec, precision="bf16x3", depth=520)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
@pytest.mark.parametrize("spec", ["ij,jk->ik", "ji,kj->ik", "ijb,jkb->ikb", "ajb,bkj->bka", "ij,jk->ki"])
@pytest.mark.parametrize("scal", [(True, False), (1.5 - 0.5j, 0.25)])
def test_packed_bf16x3(dtype, spec, scal):
    # the packed BF16x3 path (large products); ComplexF32 is realified into one real GEMM
    o = om()
    ixs, iy = parse(spec)
    sd = {"i": 300, "j": 520, "k": 260, "a": 129, "b": 2}
    r = rng(1414)
    xs = [rand(r, [sd[l] for l in ix], dtype) for ix in ixs]
    sx, sy = scal
    if dtype == np.float32 and not isinstance(sx, bool):
        sx = 1.5
    y0 = rand(r, [sd[l] for l in iy], dtype)
    o.set_einsum_backend(o.MI355XBackend(precision="bf16x3"))
    try:
        got = run(ixs, iy, xs, sd, y0=y0, sx=sx, sy=sy)
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    W = wide(dtype)
    ref = db.einsum(ixs, iy, [x.astype(W) for x in xs], sd)
    if not isinstance(sx, bool):
        ref = sx * ref + sy * y0.astype(W)
    assert_close(got, ref, dtype, 16 * 520, spec, precision="bf16x3", depth=520)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
@pytest.mark.parametrize("spec", ["ij,jk->ik", "ji,kj->ik", "ijb,jkb->ikb", "ajb,bkj->bka", "ij,jk->ki"])
@pytest.mark.parametrize("scal", [(True, False), (1.5 - 0.5j, 0.25)])
def test_packed_bf16x6(dtype, spec, scal):
    # the six-product packed path: depth 256 (512 for complex), where a kernel that dropped one
    # of its three 2^-16 products would fail the bf16x6 bound
    o = om()
    ixs, iy = parse(spec)
    sd = {"i": 1200, "j": 256, "k": 1000, "a": 300, "b": 2}
    r = rng(1418)
    xs = [rand(r, [sd[l] for l in ix], dtype) for ix in ixs]
    sx, sy = scal
    if dtype == np.float32 and not isinstance(sx, bool):
        sx = 1.5
    y0 = rand(r, [sd[l] for l in iy], dtype)
    o.set_einsum_backend(o.MI355XBackend(precision="bf16x6"))
    try:
        got = run(ixs, iy, xs, sd, y0=y0, sx=sx, sy=sy)
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    W = wide(dtype)
    ref = db.einsum(ixs, iy, [x.astype(W) for x in xs], sd)
    if not isinstance(sx, bool):
        ref = sx * ref + sy * y0.astype(W)
    assert_close(got, ref, dtype, 256, spec, precision="bf16x6", depth=256)


@pytest.mark.parametrize("dtype", [np.float32, np.complex64])
@pytest.mark.parametrize("sd", [{"i": 1201, "j": 260, "k": 999},      # ragged tiles and k tiles
                                {"i": 768, "j": 8192, "k": 384}])     # few tiles: split-K
def test_packed_bf16x6_edges(dtype, sd):
    o = om()
    ixs, iy = parse("ij,jk->ik")
    r = rng(1419)
    xs = [rand(r, [sd[l] for l in ix], dtype) for ix in ixs]
    o.set_einsum_backend(o.MI355XBackend(precision="bf16x6"))
    ctx = o._lib.context()
    ctx.set_option(o._lib.OPT_KERNEL_TIMING, 1)
    ctx.kernel_timing("gemm_packed")
    try:
        got = run(ixs, iy, xs, sd)
        _, n_packed = ctx.kernel_timing("gemm_packed")
    finally:
        ctx.set_option(o._lib.OPT_KERNEL_TIMING, 0)
        o.set_einsum_backend(o.DefaultBackend())
    assert n_packed == 1, "the product did not take the packed six-product path"
    W = wide(dtype)
    ref = xs[0].astype(W) @ xs[1].astype(W)
    assert_close(got, ref, dtype, sd["j"], "bf16x6 %s" % sd, precision="bf16x6", depth=sd["j"])
