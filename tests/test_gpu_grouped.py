"""GPU parity of the grouped launches through the C ABI: ome_grouped_gemm (uneven batched
GEMMs in one launch, every tile variant and the autotuned choice) against numpy's batched
matmul in FP64 / exact Int64 / max-plus, and ome_grouped_permute against numpy transposes
(bit-exact)."""
import numpy as np
import pytest

from helpers import assert_close, rng

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield
    from omeinsum_mi355x import _lib
    _lib.context().set_option(_lib.OPT_GROUPED_VARIANT, -1)


def om():
    import omeinsum_mi355x
    return omeinsum_mi355x


SPLIT = (3, 4, 6, 7, 9)  # the tile variants with BF16x3 products
SPLIT6 = (10,)           # the tile variant with BF16x6 products


def variant_precision(variant):
    return "bf16x3" if variant in SPLIT else "bf16x6" if variant in SPLIT6 else "fp32"


def problems(seed, n, dtype, dims=range(1, 201), batches=range(1, 5)):
    r = rng(seed)
    out = []
    for _ in range(n):
        m, k, nn = (int(r.choice(dims)) for _ in range(3))
        b = int(r.choice(batches))
        if np.dtype(dtype).kind == "c":
            A = (r.uniform(-1, 1, (b, m, k)) + 1j * r.uniform(-1, 1, (b, m, k))).astype(dtype)
            B = (r.uniform(-1, 1, (b, k, nn)) + 1j * r.uniform(-1, 1, (b, k, nn))).astype(dtype)
        elif np.dtype(dtype).kind == "i":
            A = r.integers(-8, 9, (b, m, k)).astype(dtype)
            B = r.integers(-8, 9, (b, k, nn)).astype(dtype)
        else:
            A = r.uniform(-1, 1, (b, m, k)).astype(dtype)
            B = r.uniform(-1, 1, (b, k, nn)).astype(dtype)
        C = r.uniform(-1, 1, (b, m, nn)).astype(dtype) if np.dtype(dtype).kind != "i" else \
            r.integers(-8, 9, (b, m, nn)).astype(dtype)
        out.append((A, B, C))
    return out


def run(ps, alpha, beta, tropical=False):
    o = om()
    dev = []
    for A, B, C in ps:
        t = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (A, B, C)]
        if tropical:
            t = [o.Tropical(x) for x in t]
        dev.append(tuple(t))
    o.grouped_gemm(dev, alpha, beta)
    torch.cuda.synchronize()
    return [o.to_numpy(d[2]) for d in dev]


def check(ps, got, alpha, beta, dtype, precision="fp32"):
    for (A, B, C), g in zip(ps, got):
        if np.dtype(dtype).kind == "i":
            ref = alpha * np.matmul(A, B) + beta * C
            assert np.array_equal(g, ref)
        else:
            W = np.complex128 if np.dtype(dtype).kind == "c" else np.float64
            ref = alpha * np.matmul(A.astype(W), B.astype(W)) + beta * C.astype(W)
            assert_close(g, ref, dtype, A.shape[2] + 1, precision=precision, depth=A.shape[2])


@pytest.mark.parametrize("buffers", [1, 0])
@pytest.mark.parametrize("variant", [-1, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_grouped_gemm_f32_variants(variant, buffers):
    # buffers 0 forces the pointer form of the kernels that otherwise read through buffer
    # resources (the form problems above 32-bit offsets take)
    from omeinsum_mi355x import _lib
    ctx = _lib.context()
    ctx.set_option(_lib.OPT_GROUPED_VARIANT, variant)
    ctx.set_option(_lib.OPT_GROUPED_BUFFERS, buffers)
    try:
        ps = problems(3100 + variant, 120, np.float32)
        prec = variant_precision(variant)
        got = run(ps, True, False)
        check(ps, got, 1.0, 0.0, np.float32, prec)
        got = run(ps, 1.5, -0.5)
        check(ps, got, 1.5, -0.5, np.float32, prec)
    finally:
        ctx.set_option(_lib.OPT_GROUPED_VARIANT, -1)
        ctx.set_option(_lib.OPT_GROUPED_BUFFERS, 1)


@pytest.mark.parametrize("variant", [8, 9])
def test_grouped_gemm_runs_cross_tiles_and_problems(variant):
    # about 10 000 tiles over 1536 runs: every run walks several tiles and most cross from one
    # problem (or batch entry) into the next, carrying the next tile's first K step across
    from omeinsum_mi355x import _lib
    ctx = _lib.context()
    ctx.set_option(_lib.OPT_GROUPED_VARIANT, variant)
    try:
        ps = problems(3120 + variant, 700, np.float32, dims=range(1, 141), batches=range(1, 13))
        for A, B, C in ps[:3]:
            assert A.shape[2] > 0
        prec = "bf16x3" if variant in SPLIT else "fp32"
        got = run(ps, True, False)
        check(ps, got, 1.0, 0.0, np.float32, prec)
        got = run(ps, 0.5, 2.0)
        check(ps, got, 0.5, 2.0, np.float32, prec)
        assert ctx.get_option(_lib.OPT_LAST_GROUPED_CHOICE) // 2 == variant
    finally:
        ctx.set_option(_lib.OPT_GROUPED_VARIANT, -1)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6"])
def test_grouped_gemm_c3_shapes_autotuned_twice(prec):
    # dims in {8, 16, ..., 256} and batch in 1..16 as in the uneven-batch workload; the second
    # call reuses the tuned choice and the cached descriptors; the context's precision picks
    # the FP32, the BF16x3 or the FP32-accurate kernels (BF16x6 or FP32, the faster)
    from omeinsum_mi355x import _lib
    ctx = _lib.context()
    ctx.set_option(_lib.OPT_PRECISION, _lib.PREC[prec])
    try:
        ps = problems(3110, 96, np.float32, dims=range(8, 257, 8), batches=range(1, 17))
        for _ in range(2):
            got = run(ps, True, False)
            check(ps, got, 1.0, 0.0, np.float32, prec)
            variant = ctx.get_option(_lib.OPT_LAST_GROUPED_CHOICE) // 2
            assert (variant in SPLIT) == (prec == "bf16x3"), variant
            if prec == "bf16x6":
                assert variant in SPLIT6 + (1, 8), variant
    finally:
        ctx.set_option(_lib.OPT_PRECISION, 0)


@pytest.mark.parametrize("dtype", [np.float64, np.complex64, np.complex128, np.int64])
def test_grouped_gemm_other_types(dtype):
    ps = problems(3120, 40, dtype, dims=range(1, 97))
    got = run(ps, True, False)
    check(ps, got, 1, 0, dtype)
    a, b = (3, -2) if dtype == np.int64 else (0.5, 2.0)
    got = run(ps, a, b)
    check(ps, got, a, b, dtype)


def test_grouped_gemm_tropical_exact():
    r = rng(3130)
    ps = []
    for _ in range(20):
        m, k, n = (int(x) for x in r.integers(1, 40, 3))
        b = int(r.integers(1, 3))
        A = np.round(r.uniform(-8, 8, (b, m, k)))
        B = np.round(r.uniform(-8, 8, (b, k, n)))
        ps.append((A, B, np.zeros((b, m, n))))
    got = run(ps, True, False, tropical=True)
    for (A, B, _), g in zip(ps, got):
        ref = np.max(A[:, :, :, None] + B[:, None, :, :], axis=2)
        assert np.array_equal(g, ref)


def test_grouped_gemm_empty_and_degenerate():
    ps = [(np.zeros((1, 0, 5), np.float32), np.zeros((1, 5, 3), np.float32), np.zeros((1, 0, 3), np.float32)),
          (np.ones((2, 3, 0), np.float32), np.ones((2, 0, 4), np.float32), np.ones((2, 3, 4), np.float32)),
          (np.ones((1, 1, 1), np.float32), np.full((1, 1, 1), 2, np.float32), np.zeros((1, 1, 1), np.float32))]
    got = run(ps, True, False)
    assert got[0].shape == (1, 0, 3)
    assert np.array_equal(got[1], np.zeros((2, 3, 4), np.float32))  # K = 0: sum is empty
    assert got[2][0, 0, 0] == 2.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.int64])
def test_grouped_permute(dtype):
    o = om()
    r = rng(3140)
    xs, perms, ys, refs = [], [], [], []
    for g in range(60):
        nd = int(r.integers(1, 7))
        shape = [int(s) for s in r.integers(1, 9 if nd > 3 else 70, nd)]
        perm = [int(p) for p in r.permutation(nd)]
        x = (r.uniform(-4, 4, shape) * 4).round().astype(dtype)
        xs.append(torch.from_numpy(x).cuda())
        perms.append(perm)
        ys.append(torch.empty([shape[p] for p in perm], dtype=xs[-1].dtype, device="cuda"))
        refs.append(np.transpose(x, perm))
    # the 4-D form of the uneven workload: ijkl -> ljik (perm (3, 1, 0, 2)), dims in [4, 64]
    for g in range(8):
        shape = [int(s) for s in r.integers(4, 65, 4)]
        x = (r.uniform(-4, 4, shape) * 4).round().astype(dtype)
        xs.append(torch.from_numpy(x).cuda())
        perms.append([3, 1, 0, 2])
        ys.append(torch.empty([shape[p] for p in (3, 1, 0, 2)], dtype=xs[-1].dtype, device="cuda"))
        refs.append(np.transpose(x, (3, 1, 0, 2)))
    o.grouped_permute(xs, perms, ys)
    torch.cuda.synchronize()
    for y, ref in zip(ys, refs):
        assert np.array_equal(o.to_numpy(y), ref)
    # scaled: y = 2*permute(x) - y
    prev = [o.to_numpy(y).copy() for y in ys]
    o.grouped_permute(xs, perms, ys, 2, -1)
    torch.cuda.synchronize()
    for y, ref, p in zip(ys, refs, prev):
        assert np.array_equal(o.to_numpy(y), 2 * ref - p)


@pytest.mark.parametrize("kind", ["f32", "f64", "c128", "bool"])
def test_grouped_permute_shared_dim_tiles(kind):
    # permutes whose input-leading and output-leading dims are followed by the same dim take
    # the shared-dim tile (a whole input group x a whole output group x a chunk of that dim);
    # extents chosen so that the last chunk is partial and so that 1-, 4-, 8- and 16-byte
    # elements all tile (bool through TropicalAndOr)
    o = om()
    r = rng(3150)
    # row-major shapes and torch perms; (1, 3, 2, 0) on (l, k, j, i) is the column-major
    # ijkl -> ljik of the uneven workload, where S = j follows I = i and O = l
    cases = [((3, 300, 7, 5), (1, 3, 2, 0)),
             ((4, 4, 4, 4), (1, 3, 2, 0)),
             ((60, 17, 33, 64), (1, 3, 2, 0)),
             ((9, 1000, 11), (2, 1, 0)),          # reversal: S = the middle dim
             ((2, 3, 5, 7, 11), (4, 2, 1, 3, 0)),
             ((200, 70, 3), (2, 1, 0)),
             ((5, 7, 30, 3), (3, 1, 0, 2))]       # no shared dim: the two-group tile
    xs, perms, ys, refs = [], [], [], []
    for shape, perm in cases:
        if kind == "bool":
            x = r.random(shape) < 0.5
        elif kind == "c128":
            x = (r.integers(-8, 9, shape) + 1j * r.integers(-8, 9, shape)).astype(np.complex128)
        else:
            x = r.integers(-8, 9, shape).astype(np.float32 if kind == "f32" else np.float64)
        t = torch.from_numpy(x).cuda()
        y = torch.empty([shape[p] for p in perm], dtype=t.dtype, device="cuda")
        if kind == "bool":
            t, y = o.TropicalAndOr(t), o.TropicalAndOr(y)
        xs.append(t)
        perms.append(list(perm))
        ys.append(y)
        refs.append(np.transpose(x, perm))
    o.grouped_permute(xs, perms, ys)
    torch.cuda.synchronize()
    for y, ref, (shape, perm) in zip(ys, refs, cases):
        assert np.array_equal(o.to_numpy(y), ref), (shape, perm)
    if kind != "bool":
        prev = [o.to_numpy(y).copy() for y in ys]
        o.grouped_permute(xs, perms, ys, 2, -1)
        torch.cuda.synchronize()
        for y, ref, p in zip(ys, refs, prev):
            assert np.array_equal(o.to_numpy(y), 2 * ref - p)


def test_grouped_launches_replay_in_a_graph():
    # a grouped GEMM and a grouped permute recorded once into a graph (after one eager run
    # uploaded their descriptors) and replayed several times, with unrelated eager work and
    # synchronisations in between: every replay equals the eager result bit for bit
    o = om()
    r = rng(3160)
    ps = problems(3161, 48, np.float32, dims=range(8, 161, 8), batches=range(1, 5))
    dev = [tuple(torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in p) for p in ps]
    xs, ys, perms = [], [], []
    for _ in range(24):
        shape = [int(s) for s in r.integers(4, 40, 4)]
        xs.append(torch.from_numpy(r.uniform(-1, 1, shape).astype(np.float32)).cuda())
        ys.append(torch.empty([shape[p] for p in (1, 3, 2, 0)], device="cuda"))
        perms.append((1, 3, 2, 0))
    gg = o.GroupedGemm(dev, 1.5, False)
    gp = o.GroupedPermute(xs, perms, ys)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        gg()
        gp()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    eager_c = [d[2].clone() for d in dev]
    eager_y = [y.clone() for y in ys]
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        gg()
        gp()
    junk = torch.randn(512, 512, device="cuda")
    for rep in range(4):
        for d in dev:
            d[2].fill_(float("nan"))
        for y in ys:
            y.fill_(float("nan"))
        junk = junk @ junk.T / 512  # unrelated eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for d, c in zip(dev, eager_c):
            assert torch.equal(d[2], c), rep
        for y, e in zip(ys, eager_y):
            assert torch.equal(y, e), rep
    for (A, B, _), c in zip(ps, eager_c):
        ref = 1.5 * np.matmul(A.astype(np.float64), B.astype(np.float64))
        assert_close(c.cpu().numpy(), ref, np.float32, A.shape[2] + 1, depth=A.shape[2])
    for x, y in zip(xs, eager_y):
        assert torch.equal(y, x.permute(1, 3, 2, 0))
    del graph
