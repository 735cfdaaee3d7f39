"""GPU parity of large single permutes whose dims, after merging consecutive runs, stay short
(tensor-network labels of extent 2, ranks 15-20): tensorpermute! (src/utils.jl:122-156)
through the C ABI, exact against numpy for every element size, with and without scaling.
"""
import numpy as np
import pytest

from helpers import rng

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def om():
    import omeinsum_mi355x
    return omeinsum_mi355x


def values(r, shape, dtype):
    if dtype == np.bool_:
        return r.random(shape) < 0.5
    v = r.integers(-8, 9, shape)
    if np.dtype(dtype).kind == "c":
        v = v + 1j * r.integers(-8, 9, shape)
    return v.astype(dtype)


CASES = [(15, 0), (16, 1), (18, 2), (17, 3), (20, 4)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128, np.int64,
                                   np.bool_])
@pytest.mark.parametrize("rank,seed", CASES)
def test_dim2_permutes(dtype, rank, seed):
    o = om()
    r = rng(4100 + seed)
    shape = [2] * rank
    if seed == 4:
        shape[3] = 3  # one longer dim
        shape[11] = 5
    perm = [int(p) for p in r.permutation(rank)]
    x = values(r, shape, dtype)
    ref = np.transpose(x, perm)
    t = torch.from_numpy(x).cuda()
    y = torch.empty(ref.shape, dtype=t.dtype, device="cuda")
    if dtype == np.bool_:
        t, y = o.TropicalAndOr(t), o.TropicalAndOr(y)
    o.tensorpermute(y, t, perm)
    torch.cuda.synchronize()
    assert np.array_equal(o.to_numpy(y), ref)
    if dtype not in (np.bool_,):
        y0 = values(r, ref.shape, dtype)
        yt = torch.from_numpy(np.ascontiguousarray(y0)).cuda()
        o.tensorpermute(yt, t, perm, 3, -2)
        torch.cuda.synchronize()
        assert np.array_equal(o.to_numpy(yt), 3 * ref - 2 * y0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128, np.int64,
                                   np.bool_])
@pytest.mark.parametrize("seed", range(6))
def test_power_of_two_permutes(dtype, seed):
    # extents 1, 2, 4, 8 and 64 (every one a power of two: the bit-permutation tile), 2^8 to
    # 2^20 elements, random permutations; exact, with and without scaling
    o = om()
    r = rng(4200 + seed)
    while True:
        rank = int(r.integers(2, 14))
        shape = [int(r.choice([1, 2, 2, 2, 4, 8, 64])) for _ in range(rank)]
        n = int(np.prod(shape))
        if 256 <= n <= (1 << 20):
            break
    perm = [int(p) for p in r.permutation(rank)]
    x = values(r, shape, dtype)
    ref = np.transpose(x, perm)
    t = torch.from_numpy(x).cuda()
    y = torch.empty(ref.shape, dtype=t.dtype, device="cuda")
    if dtype == np.bool_:
        t, y = o.TropicalAndOr(t), o.TropicalAndOr(y)
    o.tensorpermute(y, t, perm)
    torch.cuda.synchronize()
    assert np.array_equal(o.to_numpy(y), ref), (shape, perm)
    if dtype not in (np.bool_,):
        y0 = values(r, ref.shape, dtype)
        yt = torch.from_numpy(np.ascontiguousarray(y0)).cuda()
        o.tensorpermute(yt, t, perm, 3, -2)
        torch.cuda.synchronize()
        assert np.array_equal(o.to_numpy(yt), 3 * ref - 2 * y0), (shape, perm)


@pytest.mark.parametrize("dtype", [np.int64, np.complex64, np.float64])
@pytest.mark.parametrize("seed", range(3))
def test_gemm_with_both_operands_permuted(dtype, seed):
    # a binary contraction over extent-2 labels whose contracted labels sit between the
    # others in both operands: both are permuted before the GEMM (128 contracted points, past
    # the strided small-step route), in one paired bit-permutation launch; checked against
    # numpy's einsum (exact for Int64)
    from omeinsum_mi355x import codes
    o = om()
    r = rng(4300 + seed)
    letters = "abcdefghijklmnopqrst"
    order = [letters[i] for i in r.permutation(20)]
    o1, inner, o2 = order[:6], order[6:13], order[13:19]
    # inner labels interleaved with the free ones, in a different order in each operand
    ia = [x for pair in zip(o1, inner) for x in pair] + inner[6:]
    ib = [x for pair in zip(inner[::-1], o2) for x in pair] + inner[::-1][6:]
    iy = [o1[1], o2[0], o1[0]] + o1[2:] + o2[1:]
    code = "".join(ia) + "," + "".join(ib) + "->" + "".join(iy)
    a = values(r, [2] * len(ia), dtype)
    b = values(r, [2] * len(ib), dtype)
    ref = np.einsum(code, a, b)
    o.set_einsum_backend(o.MI355XBackend())
    try:
        got = o.to_numpy(codes.ein(code)(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()))
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    if dtype == np.int64:
        assert np.array_equal(got, ref), code
    else:
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= (2e-6 if dtype == np.complex64 else 1e-14), (code, err)
