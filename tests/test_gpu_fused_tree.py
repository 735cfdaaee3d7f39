"""GPU parity of the fused small-step launch on tree shapes that reach its limits, against
numpy restatements of the same contractions (matrix chain products):

* a chain of 80 matrices contracted left to right: one fused subtree 78 levels deep, so its
  chunk table is longer than the 64 chunks the launch holds in LDS, and 77 leaves read above
  level 1 are copied into the LDS arena by copy steps spread over several level-1 chunks;
* the same chain of 16 x 16 ComplexF64 matrices: 4 KiB per intermediate and per leaf, more
  than the 96 KiB LDS arena holds, so the later intermediates and leaves stay in global memory
  (the arena / in place), and each output is split over several lanes (16 contracted points);
* a balanced tree over 64 matrices: 32 independent steps at level 1, two chunks of 16.

Exact element types (Int64, Tropical) must match bit for bit; floats within the norm-wise
bounds of helpers.assert_close."""
import numpy as np
import pytest

from helpers import assert_close, rand, rng
from omeinsum_mi355x import codes

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


def labels(n):
    return [chr(0x4E00 + i) for i in range(n + 1)]


def chain_code(n):
    """((m0, m1), m2), ... m_{n-1} -> the labels of the chain's two ends."""
    L = labels(n)
    s = L[0] + L[1] + "," + L[1] + L[2]
    for i in range(2, n):
        s = "(" + s + ")," + L[i] + L[i + 1]
    return codes.ein(s + "->" + L[0] + L[n])


def balanced_code(n):
    L = labels(n)
    parts = [L[i] + L[i + 1] for i in range(n)]
    while len(parts) > 1:
        parts = ["(" + parts[i] + "," + parts[i + 1] + ")" if i + 1 < len(parts) else parts[i]
                 for i in range(0, len(parts), 2)]
    s = parts[0][1:-1]  # the outer pair is the root's argument list
    return codes.ein(s + "->" + L[0] + L[n])


def chain_product(xs, tropical=False):
    acc = xs[0]
    for x in xs[1:]:
        if tropical:
            acc = np.max(acc[:, :, None] + x[None, :, :], axis=1)
        else:
            acc = acc @ x
    return acc


def run(ne, xs, tropical=False, **kw):
    o = om()
    o.set_einsum_backend(o.MI355XBackend(**kw))
    dxs = [torch.from_numpy(np.array(x, order="C")).cuda() for x in xs]
    if tropical:
        dxs = [o.Tropical(t) for t in dxs]
    return o.to_numpy(ne(*dxs))


def unit_matrices(r, n, m, dtype):
    # entries of modulus about 1/sqrt(m): the chain product stays of order one
    return [rand(r, (m, m), dtype) / np.sqrt(m) for _ in range(n)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64])
def test_long_chain_float(dtype):
    n, m = 80, 4
    xs = unit_matrices(rng(3101), n, m, dtype)
    got = run(chain_code(n), xs)
    ref = chain_product([x.astype(np.complex128 if np.iscomplexobj(x) else np.float64) for x in xs])
    assert_close(got, ref, dtype, m * n, "80-matrix chain")


def test_long_chain_int64_exact():
    n, m = 80, 4
    r = rng(3102)
    # entries in {-1, 0, 1} with one nonzero per row: products stay small and exact
    xs = []
    for _ in range(n):
        x = np.zeros((m, m), np.int64)
        x[np.arange(m), r.integers(0, m, m)] = r.choice([-1, 1], m)
        xs.append(x)
    got = run(chain_code(n), xs)
    assert np.array_equal(got, chain_product(xs))


def test_long_chain_tropical_exact():
    n, m = 80, 4
    r = rng(3103)
    xs = [r.integers(-8, 9, (m, m)).astype(np.float64) for _ in range(n)]
    got = run(chain_code(n), xs, tropical=True)
    assert np.array_equal(np.asarray(got, np.float64), chain_product(xs, tropical=True))


def test_chain_beyond_the_lds_arena_complex128():
    n, m = 80, 16
    xs = unit_matrices(rng(3104), n, m, np.complex128)
    got = run(chain_code(n), xs)
    assert_close(got, chain_product(xs), np.complex128, m * n, "16x16 complex chain")


@pytest.mark.parametrize("dtype", [np.float32, np.int64])
def test_balanced_tree_wide_levels(dtype):
    n, m = 64, 4
    r = rng(3105)
    if dtype == np.int64:
        xs = [r.integers(-1, 2, (m, m)).astype(np.int64) for _ in range(n)]
        got = run(balanced_code(n), xs)
        # |entries| <= 4^63 would overflow; the ring arithmetic wraps the same way in numpy
        assert np.array_equal(got, chain_product(xs))
    else:
        xs = unit_matrices(r, n, m, dtype)
        got = run(balanced_code(n), xs)
        assert_close(got, chain_product([x.astype(np.float64) for x in xs]), dtype, m * n,
                     "balanced 64-matrix tree")


def test_chain_runs_as_one_fused_launch():
    # the 78 non-root steps are one fused subtree: the plan launches the fused kernel and the
    # root's product, so the cases above do exercise the fused launch
    from omeinsum_mi355x.tree import plan_for
    n, m = 80, 4
    o = om()
    o.set_einsum_backend(o.MI355XBackend())
    xs = unit_matrices(rng(3106), n, m, np.float64)
    dxs = [torch.from_numpy(x).cuda() for x in xs]
    info = plan_for(chain_code(n), dxs, {l: m for l in labels(n)}).info()
    assert info.n_steps == n - 1
    assert info.n_kernels <= 3, info.n_kernels
