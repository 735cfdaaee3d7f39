"""GPU parity of the CountingTropical{Float64,Float64} (OME_CTROP_F64) and TropicalAndOr
(OME_TROP_BOOL) element types through the C ABI, bit-exact against the oracle and against
combinatorial answers: counting maximum independent sets (Petersen graph: 4 and 5; random
3-regular graphs: enumeration of all subsets) through device tree plans, sliced and not, and
boolean products against integer matrix products.  Counts and tropical values are integers,
so every result is exact.  The oracle's modes are pinned in tests/test_oracle_semirings.py.
"""
import numpy as np
import pytest

from helpers import parse, rng
from oracle import loop
from omeinsum_mi355x import networks
from test_oracle_semirings import PETERSEN, brute_mis, mis_network

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


def ct_dev(z):
    return om().CountingTropical(torch.from_numpy(np.ascontiguousarray(z, dtype=np.complex128)).cuda())


def bool_dev(b):
    return om().TropicalAndOr(torch.from_numpy(np.ascontiguousarray(b, dtype=np.bool_)).cuda())


def backends():
    return [om().DefaultBackend(), om().MI355XBackend()]


def run(code, xs, sd, backend):
    o = om()
    o.set_einsum_backend(backend)
    try:
        y = o.einsum(code, xs, sd)
        torch.cuda.synchronize()
        return o.to_numpy(y)
    finally:
        o.set_einsum_backend(o.DefaultBackend())


def mis_tree(n, edges, slice_labels=()):
    o = om()
    ixs, xs = mis_network(n, edges)
    sd = {v: 2 for v in range(n)}
    code = o.optimize_code(o.EinCode(ixs, []), sd, o.GreedyMethod())
    if slice_labels:
        code = o.SlicedEinsum(list(slice_labels), code)
    return code, [ct_dev(x) for x in xs], sd


@pytest.mark.parametrize("n,seed", [(10, None), (18, 0), (18, 1), (20, 2)])
def test_counting_maximum_independent_sets(n, seed):
    edges = PETERSEN if seed is None else networks.random_regular_graph(n, 3, seed=seed)
    want = (4, 5) if seed is None else brute_mis(n, edges)
    for be in backends():
        for sl in ((), (0, 1)):
            code, xs, sd = mis_tree(n, edges, sl)
            y = run(code, xs, sd, be)
            assert (y.real, y.imag) == want, (n, seed, sl, type(be).__name__, y)


BINARY = [("ij,jk->ik", {"i": 40, "j": 50, "k": 30}),
          ("ijb,jkb->ikb", {"i": 20, "j": 17, "k": 24, "b": 3}),
          ("ij,ij->", {"i": 64, "j": 64}),
          ("ij,jk->ki", {"i": 5, "j": 3, "k": 7})]


@pytest.mark.parametrize("spec,sd", BINARY)
def test_counting_tropical_binary_rules(spec, sd):
    ixs, iy = parse(spec)
    r = rng(7301)
    xs = [r.integers(-3, 4, [sd[l] for l in ix]) + 1j * r.integers(1, 4, [sd[l] for l in ix])
          for ix in ixs]
    ref = loop.loop_einsum(ixs, iy, xs, sd, semiring="counting")
    for be in backends():
        got = run(om().EinCode(ixs, iy), [ct_dev(x) for x in xs], sd, be)
        assert np.array_equal(got, ref), (spec, type(be).__name__)


@pytest.mark.parametrize("spec,sd", BINARY + [("ij,jk,kl->il", {"i": 6, "j": 7, "k": 5, "l": 4})])
def test_and_or_products(spec, sd):
    ixs, iy = parse(spec)
    r = rng(7302)
    xs = [r.random([sd[l] for l in ix]) < 0.15 for ix in ixs]
    ref = loop.loop_einsum(ixs, iy, xs, sd, semiring="andor")
    for be in backends():
        got = run(om().EinCode(ixs, iy), [bool_dev(x) for x in xs], sd, be)
        assert got.dtype == np.bool_ and np.array_equal(got, ref), (spec, type(be).__name__)
    if spec == "ij,jk->ik":
        assert np.array_equal(ref, (xs[0].astype(np.int64) @ xs[1].astype(np.int64)) > 0)


UNARY = ["ij->", "ij->j", "ii->i", "ijk->kij", "a->aaaa", "i->ik", "illljkk->kijjcc"]


@pytest.mark.parametrize("spec", UNARY)
def test_unary_rules(spec):
    ixs, iy = parse(spec)
    sd = {l: 3 for ix in ixs + [iy] for l in ix}
    r = rng(7303)
    cx = [r.integers(-3, 4, [sd[l] for l in ix]) + 1j * r.integers(1, 4, [sd[l] for l in ix])
          for ix in ixs]
    bx = [r.random([sd[l] for l in ix]) < 0.3 for ix in ixs]
    cref = loop.loop_einsum(ixs, iy, cx, sd, semiring="counting")
    bref = loop.loop_einsum(ixs, iy, bx, sd, semiring="andor")
    for be in backends():
        assert np.array_equal(run(om().EinCode(ixs, iy), [ct_dev(x) for x in cx], sd, be), cref)
        assert np.array_equal(run(om().EinCode(ixs, iy), [bool_dev(x) for x in bx], sd, be), bref)


def test_scaling_and_permute():
    o = om()
    ixs, iy = parse("ij,jk->ik")
    sd = {"i": 24, "j": 20, "k": 18}
    r = rng(7304)
    bx = [r.random([sd[l] for l in ix]) < 0.1 for ix in ixs]
    y0 = r.random((24, 18)) < 0.5
    for be in backends():
        o.set_einsum_backend(be)
        try:
            y = bool_dev(y0)
            o.einsum_inplace(o.EinCode(ixs, iy), [bool_dev(x) for x in bx], y, True, True, sd)
            torch.cuda.synchronize()
            got = o.to_numpy(y)
        finally:
            o.set_einsum_backend(o.DefaultBackend())
        ref = loop.loop_einsum(ixs, iy, bx, sd, y=y0.copy(), sx=True, sy=True, semiring="andor")
        assert np.array_equal(got, ref)
    x = r.integers(-3, 4, (5, 6, 7)) + 1j * r.integers(1, 4, (5, 6, 7))
    y = o.CountingTropical(torch.empty((7, 5, 6), dtype=torch.complex128, device="cuda"))
    o.tensorpermute(y, ct_dev(x), [2, 0, 1])
    torch.cuda.synchronize()
    assert np.array_equal(o.to_numpy(y), np.transpose(x, (2, 0, 1)))
