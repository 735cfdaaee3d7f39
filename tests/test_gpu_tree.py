"""GPU parity of contraction trees: NestedEinsum / SlicedEinsum through the MI355X backend's
device plans (arena + graph replay) and through the per-step default backend, gradients, and
log-scale normalised full-network contraction — all against the oracle."""
import numpy as np
import pytest

from helpers import assert_close, rand, rel_err, rng
from oracle import default_backend as db
from omeinsum_mi355x import codes, contractionorder as co, networks

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


BACKENDS = ["default", "mi355x"]


def use(name, **kw):
    o = om()
    o.set_einsum_backend(o.DefaultBackend() if name == "default" else o.MI355XBackend(**kw))


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("dtype", [np.float64, np.complex64, np.int64])
def test_nested_equals_flat(backend, dtype):
    # test/einsequence.jl:28-48
    use(backend)
    r = rng(2001)
    ne = codes.ein("(ij,jk),km->im")
    xs = [rand(r, s, dtype) for s in ((5, 6), (6, 7), (7, 8))]
    got = om().to_numpy(ne(*[dev(x) for x in xs]))
    ref = db.nested_einsum(ne.to_plain(), xs, dict(zip("ijkm", (5, 6, 7, 8))))
    assert_close(got, ref, dtype, 42, "nested")
    ne2 = codes.ein("((i,ij),i),j->ik")
    xs = [rand(r, s, dtype) for s in ((4,), (4, 4), (4,), (4,))]
    sd = {"i": 4, "j": 4, "k": 3}
    got = om().to_numpy(ne2(*[dev(x) for x in xs], size_info={"k": 3}))
    ref = db.einsum([["i"], ["i", "j"], ["i"], ["j"]], ["i", "k"], xs, sd)
    assert_close(got, ref, dtype, 16, "nested with broadcast output")


@pytest.mark.parametrize("backend", BACKENDS)
def test_sliced_equals_unsliced(backend):
    # test/slicing.jl:13-32
    use(backend)
    r = rng(2002)
    se = codes.SlicedEinsum(["i", "l"], codes.ein("(ij,jk),(kl,lm)->im"))
    xs = [rand(r, s, np.float64) for s in ((2, 3), (3, 4), (4, 5), (5, 6))]
    dxs = [dev(x) for x in xs]
    expected = om().to_numpy(se.eins(*dxs))
    for _ in range(3):
        assert_close(om().to_numpy(se(*dxs)), expected, np.float64, 60, "sliced")
    y = torch.empty(2, 6, dtype=torch.float64, device="cuda")
    om().einsum_inplace(se, dxs, y, True, False)
    assert_close(om().to_numpy(y), expected, np.float64, 60, "sliced einsum!")
    ref = db.nested_einsum(se.eins.to_plain(), xs, codes.get_size_dict(se.getixsv(), xs))
    assert_close(expected, ref, np.float64, 60, "vs oracle")


def test_graph_replay_with_eager_work_between():
    # replays of one recorded graph, with unrelated eager work and synchronisations in
    # between, all match the eager evaluation
    o = om()
    r = rng(2003)
    ne = codes.ein("((ab,bc),cd),(de,ea)->")
    sd = dict(zip("abcde", (8, 9, 10, 11, 12)))
    xs = [rand(r, [sd[l] for l in ix], np.complex64) for ix in ne.getixsv()]
    dxs = [dev(x) for x in xs]
    use("mi355x", graphs=False)
    eager = o.to_numpy(ne(*dxs))
    use("mi355x", graphs=True)
    ref = db.nested_einsum(ne.to_plain(), [x.astype(np.complex128) for x in xs], sd)
    # a scalar summed over 8*9*10*11*12 products: depth 95040
    assert_close(eager, ref, np.complex64, 100, depth=95040)
    y = torch.empty((), dtype=torch.complex64, device="cuda")
    for k in range(4):
        o.einsum_inplace(ne, dxs, y, True, False)
        junk = o.einsum(o.ein("ij,jk->ik"), (dev(rand(r, (33, 40), np.float32)), dev(rand(r, (40, 7), np.float32))))
        torch.cuda.synchronize()
        assert o.to_numpy(y) == pytest.approx(eager, rel=1e-6), k
        del junk
    # sy != 0 accumulates on replay
    y2 = dev(np.array(1 + 2j, dtype=np.complex64))
    o.einsum_inplace(ne, dxs, y2, 2.0, 0.5)
    assert o.to_numpy(y2) == pytest.approx(2 * eager + 0.5 * (1 + 2j), rel=1e-5)


@pytest.mark.parametrize("backend", BACKENDS)
def test_fullerene_tropical_66(backend):
    # test/contractionorder.jl:31-70: 66 unsliced and sliced, on the device
    use(backend)
    o = om()
    edges = networks.fullerene_edges()
    ixs = [list(e) for e in edges] + [[i] for i in range(60)]
    code = codes.EinCode(ixs, [])
    sd = {i: 2 for i in range(60)}
    xs = [np.array([[-1.0, 1.0], [1.0, -1.0]])] * 90 + [np.zeros(2)] * 60
    dxs = [dev(x, tropical=True) for x in xs]
    tree = co.optimize_code(code, sd)
    assert o.to_numpy(tree(*dxs))[()] == 66.0
    sliced = co.slice_code(tree, sd, sc_target=8)
    assert sliced.slicing
    assert o.to_numpy(sliced(*dxs))[()] == 66.0


def _network(n, seed, dtype=np.complex64):
    edges = networks.random_regular_graph(n, 3, seed=seed)
    code = networks.network_code(n, edges)
    sd = {e: 2 for e in range(len(edges))}
    xs = networks.gaussian_tensors(code, sd, seed + 1, dtype)
    tree = co.optimize_code(code, sd)
    return code, tree, sd, xs


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_regular_network_vs_fp64(prec):
    # a 3-regular network (bond dimension 2, ComplexF32) contracted to a scalar
    code, tree, sd, xs = _network(40, 3001)
    ref = complex(db.nested_einsum(tree.to_plain(), [x.astype(np.complex128) for x in xs], sd)[()])
    use("mi355x", precision=prec)
    got = complex(om().to_numpy(tree(*[dev(x) for x in xs]))[()])
    assert abs(got - ref) <= 1e-4 * abs(ref), (got, ref)
    use("default")
    got2 = complex(om().to_numpy(tree(*[dev(x) for x in xs]))[()])
    assert abs(got2 - ref) <= 1e-4 * abs(ref), (got2, ref)


def test_normalized_scalar_and_sliced_fp64_accumulation():
    code, tree, sd, xs = _network(30, 3002)
    ref = complex(db.nested_einsum(tree.to_plain(), [x.astype(np.complex128) for x in xs], sd)[()])
    use("mi355x")
    from omeinsum_mi355x.tree import plan_for
    dxs = [dev(x) for x in xs]
    # ComplexF32 inputs, evaluated in FP64 after per-tensor normalisation: the FP64 oracle's
    # value of the same inputs to FP64 rounding
    v, lg = plan_for(tree, dxs, sd).execute_scalar_normalized(dxs)
    got = v * np.exp(lg)
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
    sliced = co.slice_code(tree, sd, sc_target=4)
    assert len(sliced.slicing) >= 1
    got_s = complex(om().to_numpy(sliced(*dxs))[()])
    assert abs(got_s - ref) <= 1e-4 * abs(ref), (got_s, ref)
    v2, lg2 = plan_for(sliced, dxs, sd).execute_scalar_normalized(dxs)
    assert abs(v2 * np.exp(lg2) - ref) <= 1e-12 * abs(ref)
    # inputs scaled by 2^100 (exactly, in ComplexF32): the value itself, 2^3000 times larger,
    # is far outside Float64; the normalised result is the same value with a log larger by
    # 30 * 100 ln 2
    big = [dev((x * np.float32(2.0 ** 100)).astype(np.complex64)) for x in xs]
    v3, lg3 = plan_for(tree, big, sd).execute_scalar_normalized(big)
    assert abs(v3 * np.exp(lg3 - lg - len(xs) * 100 * np.log(2.0)) - v) <= 1e-12 * abs(v)
    assert abs((lg3 - lg) - len(xs) * 100 * np.log(2.0)) <= 1e-9 * len(xs) * 100


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_cost_and_gradient(backend, dtype):
    # test/bp.jl:3-17 and the closed-form check of src/autodiff.jl:14-21
    use(backend)
    o = om()
    r = rng(2004)
    ne = codes.ein("(ij, jk), ki->")
    xs = [rand(r, s, dtype) for s in ((2, 3), (3, 4), (4, 2))]
    cost, grads = o.cost_and_gradient(ne, [dev(x) for x in xs])
    rcost, rgrads = db.cost_and_gradient(ne.to_plain(), xs)
    assert_close(o.to_numpy(cost), rcost, dtype, 12)
    for g, rg in zip(grads, rgrads):
        assert_close(o.to_numpy(g), rg, dtype, 12, "gradient")


def test_gradient_larger_tree_and_petersen():
    use("mi355x")
    o = om()
    r = rng(2005)
    code, tree, sd, xs = _network(20, 3003, np.complex128)
    cost, grads = o.cost_and_gradient(tree, [dev(x) for x in xs])
    rcost, rgrads = db.cost_and_gradient(tree.to_plain(), xs)
    assert_close(o.to_numpy(cost), rcost, np.complex128, 100)
    for g, rg in zip(grads, rgrads):
        assert_close(o.to_numpy(g), rg, np.complex128, 100)
    # Petersen: relaxing one vertex still gives no colouring (docs/src/applications.md:50-55)
    import itertools
    s = np.zeros((3, 3, 3), dtype=np.int64)
    for i, j, k in itertools.product(range(3), repeat=3):
        s[i, j, k] = int(len({i, j, k}) == 3)
    pc = codes.ein("afl,bhn,cjf,dlh,enj,ago,big,cki,dmk,eom->")
    ptree = co.optimize_code(pc, codes.uniformsize(pc, 3))
    cost, grads = o.cost_and_gradient(ptree, [dev(s)] * 10)
    assert o.to_numpy(cost)[()] == 0 and o.to_numpy(grads[0]).sum() == 0


def test_plan_freezes_planning_options():
    # a plan keeps the precision it was created with; the backend's precision selects the plan
    o = om()
    from omeinsum_mi355x import _lib
    from omeinsum_mi355x.tree import plan_for
    r = rng(2006)
    ne = codes.ein("(ij,jk),kl->il")
    sd = dict(zip("ijkl", (256, 256, 256, 256)))
    xs = [rand(r, (256, 256), np.float32) for _ in range(3)]
    dxs = [dev(x) for x in xs]
    ref = db.nested_einsum(ne.to_plain(), [x.astype(np.float64) for x in xs], sd)
    use("mi355x", precision="fp32")
    r32 = o.to_numpy(ne(*dxs))
    plan32 = plan_for(ne, dxs, sd)
    use("mi355x", precision="bf16x3")
    r3 = o.to_numpy(ne(*dxs))
    # two chained products of depth 256
    assert_close(r32, ref, np.float32, 256 * 256, "fp32 plan", depth=512)
    assert_close(r3, ref, np.float32, 256 * 256, "bf16x3 plan", precision="bf16x3", depth=512)
    assert not np.array_equal(r32, r3)  # two plans, two arithmetics
    # the fp32 plan, run while the context says bf16x3, still computes in fp32
    ctx = _lib.context()
    assert ctx.get_option(_lib.OPT_PRECISION) == _lib.PREC["bf16x3"]
    again = o.to_numpy(plan32.execute(dxs))
    assert np.array_equal(again, r32)
