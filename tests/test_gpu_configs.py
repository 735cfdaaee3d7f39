"""The benchmark workloads at their full sizes, through the C ABI, against FP64:

* ein"abcd,cdef->abef" at 64^4 (one 4096^3 product) in Float32 at fp32 / bf16x3 / fp64
  precision and ComplexF32 at fp32 / bf16x3, forward and the backward rule for both inputs,
  on sampled output rows against numpy FP64 with the norm-wise bounds of helpers.py;
* 4096 uneven ein"ijb,jkb->ikb" products in one grouped launch (fp32 and bf16x3), each against
  numpy FP64, and 1024 ein"ijkl->ljik" permutes in one grouped launch, bit-exact;
* the 100-vertex 3-regular ComplexF32 network contracted to a scalar (fp32, bf16x3, fp64
  precision, ComplexF64) against its committed FP64 value;
* the 48-request serving stream through the request loop in fp32 and bf16x3: every request
  against its committed FP64 value (tests/golden/request_stream_48.json), and each result
  the same alone and in the mix."""
import json
import os

import numpy as np
import pytest

from helpers import assert_close, norm_bound, rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield
    import omeinsum_mi355x as om
    from omeinsum_mi355x import _lib
    om.set_einsum_backend(om.DefaultBackend())
    _lib.context().set_option(_lib.OPT_PRECISION, 0)
    torch.cuda.empty_cache()


def om():
    import omeinsum_mi355x
    return omeinsum_mi355x


ROWS = np.arange(0, 4096, 64)  # sampled rows of the 4096 x 4096 matrix views


def _c4_inputs(dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if dtype == torch.complex64:
        mk = lambda: torch.complex(torch.randn(64, 64, 64, 64, device="cuda", generator=g),
                                   torch.randn(64, 64, 64, 64, device="cuda", generator=g))
    else:
        mk = lambda: torch.randn(64, 64, 64, 64, device="cuda", generator=g)
    return mk(), mk(), mk()


def _mat(t):
    W = np.complex128 if t.is_complex() else np.float64
    return t.cpu().numpy().reshape(4096, 4096).astype(W)


@pytest.mark.parametrize("dtype,prec", [(torch.float32, "fp32"), (torch.float32, "bf16x3"),
                                        (torch.float32, "fp64"), (torch.complex64, "fp32"),
                                        (torch.complex64, "bf16x3"), (torch.float32, "bf16x6"),
                                        (torch.complex64, "bf16x6")])
def test_c4_abcd_cdef_64_forward_backward(dtype, prec):
    o = om()
    a, b, dy = _c4_inputs(dtype, 4100 + 7 * (dtype == torch.complex64))
    code = o.ein("abcd,cdef->abef")
    sd = o.get_size_dict(code.ixs, (a, b))
    npdt = np.complex64 if dtype == torch.complex64 else np.float32
    o.set_einsum_backend(o.MI355XBackend(precision=prec))
    try:
        y = o.einsum(code, (a, b), sd)
        da, db_ = o.einsum_backward_rule(code, (a, b), y, sd, dy)
        torch.cuda.synchronize()
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    A, B, DY = _mat(a), _mat(b), _mat(dy)
    # row-major abcd -> (ab, cd): y = A B; dA = dy conj(B)^T; dB = conj(A)^T dy
    ref_y = A[ROWS] @ B
    assert_close(_mat(y)[ROWS].astype(npdt), ref_y, npdt, 16 * 4096, "C4 forward",
                 precision=prec, depth=4096)
    ref_da = DY[ROWS] @ np.conj(B).T
    assert_close(_mat(da)[ROWS].astype(npdt), ref_da, npdt, 16 * 4096, "C4 dA",
                 precision=prec, depth=4096)
    ref_db = np.conj(A[:, ROWS]).T @ DY
    assert_close(_mat(db_)[ROWS].astype(npdt), ref_db, npdt, 16 * 4096, "C4 dB",
                 precision=prec, depth=4096)


@pytest.fixture(scope="module")
def uneven():
    from omeinsum_mi355x.workloads import uneven_batch
    gemms, perms, info = uneven_batch()
    refs = [np.matmul(Bt.cpu().numpy().astype(np.float64), At.cpu().numpy().astype(np.float64))
            for Bt, At, _ in gemms]
    yield gemms, perms, info, refs
    del gemms, perms, info, refs
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6"])
def test_c3_grouped_gemms_full_size(uneven, prec):
    from omeinsum_mi355x import _lib
    gemms, _, _, refs = uneven
    assert len(gemms) == 4096
    ctx = _lib.context()
    ctx.set_option(_lib.OPT_PRECISION, _lib.PREC[prec])
    try:
        for _, _, Ct in gemms:
            Ct.fill_(float("nan"))
        om().GroupedGemm(gemms)()
        torch.cuda.synchronize()
    finally:
        ctx.set_option(_lib.OPT_PRECISION, 0)
    worst = 0.0
    for q, ((Bt, At, Ct), ref) in enumerate(zip(gemms, refs)):
        got = Ct.cpu().numpy()
        k = Bt.shape[2]
        assert_close(got, ref, np.float32, k + 1, "instance %d" % q, precision=prec, depth=k)
        worst = max(worst, rel_err(got, ref))
    print("C3 %s: worst norm-wise error over 4096 instances %.2e" % (prec, worst))


def test_c3_grouped_permutes_full_size(uneven):
    _, perms, _, _ = uneven
    assert len(perms) == 1024
    xs = [x for x, _ in perms]
    ys = [y for _, y in perms]
    for y in ys:
        y.fill_(float("nan"))
    om().GroupedPermute(xs, [(1, 3, 2, 0)] * len(xs), ys)()
    torch.cuda.synchronize()
    for q, (x, y) in enumerate(perms):
        assert torch.equal(y, x.permute(1, 3, 2, 0)), q
    for x, y in perms[:64]:
        assert np.array_equal(y.cpu().numpy(), np.transpose(x.cpu().numpy(), (1, 3, 2, 0)))


def _network_100():
    from omeinsum_mi355x import contractionorder as co, networks
    fx = json.load(open(os.path.join(GOLDEN, "network_100.json")))
    tmp = "/tmp/ome_golden_tree_cfg.json"
    with open(tmp, "w") as f:
        json.dump(fx["tree"], f)
    tree = co.readjson(tmp)
    edges = [tuple(e) for e in fx["edges"]]
    code = networks.network_code(fx["n"], edges)
    sd = {e: 2 for e in range(len(edges))}
    xs = networks.gaussian_tensors(code, sd, seed=fx["tensor_seed"], dtype=np.complex64)
    return tree, sd, xs, complex(*fx["value_fp64"])


# relative error of the full-network scalar against FP64, by arithmetic.  fp64 precision
# computes each step in FP64 and rounds it once to ComplexF32: the error of storing the 98
# intermediates in ComplexF32 (9e-7 here) is the floor of every ComplexF32 evaluation.  fp32
# adds the FP32 accumulation of each product.  bf16x3 adds the split's ~2^-16.5 relative
# representation error of every operand of the split products, which the network's
# conditioning amplifies (2.1e-5 measured here; the bound allows 2.5x).  bf16x6 keeps ~24 bits
# of every operand: the fp32 bound.  ComplexF64 inputs stay in FP64.
C2_BOUNDS = {"fp32": 1e-5, "bf16x3": 5e-5, "fp64": 2e-6, "c128": 1e-12, "bf16x6": 1e-5}


@pytest.mark.parametrize("kind", ["fp32", "bf16x3", "fp64", "c128", "bf16x6"])
def test_c2_network_100_scalar(kind):
    o = om()
    tree, sd, xs, ref = _network_100()
    if kind == "c128":
        xs = [x.astype(np.complex128) for x in xs]
    dxs = [torch.from_numpy(x).cuda() for x in xs]
    o.set_einsum_backend(o.MI355XBackend(precision="fp32" if kind == "c128" else kind))
    try:
        got = complex(o.to_numpy(tree(*dxs))[()])
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    e = abs(got - ref) / abs(ref)
    print("C2 %s: error vs FP64 %.2e" % (kind, e))
    assert e <= C2_BOUNDS[kind], (kind, e)


def _stream(sc_budget=None):
    from omeinsum_mi355x import contractionorder as co, networks
    fx = json.load(open(os.path.join(GOLDEN, "request_stream_48.json")))
    reqs = []
    for q, (spec, fq) in enumerate(zip(networks.request_stream(48, fx["seed"]), fx["requests"])):
        assert spec["graph_seed"] == fq["graph_seed"] and spec["n"] == fq["n"], q
        code = networks.network_code(spec["n"], spec["edges"])
        sd = {e: 2 for e in range(len(spec["edges"]))}
        tree = co.optimize_code(code, sd)
        assert co.flop(tree, sd) == fq["flop"], q   # the same tree as the fixture's
        if sc_budget is not None and co.contraction_complexity(tree, sd)[1] > sc_budget:
            tree = co.slice_code(tree, sd, sc_target=sc_budget)
        xs = networks.gaussian_tensors(code, sd, spec["tensor_seed"], np.complex64)
        reqs.append((tree, sd, [torch.from_numpy(x).cuda() for x in xs],
                     complex(*fq["value_fp64"])))
    return reqs


# relative error of one request's scalar against FP64, by the loop's arithmetic.  FP32: the
# ComplexF32 storage of the intermediates alone costs up to 2.6e-6 on these networks (request
# 28, 134 vertices: every step in FP64, rounded to ComplexF32); products sum at most 8192 real
# terms per FP32 accumulator.  BF16x3: the split's representation error, amplified by each
# network's conditioning: up to 1.3e-4 measured (request 14), within 3e-4.  BF16x6: the fp32
# bound (its large products keep ~24 bits of every operand and accumulate in FP32).
C5_BOUNDS = {"fp32": 1e-5, "bf16x3": 3e-4, "bf16x6": 1e-5}


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6"])
def test_c5_request_stream_alone_and_mixed(prec):
    o = om()
    from omeinsum_mi355x.server import Server
    # the requests go in unsliced: those whose plans need more than the 1 GiB per-request
    # limit are sliced by the loop at submission
    reqs = _stream()
    be = o.MI355XBackend(precision=prec)
    kw = dict(hbm_budget=16 << 30, backend=be, request_limit=1 << 30)
    with Server(max_active=1, **kw) as srv:
        alone = []
        for tree, sd, dxs, _ in reqs:
            y, _ = srv.wait(srv.submit(srv.plan(tree, dxs, sd), dxs))
            alone.append(complex(o.to_numpy(y)[()]))
        assert srv.stats()["sliced"] >= 4
    with Server(max_active=8, **kw) as srv:
        plans = [srv.plan(tree, dxs, sd) for tree, sd, dxs, _ in reqs]
        tickets = [srv.submit(p, r[2]) for p, r in zip(plans, reqs)]
        mixed = [complex(o.to_numpy(srv.wait(t)[0])[()]) for t in tickets]
        print("C5 %s loop stats: %s" % (prec, srv.stats()))
    errs = []
    for q, ((_, _, _, ref), a, m) in enumerate(zip(reqs, alone, mixed)):
        assert abs(m - a) <= 1e-6 * abs(a), (q, m, a)
        errs.append(abs(m - ref) / abs(ref))
    print("C5 %s: error vs FP64 over 48 requests: median %.2e max %.2e (request %d)" %
          (prec, float(np.median(errs)), max(errs), int(np.argmax(errs))))
    assert max(errs) <= C5_BOUNDS[prec], (prec, errs)


def test_c2_network_100_normalized_fp64_scalar():
    # per-tensor and per-step log-scale normalisation, every step and the final scalar in
    # FP64: the FP64 value of the committed ComplexF32 inputs, to FP64 rounding
    o = om()
    from omeinsum_mi355x.tree import plan_for
    tree, sd, xs, ref = _network_100()
    dxs = [torch.from_numpy(x).cuda() for x in xs]
    o.set_einsum_backend(o.MI355XBackend(precision="fp32"))
    try:
        v, lg = plan_for(tree, dxs, sd).execute_scalar_normalized(dxs)
    finally:
        o.set_einsum_backend(o.DefaultBackend())
    e = abs(v * np.exp(lg) - ref) / abs(ref)
    print("C2 normalised FP64: error vs FP64 %.2e" % e)
    assert e <= 1e-10, e
