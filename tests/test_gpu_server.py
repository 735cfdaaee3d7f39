"""GPU tests of the continuous-batching request loop (ome_server_*) through the C ABI:
results through the loop equal the oracle and equal the same request evaluated alone, when
requests of different sizes are mixed, queued behind the HBM budget and the slot count, and
sliced."""
import numpy as np
import pytest

from oracle import default_backend as db
from omeinsum_mi355x import codes, contractionorder as co, networks

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield


def om():
    import omeinsum_mi355x
    return omeinsum_mi355x


def request(n, seed, dtype=np.complex64):
    edges = networks.random_regular_graph(n, 3, seed=seed)
    code = networks.network_code(n, edges)
    sd = {e: 2 for e in range(len(edges))}
    xs = networks.gaussian_tensors(code, sd, seed + 1, dtype)
    tree = co.optimize_code(code, sd)
    return tree, sd, xs, [torch.from_numpy(x).cuda() for x in xs]


def fp64(tree, sd, xs):
    return complex(db.nested_einsum(tree.to_plain(), [x.astype(np.complex128) for x in xs], sd)[()])


def test_single_request_matches_oracle_and_direct_plan():
    from omeinsum_mi355x.server import Server
    tree, sd, xs, dxs = request(40, 5001)
    ref = fp64(tree, sd, xs)
    with Server(hbm_budget=1 << 30, max_active=2) as srv:
        plan = srv.plan(tree, dxs, sd)
        y, lat = srv.wait(srv.submit(plan, dxs))
        got = complex(om().to_numpy(y)[()])
        assert abs(got - ref) <= 1e-4 * abs(ref), (got, ref)
        assert lat > 0
        # the same plan called directly (it allocates its own memory on first use)
        direct = complex(om().to_numpy(plan.execute(dxs))[()])
        assert direct == got


def test_mixed_requests_equal_alone():
    # requests of different sizes: each result through a busy loop (4 slots, a budget that
    # holds only a few plans at a time) equals the same request run alone
    from omeinsum_mi355x.server import Server, plan_bytes
    reqs = [request(n, 5100 + i) for i, n in enumerate([20, 44, 30, 60, 24, 52, 36, 28, 48, 40,
                                                         22, 56, 32, 26, 46, 38])]
    with Server(hbm_budget=1 << 30, max_active=1) as alone_srv:
        alone = []
        for tree, sd, xs, dxs in reqs:
            plan = alone_srv.plan(tree, dxs, sd)
            y, _ = alone_srv.wait(alone_srv.submit(plan, dxs))
            alone.append(complex(om().to_numpy(y)[()]))
    plans = []
    with Server(hbm_budget=1 << 30, max_active=4) as probe:
        plans = [probe.plan(tree, dxs, sd) for tree, sd, xs, dxs in reqs]
    budget = 3 * max(plan_bytes(p) for p in plans)
    with Server(hbm_budget=budget, max_active=4) as srv:
        plans = [srv.plan(tree, dxs, sd) for tree, sd, xs, dxs in reqs]
        tickets = [srv.submit(p, r[3]) for p, r in zip(plans, reqs)]
        states = [srv.poll(t) for t in tickets]
        assert all(s in (0, 1, 2) for s in states)
        mixed = [complex(om().to_numpy(srv.wait(t)[0])[()]) for t in tickets]
    for (tree, sd, xs, _), a, m in zip(reqs, alone, mixed):
        assert abs(m - a) <= 1e-6 * abs(a), (m, a)
        ref = fp64(tree, sd, xs)
        assert abs(m - ref) <= 1e-4 * abs(ref), (m, ref)


def test_budget_admission_and_rejection():
    from omeinsum_mi355x import _lib
    from omeinsum_mi355x.server import Server, plan_bytes
    tree, sd, xs, dxs = request(48, 5201)
    with Server(hbm_budget=1 << 30, max_active=2) as probe:
        need = plan_bytes(probe.plan(tree, dxs, sd))
    # a budget that holds exactly one copy: eight submissions run one after another
    with Server(hbm_budget=need, max_active=4) as srv:
        plans = [srv.plan(tree, dxs, sd) for _ in range(8)]
        tickets = [srv.submit(p, dxs) for p in plans]
        outs = [complex(om().to_numpy(srv.wait(t)[0])[()]) for t in tickets]
        assert len(set(outs)) == 1
    # a plan larger than the whole budget is sliced at submission until it fits; a budget
    # no slicing fits refuses the request
    ref = fp64(tree, sd, xs)
    with Server(hbm_budget=max(need // 2, 1), max_active=2) as srv:
        plan = srv.plan(tree, dxs, sd)
        got = complex(om().to_numpy(srv.wait(srv.submit(plan, dxs))[0])[()])
        assert srv.stats()["sliced"] == 1
    assert abs(got - ref) <= 1e-5 * abs(ref), (got, ref)
    with Server(hbm_budget=4096, max_active=2) as srv:
        plan = srv.plan(tree, dxs, sd)
        with pytest.raises(MemoryError):
            srv.submit(plan, dxs)
    assert isinstance(_lib.OPT_TREE_DEFER_MEMORY, int)


def test_request_limit_slices_at_submission_and_blocks_return_to_their_plan():
    # a per-request limit below a plan's memory: the loop admits a sliced twin (FP64
    # cross-slice sum) whose memory fits; resubmitting a plan lends it the block it used
    # before, and its recorded graph replays with the same result
    from omeinsum_mi355x.server import Server, plan_bytes
    reqs = [request(n, 5500 + i) for i, n in enumerate([56, 64, 44, 60])]
    with Server(hbm_budget=1 << 30, max_active=2) as probe:
        sizes = [plan_bytes(probe.plan(tree, dxs, sd)) for tree, sd, xs, dxs in reqs]
    limit = sorted(sizes)[1]  # the two largest plans exceed it
    with Server(hbm_budget=1 << 30, max_active=2, request_limit=limit) as srv:
        plans = [srv.plan(tree, dxs, sd) for tree, sd, xs, dxs in reqs]
        first = [complex(om().to_numpy(srv.wait(srv.submit(p, r[3]))[0])[()])
                 for p, r in zip(plans, reqs)]
        st = srv.stats()
        assert st["sliced"] == sum(1 for b in sizes if b > limit) >= 1
        again = [complex(om().to_numpy(srv.wait(srv.submit(p, r[3]))[0])[()])
                 for p, r in zip(plans, reqs)]
        assert srv.stats()["blocks_reused"] >= 1
    for (tree, sd, xs, _), a, b in zip(reqs, first, again):
        ref = fp64(tree, sd, xs)
        assert abs(a - ref) <= 1e-5 * abs(ref), (a, ref)
        assert a == b


def test_sliced_request_through_the_loop():
    from omeinsum_mi355x.server import Server
    tree, sd, xs, dxs = request(40, 5301)
    sliced = co.slice_code(tree, sd, sc_target=5)
    assert sliced.slicing
    ref = fp64(tree, sd, xs)
    with Server(hbm_budget=1 << 30, max_active=2) as srv:
        p1, p2 = srv.plan(tree, dxs, sd), srv.plan(sliced, dxs, sd)
        t1, t2 = srv.submit(p1, dxs), srv.submit(p2, dxs)
        a = complex(om().to_numpy(srv.wait(t1)[0])[()])
        b = complex(om().to_numpy(srv.wait(t2)[0])[()])
    assert abs(a - ref) <= 1e-4 * abs(ref)
    assert abs(b - ref) <= 1e-4 * abs(ref)


def test_admission_waits_for_reused_pool_memory():
    # A request queued behind a busy slot is admitted later and may be lent a pool block that
    # the caller freed right after enqueuing writes to it on the context's stream (the pool
    # reuses blocks in the order of that stream).  Admission orders the request after the
    # context's stream, so those writes land before the request uses the block: its result
    # equals the same request run alone.
    import ctypes
    from omeinsum_mi355x import _lib
    from omeinsum_mi355x.server import Server, plan_bytes
    tree0, sd0, xs0, dxs0 = request(60, 5401)
    tree, sd, xs, dxs = request(44, 5402)
    g = torch.Generator(device="cuda").manual_seed(5403)
    big = torch.randn(4096, 4096, device="cuda", generator=g)
    with Server(hbm_budget=1 << 30, max_active=1) as srv:
        p0 = srv.plan(tree0, dxs0, sd0)
        p = srv.plan(tree, dxs, sd)
        alone = complex(om().to_numpy(srv.wait(srv.submit(p, dxs))[0])[()])
        ctx = srv.ctx
        nb = plan_bytes(p)
        garbage = torch.full((nb // 4 + 1,), float("nan"), device="cuda")
        blk = ctypes.c_void_p()
        _lib.check(ctx.lib.ome_alloc(ctx.h, nb, ctypes.byref(blk)), ctx.h)
        t0 = srv.submit(p0, dxs0)
        t1 = srv.submit(p, dxs)  # queued behind t0: one slot
        acc = torch.zeros_like(big)
        for _ in range(20):  # milliseconds of work on the context's stream, then the writes
            acc += big @ big
            _lib.check(ctx.lib.ome_memcpy_d2d(ctx.h, blk, ctypes.c_void_p(garbage.data_ptr()),
                                              nb), ctx.h)
        _lib.check(ctx.lib.ome_free(ctx.h, blk), ctx.h)
        y0, _ = srv.wait(t0)
        y1, _ = srv.wait(t1)
        torch.cuda.synchronize()
        got = complex(om().to_numpy(y1)[()])
        assert abs(got - alone) <= 1e-6 * abs(alone), (got, alone)
        ref0 = fp64(tree0, sd0, xs0)
        got0 = complex(om().to_numpy(y0)[()])
        assert abs(got0 - ref0) <= 1e-4 * abs(ref0), (got0, ref0)


def test_cross_request_batching_matches_unbatched():
    # requests admitted in one round run their fused small-step subtrees in one shared launch;
    # every result equals the same request's result with batching off, bit for bit
    from omeinsum_mi355x.server import Server
    reqs = [request(n, 5600 + i) for i, n in enumerate([20, 24, 28, 30, 32, 36, 40, 44, 26, 22,
                                                         34, 38])]
    outs = {}
    for batching in (False, True):
        with Server(hbm_budget=1 << 30, max_active=8, batching=batching) as srv:
            plans = [srv.plan(tree, dxs, sd) for tree, sd, xs, dxs in reqs]
            for _ in range(2):  # the second pass replays graphs recorded without fused steps
                # submitted under hold: admitted in rounds of up to 8 whatever the timing
                srv.hold(True)
                tickets = [srv.submit(p, r[3]) for p, r in zip(plans, reqs)]
                srv.hold(False)
                outs[batching] = [complex(om().to_numpy(srv.wait(t)[0])[()]) for t in tickets]
            st = srv.stats()
        if batching:
            assert st["batched_launches"] >= 1
            assert st["batched_requests"] > st["batched_launches"]
        else:
            assert st["batched_launches"] == 0
    assert outs[True] == outs[False]
    for (tree, sd, xs, _), v in zip(reqs, outs[True]):
        ref = fp64(tree, sd, xs)
        assert abs(v - ref) <= 1e-5 * abs(ref), (v, ref)


def test_injected_status_fails_one_request_alone():
    # eight requests share four slots; the fourth writes its slot's status word after its steps
    # (every kernel runs to its end): it fails with a device error naming the status, the seven
    # beside it complete with their results from a loop without the injection, and the failed
    # plan runs cleanly when submitted again
    from omeinsum_mi355x._lib import OMEError
    from omeinsum_mi355x.server import Server
    reqs = [request(n, 5700 + i) for i, n in enumerate([20, 30, 24, 28, 22, 26, 32, 34])]
    with Server(hbm_budget=1 << 30, max_active=4) as srv:
        plans = [srv.plan(tree, dxs, sd) for tree, sd, xs, dxs in reqs]
        clean = [complex(om().to_numpy(srv.wait(srv.submit(p, r[3]))[0])[()])
                 for p, r in zip(plans, reqs)]
        tickets = []
        for i, (p, r) in enumerate(zip(plans, reqs)):
            if i == 3:
                srv.inject_status(7)
            tickets.append(srv.submit(p, r[3]))
        for i, t in enumerate(tickets):
            if i == 3:
                with pytest.raises(OMEError, match="status 7"):
                    srv.wait(t)
            else:
                assert complex(om().to_numpy(srv.wait(t)[0])[()]) == clean[i]
        again = complex(om().to_numpy(srv.wait(srv.submit(plans[3], reqs[3][3]))[0])[()])
        assert again == clean[3]
        st = srv.stats()
        assert st["submitted"] == 2 * len(reqs) + 1
        assert st["completed"] == 2 * len(reqs)


def test_trace_ranges_leave_results_unchanged():
    # roctx ranges on (tree steps, graph replays, grouped launches, requests): the same tree
    # through a direct plan and through the loop gives the results it gives with ranges off
    from omeinsum_mi355x import _lib
    from omeinsum_mi355x.server import Server
    ctx = _lib.context()
    tree, sd, xs, dxs = request(36, 5800)
    outs = {}
    for on in (0, 1):
        ctx.set_option(_lib.OPT_TRACE_RANGES, on)
        try:
            assert ctx.get_option(_lib.OPT_TRACE_RANGES) == on
            with Server(hbm_budget=1 << 30, max_active=2) as srv:
                plan = srv.plan(tree, dxs, sd)
                looped = [complex(om().to_numpy(srv.wait(srv.submit(plan, dxs))[0])[()])
                          for _ in range(2)]
                direct = complex(om().to_numpy(plan.execute(dxs))[()])
            outs[on] = (looped, direct)
        finally:
            ctx.set_option(_lib.OPT_TRACE_RANGES, 0)
    assert outs[0] == outs[1]
    assert outs[1][0][0] == outs[1][0][1]
