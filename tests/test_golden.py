"""The committed golden fixtures (tests/golden/) are reproducible: the seeded inputs, the
oracle's outputs, the 100-vertex graph, its greedy tree and the tree's FP64 value all come
back identical from their generators (CPU only)."""
import json
import os

import numpy as np

from oracle import default_backend as db
from omeinsum_mi355x import contractionorder as co, networks

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(HERE, "make_fixtures.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_einsum_cases_reproduce():
    m = _load()
    meta = json.load(open(os.path.join(HERE, "einsum_cases.json")))
    arr = np.load(os.path.join(HERE, "einsum_cases.npz"))
    assert len(meta) == len(m.CASES)
    for c, case in enumerate(meta):
        ixs, xs = m.inputs(case["spec"], case["sizes"], case["dtype"], case["seed"])
        for k, x in enumerate(xs):
            assert np.array_equal(arr["c%d_x%d" % (c, k)], x)
        y = m.expected(case["spec"], case["sizes"], case["dtype"], case["tropical"], xs)
        assert np.array_equal(np.asarray(y), arr["c%d_y" % c]), case["spec"]


def test_network_fixture_reproduces():
    fx = json.load(open(os.path.join(HERE, "network_100.json")))
    edges = networks.random_regular_graph(fx["n"], 3, seed=fx["graph_seed"])
    assert [list(e) for e in edges] == fx["edges"]
    code = networks.network_code(fx["n"], edges)
    sd = {e: 2 for e in range(len(edges))}
    tree = co.optimize_code(code, sd)
    assert list(co.contraction_complexity(tree, sd)) == fx["complexity"]
    assert tree.to_plain() == _plain_from_json(fx)
    xs = networks.gaussian_tensors(code, sd, seed=fx["tensor_seed"], dtype=np.complex64)
    val = complex(db.nested_einsum(tree.to_plain(), [x.astype(np.complex128) for x in xs], sd)[()])
    ref = complex(*fx["value_fp64"])
    assert abs(val - ref) <= 1e-12 * abs(ref)


def _plain_from_json(fx, tmp="/tmp/ome_golden_tree.json"):
    with open(tmp, "w") as f:
        json.dump(fx["tree"], f)
    return co.readjson(tmp).to_plain()


def test_request_stream_fixture_reproduces():
    # the 48 requests come back from their seeds with the same graphs, greedy trees (flop
    # counts) and complexities; the FP64 values of the six smallest are recomputed through the
    # oracle (the whole stream, 5.2e12 flop, is left to `make_fixtures.py --stream`)
    m = _load()
    fx = json.load(open(os.path.join(HERE, "request_stream_48.json")))
    stream = networks.request_stream(48, fx["seed"])
    assert len(fx["requests"]) == 48
    for q, fq in zip(stream, fx["requests"]):
        assert (q["n"], q["graph_seed"], q["tensor_seed"]) == (fq["n"], fq["graph_seed"], fq["tensor_seed"])
        code = networks.network_code(q["n"], q["edges"])
        sd = {e: 2 for e in range(len(q["edges"]))}
        tree = co.optimize_code(code, sd)
        assert co.flop(tree, sd) == fq["flop"]
        assert list(co.contraction_complexity(tree, sd))[:2] == fq["complexity"]
    for i in sorted(range(48), key=lambda i: fx["requests"][i]["flop"])[:6]:
        _, _, _, val = m.stream_request_value(stream[i])
        ref = complex(*fx["requests"][i]["value_fp64"])
        assert abs(val - ref) <= 1e-12 * abs(ref), i
