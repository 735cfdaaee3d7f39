"""The order and slicing planner against the values the reference's tests pin
(test/contractionorder.jl, test/einsequence.jl:65-73), on the CPU: complexity measures, the
greedy order, TreeSA on the C60 fullerene, the slicer, the max-plus value 66 of the sliced
and unsliced fullerene trees (through the oracle), and the writejson / readjson round trip."""
import math

import numpy as np

from oracle import default_backend as db
from omeinsum_mi355x import codes, contractionorder as co, networks

# test/contractionorder.jl:6-11: labels a..f of sizes 2, 4, ..., 64
SIZES = {c: 1 << (i + 1) for i, c in enumerate("abcdef")}
TC_MAX = math.log2(2.0 ** 10 + 2.0 ** 16 + 2.0 ** 15 + 2.0 ** 9 + 1e-8)


def test_tree_greedy_pins():
    # test/contractionorder.jl:5-23
    code = codes.ein("ab,acd,bcef,e,df->")
    opt = co.optimize_code(code, SIZES, co.GreedyMethod())
    tc, sc, _ = co.contraction_complexity(opt, SIZES)
    assert math.isclose(tc, math.log2(co.flop(opt, SIZES)))
    assert co.flop(codes.ein("i->"), {"i": 4}) == 4
    assert 16 <= tc <= TC_MAX
    assert sc == 11
    assert opt.flatten() == code
    code3 = codes.ein("(ab,acd),bcef,e,df->")
    opt3 = co.optimize_code(code3, SIZES, co.GreedyMethod())
    tc3, _, _ = co.contraction_complexity(opt3, SIZES)
    assert 16 <= tc3 <= TC_MAX
    opt4 = co.optimize_code(code3, SIZES, co.TreeSA(ntrials=2, seed=2))
    tc4, sc4, _ = co.contraction_complexity(opt4, SIZES)
    assert 16 <= tc4 <= TC_MAX and sc4 <= 11
    assert opt4.flatten() == code3.flatten()


def test_space_complexity_counts_written_tensors():
    # an input of 2^16 elements (bcef) does not count: sc is the largest tensor a step
    # writes, so the flat code to a scalar has sc 0 (test/contractionorder.jl:53-55)
    code = codes.ein("ab,acd,bcef,e,df->")
    tc, sc, _ = co.contraction_complexity(code, SIZES)
    assert tc == math.log2(2 * 4 * 8 * 16 * 32 * 64) and sc == 0
    ne = codes.ein("(ij,jkc),klc->il")
    sd = {l: 10 for l in "ijklc"}
    tc, sc, rw = co.contraction_complexity(ne, sd)
    assert math.isclose(tc, math.log2(20000)) and math.isclose(sc, math.log2(1000))
    assert math.isclose(rw, math.log2(4200)) and co.peak_memory(ne, sd) == 2100


def _fullerene():
    edges = networks.fullerene_edges()
    ixs = [list(e) for e in edges] + [[i] for i in range(60)]
    return codes.EinCode(ixs, []), {i: 2 for i in range(60)}


def test_fullerene_treesa_and_slicer():
    # test/contractionorder.jl:31-70
    code, sd = _fullerene()
    tc, sc, _ = co.contraction_complexity(code, sd)
    assert tc == 60 and sc == 0
    opt = co.optimize_code(code, sd, co.TreeSA(ntrials=1))
    tc2, sc2, _ = co.contraction_complexity(opt, sd)
    assert sc2 == 10
    greedy = co.optimize_code(code, sd, co.GreedyMethod())
    assert tc2 <= co.contraction_complexity(greedy, sd)[0]
    assert opt.flatten() == code and codes.flatten(code) == code
    xs = [np.array([[-1.0, 1.0], [1.0, -1.0]])] * 90 + [np.zeros(2)] * 60
    assert db.nested_einsum(opt.to_plain(), xs, sd, semiring="max_plus")[()] == 66.0
    slicer = co.TreeSASlicer(score=co.ScoreFunction(sc_target=8))
    scode = co.slice_code(opt, sd, slicer)
    assert isinstance(scode, codes.SlicedEinsum)
    assert co.contraction_complexity(scode, sd)[1] == 8
    val = db.sliced_einsum(scode.slicing, scode.eins.to_plain(), xs, sd, semiring="max_plus")
    assert val[()] == 66.0


def test_regression_codes():
    # test/contractionorder.jl:73-92
    r = np.random.Generator(np.random.PCG64(11))
    for code, sd in [(codes.ein("i->"), {"i": 3}), (codes.ein("i,j->"), {"i": 3, "j": 3}),
                     (codes.ein("ij,jk,kl->ijl"), {l: 3 for l in "ijkl"})]:
        opt = co.optimize_code(code, sd, co.GreedyMethod())
        assert isinstance(opt, codes.NestedEinsum)
        xs = [r.normal(size=[sd[l] for l in ix]) for ix in code.ixs]
        assert np.allclose(db.nested_einsum(opt.to_plain(), xs, sd),
                           db.einsum(code.ixs, code.iy, xs, sd))


def test_save_load(tmp_path):
    # test/contractionorder.jl:101-115: writejson / readjson give back an equal code, for
    # Int and Char labels, plain and sliced
    for code in (codes.EinCode([[1, 2], [2, 3], [3, 4]], [1, 4]),
                 codes.EinCode([["a", "b"], ["b", "c"], ["c", "d"]], ["a", "d"])):
        sd = codes.uniformsize(code, 2)
        g = co.optimize_code(code, sd, co.GreedyMethod())
        s = co.slice_code(g, sd, sc_target=1)
        assert s.slicing
        for opt in (g, s):
            f = str(tmp_path / "tree.json")
            co.writejson(f, opt)
            assert co.readjson(f) == opt
    import json
    obj = json.load(open(f))
    assert obj["label-type"] == "Char" and obj["output"] == ["a", "d"]
    assert obj["tree"]["isleaf"] is False and "slices" in obj
    leaves = []

    def walk(d):
        if d["isleaf"]:
            leaves.append(d["tensorindex"])
        else:
            for a in d["args"]:
                walk(a)
    walk(obj["tree"])
    assert sorted(leaves) == [1, 2, 3]  # Julia's 1-based tensor indices
