"""Writes the golden fixtures of this directory (run from the repository root:
``python tests/golden/make_fixtures.py``).

The reference stores no golden vectors (its random tests are unseeded), so the fixtures are
made here from seeded PCG64 inputs and the oracle (the C restatement of loop_einsum in FP64 /
exact Int64 / max-plus, and the numpy DefaultBackend restatement for the tree):

  einsum_cases.json / einsum_cases.npz   small flat einsum cases: code, sizes, inputs, output
  network_100.json                       the 100-vertex 3-regular graph (seed 2000), its greedy
                                         contraction tree (writejson format) and the FP64 value
                                         of the full contraction of its seeded ComplexF32 tensors
  request_stream_48.json                 the serving workload's 48 requests (networks.request_
                                         stream, seed 5000): per request the vertex count, graph
                                         and tensor seeds, the greedy tree's complexity and flop
                                         count, and the FP64 value of its full contraction

``python tests/golden/make_fixtures.py --stream`` writes only the request-stream fixture (the
FP64 contraction of all 48 networks, 5.2e12 flop, takes minutes on the host).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "omeinsum.jl_amd"))

from oracle import default_backend as db  # noqa: E402
from oracle import loop  # noqa: E402
from omeinsum_mi355x import contractionorder as co, networks  # noqa: E402

CASES = [
    # (spec, sizes, dtype, tropical)
    ("ij,jk->ik", {"i": 7, "j": 9, "k": 5}, "float64", False),
    ("ijk,jkl->il", {"i": 6, "j": 5, "k": 4, "l": 3}, "float32", False),
    ("ijk,jkl->il", {"i": 6, "j": 5, "k": 4, "l": 3}, "int64", False),
    ("abcd,cdef->abef", {l: 4 for l in "abcdef"}, "complex64", False),
    ("ijb,jkb->ikb", {"i": 5, "j": 6, "k": 7, "b": 3}, "complex128", False),
    ("ii->i", {"i": 6}, "float64", False),
    ("ijk->kij", {"i": 3, "j": 4, "k": 5}, "int64", False),
    ("ij->", {"i": 8, "j": 9}, "float32", False),
    ("i->ik", {"i": 4, "k": 3}, "float64", False),
    ("ij,jk,kl->il", {"i": 4, "j": 5, "k": 6, "l": 3}, "float64", False),
    ("ij,ik,il->jkl", {"i": 5, "j": 3, "k": 4, "l": 2}, "int64", False),
    ("ijl,jkl->ikl", {"i": 4, "j": 5, "k": 3, "l": 2}, "float64", True),
    ("ij,jk->ik", {"i": 6, "j": 7, "k": 5}, "float32", True),
]


def inputs(spec, sizes, dtype, seed):
    r = np.random.Generator(np.random.PCG64(seed))
    ixs = [list(w) for w in spec.split("->")[0].split(",")]
    xs = []
    for ix in ixs:
        shape = [sizes[l] for l in ix]
        if dtype == "int64":
            x = r.integers(-8, 9, size=shape)
        elif dtype.startswith("complex"):
            x = r.uniform(-1, 1, shape) + 1j * r.uniform(-1, 1, shape)
        else:
            x = r.uniform(-1, 1, shape)
        xs.append(np.asarray(x, dtype=dtype))
    return ixs, xs


def expected(spec, sizes, dtype, tropical, xs):
    ixs = [list(w) for w in spec.split("->")[0].split(",")]
    iy = list(spec.split("->")[1])
    wide = {"float32": np.float64, "complex64": np.complex128}.get(dtype, np.dtype(dtype).type)
    if tropical:
        wide = np.dtype(dtype).type
    return loop.loop_einsum(ixs, iy, [x.astype(wide) for x in xs], sizes, tropical=tropical)


def main():
    meta, arrays = [], {}
    for c, (spec, sizes, dtype, trop) in enumerate(CASES):
        ixs, xs = inputs(spec, sizes, dtype, 7000 + c)
        y = expected(spec, sizes, dtype, trop, xs)
        meta.append({"spec": spec, "sizes": sizes, "dtype": dtype, "tropical": trop,
                     "seed": 7000 + c})
        for k, x in enumerate(xs):
            arrays["c%d_x%d" % (c, k)] = x
        arrays["c%d_y" % c] = np.asarray(y)
    with open(os.path.join(HERE, "einsum_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "einsum_cases.npz"), **arrays)

    n = 100
    edges = networks.random_regular_graph(n, 3, seed=2000)
    code = networks.network_code(n, edges)
    sd = {e: 2 for e in range(len(edges))}
    xs = networks.gaussian_tensors(code, sd, seed=2001, dtype=np.complex64)
    tree = co.optimize_code(code, sd)
    val = complex(db.nested_einsum(tree.to_plain(), [x.astype(np.complex128) for x in xs], sd)[()])
    tmp = os.path.join(HERE, "_tree.json")
    co.writejson(tmp, tree)
    with open(tmp) as f:
        tree_obj = json.load(f)
    os.remove(tmp)
    with open(os.path.join(HERE, "network_100.json"), "w") as f:
        json.dump({"n": n, "graph_seed": 2000, "tensor_seed": 2001, "edges": edges,
                   "tree": tree_obj, "value_fp64": [val.real, val.imag],
                   "complexity": list(co.contraction_complexity(tree, sd))}, f)


STREAM_SLICE_SC = 24   # host memory bound for the FP64 contraction: 2^24 complex128 = 256 MiB


def stream_request_value(q):
    """FP64 value of one request's full contraction through the oracle (complex128): its
    greedy tree, sliced to 2^STREAM_SLICE_SC elements when larger, cross-slice sum in FP64."""
    code = networks.network_code(q["n"], q["edges"])
    sd = {e: 2 for e in range(len(q["edges"]))}
    tree = co.optimize_code(code, sd)
    xs = [x.astype(np.complex128) for x in
          networks.gaussian_tensors(code, sd, q["tensor_seed"], np.complex64)]
    tc, sc, _ = co.contraction_complexity(tree, sd)
    if sc > STREAM_SLICE_SC:
        st = co.slice_code(tree, sd, sc_target=STREAM_SLICE_SC)
        val = db.sliced_einsum(st.slicing, st.eins.to_plain(), xs, sd)
    else:
        val = db.nested_einsum(tree.to_plain(), xs, sd)
    return tree, sd, [tc, sc], complex(np.asarray(val)[()])


def stream_fixture(path=os.path.join(HERE, "request_stream_48.json"), nreq=48, seed=5000):
    out = []
    for q in networks.request_stream(nreq, seed):
        tree, sd, cx, val = stream_request_value(q)
        out.append({"n": q["n"], "graph_seed": q["graph_seed"], "tensor_seed": q["tensor_seed"],
                    "complexity": cx, "flop": co.flop(tree, sd),
                    "value_fp64": [val.real, val.imag]})
        print("request %d: n %d, sc %d, value %r" % (len(out) - 1, q["n"], cx[1], val), flush=True)
    with open(path, "w") as f:
        json.dump({"seed": seed, "requests": out}, f, indent=0)


if __name__ == "__main__":
    if "--stream" in sys.argv:
        stream_fixture()
    else:
        main()
