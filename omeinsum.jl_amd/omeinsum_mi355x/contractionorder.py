"""Contraction order, slicing and complexity measures (host side).

OMEinsum delegates these to the unvendored OMEinsumContractionOrders 1.1
(src/contractionorder.jl:1-51).  This module is this backend's own planner, written for the
tensor-network workloads (greedy pair selection and a budget slicer); its complexity
measures follow the values the reference's tests pin:
  flop(ein"i->", i=4) == 4                                  test/contractionorder.jl:17
  (ij,jkc),klc->il at size 10: tc = log2(20000), sc = log2(1000),
  rw = log2(4200), peak_memory == 2100                      test/einsequence.jl:65-73
"""
import json
import math

from .codes import EinCode, NestedEinsum, SlicedEinsum, _unique


def _size(labels, size_dict):
    p = 1
    for l in _unique(labels):
        p *= size_dict[l]
    return p


def _mem(labels, size_dict):
    p = 1
    for l in labels:
        p *= size_dict[l]
    return p


def flop(code, size_dict):
    """Sum over contraction steps of the product of the sizes of the step's distinct labels
    (OMEinsumContractionOrders' convention: no factor 2)."""
    if isinstance(code, SlicedEinsum):
        sd = dict(size_dict)
        n = 1
        for l in code.slicing:
            n *= size_dict[l]
            sd[l] = 1
        return n * flop(code.eins, sd)
    if isinstance(code, EinCode):
        return _size([l for ix in code.ixs for l in ix] + code.iy, size_dict)
    total = 0
    for node in code.internal_nodes():
        e = node.eins
        total += _size([l for ix in e.ixs for l in ix] + e.iy, size_dict)
    return total


def peak_memory(code, size_dict):
    """Peak number of live elements when evaluating the tree: within a node, the results
    of earlier siblings stay live while the next one is evaluated, and the node's output
    coexists with all its inputs.  Pinned by test/einsequence.jl:65-73 (== 2100)."""
    if isinstance(code, SlicedEinsum):
        sd = dict(size_dict)
        for l in code.slicing:
            sd[l] = 1
        return peak_memory(code.eins, sd)
    if isinstance(code, EinCode):
        return sum(_mem(ix, size_dict) for ix in code.ixs) + _mem(code.iy, size_dict)
    largest, temp = 0, 0
    for arg, ix in zip(code.args, code.eins.ixs):
        sub = _mem(ix, size_dict) if arg.isleaf() else peak_memory(arg, size_dict)
        largest = max(largest, sub + temp)
        temp += _mem(ix, size_dict)
    return max(largest, temp + _mem(code.eins.iy, size_dict))


def contraction_complexity(code, size_dict):
    """(tc, sc, rw): log2 of flop, of the largest tensor (inputs and intermediates), and of
    the read-write volume (sum over steps of input + output sizes)."""
    if isinstance(code, SlicedEinsum):
        sd = dict(size_dict)
        n = 1
        for l in code.slicing:
            n *= size_dict[l]
            sd[l] = 1
        tc, sc, rw = contraction_complexity(code.eins, sd)
        return tc + math.log2(n), sc, rw + math.log2(n)
    if isinstance(code, EinCode):
        nodes = [code]
    else:
        nodes = [n.eins for n in code.internal_nodes()]
    f = sum(_size([l for ix in e.ixs for l in ix] + e.iy, size_dict) for e in nodes)
    sc = max([_mem(ix, size_dict) for e in nodes for ix in e.ixs] +
             [_mem(e.iy, size_dict) for e in nodes])
    rw = sum(sum(_mem(ix, size_dict) for ix in e.ixs) + _mem(e.iy, size_dict) for e in nodes)
    lg = lambda v: math.log2(v) if v > 0 else -math.inf
    return lg(f), lg(sc), lg(rw)


class GreedyMethod:
    """Greedy pairwise order: at every step contract the pair of live tensors (sharing a
    label when possible) whose output is smallest relative to its inputs,
    cost = |out| - alpha*(|a| + |b|), ties broken by the lowest tensor positions, so the
    result is deterministic."""

    def __init__(self, alpha=0.0):
        self.alpha = alpha


def optimize_code(code, size_dict, optimizer=None):
    """optimize_code (src/contractionorder.jl:30-32) for GreedyMethod: returns a binary
    NestedEinsum whose flatten() equals ``code``.  Candidate pairs are the pairs of live
    tensors that share a label; disconnected components are joined by outer products at
    the end (smallest first)."""
    optimizer = optimizer or GreedyMethod()
    if isinstance(code, NestedEinsum):
        code = code.flatten()
    ixs = [list(ix) for ix in code.ixs]
    iy = list(code.iy)
    n = len(ixs)
    if n == 1:
        return NestedEinsum([NestedEinsum.leaf(0)], EinCode([ixs[0]], iy))
    live = {i: (NestedEinsum.leaf(i), ixs[i]) for i in range(n)}
    holders = {}
    for i, ls in enumerate(ixs):
        for l in _unique(ls):
            holders.setdefault(l, set()).add(i)
    iyset = set(iy)
    next_id = n

    def out_labels(la, lb, a, b):
        out = []
        for l in _unique(la + lb):
            if l in iyset or any(h not in (a, b) for h in holders[l]):
                out.append(l)
        return out

    while len(live) > 1:
        best = None
        seen = set()
        for l, hs in holders.items():
            if len(hs) < 2:
                continue
            hs = sorted(hs)
            for x in range(len(hs)):
                for y in range(x + 1, len(hs)):
                    a, b = hs[x], hs[y]
                    if (a, b) in seen:
                        continue
                    seen.add((a, b))
                    la, lb = live[a][1], live[b][1]
                    out = out_labels(la, lb, a, b)
                    cost = _mem(out, size_dict) - optimizer.alpha * (
                        _mem(la, size_dict) + _mem(lb, size_dict))
                    key = (cost, a, b)
                    if best is None or key < best[0]:
                        best = (key, a, b, out)
        if best is None:
            # no shared labels left: outer product of the two smallest tensors
            ids = sorted(live, key=lambda k: (_mem(live[k][1], size_dict), k))
            a, b = sorted(ids[:2])
            best = (None, a, b, out_labels(live[a][1], live[b][1], a, b))
        _, a, b, out = best
        if len(live) == 2:
            out = iy
        na, la = live.pop(a)
        nb, lb = live.pop(b)
        node = NestedEinsum([na, nb], EinCode([la, lb], out))
        for l in _unique(la + lb):
            holders[l].discard(a)
            holders[l].discard(b)
        for l in _unique(out):
            holders.setdefault(l, set()).add(next_id)
        live[next_id] = (node, list(out))
        next_id += 1
    (root, ls), = live.values()
    if root.isleaf() or list(root.eins.iy) != iy:
        root = NestedEinsum([root], EinCode([ls], iy))
    return root


class TreeSASlicer:
    """Placeholder name kept for API parity; see ``slice_code``."""

    def __init__(self, sc_target=None, ntrials=1, niters=0):
        self.sc_target = sc_target


def slice_code(code, size_dict, slicer=None, sc_target=None, max_elements=None):
    """slice_code (src/contractionorder.jl:33-35): choose labels to slice so that the
    largest intermediate fits ``max_elements`` (or 2**sc_target).  Greedy: repeatedly slice
    the label carried by the most elements of over-budget tensors.  Returns a
    SlicedEinsum sharing the tree."""
    if isinstance(code, SlicedEinsum):
        code = code.eins
    if sc_target is None and slicer is not None:
        sc_target = slicer.sc_target
    if max_elements is None:
        max_elements = 2 ** sc_target if sc_target is not None else None
    if max_elements is None:
        return SlicedEinsum([], code)
    sd = dict(size_dict)
    slicing = []
    nodes = [n.eins for n in code.internal_nodes()]
    out_labels = set(code.getiyv())
    while True:
        tensors = [ix for e in nodes for ix in e.ixs] + [e.iy for e in nodes]
        over = [t for t in tensors if _mem(t, sd) > max_elements]
        if not over:
            break
        score = {}
        for t in over:
            w = _mem(t, sd)
            for l in t:
                if sd[l] > 1 and l not in out_labels:
                    score[l] = score.get(l, 0) + w
        if not score:
            break
        l = max(score, key=lambda k: (score[k], -list(score).index(k)))
        slicing.append(l)
        sd[l] = 1
    return SlicedEinsum(slicing, code)


# ---- wire format (writejson/readjson, src/contractionorder.jl:47-51) -----------------


def _lab(l):
    return l


def writejson(filename, code):
    """Write a NestedEinsum or SlicedEinsum as JSON: {"slicing": [...], "tree": plain}.
    Labels are stored as JSON values (strings or ints)."""
    if isinstance(code, SlicedEinsum):
        obj = {"slicing": code.slicing, "tree": code.eins.to_plain()}
    else:
        obj = {"slicing": None, "tree": code.to_plain()}
    with open(filename, "w") as f:
        json.dump(obj, f)


def _from_json_plain(p):
    if isinstance(p, int):
        return p
    args, ixs, iy = p
    return ([_from_json_plain(a) for a in args], ixs, iy)


def readjson(filename):
    with open(filename) as f:
        obj = json.load(f)
    tree = NestedEinsum.from_plain(_from_json_plain(obj["tree"]))
    if obj.get("slicing") is None:
        return tree
    return SlicedEinsum(obj["slicing"], tree)
