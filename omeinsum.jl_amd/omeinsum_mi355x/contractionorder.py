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
    """(tc, sc, rw): log2 of flop, of the largest tensor a contraction step writes (the
    intermediates and the output; the inputs are not counted), and of the read-write volume
    (sum over steps of input + output sizes).  Pinned by test/contractionorder.jl:14-19
    (sc == 11 while an input holds 2^16 elements) and :53-55 (a flat code to a scalar:
    tc == 60, sc == 0)."""
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
    sc = max(_mem(e.iy, size_dict) for e in nodes)
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
    if isinstance(optimizer, TreeSA):
        return optimizer.optimize(code, size_dict)
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
    """The slicer argument of slice_code: ``sc_target`` (log2 of the largest intermediate
    allowed after slicing).  The labels are chosen by slice_code's greedy rule."""

    def __init__(self, sc_target=None, ntrials=1, niters=0, score=None):
        if score is not None and sc_target is None:
            sc_target = score.sc_target
        self.sc_target = sc_target


class ScoreFunction:
    """The objective of TreeSA (OMEinsumContractionOrders' ScoreFunction):
    tc_weight * tc + sc_weight * max(0, sc - sc_target) + rw_weight * rw."""

    def __init__(self, tc_weight=1.0, sc_weight=1.0, rw_weight=0.0, sc_target=20.0):
        self.tc_weight = float(tc_weight)
        self.sc_weight = float(sc_weight)
        self.rw_weight = float(rw_weight)
        self.sc_target = float(sc_target)


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
        tensors = [e.iy for e in nodes]  # the intermediates and the output (as sc counts)
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
#
# The layout of OMEinsumContractionOrders 1.x's writejson (the unvendored dependency that
# OMEinsum's writejson forwards to): a JSON object with "label-type" (the Julia label type:
# "Char" for one-character labels, "Int64" for integers, else "String"), "inputs" and
# "output" (the flat code's labels), "tree" (nodes {"isleaf": true, "tensorindex": k} with k
# 1-based as in Julia, or {"isleaf": false, "eins": {"ixs": ..., "iy": ...}, "args": [...]})
# and, for a SlicedEinsum, "slices".  readjson also reads this package's earlier layout
# {"slicing": [...] | null, "tree": plain}.


def _label_type(labels):
    if all(isinstance(l, bool) is False and isinstance(l, int) for l in labels):
        return "Int64"
    if all(isinstance(l, str) and len(l) == 1 for l in labels):
        return "Char"
    return "String"


def _tree_todict(ne):
    if ne.isleaf():
        return {"isleaf": True, "tensorindex": ne.tensorindex + 1}
    return {"isleaf": False, "args": [_tree_todict(a) for a in ne.args],
            "eins": {"ixs": [list(ix) for ix in ne.eins.ixs], "iy": list(ne.eins.iy)}}


def _tree_fromdict(d):
    if d["isleaf"]:
        return NestedEinsum.leaf(int(d["tensorindex"]) - 1)
    return NestedEinsum([_tree_fromdict(a) for a in d["args"]],
                        EinCode([list(ix) for ix in d["eins"]["ixs"]], list(d["eins"]["iy"])))


def writejson(filename, code):
    """writejson(filename, code) (src/contractionorder.jl:47-51) for a NestedEinsum or a
    SlicedEinsum, in the layout described above."""
    tree = code.eins if isinstance(code, SlicedEinsum) else code
    labels = tree.uniquelabels()
    obj = {"label-type": _label_type(labels), "inputs": [list(ix) for ix in tree.getixsv()],
           "output": list(tree.getiyv()), "tree": _tree_todict(tree)}
    if isinstance(code, SlicedEinsum):
        obj["slices"] = list(code.slicing)
    with open(filename, "w") as f:
        json.dump(obj, f)


def _from_json_plain(p):
    if isinstance(p, int):
        return p
    args, ixs, iy = p
    return ([_from_json_plain(a) for a in args], ixs, iy)


def readjson(filename):
    """readjson(filename) (src/contractionorder.jl:47-51): the NestedEinsum or SlicedEinsum
    written by writejson."""
    with open(filename) as f:
        obj = json.load(f)
    if isinstance(obj.get("tree"), dict):
        tree = _tree_fromdict(obj["tree"])
        slices = obj.get("slices")
        return SlicedEinsum(list(slices), tree) if slices is not None else tree
    tree = NestedEinsum.from_plain(_from_json_plain(obj["tree"]))
    if obj.get("slicing") is None:
        return tree
    return SlicedEinsum(obj["slicing"], tree)


# ---- TreeSA: simulated annealing over binary contraction trees ------------------------------


class TreeSA:
    """Simulated annealing over binary trees (the method OMEinsumContractionOrders calls
    TreeSA, after Kalachev, Panteleev and Yung, arXiv:2108.05665): from a greedy tree, each
    sweep visits every internal node and proposes one local rotation
    (A, (C, D)) -> ((A, C), D) or ((A, D), C), and the mirror forms, accepted with the
    Metropolis rule exp(-beta * d_score) under the ScoreFunction objective while beta rises
    over ``betas``.  ntrials independent runs; the best tree is returned.  Codes with a label
    repeated within one tensor are returned as the greedy tree."""

    def __init__(self, sc_target=20, betas=None, ntrials=3, niters=20, sc_weight=1.0,
                 rw_weight=0.0, score=None, seed=0):
        self.score = score or ScoreFunction(sc_weight=sc_weight, rw_weight=rw_weight,
                                            sc_target=sc_target)
        self.betas = list(betas) if betas is not None else [0.01 + 0.1 * i for i in range(150)]
        self.ntrials = int(ntrials)
        self.niters = int(niters)
        self.seed = seed

    def optimize(self, code, size_dict):
        import numpy as np
        greedy = optimize_code(code, size_dict, GreedyMethod())
        ixs = [list(ix) for ix in code.ixs]
        if len(ixs) < 3 or any(len(set(ix)) != len(ix) for ix in ixs):
            return greedy
        labels = _unique([l for ix in ixs for l in ix] + list(code.iy))
        lid = {l: i for i, l in enumerate(labels)}
        l2 = [math.log2(size_dict[l]) if size_dict[l] > 0 else -math.inf for l in labels]
        best, best_score = greedy, None
        for trial in range(self.ntrials):
            rng = np.random.Generator(np.random.PCG64(hash((self.seed, trial)) & 0xFFFFFFFF))
            st = _SATree(greedy, ixs, list(code.iy), lid, l2, self.score)
            if best_score is None:
                best_score = st.score()  # the greedy tree, unless a run does better
            st.anneal(self.betas, self.niters, rng)
            sc_ = st.score()
            if sc_ < best_score:
                best_score = sc_
                best = st.to_nested(labels)
        return best


class _SANode:
    __slots__ = ("l", "r", "p", "out", "leaf")

    def __init__(self, l=None, r=None, out=0, leaf=-1):
        self.l, self.r, self.p, self.out, self.leaf = l, r, None, out, leaf


class _SATree:
    """A binary tree with label bitmasks: a node's ``out`` holds the labels its result
    carries, its loop labels are out(l) | out(r)."""

    def __init__(self, tree, ixs, iy, lid, l2, score):
        self.ixs, self.iy, self.lid, self.l2, self.sf = ixs, iy, lid, l2, score
        self._lg = {}
        mask = lambda ls: sum(1 << lid[l] for l in set(ls))
        self.iymask = mask(iy)

        def build(n):
            if n.isleaf():
                return _SANode(out=mask(ixs[n.tensorindex]), leaf=n.tensorindex)
            kids = [build(a) for a in n.args]
            while len(kids) > 2:  # n-ary nodes (not produced by the greedy) become left-deep
                kids = [self._join(kids[0], kids[1], None)] + kids[2:]
            if len(kids) == 1:
                return kids[0]
            node = self._join(kids[0], kids[1], mask(n.eins.iy))
            return node
        self.root = build(tree)
        self.root.out = self.iymask
        # internal outputs recomputed from the leaves: a label leaves a subtree when it is used
        # outside it or in the final output
        self._fix_outs()
        self.nodes = []
        self._collect(self.root)
        self.flops = {id(n): self.lg(n.l.out | n.r.out) for n in self.nodes}
        self.tc_sum = sum(2.0 ** v for v in self.flops.values())
        self.scs = sorted(self.lg(n.out) for n in self.nodes)

    @staticmethod
    def _join(a, b, out):
        n = _SANode(a, b, out if out is not None else (a.out | b.out))
        a.p = n
        b.p = n
        return n

    def lg(self, m):
        v = self._lg.get(m)
        if v is None:
            v, i, r = 0.0, 0, m
            while r:
                if r & 1:
                    v += self.l2[i]
                r >>= 1
                i += 1
            self._lg[m] = v
        return v

    def _collect(self, n):
        if n.leaf >= 0:
            return
        self._collect(n.l)
        self._collect(n.r)
        self.nodes.append(n)

    def _fix_outs(self):
        count = {}
        for ix in self.ixs:
            for l in set(ix):
                count[self.lid[l]] = count.get(self.lid[l], 0) + 1

        def rec(n):
            if n.leaf >= 0:
                c = {}
                for l in set(self.ixs[n.leaf]):
                    c[self.lid[l]] = 1
                return c
            a, b = rec(n.l), rec(n.r)
            for k, v in b.items():
                a[k] = a.get(k, 0) + v
            if n is not self.root:
                m = 0
                for k, v in a.items():
                    if v < count[k] or (self.iymask >> k) & 1:
                        m |= 1 << k
                n.out = m
            return a
        rec(self.root)

    def score(self):
        tc = math.log2(self.tc_sum) if self.tc_sum > 0 else 0.0
        sc = self.scs[-1] if self.scs else 0.0
        return self.sf.tc_weight * tc + self.sf.sc_weight * max(0.0, sc - self.sf.sc_target)

    def anneal(self, betas, niters, rng):
        import bisect
        for beta in betas:
            for _ in range(niters):
                order = rng.permutation(len(self.nodes))
                for q in order:
                    n = self.nodes[q]
                    cand = []
                    if n.r.leaf < 0:
                        cand += [(0, 0), (0, 1)]
                    if n.l.leaf < 0:
                        cand += [(1, 0), (1, 1)]
                    if not cand:
                        continue
                    side, pick = cand[int(rng.integers(len(cand)))]
                    # rotation: n = (X, (C, D)) -> n = ((X, C), D) [pick 0] or ((X, D), C)
                    inner = n.r if side == 0 else n.l
                    x = n.l if side == 0 else n.r
                    c, d = (inner.l, inner.r) if pick == 0 else (inner.r, inner.l)
                    new_out = (x.out | c.out) & (d.out | n.out)
                    old_f = self.flops[id(inner)]
                    new_f = self.lg(x.out | c.out)
                    old_s = self.lg(inner.out)
                    new_s = self.lg(new_out)
                    # n's own loop labels change from out(x)|out(inner) to new_out|out(d)
                    old_fn = self.flops[id(n)]
                    new_fn = self.lg(new_out | d.out)
                    # the move is scored locally, on the two nodes it changes: the log2 of
                    # their summed flops, and the size of the node it rebuilds when that
                    # passes sc_target
                    dtc = (math.log2(2.0 ** new_f + 2.0 ** new_fn) -
                           math.log2(2.0 ** old_f + 2.0 ** old_fn))
                    dsc = new_s - old_s
                    d_score = self.sf.tc_weight * dtc + (
                        self.sf.sc_weight * dsc if max(old_s, new_s) > self.sf.sc_target else 0.0)
                    tc_new = self.tc_sum - 2.0 ** old_f - 2.0 ** old_fn + 2.0 ** new_f + 2.0 ** new_fn
                    scs = self.scs
                    i = bisect.bisect_left(scs, old_s)
                    scs.pop(i)
                    bisect.insort(scs, new_s)
                    tc_old = self.tc_sum
                    self.tc_sum = tc_new
                    if d_score <= 0 or rng.random() < math.exp(-beta * d_score):
                        # apply: inner becomes (x, c), n becomes (inner, d) on the same side
                        inner.l, inner.r = x, c
                        x.p = inner
                        c.p = inner
                        inner.out = new_out
                        if side == 0:
                            n.l, n.r = inner, d
                        else:
                            n.r, n.l = inner, d
                        d.p = n
                        self.flops[id(inner)] = new_f
                        self.flops[id(n)] = new_fn
                    else:
                        self.tc_sum = tc_old
                        i = bisect.bisect_left(scs, new_s)
                        scs.pop(i)
                        bisect.insort(scs, old_s)

    def to_nested(self, labels):
        order = {i: k for k, i in enumerate(range(len(labels)))}

        def lab(m):
            return [labels[i] for i in sorted((i for i in range(len(labels)) if (m >> i) & 1),
                                              key=order.get)]

        def rec(n):
            if n.leaf >= 0:
                return NestedEinsum.leaf(n.leaf), list(self.ixs[n.leaf])
            (a, la), (b, lb) = rec(n.l), rec(n.r)
            iy = list(self.iy) if n is self.root else lab(n.out)
            return NestedEinsum([a, b], EinCode([la, lb], iy)), iy
        if self.root.leaf >= 0:
            return NestedEinsum([NestedEinsum.leaf(self.root.leaf)],
                                EinCode([list(self.ixs[self.root.leaf])], list(self.iy)))
        return rec(self.root)[0]
