"""TEST INFRASTRUCTURE ONLY — numpy restatement of OMEinsum's DefaultBackend algorithm.

This is what OMEinsum does for CPU ``Array``s, restated function by function (file:line of
the reference in each docstring).  It serves two purposes:
  * the timed CPU baseline (OpenBLAS GEMM through numpy, permutes as numpy copies), and
  * a second, independent checker at sizes the brute-force ``loop_einsum`` cannot reach.

Trees are given in a plain form so this module shares no code with the product:
  tree  = int                     a leaf: the tensor index (0-based)
        | (args, ixs, iy)         an internal node: args a list of trees
  sliced = (slicing_labels, tree)

Semirings: ``semiring="plus_times"`` (floats, complex, Int64 with wrapping) or
``semiring="max_plus"`` (TropicalNumbers: + is max, * is +, zero -Inf, one 0).
"""
import itertools

import numpy as np

from . import loop

# ---------------------------------------------------------------------------------------
# semirings


class _PlusTimes:
    tropical = False

    @staticmethod
    def zero(dt):
        return np.zeros((), dtype=dt)[()]

    @staticmethod
    def one(dt):
        return np.ones((), dtype=dt)[()]

    @staticmethod
    def mul(a, b):
        return a * b

    @staticmethod
    def add(a, b):
        return a + b

    @staticmethod
    def is_zero(s):
        return s == 0

    @staticmethod
    def is_one(s):
        return s == 1

    @staticmethod
    def reduce(x, axes):
        with np.errstate(over="ignore"):
            return np.sum(x, axis=axes, dtype=x.dtype)

    @staticmethod
    def bmm(a, b):
        """(i,j,l) x (j,k,l) -> (i,k,l); one BLAS GEMM per batch slice, as
        BatchedRoutines.batched_gemm! does (src/utils.jl:162-165)."""
        if a.shape[2] == 1:
            return np.asfortranarray((a[:, :, 0] @ b[:, :, 0])[:, :, None])
        out = np.empty((a.shape[0], b.shape[1], a.shape[2]), dtype=a.dtype, order="F")
        for l in range(a.shape[2]):
            out[:, :, l] = a[:, :, l] @ b[:, :, l]
        return out


class _MaxPlus:
    tropical = True

    @staticmethod
    def zero(dt):
        return np.array(-np.inf, dtype=dt)[()]

    @staticmethod
    def one(dt):
        return np.zeros((), dtype=dt)[()]

    @staticmethod
    def mul(a, b):
        return a + b

    @staticmethod
    def add(a, b):
        return np.maximum(a, b)

    @staticmethod
    def is_zero(s):
        return s == -np.inf

    @staticmethod
    def is_one(s):
        return s == 0

    @staticmethod
    def reduce(x, axes):
        if x.size == 0:
            shape = [d for i, d in enumerate(x.shape) if i not in axes]
            return np.full(shape, -np.inf, dtype=x.dtype)
        return np.max(x, axis=axes)

    @staticmethod
    def bmm(a, b):
        # generic per-batch mul! loop for non-BLAS eltypes (src/utils.jl:167-177)
        out = np.full((a.shape[0], b.shape[1], a.shape[2]), -np.inf, dtype=a.dtype, order="F")
        for l in range(a.shape[2]):
            if a.shape[1] == 0:
                continue
            out[:, :, l] = np.max(a[:, :, None, l] + b[None, :, :, l], axis=1)
        return out


def _sr(semiring):
    return _MaxPlus if semiring == "max_plus" else _PlusTimes


def _F(x):
    x = np.asarray(x)
    return x if x.ndim == 0 else np.asfortranarray(x)


# ---------------------------------------------------------------------------------------
# utils (src/utils.jl)


def _unique(ix):
    """_unique: first-appearance dedup (src/utils.jl:104-106)."""
    out = []
    for l in ix:
        if l not in out:
            out.append(l)
    return out


def addmul(y, sy, v, sx, sr):
    """@addmul!/@flatten_addmul! sy*y + sx*v (src/utils.jl:1-37): b==0 -> lmul!(a,y);
    a==0 -> y .= b*v; a==1 -> y .+= b*v; else y .= a*y .+ b*v.  Returns the new y."""
    with np.errstate(over="ignore", invalid="ignore"):
        if sr.is_zero(sx):
            y[...] = sr.mul(sy, y)
        elif sr.is_zero(sy):
            y[...] = np.broadcast_to(sr.mul(sx, v), y.shape)
        elif sr.is_one(sy):
            y[...] = sr.add(y, sr.mul(sx, v))
        else:
            y[...] = sr.add(sr.mul(sy, y), sr.mul(sx, v))
    return y


def tensorpermute(C, A, perm, sx, sy, sr):
    """tensorpermute!(C, A, perm, sx, sy) (src/utils.jl:122-156).  Grouping runs of
    consecutive perm entries only changes the shapes handed to permutedims!, not values,
    so the restatement permutes directly."""
    if A.ndim == 0:
        return addmul(C, sy, A, sx, sr)
    P = np.transpose(A, perm)
    if sr.is_zero(sy):
        with np.errstate(over="ignore"):
            C[...] = P if sr.is_one(sx) else sr.mul(sx, P)
        return C
    return addmul(C, sy, P, sx, sr)


# ---------------------------------------------------------------------------------------
# unary rules (src/unaryrules.jl) and the unary pipeline (src/einsum.jl:52-101)


def unary_pipeline(ix, iy):
    """src/einsum.jl:60-82: Diag -> (Sum | Permutedims) -> Repeat -> Duplicate."""
    ix_unique = _unique(ix)
    iy_unique = _unique(iy)
    iy_a = [i for i in iy_unique if i in ix]
    ops = []
    if len(ix_unique) != len(ix):
        ops.append(("Diag", list(ix), ix_unique))
    if len(ix_unique) != len(iy_a):
        ops.append(("Sum", ix_unique, iy_a))
    elif ix_unique != iy_a:
        ops.append(("Permutedims", ix_unique, iy_a))
    if len(iy_a) != len(iy_unique):
        ops.append(("Repeat", iy_a, iy_unique))
    if len(iy_unique) != len(iy):
        ops.append(("Duplicate", iy_unique, list(iy)))
    return ops


def _unary_rule(rule, ix, iy, x, y, sx, sy, sr):
    if rule == "Diag":
        # compactify! (src/unaryrules.jl:62-91): y[ci] = x[subindex(indexer, ci)]
        locs = [iy.index(l) for l in ix]
        grids = np.indices(y.shape, sparse=True) if y.ndim else ()
        v = x[tuple(grids[q] for q in locs)] if x.ndim else x
        return addmul(y, sy, v, sx, sr)
    if rule == "Sum":
        # src/unaryrules.jl:30-40
        dims = tuple(p for p, l in enumerate(ix) if l not in iy)
        res = sr.reduce(x, dims)
        ix1f = [l for l in ix if l in iy]
        if ix1f != list(iy):
            return _unary_rule("Permutedims", ix1f, iy, res, y, sx, sy, sr)
        return addmul(y, sy, res, sx, sr)
    if rule == "Permutedims":
        # src/unaryrules.jl:125-129
        perm = [ix.index(l) for l in iy]
        return tensorpermute(y, x, perm, sx, sy, sr)
    if rule == "Repeat":
        # src/unaryrules.jl:44-58
        ix1f = [l for l in iy if l in ix]
        if ix1f != list(ix):
            y1 = np.empty([y.shape[iy.index(l)] for l in ix1f], dtype=x.dtype, order="F")
            y1 = _unary_rule("Permutedims", ix, ix1f, x, y1, sr.one(x.dtype), sr.zero(x.dtype), sr)
        else:
            y1 = x
        shape2 = [s if l in ix else 1 for l, s in zip(iy, y.shape)]
        return addmul(y, sy, np.broadcast_to(y1.reshape(shape2, order="F"), y.shape), sx, sr)
    if rule == "Duplicate":
        # duplicate! (src/unaryrules.jl:93-121)
        locs = [ix.index(l) for l in iy]
        grids = np.indices(x.shape, sparse=True) if x.ndim else ()
        idx = tuple(grids[q] for q in locs) if y.ndim else ()
        with np.errstate(over="ignore", invalid="ignore"):
            if sr.is_zero(sy):
                y[...] = sr.zero(y.dtype)
                y[idx] = sr.mul(sx, x)
            else:
                y[...] = sr.mul(sy, y)
                y[idx] = sr.add(sr.mul(sx, x), y[idx])
        return y
    raise ValueError(rule)


def _unary_einsum(ix, iy, x, y, sx, sy, size_dict, sr):
    """einsum!(ixs, iy, xs::NTuple{1}, ...) (src/einsum.jl:84-101)."""
    pipeline = unary_pipeline(list(ix), list(iy))
    last = x
    for k, (rule, oix, oiy) in enumerate(pipeline):
        if k == len(pipeline) - 1:
            _unary_rule(rule, oix, oiy, last, y, sx, sy, sr)
        else:
            cache = np.empty([size_dict[l] for l in oiy], dtype=y.dtype, order="F")
            _unary_rule(rule, oix, oiy, last, cache, sr.one(y.dtype), sr.zero(y.dtype), sr)
            last = cache
    if not pipeline:
        addmul(y, sy, last, sx, sr)
    return y


# ---------------------------------------------------------------------------------------
# binary path (src/einsum.jl:104-211, src/binaryrules.jl)


def _analyze_binary_input(ix1, ix2, iy):
    """src/einsum.jl:187-211."""
    batch, inner, outer1 = [], [], []
    for l1 in ix1:
        if l1 in ix2:
            if l1 in iy:
                if l1 not in batch:
                    batch.append(l1)
            elif l1 not in inner:
                inner.append(l1)
        elif l1 in iy and l1 not in outer1:
            outer1.append(l1)
    outer2 = []
    for l2 in ix2:
        if l2 not in ix1 and l2 in iy and l2 not in outer2:
            outer2.append(l2)
    return inner, outer1, outer2, batch


def analyze_binary(ix1, ix2, iy, size_dict):
    """src/einsum.jl:139-185: returns c1, c2, cy, s1, s2, s3, i1, i2, iyb."""
    inner, outer1, outer2, batch = _analyze_binary_input(ix1, ix2, iy)
    c1 = outer1 + inner + batch
    c2 = inner + outer2 + batch
    cy = outer1 + outer2 + batch
    prod = lambda ls: int(np.prod([size_dict[l] for l in ls], dtype=np.int64))
    si, sj, sk, sl = prod(outer1), prod(inner), prod(outer2), prod(batch)
    i1, i2, iyb, s1, s2, s3 = [], [], [], [], [], []
    if outer1:
        i1.append("i"); iyb.append("i"); s1.append(si); s3.append(si)
    if inner:
        i1.append("j"); i2.append("j"); s1.append(sj); s2.append(sj)
    if outer2:
        i2.append("k"); iyb.append("k"); s2.append(sk); s3.append(sk)
    if batch:
        i1.append("l"); i2.append("l"); iyb.append("l")
        s1.append(sl); s2.append(sl); s3.append(sl)
    return c1, c2, cy, s1, s2, s3, i1, i2, iyb


def _simplifyto(ix1, c1, x1, size_dict, sr):
    """src/einsum.jl:127-134."""
    if list(c1) != list(ix1):
        xs1 = np.empty([size_dict[l] for l in c1], dtype=x1.dtype, order="F")
        return _unary_einsum(ix1, c1, x1, xs1, sr.one(x1.dtype), sr.zero(x1.dtype), size_dict, sr)
    return x1


def _binary_rule(i1, i2, iyb, x1, x2, y, sx, sy, sr):
    """binary_einsum!(::SimpleBinaryRule, ...) (src/binaryrules.jl:18-142).  Every form that
    analyze_binary produces is an instance of the batched product
    y[i,k,l] = sy*y + sx*sum_j x1[i,j,l]*x2[j,k,l] with absent letters of size 1; the
    reference dispatches each instance to mul!/batched_gemm!/broadcast, which compute the
    same sums (summation order inside BLAS differs, as it does between BLAS builds)."""
    def expand(x, have, want):
        shape = []
        it = iter(x.shape)
        for c in want:
            shape.append(next(it) if c in have else 1)
        return x.reshape(shape, order="F")
    a = expand(x1, i1, "ijl")
    b = expand(x2, i2, "jkl")
    yy = expand(y, iyb, "ikl")
    with np.errstate(over="ignore", invalid="ignore"):
        prod = sr.bmm(a, b)
    addmul(yy, sy, prod, sx, sr)
    return y


def _binary_einsum(ix1, ix2, iy, x1, x2, y, sx, sy, size_dict, sr):
    """einsum!(ixs, iy, xs::NTuple{2}, ...) (src/einsum.jl:104-124)."""
    c1, c2, cy, s1, s2, s3, i1, i2, iyb = analyze_binary(list(ix1), list(ix2), list(iy), size_dict)
    xs1 = _simplifyto(ix1, c1, x1, size_dict, sr)
    xs2 = _simplifyto(ix2, c2, x2, size_dict, sr)
    x1_ = xs1.reshape(s1, order="F")
    x2_ = xs2.reshape(s2, order="F")
    if list(cy) != list(iy):
        y_ = np.empty(s3, dtype=y.dtype, order="F")
        _binary_rule(i1, i2, iyb, x1_, x2_, y_, sr.one(y.dtype), sr.zero(y.dtype), sr)
        y_ = y_.reshape([size_dict[l] for l in cy], order="F")
        return _unary_einsum(cy, iy, y_, y, sx, sy, size_dict, sr)
    _binary_rule(i1, i2, iyb, x1_, x2_, y.reshape(s3, order="F"), sx, sy, sr)
    return y


# ---------------------------------------------------------------------------------------
# entry points


def get_size_dict(ixs, xs, size_info=None):
    """get_size_dict! (src/interfaces.jl:67-106): arity/rank -> ArgumentError (ValueError
    here), size disagreement -> DimensionMismatch (ValueError here)."""
    d = dict(size_info or {})
    if len(ixs) != len(xs):
        raise ValueError("%d tensors labelled by %d indices" % (len(xs), len(ixs)))
    for ix, x in zip(ixs, xs):
        if len(ix) != np.ndim(x):
            raise ValueError("indices %r invalid for tensor with ndims = %d" % (ix, np.ndim(x)))
        for l, s in zip(ix, np.shape(x)):
            if d.setdefault(l, s) != s:
                raise ValueError("DimensionMismatch: %r = %d or %d" % (l, d[l], s))
    return d


def einsum_inplace(ixs, iy, xs, y, sx, sy, size_dict, semiring="plus_times"):
    """einsum!(ixs, iy, xs, y, sx, sy, size_dict) (src/einsum.jl:43-124)."""
    sr = _sr(semiring)
    xs = [_F(np.asarray(x)) for x in xs]
    if len(xs) == 1:
        return _unary_einsum(list(ixs[0]), list(iy), xs[0], y, sx, sy, size_dict, sr)
    if len(xs) == 2:
        return _binary_einsum(list(ixs[0]), list(ixs[1]), list(iy), xs[0], xs[1], y, sx, sy,
                              size_dict, sr)
    return loop.loop_einsum(ixs, iy, xs, size_dict, y=y, sx=sx, sy=sy, tropical=sr.tropical)


def einsum(ixs, iy, xs, size_dict=None, semiring="plus_times"):
    """einsum(code, xs, size_dict) (src/einsum.jl:24-27): y = get_output_array(...);
    einsum!(code, xs, y, true, false, size_dict)."""
    sr = _sr(semiring)
    xs = [_F(np.asarray(x)) for x in xs]
    size_dict = get_size_dict(ixs, xs, size_dict)
    T = np.result_type(*[x.dtype for x in xs])
    xs = [x.astype(T, copy=False) for x in xs]
    y = np.empty([size_dict[l] for l in iy], dtype=T, order="F")
    return einsum_inplace(ixs, iy, xs, y, sr.one(T), sr.zero(T), size_dict, semiring)


def loop_einsum(ixs, iy, xs, size_dict=None, semiring="plus_times"):
    """loop_einsum(code, xs, size_dict) (src/loop_einsum.jl:8-13)."""
    return loop.loop_einsum(ixs, iy, xs, size_dict, tropical=(semiring == "max_plus"))


# ---- trees (src/einsequence.jl:257-287, src/slicing.jl:10-137)


def tree_leaves_ixs(tree, out=None):
    """collect_ixs! (src/einsequence.jl:637-648): {tensorindex: labels}."""
    out = {} if out is None else out
    args, ixs, _ = tree
    for a, ix in zip(args, ixs):
        if isinstance(a, int):
            out[a] = list(ix)
        else:
            tree_leaves_ixs(a, out)
    return out


def nested_einsum(tree, xs, size_dict, semiring="plus_times"):
    """einsum(::NestedEinsum, xs, size_dict) (src/einsequence.jl:279-287): post-order,
    left to right."""
    args, ixs, iy = tree
    mxs = [xs[a] if isinstance(a, int) else nested_einsum(a, xs, size_dict, semiring)
           for a in args]
    return einsum(ixs, iy, mxs, size_dict, semiring)


def drop_slicedim(tree, slices):
    """src/slicing.jl:139-145."""
    if isinstance(tree, int):
        return tree
    args, ixs, iy = tree
    return ([drop_slicedim(a, slices) for a in args],
            [[l for l in ix if l not in slices] for ix in ixs],
            [l for l in iy if l not in slices])


def sliced_einsum(slicing, tree, xs, size_dict, semiring="plus_times", acc_dtype=None):
    """einsum(::SlicedEinsum, xs, size_dict) (src/slicing.jl:123-137): zero output, loop
    the SliceIterator (column-major over the sliced labels, :37-71), take_slice (:78-85),
    evaluate the reduced tree, fill_slice! (:86-94) accumulate.  acc_dtype (optional)
    widens the accumulator; None keeps the reference's output type."""
    sr = _sr(semiring)
    if not slicing:
        return nested_einsum(tree, xs, size_dict, semiring)
    leaves = tree_leaves_ixs(tree)
    ixsv = [leaves[i] for i in range(len(xs))]
    iyv = list(tree[2])
    T = np.result_type(*[np.asarray(x).dtype for x in xs])
    res = np.full([size_dict[l] for l in iyv], sr.zero(T), dtype=acc_dtype or T, order="F")
    sd = dict(size_dict)
    for l in slicing:
        sd[l] = 1
    inner = drop_slicedim(tree, slicing)
    ranges = [range(size_dict[l]) for l in slicing]
    # CartesianIndices order: first sliced label fastest
    for combo in itertools.product(*reversed(ranges)):
        sm = dict(zip(reversed(slicing), combo))
        xsi = []
        for x, ix in zip(xs, ixsv):
            idx = tuple(sm[l] if l in sm else slice(None) for l in ix)
            xsi.append(np.asarray(x)[idx])
        resi = nested_einsum(inner, xsi, {l: s for l, s in sd.items()}, semiring)
        idx = tuple(sm[l] if l in sm else slice(None) for l in iyv)
        with np.errstate(over="ignore"):
            res[idx] = sr.add(res[idx], resi)
    return res


# ---- gradients (src/autodiff.jl:24-33, src/bp.jl)


def einsum_grad(ixs, xs, iy, size_dict, dy, i):
    """src/autodiff.jl:24-30: nixs = ixs[i <- iy]; nxs = xs[i <- conj(dy)];
    y = einsum(nixs -> ixs[i]); ProjectTo(xs[i])(conj(y)) (real input keeps the real part)."""
    nixs = list(ixs)
    nixs[i] = list(iy)
    nxs = list(xs)
    nxs[i] = np.conj(dy)
    y = einsum(nixs, list(ixs[i]), nxs, size_dict)
    g = np.conj(y)
    if not np.iscomplexobj(np.asarray(xs[i])):
        g = np.real(g)
    return g.astype(np.asarray(xs[i]).dtype, copy=False)


def cost_and_gradient(tree, xs, ybar=None):
    """cost_and_gradient (src/bp.jl:122-130) with cached_einsum (:33-48), back_propagate
    (:79-96), init_gradient (:133-138) and extract_leaves (:144-160)."""
    leaves = tree_leaves_ixs(tree)
    ixsv = [leaves[i] for i in range(len(xs))]
    size_dict = get_size_dict(ixsv, xs)

    def cached(t):
        if isinstance(t, int):
            return (np.asarray(xs[t]), [])
        args, ixs, iy = t
        caches = [cached(a) for a in args]
        y = einsum(ixs, iy, [c[0] for c in caches], size_dict)
        return (y, caches)

    cache = cached(tree)
    if ybar is None:
        ybar = np.ones_like(cache[0])
        assert ybar.ndim == 0, "The output must be a scalar"
    grads = [None] * len(xs)

    def back(t, c, dy):
        if isinstance(t, int):
            grads[t] = dy
            return
        args, ixs, iy = t
        cxs = [cc[0] for cc in c[1]]
        for k, (a, cc) in enumerate(zip(args, c[1])):
            back(a, cc, einsum_grad(ixs, cxs, iy, size_dict, dy, k))

    back(tree, cache, np.asarray(ybar))
    return cache[0], grads
