"""Gradients (mirrors src/autodiff.jl and src/bp.jl).

* ``einsum_grad(ixs, xs, iy, size_dict, dy, i)`` — src/autodiff.jl:24-30: the gradient of
  input i is another einsum with ixs[i] replaced by iy and xs[i] by conj(dy), conjugated;
  real inputs keep the real part.
* ``cost_and_gradient(code, xs, ybar=None)`` — src/bp.jl:122-130: with the MI355X backend the
  cached forward and the reverse sweep both run inside the library on one device arena
  (ome_tree_gradient); with the default backend the reference's Python-level recursion is
  restated over per-step device einsums.
* ``einsum_autograd(code, *xs)`` — the ChainRules rrule (src/autodiff.jl:35-44) as a
  torch.autograd.Function.
"""
import ctypes

from . import _lib
from .codes import EinCode, NestedEinsum, SlicedEinsum, get_size_dict
from .device import common_kind, dtype_code, raw, scalar, to_device, wrap


def _conj_copy(t):
    """conj(t) on the device through ome_addmul (a copy for real tensors)."""
    import torch
    out = torch.empty_like(t)
    if t.numel() == 0:
        return out
    ctx = _lib.context(t.device.index)
    code = dtype_code(t.dtype)
    one, zero = scalar(True, code), scalar(False, code)
    _lib.check(ctx.lib.ome_addmul(ctx.h, code, t.numel(), ctypes.c_void_p(t.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), ctypes.byref(one),
                                  ctypes.byref(zero), 1 if t.is_complex() else 0), ctx.h)
    return out


def einsum_grad(ixs, xs, iy, size_dict, dy, i):
    """src/autodiff.jl:24-30.  conj is the identity on real tensors, so real operands are
    used as they are instead of copied."""
    from .ein_ops import flat_einsum
    nixs = [list(ix) for ix in ixs]
    nixs[i] = list(iy)
    nxs = [to_device(x) for x in xs]
    d = to_device(dy)
    if common_kind(list(nxs) + [d]) is not None:
        # wrapped element types (Tropical, Double64) are real: conj is the identity
        nxs[i] = d
        return flat_einsum(EinCode(nixs, list(ixs[i])), tuple(nxs), size_dict)
    nxs[i] = _conj_copy(d) if d.is_complex() else d
    y = flat_einsum(EinCode(nixs, list(ixs[i])), tuple(nxs), size_dict)
    g = _conj_copy(y) if y.is_complex() else y
    x = to_device(xs[i])
    if not x.is_complex() and g.is_complex():
        g = g.real.contiguous()  # ProjectTo(x) keeps the real part
    return g


def einsum_backward_rule(eins, xs, y, size_dict, dy):
    """src/autodiff.jl:31-33."""
    return [einsum_grad(eins.ixs, xs, eins.iy, size_dict, dy, i) for i in range(len(xs))]


def _cost_and_gradient_steps(ne, xs, ybar):
    """cached_einsum + back_propagate + extract_leaves (src/bp.jl:33-160) over per-step
    device einsums."""
    from .ein_ops import flat_einsum
    size_dict = get_size_dict(ne.getixsv(), xs)

    def cached(n):
        if n.isleaf():
            return (to_device(xs[n.tensorindex]), [])
        caches = [cached(a) for a in n.args]
        return (flat_einsum(n.eins, tuple(c[0] for c in caches), size_dict), caches)

    cache = cached(ne)
    import torch
    if ybar is None:
        if raw(cache[0]).dim() != 0:
            raise ValueError("The output must be a scalar! Or you need to feed the gradient manually.")
        kind = common_kind([cache[0]])
        ybar = wrap(kind, torch.full_like(raw(cache[0]), kind.ONE if kind else 1))
    grads = [None] * len(xs)

    def back(n, c, dy):
        if n.isleaf():
            grads[n.tensorindex] = dy
            return
        cxs = [cc[0] for cc in c[1]]
        dxs = einsum_backward_rule(n.eins, cxs, c[0], size_dict, dy)
        for a, cc, dx in zip(n.args, c[1], dxs):
            back(a, cc, dx)

    back(ne, cache, to_device(ybar))
    return cache[0], grads


def cost_and_gradient(code, xs, ybar=None):
    """cost_and_gradient(code, xs, ybar) (src/bp.jl:122-130)."""
    from .backends import MI355XBackend, get_einsum_backend
    xs = tuple(xs)
    if isinstance(code, EinCode):
        code = NestedEinsum([NestedEinsum.leaf(i) for i in range(len(code.ixs))], code)
    if isinstance(code, SlicedEinsum):
        # slicing is not supported for caching: use the unsliced tree (src/bp.jl:26-29)
        code = code.eins
    if isinstance(get_einsum_backend(), MI355XBackend):
        from .tree import plan_for
        size_dict = get_size_dict(code.getixsv(), xs)
        get_einsum_backend().apply(_lib.context())
        return plan_for(code, xs, size_dict).gradient(xs, ybar)
    return _cost_and_gradient_steps(code, xs, ybar)


def einsum_autograd(code, *xs, size_info=None):
    """einsum with a torch autograd rule (the rrule of src/autodiff.jl:35-44)."""
    import torch
    from .ein_ops import einsum
    size_dict = get_size_dict(code.getixsv(), xs, size_info)

    class _Ein(torch.autograd.Function):
        @staticmethod
        def forward(ctx, *ts):
            ctx.save_for_backward(*ts)
            with torch.no_grad():
                return einsum(code, tuple(t.detach() for t in ts), size_dict)

        @staticmethod
        def backward(ctx, dy):
            ts = ctx.saved_tensors
            if isinstance(code, EinCode):
                return tuple(einsum_grad(code.ixs, ts, code.iy, size_dict, dy.contiguous(), i)
                             for i in range(len(ts)))
            _, grads = cost_and_gradient(code, ts, dy.contiguous())
            return tuple(grads)

    return _Ein.apply(*[to_device(x) for x in xs])
