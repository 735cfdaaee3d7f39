"""Device plans for contraction trees (NestedEinsum / SlicedEinsum) over ome_tree_*.

A plan is compiled once per (tree, label sizes, element type, planning options) and cached;
evaluation replays a recorded hipGraph when the same input/output buffers come back (see
csrc/tree.hip).
"""
import collections
import ctypes

from . import _lib
from .codes import NestedEinsum, SlicedEinsum
from .device import common_kind, dtype_code, raw, scalar, to_device, wrap


def tree_desc(ne, ids):
    """The post-order int32 description of ome_tree_create.  Every label list is reversed
    because row-major tensors are column-major tensors with reversed labels."""
    nodes = ne.internal_nodes()
    index = {id(n): k for k, n in enumerate(nodes)}
    leaves = ne.getixsv()
    desc = [len(leaves), len(nodes)]
    for n in nodes:
        desc.append(len(n.args))
        for a, ix in zip(n.args, n.eins.ixs):
            ref = a.tensorindex if a.isleaf() else -(index[id(a)] + 1)
            desc += [ref, len(ix)] + [ids[l] for l in reversed(list(ix))]
        desc += [len(n.eins.iy)] + [ids[l] for l in reversed(list(n.eins.iy))]
    return desc


class TreePlan:
    """A device plan of one tree.  defer_memory=True sizes the plan's device memory without
    allocating it (the request loop lends memory when it admits the plan; a direct call
    allocates it on first use)."""

    def __init__(self, code, size_dict, code_dt, device, defer_memory=False):
        if isinstance(code, SlicedEinsum):
            ne, slicing = code.eins, list(code.slicing)
        else:
            ne, slicing = code, []
        if ne.isleaf():
            raise ValueError("a leaf is not a contraction")
        self.code = code
        self.ne = ne
        self.dt = code_dt
        self.device = device
        self.ctx = _lib.context(device)
        ids = {}
        for ix in ne.getixsv() + [ne.getiyv()]:
            for l in ix:
                ids.setdefault(l, len(ids))
        for n in ne.internal_nodes():
            for ix in list(n.eins.ixs) + [n.eins.iy]:
                for l in ix:
                    ids.setdefault(l, len(ids))
        self.ids = ids
        self.iy = ne.getiyv()
        self.n_leaves = len(ne.getixsv())
        self.out_shape = tuple(int(size_dict[l]) for l in self.iy)
        desc = tree_desc(ne, ids)
        cdesc = (ctypes.c_int32 * len(desc))(*desc)
        sizes = (ctypes.c_int64 * max(len(ids), 1))()
        for l, i in ids.items():
            sizes[i] = int(size_dict[l])
        csl = (ctypes.c_int32 * max(len(slicing), 1))(*[ids[l] for l in slicing])
        h = ctypes.c_void_p()
        if defer_memory:
            self.ctx.set_option(_lib.OPT_TREE_DEFER_MEMORY, 1)
        try:
            _lib.check(self.ctx.lib.ome_tree_create(self.ctx.h, code_dt, cdesc, len(desc),
                                                    len(ids), sizes, len(slicing), csl,
                                                    ctypes.byref(h)), self.ctx.h)
        finally:
            if defer_memory:
                self.ctx.set_option(_lib.OPT_TREE_DEFER_MEMORY, 0)
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                self.ctx.lib.ome_tree_destroy(h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass

    def info(self):
        info = _lib.TreeInfo()
        _lib.check(self.ctx.lib.ome_get_tree_info(self.h, ctypes.byref(info)), self.ctx.h)
        return info

    def _ptrs(self, xs, ts=None):
        """Device tensors of xs (ts when the caller has them already) and their pointers."""
        if ts is None:
            ts = [to_device(raw(x)) for x in xs]
        return ts, (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def _out(self, ts, kind):
        import torch
        y = torch.empty(self.out_shape, dtype=ts[0].dtype, device=ts[0].device)
        return (wrap(kind, y)), y

    def execute(self, xs, y=None, sx=True, sy=False, ts=None):
        """einsum / einsum! of the tree on the device (ts: xs already as device tensors)."""
        self.ctx.sync_stream()
        kind = common_kind(xs)
        ts, ptrs = self._ptrs(xs, ts)
        if y is None:
            res, yt = self._out(ts, kind)
        else:
            res, yt = y, (raw(y))
        s_x, s_y = scalar(sx, self.dt), scalar(sy, self.dt)
        _lib.check(self.ctx.lib.ome_tree_execute(self.h, ptrs, ctypes.c_void_p(yt.data_ptr()),
                                                 ctypes.byref(s_x), ctypes.byref(s_y)), self.ctx.h)
        self._keep = ts  # inputs stay referenced until the next call on this plan
        return res

    def execute_scalar_normalized(self, xs):
        """Full contraction to a scalar with per-step log-scale normalisation:
        returns (value, log_scale) with result = value * exp(log_scale)."""
        self.ctx.sync_stream()
        ts, ptrs = self._ptrs(xs)
        v = (ctypes.c_double * 2)()
        lg = ctypes.c_double()
        _lib.check(self.ctx.lib.ome_tree_execute_scalar_normalized(self.h, ptrs, v, ctypes.byref(lg)),
                   self.ctx.h)
        return complex(v[0], v[1]), lg.value

    def gradient(self, xs, ybar=None):
        """cost_and_gradient on the device: returns (y, [xbar_i])."""
        import torch
        self.ctx.sync_stream()
        kind = common_kind(xs)
        ts, ptrs = self._ptrs(xs)
        res, yt = self._out(ts, kind)
        if ybar is None:
            if self.out_shape != ():
                raise ValueError("The output must be a scalar! Or you need to feed the gradient manually.")
            yb = torch.ones((), dtype=ts[0].dtype, device=ts[0].device)
        else:
            yb = to_device(raw(ybar))
        grads = [torch.empty_like(t) for t in ts]
        gp = (ctypes.c_void_p * len(grads))(*[g.data_ptr() for g in grads])
        _lib.check(self.ctx.lib.ome_tree_gradient(self.h, ptrs, ctypes.c_void_p(yb.data_ptr()),
                                                  ctypes.c_void_p(yt.data_ptr()), gp), self.ctx.h)
        return res, grads


_CACHE = collections.OrderedDict()
_CACHE_MAX = 64


def element_code(xs):
    """The C ABI element type of a tree's inputs (they must share one type)."""
    kind = common_kind(xs)
    ts = [to_device(raw(x)) for x in xs]
    dts = {t.dtype for t in ts}
    if len(dts) != 1:
        raise TypeError("tree inputs must share one element type (promote first)")
    return dtype_code(dts.pop(), kind), ts


def plan_for(code, xs, size_dict, with_tensors=False):
    """The cached plan of code for inputs like xs; with_tensors=True also returns xs as device
    tensors (for execute(..., ts=...))."""
    code_dt, ts = element_code(xs)
    labels = code.uniquelabels() if not isinstance(code, SlicedEinsum) else code.eins.uniquelabels()
    # a plan freezes the context's planning options when it is created
    ctx = _lib.context(ts[0].device.index)
    opts = tuple(ctx.get_option(k) for k in _lib.PLAN_OPTIONS)
    key = (code, tuple((l, size_dict[l]) for l in labels), code_dt, ts[0].device.index, opts)
    plan = _CACHE.get(key)
    if plan is None:
        plan = TreePlan(code, size_dict, code_dt, ts[0].device.index)
        _CACHE[key] = plan
        if len(_CACHE) > _CACHE_MAX:
            _CACHE.popitem(last=False)
    else:
        _CACHE.move_to_end(key)
    return (plan, ts) if with_tensors else plan
