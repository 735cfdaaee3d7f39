"""einsum / einsum! / loop_einsum front end (mirrors src/einsum.jl:24-50,
src/loop_einsum.jl:8-32, src/utils.jl:122-177) over the C ABI.

Every contraction is enqueued on the GPU by libomeinsum_mi355x.so on torch's current stream.
"""
import collections
import ctypes

import numpy as np

from . import _lib
from .codes import EinCode, NestedEinsum, SlicedEinsum, get_size_dict
from .device import _AS_IS, common_kind, dtype_code, kind_of, raw, scalar, to_device, wrap


def _torch():
    import torch
    return torch


def _unwrap(xs):
    """-> (device tensors, wrapper class of the elements or None, common torch dtype)."""
    torch = _torch()
    # plain contiguous tensors of one supported type on the current device pass as they are
    t0 = xs[0] if xs else None
    if type(t0) is torch.Tensor:
        T, dev = t0.dtype, t0.device
        if dev.type == "cuda" and dev.index == torch.cuda.current_device() and T in _AS_IS and \
                all(type(x) is torch.Tensor and x.dtype == T and x.device == dev and
                    x.is_contiguous() for x in xs):
            return list(xs), None, T
    kind = common_kind(xs)
    ts = [to_device(raw(x)) for x in xs]
    T = ts[0].dtype
    for t in ts[1:]:
        T = torch.promote_types(T, t.dtype)
    # promote_type of get_output_array (src/loop_einsum.jl:48-61); align_eltypes converts
    ts = [t if t.dtype == T else t.to(T) for t in ts]
    return ts, kind, T


def _labels(ixs, iy):
    ids = {}
    for ix in list(ixs) + [iy]:
        for l in ix:
            if l not in ids:
                ids[l] = len(ids)
    return ids


_ARGS = collections.OrderedDict()  # (ixs, iy, sizes) -> the ctypes label / size arrays
_ARGS_MAX = 256
_IDS = {}  # (ixs, iy) -> labels in first-appearance order


def _flat_args(ixs, iy, size_dict):
    """(nix, cixs, ciy, nlabels, sizes) ctypes arrays of one flat code, cached per code and
    sizes (the arrays are read-only inputs of the C call)."""
    key = (tuple(map(tuple, ixs)), tuple(iy))
    ids = _IDS.get(key)
    if ids is None:
        ids = _labels(ixs, iy)
        if len(_IDS) > 4 * _ARGS_MAX:
            _IDS.clear()
        _IDS[key] = ids
    try:
        key = key + (tuple([int(size_dict[l]) for l in ids]),)
    except KeyError as e:
        raise KeyError("size of label %r is unknown (pass size_info)" % (e.args[0],))
    hit = _ARGS.get(key)
    if hit is not None:
        _ARGS.move_to_end(key)
        return hit
    sizes = (ctypes.c_int64 * max(len(ids), 1))(*key[2])
    n = len(ixs)
    nix = (ctypes.c_int32 * n)(*[len(ix) for ix in ixs])
    flat = [ids[l] for ix in ixs for l in reversed(list(ix))]
    cixs = (ctypes.c_int32 * max(len(flat), 1))(*flat)
    ciy = (ctypes.c_int32 * max(len(iy), 1))(*[ids[l] for l in reversed(list(iy))])
    hit = (nix, cixs, ciy, len(ids), sizes)
    _ARGS[key] = hit
    if len(_ARGS) > _ARGS_MAX:
        _ARGS.popitem(last=False)
    return hit


_PTR_ARRAYS = {}   # n -> ctypes array type of n pointers
_SCALARS = {}      # (type, value, element code) -> ome_scalar (a read-only input)


def _scalar(v, code):
    key = (type(v), v, code)
    try:
        hit = _SCALARS.get(key)
    except TypeError:  # unhashable number type
        return scalar(v, code)
    if hit is None:
        hit = scalar(v, code)
        if len(_SCALARS) > 1024:
            _SCALARS.clear()
        _SCALARS[key] = hit
    return hit


def _call_flat(ixs, iy, ts, yt, code, sx, sy, size_dict, force_loop=False):
    """One flat einsum on the device.  Row-major tensors are passed with reversed label
    lists (see device.py)."""
    ctx = _lib.context(yt.device.index)
    nix, cixs, ciy, nlabels, sizes = _flat_args(ixs, iy, size_dict)
    n = len(ixs)
    arr = _PTR_ARRAYS.get(n)
    if arr is None:
        arr = _PTR_ARRAYS[n] = ctypes.c_void_p * n
    ptrs = arr(*[t.data_ptr() for t in ts])
    fn = ctx.lib.ome_loop_einsum if force_loop else ctx.lib.ome_einsum
    _lib.check(fn(ctx.h, code, n, ptrs, nix, cixs, yt.data_ptr(), len(iy), ciy, nlabels, sizes,
                  ctypes.byref(_scalar(sx, code)), ctypes.byref(_scalar(sy, code))), ctx.h)


def _flat_einsum_inplace(ixs, iy, xs, y, sx, sy, size_dict, force_loop=False):
    ts, kind, T = _unwrap(xs)
    yt = raw(y)
    if kind != kind_of(y):
        raise TypeError("output and inputs must have the same element kind")
    if yt.dtype != T:
        ts = [t.to(yt.dtype) for t in ts]
    if not yt.is_contiguous():
        raise ValueError("output tensor must be contiguous")
    for ix, t in zip(ixs, ts):
        if t.dim() != len(ix):
            from .codes import ArgumentError
            raise ArgumentError("indices %r invalid for tensor with ndims = %d" % (tuple(ix), t.dim()))
    code = dtype_code(yt.dtype, kind)
    _call_flat(ixs, iy, ts, yt, code, sx, sy, size_dict, force_loop)
    return y


def _new_output(xs, shape):
    ts, kind, T = _unwrap(xs)
    torch = _torch()
    y = torch.empty(tuple(int(s) for s in shape), dtype=T, device=ts[0].device)
    return wrap(kind, y)


def flat_einsum(code, xs, size_dict):
    """einsum(code::EinCode, xs, size_dict) (src/einsum.jl:24-27)."""
    y = _new_output(xs, [size_dict[l] for l in code.iy])
    return _flat_einsum_inplace(code.ixs, code.iy, xs, y, True, False, size_dict)


_GET_BACKEND = []


def _active_backend():
    """get_einsum_backend(), with backends imported once (it imports this module)."""
    if not _GET_BACKEND:
        from .backends import get_einsum_backend
        _GET_BACKEND.append(get_einsum_backend)
    return _GET_BACKEND[0]()


def einsum(code, xs, size_dict=None, **kw):
    """einsum(code, xs, size_dict): EinCode, NestedEinsum or SlicedEinsum, dispatched on the
    active backend (src/backends.jl:108,219)."""
    xs = tuple(xs)
    if size_dict is None:
        size_dict = get_size_dict(code.getixsv(), xs)
    return _active_backend().einsum(code, xs, size_dict, **kw)


def einsum_inplace(code, xs, y, sx=True, sy=False, size_dict=None):
    """einsum!(code, xs, y, sx, sy, size_dict) (src/einsum.jl:43-50, src/einsequence.jl:270-277,
    src/slicing.jl:112-121)."""
    xs = tuple(xs)
    if size_dict is None:
        size_dict = get_size_dict(code.getixsv(), xs)
    return _active_backend().einsum_inplace(code, xs, y, sx, sy, size_dict)


def loop_einsum(code, xs, size_dict=None):
    """loop_einsum(code, xs, size_dict) (src/loop_einsum.jl:8-13): the index walk."""
    xs = tuple(xs)
    if size_dict is None:
        size_dict = get_size_dict(code.ixs, xs)
    y = _new_output(xs, [size_dict[l] for l in code.iy])
    return _flat_einsum_inplace(code.ixs, code.iy, xs, y, True, False, size_dict, force_loop=True)


def loop_einsum_inplace(ixs, iy, xs, y, sx, sy, size_dict):
    """loop_einsum!(ixs, iy, xs, y, sx, sy, size_dict) (src/loop_einsum.jl:21-32)."""
    return _flat_einsum_inplace(ixs, iy, tuple(xs), y, sx, sy, size_dict, force_loop=True)


def allow_loops(flag):
    """allow_loops(flag) (src/loop_einsum.jl:63-73)."""
    _lib.context().set_option(_lib.OPT_ALLOW_LOOPS, 1 if flag else 0)
    return flag


def tensorpermute(C, A, perm, sx=True, sy=False):
    """tensorpermute!(C, A, perm, sx, sy) (src/utils.jl:122-156) with 0-based perm:
    C = permutedims(A, perm) scaled."""
    ts, kind, T = _unwrap([A])
    a = ts[0]
    ct = raw(C)
    ctx = _lib.context(ct.device.index)
    code = dtype_code(ct.dtype, kind)
    nd = a.dim()
    # row-major axis p is column-major dim nd-1-p
    dims = (ctypes.c_int64 * max(nd, 1))(*[a.shape[nd - 1 - q] for q in range(nd)])
    cperm = (ctypes.c_int32 * max(nd, 1))(*[nd - 1 - perm[nd - 1 - q] for q in range(nd)])
    s_x, s_y = scalar(sx, code), scalar(sy, code)
    _lib.check(ctx.lib.ome_permute(ctx.h, code, nd, dims, cperm, ctypes.c_void_p(a.data_ptr()),
                                   ctypes.c_void_p(ct.data_ptr()), ctypes.byref(s_x),
                                   ctypes.byref(s_y)), ctx.h)
    return C


def batched_gemm(ta, tb, alpha, A, B, beta, C):
    """_batched_gemm!(C1, C2, alpha, A, B, beta, C) (src/utils.jl:162-177) in torch's row-major
    convention: A is (batch, m, k) ('N') or (batch, k, m) ('T'), B is (batch, k, n) or
    (batch, n, k), C is (batch, m, n); C = alpha*op(A) op(B) + beta*C."""
    ts, kind, T = _unwrap([A, B])
    a, b = ts
    ct = raw(C)
    ctx = _lib.context(ct.device.index)
    code = dtype_code(ct.dtype, kind)
    batch, m, n = ct.shape
    k = a.shape[2] if ta == "N" else a.shape[1]
    al, be = scalar(alpha, code), scalar(beta, code)
    # row-major C(b,m,n) is column-major C^T (n,m,b) = op(B)^T op(A)^T
    _lib.check(ctx.lib.ome_batched_gemm(ctx.h, code, tb.encode(), ta.encode(), n, m, k, batch,
                                        ctypes.byref(al), ctypes.c_void_p(b.data_ptr()),
                                        ctypes.c_void_p(a.data_ptr()), ctypes.byref(be),
                                        ctypes.c_void_p(ct.data_ptr())), ctx.h)
    return C


class GroupedGemm:
    """Many independent batched GEMMs of uneven shapes in one launch (ome_grouped_gemm).
    problems: list of (A, B, C) row-major tensors with A (batch, m, k), B (batch, k, n),
    C (batch, m, n) (torch convention); calling the object computes C = alpha*A@B + beta*C
    for every problem.  The argument arrays are built once, so a repeated call costs one
    library call (the library caches the device descriptors of the same problem list)."""

    def __init__(self, problems, alpha=True, beta=False):
        self.n = len(problems)
        if not self.n:
            return
        ts, kind, T = _unwrap([p[0] for p in problems] + [p[1] for p in problems])
        na = self.n
        ct0 = problems[0][2]
        ct0 = raw(ct0)
        self.ctx = _lib.context(ct0.device.index)
        self.code = dtype_code(T, kind)
        self.dims = (ctypes.c_int64 * (4 * na))()
        self.pa = (ctypes.c_void_p * na)()
        self.pb = (ctypes.c_void_p * na)()
        self.pc = (ctypes.c_void_p * na)()
        self.keep = []
        for g, (A, B, C) in enumerate(problems):
            a, b = ts[g], ts[na + g]
            ct = raw(C)
            batch, m, k = a.shape
            n = b.shape[2]
            if tuple(b.shape) != (batch, k, n) or tuple(ct.shape) != (batch, m, n):
                raise ValueError("problem %d: shapes %s x %s -> %s do not match" %
                                 (g, tuple(a.shape), tuple(b.shape), tuple(ct.shape)))
            if not ct.is_contiguous() or ct.dtype != T:
                raise ValueError("problem %d: C must be a contiguous %s tensor" % (g, T))
            self.keep.append((a, b, ct))
            # column-major view: C^T (n, m, batch) = B^T (n, k) * A^T (k, m)
            self.dims[4 * g:4 * g + 4] = [n, m, k, batch]
            self.pa[g] = b.data_ptr()
            self.pb[g] = a.data_ptr()
            self.pc[g] = ct.data_ptr()
        self.al, self.be = scalar(alpha, self.code), scalar(beta, self.code)

    def __call__(self):
        if not self.n:
            return
        self.ctx.sync_stream()
        _lib.check(self.ctx.lib.ome_grouped_gemm(self.ctx.h, self.code, self.n, self.dims, self.pa,
                                                 self.pb, self.pc, ctypes.byref(self.al),
                                                 ctypes.byref(self.be)), self.ctx.h)


def grouped_gemm(problems, alpha=True, beta=False):
    """One call of GroupedGemm(problems, alpha, beta)."""
    GroupedGemm(problems, alpha, beta)()


class GroupedPermute:
    """Many tensorpermute! calls (src/utils.jl:122-156) in one launch (ome_grouped_permute):
    calling the object sets ys[g] = sx * xs[g].permute(perms[g]) + sy * ys[g] (torch's axis
    convention: output axis p is input axis perms[g][p])."""

    def __init__(self, xs, perms, ys, sx=True, sy=False):
        self.n = len(xs)
        if not self.n:
            return
        ts, kind, T = _unwrap(list(xs))
        self.ctx = _lib.context(ts[0].device.index)
        self.code = dtype_code(T, kind)
        n = self.n
        self.nd = (ctypes.c_int32 * n)()
        dims, cperm = [], []
        self.pa = (ctypes.c_void_p * n)()
        self.pc = (ctypes.c_void_p * n)()
        self.keep = []
        for g, (a, perm, y) in enumerate(zip(ts, perms, ys)):
            yt = raw(y)
            r = a.dim()
            if sorted(perm) != list(range(r)):
                raise ValueError("problem %d: %s is not a permutation of %d axes" % (g, perm, r))
            if tuple(yt.shape) != tuple(a.shape[p] for p in perm) or not yt.is_contiguous() \
                    or yt.dtype != T:
                raise ValueError("problem %d: output must be a contiguous %s tensor of the "
                                 "permuted shape" % (g, T))
            self.keep.append((a, yt))
            self.nd[g] = r
            # row-major axis p is column-major dim r-1-p
            dims += [a.shape[r - 1 - q] for q in range(r)]
            cperm += [r - 1 - perm[r - 1 - q] for q in range(r)]
            self.pa[g] = a.data_ptr()
            self.pc[g] = yt.data_ptr()
        self.dims = (ctypes.c_int64 * max(len(dims), 1))(*dims)
        self.perm = (ctypes.c_int32 * max(len(cperm), 1))(*cperm)
        self.sx, self.sy = scalar(sx, self.code), scalar(sy, self.code)

    def __call__(self):
        if not self.n:
            return
        self.ctx.sync_stream()
        _lib.check(self.ctx.lib.ome_grouped_permute(self.ctx.h, self.code, self.n, self.nd,
                                                    self.dims, self.perm, self.pa, self.pc,
                                                    ctypes.byref(self.sx), ctypes.byref(self.sy)),
                   self.ctx.h)


def grouped_permute(xs, perms, ys, sx=True, sy=False):
    """One call of GroupedPermute(xs, perms, ys, sx, sy); returns ys."""
    GroupedPermute(xs, perms, ys, sx, sy)()
    return ys
