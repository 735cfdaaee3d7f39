"""TEST INFRASTRUCTURE ONLY — ctypes binding of the C restatement of OMEinsum.loop_einsum!.

Restates src/loop_einsum.jl:21-45 (loop_einsum!, reduce_einarray!) over
src/Core.jl:122-145,201-268 (EinIndexer, einarray, indices_and_locs, map_prod).  The C source
is ``oracle/loop_einsum.c``; ``make -C oracle`` builds ``oracle/liboracle.so``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# element-type codes shared with the C oracle and with include/omeinsum_mi355x.h
F32, F64, C32, C64, I64, TF32, TF64, DD, CT64, TBOOL = range(10)

# numpy storage of each element type.  Double-word numbers (DoubleFloats' Double64) are held
# as complex128 arrays whose real part is hi and imaginary part lo: the same interleaved
# 16-byte elements as the C ABI, so numpy's element order and strides apply unchanged.
# CountingTropical{Float64,Float64} is held the same way (real part n, imaginary part c), and
# TropicalAndOr as numpy bool.
NP_DTYPE = {F32: np.float32, F64: np.float64, C32: np.complex64, C64: np.complex128,
            I64: np.int64, TF32: np.float32, TF64: np.float64, DD: np.complex128,
            CT64: np.complex128, TBOOL: np.bool_}


class _Scalar(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double), ("i", ctypes.c_int64)]


def build():
    """Compile liboracle.so from loop_einsum.c (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.oracle_loop_einsum.restype = ctypes.c_int
        lib.oracle_loop_einsum.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p,
            ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.c_int,
            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_Scalar), ctypes.POINTER(_Scalar)]
        _LIB = lib
    return _LIB


def dtype_code(arr_dtype, tropical=False, dd=False, semiring=None):
    """semiring: None, "tropical", "dd", "counting" (CountingTropical) or "andor"."""
    d = np.dtype(arr_dtype)
    if semiring == "counting":
        if d != np.dtype(np.complex128):
            raise TypeError("CountingTropical arrays are stored as complex128 (n + i*c)")
        return CT64
    if semiring == "andor":
        if d != np.dtype(np.bool_):
            raise TypeError("TropicalAndOr arrays are numpy bool")
        return TBOOL
    dd = dd or semiring == "dd"
    tropical = tropical or semiring == "tropical"
    if dd:
        if d != np.dtype(np.complex128):
            raise TypeError("double-word arrays are stored as complex128 (hi + i*lo)")
        return DD
    if tropical:
        return {np.dtype(np.float32): TF32, np.dtype(np.float64): TF64}[d]
    return {np.dtype(np.float32): F32, np.dtype(np.float64): F64,
            np.dtype(np.complex64): C32, np.dtype(np.complex128): C64,
            np.dtype(np.int64): I64}[d]


def scalar(v, code):
    """Julia Bools promote to the element type's one/zero (true*x == x, false -> zero); for
    tropical elements a number is the tropical content, so true -> 0 and false -> -Inf."""
    s = _Scalar()
    if isinstance(v, (bool, np.bool_)) and code in (TF32, TF64):
        s.re = 0.0 if v else -np.inf
        return s
    if code == CT64:
        if isinstance(v, (bool, np.bool_)):
            s.re, s.im = (0.0, 1.0) if v else (-np.inf, 0.0)
        else:
            c = complex(v)
            s.re, s.im = c.real, c.imag
        return s
    if code == TBOOL:
        s.i = 1 if v else 0
        return s
    if code == I64:
        s.i = int(v)
    elif code == DD:
        s.re = float(v)
    elif code in (C32, C64):
        c = complex(v)
        s.re, s.im = c.real, c.imag
    else:
        s.re = float(v)
    return s


def label_ids(ixs, iy):
    """Dense int ids in first-appearance order over ixs then iy."""
    ids = {}
    for ix in list(ixs) + [iy]:
        for l in ix:
            if l not in ids:
                ids[l] = len(ids)
    return ids


def loop_einsum(ixs, iy, xs, size_dict=None, y=None, sx=True, sy=False, tropical=False,
                dd=False, semiring=None):
    """loop_einsum!(ixs, iy, xs, y, sx, sy, size_dict) on numpy arrays (axis p <-> label ix[p]).

    All inputs are promoted to a common dtype first, as get_output_array does
    (src/loop_einsum.jl:48-61).  Returns y (a new array unless one is given).  dd=True: the
    complex128 arrays hold double-word numbers (real part hi, imaginary part lo)."""
    ixs = [tuple(ix) for ix in ixs]
    iy = tuple(iy)
    if len(ixs) != len(xs):
        raise ValueError("%d tensors labelled by %d indices" % (len(xs), len(ixs)))
    size_dict = dict(size_dict or {})
    for ix, x in zip(ixs, xs):
        x = np.asarray(x)
        if x.ndim != len(ix):
            raise ValueError("indices %s invalid for tensor with ndims = %d" % (ix, x.ndim))
        for l, s in zip(ix, x.shape):
            if size_dict.setdefault(l, s) != s:
                raise ValueError("dimension mismatch for label %r" % (l,))
    T = np.result_type(*[np.asarray(x).dtype for x in xs])
    if y is not None:
        T = y.dtype
    code = dtype_code(T, tropical, dd, semiring)
    fx = [np.asarray(x, dtype=T) if np.ndim(x) == 0 else np.asfortranarray(np.asarray(x, dtype=T)) for x in xs]
    ids = label_ids(ixs, iy)
    sizes = np.zeros(max(len(ids), 1), dtype=np.int64)
    for l, i in ids.items():
        sizes[i] = size_dict[l]
    yshape = tuple(int(size_dict[l]) for l in iy)
    if y is None:
        out = np.zeros(yshape, dtype=T, order="F")
    else:
        assert y.shape == yshape
        out = y.copy() if y.ndim == 0 else np.asfortranarray(y.copy())
    nix = (ctypes.c_int32 * len(ixs))(*[len(ix) for ix in ixs])
    flat = [ids[l] for ix in ixs for l in ix]
    cixs = (ctypes.c_int32 * max(len(flat), 1))(*flat)
    ciy = (ctypes.c_int32 * max(len(iy), 1))(*[ids[l] for l in iy])
    ptrs = (ctypes.c_void_p * len(fx))(*[x.ctypes.data for x in fx])
    csizes = sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    sxs, sys_ = scalar(sx, code), scalar(sy, code)
    rc = _lib().oracle_loop_einsum(code, len(fx), ptrs, nix, cixs, out.ctypes.data, len(iy),
                                   ciy, len(ids), csizes, ctypes.byref(sxs), ctypes.byref(sys_))
    if rc != 0:
        raise ValueError("oracle_loop_einsum: invalid arguments")
    if y is not None:
        y[...] = out
        return y
    return out


def dw_op(op, x, y):
    """The oracle's double-word + (op "+") or * (op "*") elementwise on complex128 arrays
    (hi + i*lo) of one shape."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.ascontiguousarray(y, dtype=np.complex128)
    assert x.shape == y.shape
    z = np.empty_like(x)
    lib = _lib()
    lib.oracle_dw_op.restype = ctypes.c_int
    lib.oracle_dw_op.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
    lib.oracle_dw_op(0 if op == "+" else 1, x.size, x.ctypes.data, y.ctypes.data, z.ctypes.data)
    return z
