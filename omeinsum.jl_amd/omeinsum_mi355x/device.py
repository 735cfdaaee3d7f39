"""Device tensors and element types.

Tensors are torch tensors on the GPU (torch is used for device memory and streams only).  A
row-major (C-contiguous) tensor with axis labels (l0, ..., l{r-1}) occupies exactly the
memory of a column-major tensor with labels (l{r-1}, ..., l0), so calls into the column-major
C ABI pass every label list reversed.

Element types torch has no dtype for are tensors wrapped in a subclass of ``Wrapped``; the
wrapped ``data`` tensor holds one element per entry, with the C ABI's element layout:
  * ``Tropical``: TropicalNumbers' Tropical{T} (+ is max, * is +), data float32/float64
    holding the tropical numbers' contents;
  * ``Double64``: DoubleFloats' Double64 (value hi + lo), data complex128 whose real part is
    hi and imaginary part lo (the interleaved 16-byte pairs of the C ABI);
  * ``CountingTropical``: TropicalNumbers' CountingTropical{Float64,Float64} (tropical value
    n with a count c), data complex128 with real part n and imaginary part c;
  * ``TropicalAndOr``: TropicalNumbers' boolean semiring (+ is or, * is and), data bool.
"""
import numpy as np

from . import _lib


class Wrapped:
    """A device tensor whose elements are not plain torch numbers; ``data`` holds them.
    Subclasses give codes() (torch dtype of data -> element type code) and ZERO / ONE (the
    fill values of data for the semiring's zero and one)."""

    ZERO = 0.0
    ONE = 1.0

    def __init__(self, data):
        import torch
        if isinstance(data, Wrapped):
            data = data.data
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(np.asarray(data))
        self.data = self._coerce(data)

    @staticmethod
    def _coerce(data):
        return data

    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def dtype(self):
        return self.data.dtype

    def cpu(self):
        return type(self)(self.data.cpu())

    def numpy(self):
        return self.data.detach().cpu().numpy()

    def __repr__(self):
        return "%s(%r)" % (type(self).__name__, self.data)


class Tropical(Wrapped):
    """A tensor of tropical numbers; ``data`` holds their contents (float32 or float64)."""

    ZERO = float("-inf")
    ONE = 0.0

    @staticmethod
    def _coerce(data):
        import torch
        return data if data.dtype in (torch.float32, torch.float64) else data.to(torch.float64)

    @classmethod
    def codes(cls):
        import torch
        return {torch.float32: _lib.TF32, torch.float64: _lib.TF64}


class Double64(Wrapped):
    """A tensor of double-word numbers (DoubleFloats' Double64): each value is hi + lo with
    |lo| <= ulp(hi) / 2.  ``data`` is complex128 with real part hi and imaginary part lo;
    torch's complex arithmetic on it is not double-word arithmetic, so use the hi / lo views
    and the contraction entry points.  Double64(hi) or Double64(hi, lo) builds one from
    float64 tensors."""

    def __init__(self, hi, lo=None):
        import torch
        if lo is None and isinstance(hi, (Double64, torch.Tensor)) and \
                (isinstance(hi, Double64) or hi.dtype == torch.complex128):
            super().__init__(hi)
            return
        if isinstance(hi, np.ndarray) and hi.dtype == np.complex128 and lo is None:
            super().__init__(torch.from_numpy(np.ascontiguousarray(hi)))
            return
        h = torch.as_tensor(hi, dtype=torch.float64)
        l = torch.zeros_like(h) if lo is None else torch.as_tensor(lo, dtype=torch.float64,
                                                                   device=h.device)
        super().__init__(torch.complex(h, l))

    @staticmethod
    def _coerce(data):
        import torch
        if data.dtype != torch.complex128:
            raise TypeError("Double64 data is complex128 (real part hi, imaginary part lo)")
        return data

    @classmethod
    def codes(cls):
        import torch
        return {torch.complex128: _lib.DD}

    @property
    def hi(self):
        return self.data.real

    @property
    def lo(self):
        return self.data.imag

    def to_float(self):
        """hi + lo rounded to float64."""
        return self.data.real + self.data.imag


class CountingTropical(Wrapped):
    """A tensor of CountingTropical{Float64,Float64} numbers: (n, c) with
    a * b = (a.n + b.n, a.c * b.c) and a + b = the larger n with its count (counts added on
    equal n).  ``data`` is complex128 with real part n and imaginary part c;
    CountingTropical(n, c) builds one from float64 tensors (c defaults to ones)."""

    ZERO = complex(float("-inf"), 0.0)
    ONE = complex(0.0, 1.0)

    def __init__(self, n, c=None):
        import torch
        if c is None and isinstance(n, (CountingTropical, torch.Tensor)) and \
                (isinstance(n, CountingTropical) or n.dtype == torch.complex128):
            super().__init__(n)
            return
        nt = torch.as_tensor(n, dtype=torch.float64)
        ct = torch.ones_like(nt) if c is None else torch.as_tensor(c, dtype=torch.float64,
                                                                   device=nt.device)
        super().__init__(torch.complex(nt, ct))

    @staticmethod
    def _coerce(data):
        import torch
        if data.dtype != torch.complex128:
            raise TypeError("CountingTropical data is complex128 (real part n, imaginary part c)")
        return data

    @classmethod
    def codes(cls):
        import torch
        return {torch.complex128: _lib.CT64}

    @property
    def n(self):
        return self.data.real

    @property
    def c(self):
        return self.data.imag


class TropicalAndOr(Wrapped):
    """A tensor of TropicalAndOr numbers (the boolean semiring: + is or, * is and);
    ``data`` is a bool tensor."""

    ZERO = False
    ONE = True

    @staticmethod
    def _coerce(data):
        import torch
        return data if data.dtype == torch.bool else data != 0

    @classmethod
    def codes(cls):
        import torch
        return {torch.bool: _lib.TBOOL}


def kind_of(x):
    """The wrapper class of x, or None for a plain tensor / array / number."""
    return type(x) if isinstance(x, Wrapped) else None


def wrap(kind, t):
    return kind(t) if kind is not None else t


def raw(x):
    """The tensor behind x (x itself when it is not wrapped)."""
    return x.data if isinstance(x, Wrapped) else x


def common_kind(xs):
    """The shared wrapper class of xs (None: all plain); mixing kinds is an error."""
    kinds = {kind_of(x) for x in xs}
    if len(kinds) > 1:
        names = sorted("plain" if k is None else k.__name__ for k in kinds)
        raise TypeError("cannot mix %s tensors in one contraction" % " and ".join(names))
    return kinds.pop() if kinds else None


def _torch():
    import torch
    return torch


_PLAIN_CODES = {}  # torch dtype -> C ABI element type of plain tensors


def dtype_code(t, kind=None):
    """C ABI element type of torch dtype t for plain tensors (kind None) or for the data of a
    wrapper class (a True kind means Tropical)."""
    torch = _torch()
    if kind is True:
        kind = Tropical
    if kind:
        try:
            return kind.codes()[t]
        except KeyError:
            raise TypeError("%s tensors of %s are not supported" % (kind.__name__, t))
    if not _PLAIN_CODES:
        _PLAIN_CODES.update({torch.float32: _lib.F32, torch.float64: _lib.F64,
                             torch.complex64: _lib.C32, torch.complex128: _lib.C64,
                             torch.int64: _lib.I64})
    try:
        return _PLAIN_CODES[t]
    except KeyError:
        raise TypeError("element type %s is not supported by the MI355X backend" % t)


_AS_IS_NAMES = ("float32", "float64", "complex64", "complex128", "int64", "bool")


class _DtypeSet:
    """torch dtypes passed through to_device unchanged (resolved on first use)."""

    def __init__(self):
        self.s = None

    def __contains__(self, dt):
        if self.s is None:
            import torch
            self.s = {getattr(torch, n) for n in _AS_IS_NAMES}
        return dt in self.s


_AS_IS = _DtypeSet()


def to_device(x, device=None):
    """numpy array / python scalar / torch tensor -> contiguous torch tensor on the GPU."""
    torch = _torch()
    if isinstance(x, Wrapped):
        return type(x)(to_device(x.data, device))
    if device is None:
        cur = torch.cuda.current_device()
        # already a contiguous tensor of a supported type on the current device: as it is
        if type(x) is torch.Tensor and x.get_device() == cur and x.dtype in _AS_IS and \
                x.is_contiguous():
            return x
        device = torch.device("cuda", cur)
    if isinstance(x, torch.Tensor):
        t = x
    else:
        a = np.asarray(x)
        if a.dtype == np.bool_:
            a = a.astype(np.int64)
        t = torch.from_numpy(np.ascontiguousarray(a))
    if t.dtype == torch.int32:
        t = t.to(torch.int64)
    if t.dtype == torch.float16 or t.dtype == torch.bfloat16:
        t = t.to(torch.float32)
    return t.to(device).contiguous()


def to_numpy(x):
    if isinstance(x, Wrapped):
        return x.numpy()
    return x.detach().cpu().numpy()


def asarray(x, like=None):
    """asarray (src/utils.jl:45-48): wrap a scalar into a 0-d device tensor."""
    torch = _torch()
    if isinstance(x, (torch.Tensor, Wrapped)):
        return x
    return to_device(np.asarray(x))


def asscalar(x):
    """asscalar (src/utils.jl:49-50): the single element of a 0-d tensor, as a Python number."""
    if isinstance(x, Wrapped):
        return x.data.reshape(()).item()
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.reshape(()).item()
    return x


def scalar(v, code):
    """Python number / bool -> ome_scalar for element type `code`.  Bools follow Julia:
    true is the element type's one and false its zero (tropical: 0 and -Inf)."""
    s = _lib.Scalar()
    if code == _lib.TBOOL:
        s.i = 1 if v else 0
        return s
    if code == _lib.CT64:
        if isinstance(v, (bool, np.bool_)):
            s.re, s.im = (0.0, 1.0) if v else (float("-inf"), 0.0)
        else:
            c = complex(v)
            s.re, s.im = c.real, c.imag
        return s
    if isinstance(v, (bool, np.bool_)):
        if code in (_lib.TF32, _lib.TF64):
            s.re = 0.0 if v else -np.inf
        elif code == _lib.I64:
            s.i = 1 if v else 0
        else:
            s.re = 1.0 if v else 0.0
        return s
    if code == _lib.I64:
        s.i = int(v)
    elif code == _lib.DD:
        s.re = float(v)
    elif code in (_lib.C32, _lib.C64):
        c = complex(v)
        s.re, s.im = c.real, c.imag
    else:
        s.re = float(np.real(v))
    return s
