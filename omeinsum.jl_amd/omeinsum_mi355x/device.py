"""Device tensors and element types.

Tensors are torch tensors on the GPU (torch is used for device memory and streams only).  A
row-major (C-contiguous) tensor with axis labels (l0, ..., l{r-1}) occupies exactly the
memory of a column-major tensor with labels (l{r-1}, ..., l0), so calls into the column-major
C ABI pass every label list reversed.  ``Tropical`` wraps a float tensor whose entries are
tropical numbers (TropicalNumbers' Tropical{T}: + is max, * is +).
"""
import numpy as np

from . import _lib


class Tropical:
    """A tensor of tropical numbers; ``data`` holds their contents (float32 or float64)."""

    def __init__(self, data):
        import torch
        if isinstance(data, Tropical):
            data = data.data
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(np.asarray(data))
        if data.dtype not in (torch.float32, torch.float64):
            data = data.to(torch.float64)
        self.data = data

    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def dtype(self):
        return self.data.dtype

    def cpu(self):
        return Tropical(self.data.cpu())

    def numpy(self):
        return self.data.detach().cpu().numpy()

    def __repr__(self):
        return "Tropical(%r)" % (self.data,)


def _torch():
    import torch
    return torch


def dtype_code(t, tropical=False):
    torch = _torch()
    if tropical:
        return {torch.float32: _lib.TF32, torch.float64: _lib.TF64}[t]
    try:
        return {torch.float32: _lib.F32, torch.float64: _lib.F64, torch.complex64: _lib.C32,
                torch.complex128: _lib.C64, torch.int64: _lib.I64}[t]
    except KeyError:
        raise TypeError("element type %s is not supported by the MI355X backend" % t)


def to_device(x, device=None):
    """numpy array / python scalar / torch tensor -> contiguous torch tensor on the GPU."""
    torch = _torch()
    if isinstance(x, Tropical):
        return Tropical(to_device(x.data, device))
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
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
    if isinstance(x, Tropical):
        return x.numpy()
    return x.detach().cpu().numpy()


def asarray(x, like=None):
    """asarray (src/utils.jl:45-48): wrap a scalar into a 0-d device tensor."""
    torch = _torch()
    if isinstance(x, (torch.Tensor, Tropical)):
        return x
    return to_device(np.asarray(x))


def asscalar(x):
    """asscalar (src/utils.jl:49-50): the single element of a 0-d tensor, as a Python number."""
    if isinstance(x, Tropical):
        return x.data.reshape(()).item()
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.reshape(()).item()
    return x


def scalar(v, code):
    """Python number / bool -> ome_scalar for element type `code`.  Bools follow Julia:
    true is the element type's one and false its zero (tropical: 0 and -Inf)."""
    s = _lib.Scalar()
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
    elif code in (_lib.C32, _lib.C64):
        c = complex(v)
        s.re, s.im = c.real, c.imag
    else:
        s.re = float(np.real(v))
    return s
