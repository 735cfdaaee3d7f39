"""Backend selection (mirrors src/backends.jl:24-233).

* ``DefaultBackend`` — the reference's per-step structure on the device: one flat einsum per
  tree node, evaluated post-order from Python with a fresh output per node
  (src/einsequence.jl:279-287, ext/AMDGPUExt.jl:130-142), slices looped from Python
  (src/slicing.jl:123-137).  Every step still runs on this library's kernels; this is the
  "plain per-step GPU path" the MI355X backend is measured against.
* ``MI355XBackend`` — contraction trees are compiled once into a device plan: intermediates
  in one arena laid out by a static liveness plan, the step sequence recorded into a
  hipGraph and replayed, sliced trees summed in FP64 on the device.  Float32/ComplexF32
  GEMM-shaped steps use ``precision``: "fp32" (FP32 MFMA), "bf16x3" (three products of bf16
  hi/lo planes), "bf16x6" (six products of three bf16 planes: FP32-level error; large
  products only, the rest at FP32) or "fp64" (widened).
"""
from . import _lib
from . import ein_ops as E
from .codes import EinCode, NestedEinsum, SlicedEinsum, drop_slicedim, get_size_dict


class EinsumBackend:
    """Abstract backend (src/backends.jl:24)."""

    def apply(self, ctx):
        pass


class DefaultBackend(EinsumBackend):
    """src/backends.jl:53: per-step evaluation."""

    def __repr__(self):
        return "DefaultBackend()"

    def __eq__(self, other):
        return isinstance(other, DefaultBackend)

    def apply(self, ctx):
        ctx.set_option(_lib.OPT_PRECISION, 0)

    # -- flat codes
    def einsum(self, code, xs, size_dict, **kw):
        self.apply(_lib.context())
        if isinstance(code, EinCode):
            return E.flat_einsum(code, xs, size_dict)
        if isinstance(code, SlicedEinsum):
            return self._sliced(code, xs, size_dict)
        return self._nested(code, xs, size_dict)

    def einsum_inplace(self, code, xs, y, sx, sy, size_dict):
        self.apply(_lib.context())
        if isinstance(code, EinCode):
            return E._flat_einsum_inplace(code.ixs, code.iy, xs, y, sx, sy, size_dict)
        if isinstance(code, SlicedEinsum):
            res = self._sliced(code, xs, size_dict)
            return E._flat_einsum_inplace([code.getiyv()], code.getiyv(), (res,), y, sx, sy, size_dict)
        mxs = [xs[a.tensorindex] if a.isleaf() else self._nested(a, xs, size_dict)
               for a in code.args]
        return E._flat_einsum_inplace(code.eins.ixs, code.eins.iy, tuple(mxs), y, sx, sy, size_dict)

    def _nested(self, ne, xs, size_dict):
        """einsum(::NestedEinsum, xs, size_dict) (src/einsequence.jl:279-287)."""
        mxs = [xs[a.tensorindex] if a.isleaf() else self._nested(a, xs, size_dict) for a in ne.args]
        return E.flat_einsum(ne.eins, tuple(mxs), size_dict)

    def _sliced(self, se, xs, size_dict):
        """einsum(::SlicedEinsum, xs, size_dict) (src/slicing.jl:123-137): zero output, loop
        the slices column-major over the sliced labels, take_slice, evaluate, fill_slice!."""
        import itertools
        from .device import kind_of, raw, wrap
        if not se.slicing:
            return self.einsum(se.eins, xs, size_dict)
        ixsv = se.getixsv()
        iyv = se.getiyv()
        res = E._new_output(xs, [size_dict[l] for l in iyv])
        rt = raw(res)
        kind = kind_of(res)
        rt.fill_(kind.ZERO if kind else 0)
        sd = dict(size_dict)
        for l in se.slicing:
            sd[l] = 1
        inner = drop_slicedim(se.eins, se.slicing)
        ranges = [range(size_dict[l]) for l in se.slicing]
        for combo in itertools.product(*reversed(ranges)):
            sm = dict(zip(reversed(se.slicing), combo))
            xsi = []
            for x, ix in zip(xs, ixsv):
                xt = E.to_device(raw(x))
                idx = tuple(sm[l] if l in sm else slice(None) for l in ix)
                xi = xt[idx].contiguous()
                xsi.append(wrap(kind, xi))
            resi = self.einsum(inner, tuple(xsi), sd)
            ri = raw(resi)
            idx = tuple(sm[l] if l in sm else slice(None) for l in iyv)
            if kind is None:
                rt[idx] += ri
            else:
                # the semiring's own sum: y <- one*y + one*slice through the library
                acc = rt[idx].contiguous()
                E.tensorpermute(kind(acc), kind(ri), list(range(ri.dim())), True, True)
                rt[idx] = acc
        return res


class MI355XBackend(EinsumBackend):
    """The MI355X backend: device tree plans with graph replay (see module docstring)."""

    def __init__(self, precision="fp32", complex_algo="4m", graphs=True, direct_gemm=True):
        if precision not in _lib.PREC:
            raise ValueError("precision must be one of %s" % list(_lib.PREC))
        if complex_algo not in ("4m", "3m"):
            raise ValueError("complex_algo must be '4m' or '3m'")
        self.precision = precision
        self.complex_algo = complex_algo
        self.graphs = graphs
        self.direct_gemm = direct_gemm

    def __repr__(self):
        return "MI355XBackend(precision=%r, complex_algo=%r, graphs=%r)" % (
            self.precision, self.complex_algo, self.graphs)

    def __eq__(self, other):
        return isinstance(other, MI355XBackend) and vars(self) == vars(other)

    def apply(self, ctx):
        ctx.set_option(_lib.OPT_PRECISION, _lib.PREC[self.precision])
        ctx.set_option(_lib.OPT_COMPLEX_ALGO, 1 if self.complex_algo == "3m" else 0)
        ctx.set_option(_lib.OPT_USE_GRAPHS, 1 if self.graphs else 0)
        ctx.set_option(_lib.OPT_DIRECT_GEMM, 1 if self.direct_gemm else 0)

    def einsum(self, code, xs, size_dict, **kw):
        self.apply(_lib.context())
        if isinstance(code, EinCode):
            return E.flat_einsum(code, xs, size_dict)
        from .tree import plan_for
        plan, ts = plan_for(code, xs, size_dict, with_tensors=True)
        return plan.execute(xs, ts=ts)

    def einsum_inplace(self, code, xs, y, sx, sy, size_dict):
        self.apply(_lib.context())
        if isinstance(code, EinCode):
            return E._flat_einsum_inplace(code.ixs, code.iy, xs, y, sx, sy, size_dict)
        from .tree import plan_for
        plan, ts = plan_for(code, xs, size_dict, with_tensors=True)
        return plan.execute(xs, y=y, sx=sx, sy=sy, ts=ts)


_BACKEND = [DefaultBackend()]


def check_backend_availability(backend):
    """(available, message) (src/backends.jl:170,186-191)."""
    if isinstance(backend, MI355XBackend):
        try:
            ctx = _lib.context()
            arch = ctx.device_info().arch.decode()
            return (arch.startswith("gfx950"), "" if arch.startswith("gfx950") else arch)
        except Exception as e:  # noqa: BLE001 - report why the backend is unavailable
            return (False, str(e))
    return (True, "")


def set_einsum_backend(backend):
    """set_einsum_backend!(backend) (src/backends.jl:164-167,193-197)."""
    if not isinstance(backend, EinsumBackend):
        raise TypeError("expected an EinsumBackend")
    if isinstance(backend, MI355XBackend):
        ok, msg = check_backend_availability(backend)
        if not ok:
            raise RuntimeError("MI355XBackend unavailable: %s" % msg)
    _BACKEND[0] = backend
    return backend


def get_einsum_backend():
    """get_einsum_backend() (src/backends.jl:219)."""
    return _BACKEND[0]
