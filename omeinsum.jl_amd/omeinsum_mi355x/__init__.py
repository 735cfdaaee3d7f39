"""omeinsum_mi355x — MI355X (gfx950) contraction backend with OMEinsum's interface.

The host-side surface mirrors OMEinsum (src/OMEinsum.jl:10-36): ``ein``, ``EinCode``,
``NestedEinsum``, ``SlicedEinsum``, ``einsum``, ``einsum_inplace`` (einsum!), ``loop_einsum``,
``set_einsum_backend``/``get_einsum_backend``, ``cost_and_gradient`` and the
contraction-order helpers.  Every contraction runs on the GPU through the C ABI of
``libomeinsum_mi355x.so`` (include/omeinsum_mi355x.h); there is no CPU fallback.
"""
from .codes import (ArgumentError, DimensionMismatch, EinCode, DynamicEinCode, StaticEinCode,
                    NestedEinsum, DynamicNestedEinsum, StaticNestedEinsum, SlicedEinsum, ein,
                    parse_nested, get_size_dict, getixsv, getiyv, uniquelabels, labeltype,
                    flatten, uniformsize, drop_slicedim)
from .contractionorder import (GreedyMethod, TreeSA, ScoreFunction, TreeSASlicer, optimize_code,
                               slice_code, flop, peak_memory, contraction_complexity, writejson,
                               readjson)

_LAZY = {
    "einsum": "ein_ops", "einsum_inplace": "ein_ops", "loop_einsum": "ein_ops",
    "loop_einsum_inplace": "ein_ops", "allow_loops": "ein_ops", "tensorpermute": "ein_ops",
    "batched_gemm": "ein_ops", "grouped_gemm": "ein_ops", "grouped_permute": "ein_ops",
    "GroupedGemm": "ein_ops", "GroupedPermute": "ein_ops", "asarray": "device",
    "asscalar": "device", "Tropical": "device", "Double64": "device", "Wrapped": "device",
    "CountingTropical": "device", "TropicalAndOr": "device", "to_device": "device", "to_numpy": "device",
    "EinsumBackend": "backends", "DefaultBackend": "backends", "MI355XBackend": "backends",
    "set_einsum_backend": "backends", "get_einsum_backend": "backends",
    "einsum_grad": "autodiff", "cost_and_gradient": "autodiff", "einsum_backward_rule": "autodiff",
    "einsum_autograd": "autodiff",
    "Server": "server", "library": "_lib",
}


def __getattr__(name):
    mod = _LAZY.get(name)
    if mod is None:
        raise AttributeError(name)
    import importlib
    m = importlib.import_module("." + mod, __name__)
    value = getattr(m, name)
    globals()[name] = value  # later lookups are plain attribute reads
    return value
