"""ctypes binding of libomeinsum_mi355x.so (include/omeinsum_mi355x.h).

The library is the only compute path: when it is missing, or when no gfx950 device is
present, every call raises — there is no CPU fallback.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libomeinsum_mi355x.so")

OME_OK, OME_ERR_INVALID_ARG, OME_ERR_DIM_MISMATCH, OME_ERR_OOM = 0, 1, 2, 3
OME_ERR_LOOPS_FORBIDDEN, OME_ERR_DEVICE, OME_ERR_UNSUPPORTED = 4, 5, 6
F32, F64, C32, C64, I64, TF32, TF64, DD, CT64, TBOOL = range(10)
OPT_PRECISION, OPT_COMPLEX_ALGO, OPT_ALLOW_LOOPS, OPT_DIRECT_GEMM = 0, 1, 2, 3
OPT_USE_GRAPHS, OPT_AUTOTUNE, OPT_GEMM_KERNEL, OPT_KERNEL_TIMING = 4, 5, 6, 7
OPT_GROUPED_VARIANT, OPT_LAST_GROUPED_CHOICE, OPT_TREE_DEFER_MEMORY = 8, 9, 10
OPT_GROUPED_BUFFERS = 11
OPT_TRACE_RANGES = 12
# the options a tree plan freezes at creation
PLAN_OPTIONS = (OPT_PRECISION, OPT_COMPLEX_ALGO, OPT_ALLOW_LOOPS, OPT_DIRECT_GEMM, OPT_GEMM_KERNEL)
# kernel classes of ome_get_kernel_timing
KC = {"gemm_packed": 0, "pack": 1, "gemm_mfma": 2, "gemm_valu": 3, "permute": 4, "walk": 5,
      "grouped": 6, "other": 7}
PREC = {"fp32": 0, "bf16x3": 1, "fp64": 2, "bf16x6": 3}

c_int, c_i32, c_i64, c_void_p, c_size_t, c_double, c_char = (
    ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t,
    ctypes.c_double, ctypes.c_char)
P = ctypes.POINTER


class Scalar(ctypes.Structure):
    _fields_ = [("re", c_double), ("im", c_double), ("i", c_i64)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", c_char * 64), ("arch", c_char * 32), ("compute_units", c_i32),
                ("xcds", c_i32), ("total_mem", c_i64), ("lds_per_cu", c_i64),
                ("clock_mhz", c_i32), ("device", c_i32)]


class PoolStats(ctypes.Structure):
    _fields_ = [("reserved_bytes", c_i64), ("in_use_bytes", c_i64),
                ("peak_in_use_bytes", c_i64), ("n_allocs", c_i64), ("n_device_mallocs", c_i64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", c_i64), ("completed", c_i64), ("sliced", c_i64),
                ("blocks_reused", c_i64), ("kept_bytes", c_i64), ("batched_launches", c_i64),
                ("batched_requests", c_i64)]


class TreeInfo(ctypes.Structure):
    _fields_ = [("n_steps", c_i32), ("n_leaves", c_i32), ("n_slices", c_i64),
                ("flops", c_double), ("bytes", c_double), ("arena_bytes", c_i64),
                ("peak_live_elems", c_i64), ("n_kernels", c_i64)]


# name -> (restype, argtypes); every declared symbol of the header
SIGNATURES = {
    "ome_abi_version": (c_int, []),
    "ome_init": (c_int, [c_int, c_size_t, c_void_p, P(c_void_p)]),
    "ome_shutdown": (c_int, [c_void_p]),
    "ome_set_stream": (c_int, [c_void_p, c_void_p]),
    "ome_last_error": (ctypes.c_char_p, [c_void_p]),
    "ome_set_option": (c_int, [c_void_p, c_int, c_i64]),
    "ome_get_option": (c_int, [c_void_p, c_int, P(c_i64)]),
    "ome_get_device_info": (c_int, [c_void_p, P(DeviceInfo)]),
    "ome_synchronize": (c_int, [c_void_p]),
    "ome_get_kernel_timing": (c_int, [c_void_p, c_int, P(c_double), P(c_i64)]),
    "ome_alloc": (c_int, [c_void_p, c_size_t, P(c_void_p)]),
    "ome_free": (c_int, [c_void_p, c_void_p]),
    "ome_get_pool_stats": (c_int, [c_void_p, P(PoolStats)]),
    "ome_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "ome_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "ome_memcpy_d2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "ome_einsum": (c_int, [c_void_p, c_int, c_int, P(c_void_p), P(c_i32), P(c_i32), c_void_p,
                           c_int, P(c_i32), c_int, P(c_i64), P(Scalar), P(Scalar)]),
    "ome_einsum_noalloc": (c_int, [c_void_p, c_int, c_int, P(c_void_p), P(c_i32), P(c_i32),
                                   c_void_p, c_int, P(c_i32), c_int, P(c_i64), P(Scalar),
                                   P(Scalar)]),
    "ome_loop_einsum": (c_int, [c_void_p, c_int, c_int, P(c_void_p), P(c_i32), P(c_i32),
                                c_void_p, c_int, P(c_i32), c_int, P(c_i64), P(Scalar),
                                P(Scalar)]),
    "ome_permute": (c_int, [c_void_p, c_int, c_int, P(c_i64), P(c_i32), c_void_p, c_void_p,
                            P(Scalar), P(Scalar)]),
    "ome_addmul": (c_int, [c_void_p, c_int, c_i64, c_void_p, c_void_p, P(Scalar), P(Scalar),
                           c_int]),
    "ome_batched_gemm": (c_int, [c_void_p, c_int, c_char, c_char, c_i64, c_i64, c_i64, c_i64,
                                 P(Scalar), c_void_p, c_void_p, P(Scalar), c_void_p]),
    "ome_grouped_gemm": (c_int, [c_void_p, c_int, c_int, P(c_i64), P(c_void_p), P(c_void_p),
                                 P(c_void_p), P(Scalar), P(Scalar)]),
    "ome_grouped_permute": (c_int, [c_void_p, c_int, c_int, P(c_i32), P(c_i64), P(c_i32),
                                    P(c_void_p), P(c_void_p), P(Scalar), P(Scalar)]),
    "ome_tree_create": (c_int, [c_void_p, c_int, P(c_i32), c_i64, c_int, P(c_i64), c_int,
                                P(c_i32), P(c_void_p)]),
    "ome_get_tree_info": (c_int, [c_void_p, P(TreeInfo)]),
    "ome_tree_execute": (c_int, [c_void_p, P(c_void_p), c_void_p, P(Scalar), P(Scalar)]),
    "ome_tree_execute_scalar_normalized": (c_int, [c_void_p, P(c_void_p), P(c_double),
                                                   P(c_double)]),
    "ome_tree_gradient": (c_int, [c_void_p, P(c_void_p), c_void_p, c_void_p, P(c_void_p)]),
    "ome_tree_destroy": (c_int, [c_void_p]),
    "ome_server_create": (c_int, [c_void_p, c_size_t, c_int, P(c_void_p)]),
    "ome_server_submit": (c_int, [c_void_p, c_void_p, P(c_void_p), c_void_p, P(c_i64)]),
    "ome_server_poll": (c_int, [c_void_p, c_i64, P(c_int)]),
    "ome_server_wait": (c_int, [c_void_p, c_i64, P(c_double)]),
    "ome_server_destroy": (c_int, [c_void_p]),
    "ome_server_set_request_limit": (c_int, [c_void_p, c_size_t]),
    "ome_server_get_stats": (c_int, [c_void_p, P(ServerStats)]),
    "ome_server_set_batching": (c_int, [c_void_p, c_int]),
    "ome_server_hold": (c_int, [c_void_p, c_int]),
    "ome_server_inject_status": (c_int, [c_void_p, c_int]),
}

_lib = None
_lock = threading.RLock()  # context() may call load() while holding it


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "libomeinsum_mi355x.so is not built (run `python omeinsum.jl_amd/build.py` or "
                    "__graft_entry__.build()); the MI355X backend has no CPU fallback")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(lib, name)
                f.restype = res
                f.argtypes = args
            _lib = lib
    return _lib


class OMEError(RuntimeError):
    pass


class LoopsForbidden(OMEError):
    """Julia's ErrorException from a forbidden loop_einsum (src/loop_einsum.jl:24)."""


def check(rc, ctx=None):
    if rc == OME_OK:
        return
    from .codes import ArgumentError, DimensionMismatch
    msg = (load().ome_last_error(ctx) or b"").decode(errors="replace")
    if rc == OME_ERR_INVALID_ARG:
        raise ArgumentError(msg)
    if rc == OME_ERR_DIM_MISMATCH:
        raise DimensionMismatch(msg)
    if rc == OME_ERR_OOM:
        raise MemoryError(msg)
    if rc == OME_ERR_LOOPS_FORBIDDEN:
        raise LoopsForbidden(msg)
    raise OMEError("omeinsum_mi355x error %d: %s" % (rc, msg))


class Context:
    """One library context per device, bound to torch's current stream of that device."""

    def __init__(self, device=0, pool_bytes=0):
        import torch
        self.lib = load()
        self.device = device
        self.stream = torch.cuda.current_stream(device).cuda_stream
        h = c_void_p()
        with torch.cuda.device(device):
            check(self.lib.ome_init(device, pool_bytes, c_void_p(self.stream), ctypes.byref(h)))
        self.h = h
        self._opts = {}  # option values set through this object (skips repeated calls)

    def sync_stream(self):
        """Follow torch's current stream (the pool reuses blocks in stream order)."""
        import torch
        raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        s = raw_stream(self.device) if raw_stream is not None else \
            torch.cuda.current_stream(self.device).cuda_stream
        if s != self.stream:
            check(self.lib.ome_set_stream(self.h, c_void_p(s)), self.h)
            self.stream = s

    def set_option(self, key, value):
        value = int(value)
        if self._opts.get(key) == value:
            return
        self._opts.pop(key, None)
        check(self.lib.ome_set_option(self.h, key, value), self.h)
        self._opts[key] = value

    def get_option(self, key):
        if key in self._opts:  # set through this object: the library holds that value
            return self._opts[key]
        v = c_i64()
        check(self.lib.ome_get_option(self.h, key, ctypes.byref(v)), self.h)
        return v.value

    def device_info(self):
        info = DeviceInfo()
        check(self.lib.ome_get_device_info(self.h, ctypes.byref(info)), self.h)
        return info

    def pool_stats(self):
        s = PoolStats()
        check(self.lib.ome_get_pool_stats(self.h, ctypes.byref(s)), self.h)
        return s

    def synchronize(self):
        check(self.lib.ome_synchronize(self.h), self.h)

    def kernel_timing(self, kclass):
        """(total milliseconds, launches) of the timed launches of one kernel class since the
        last call (OPT_KERNEL_TIMING must be on while they ran)."""
        ms, n = c_double(), c_i64()
        check(self.lib.ome_get_kernel_timing(self.h, KC[kclass], ctypes.byref(ms),
                                             ctypes.byref(n)), self.h)
        return ms.value, n.value


_contexts = {}


def context(device=None):
    import torch
    if device is None:
        if not _contexts and not torch.cuda.is_available():
            raise RuntimeError("omeinsum_mi355x needs an MI355X (gfx950) GPU; none is visible")
        device = torch.cuda.current_device()
    ctx = _contexts.get(device)
    if ctx is None:
        if not torch.cuda.is_available():
            raise RuntimeError("omeinsum_mi355x needs an MI355X (gfx950) GPU; none is visible")
        with _lock:
            ctx = _contexts.get(device)
            if ctx is None:
                ctx = Context(device)
                _contexts[device] = ctx
    ctx.sync_stream()
    return ctx


def library():
    return load()
