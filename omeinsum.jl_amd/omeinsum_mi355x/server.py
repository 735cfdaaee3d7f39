"""Continuous-batching request loop over ome_server_* (no counterpart in OMEinsum).

    srv = Server(hbm_budget=8 << 30, max_active=8, backend=MI355XBackend())
    plan = srv.plan(tree, xs, size_dict)      # device plan, memory sized but not allocated
    t = srv.submit(plan, xs)                  # returns at once; inputs stay on the device
    y, latency_ms = srv.wait(t)

A native scheduler thread admits queued requests, in arrival order, whenever an execution
slot is free and the plan's device memory fits the remaining budget; each slot runs on its
own stream, so admitted requests overlap on the device and requests that arrive later are
admitted as earlier ones retire.  A plan above ``request_limit`` bytes (default: the budget)
is sliced at submission until it fits; a plan lent the memory block it used before replays
its recorded graph; requests admitted together run their fused small-step subtrees in one
shared launch (``batching``, on by default).  Latency is measured from submission to the
moment the scheduler sees the request's completion event.
"""
import ctypes

from . import _lib
from .backends import MI355XBackend
from .device import common_kind, raw
from .tree import TreePlan, element_code


class Server:
    def __init__(self, hbm_budget=8 << 30, max_active=8, backend=None, device=None,
                 request_limit=None, batching=True):
        import torch
        self.ctx = _lib.context(device)
        self.backend = backend if backend is not None else MI355XBackend()
        h = ctypes.c_void_p()
        _lib.check(self.ctx.lib.ome_server_create(self.ctx.h, int(hbm_budget), int(max_active),
                                                  ctypes.byref(h)), self.ctx.h)
        self.h = h
        self.budget = int(hbm_budget)
        if not batching:
            _lib.check(self.ctx.lib.ome_server_set_batching(self.h, 0), self.ctx.h)
        if request_limit is not None:
            _lib.check(self.ctx.lib.ome_server_set_request_limit(self.h, int(request_limit)),
                       self.ctx.h)
        self._live = {}
        self._torch = torch

    def plan(self, code, xs, size_dict):
        """A device plan of `code` (NestedEinsum or SlicedEinsum) whose memory the loop lends
        at admission."""
        self.backend.apply(self.ctx)
        code_dt, _ = element_code(xs)
        return TreePlan(code, size_dict, code_dt, self.ctx.device, defer_memory=True)

    def submit(self, plan, xs, y=None):
        """Queue one evaluation of `plan` on the device tensors xs; returns a ticket.  The
        output (allocated when y is None) and the inputs stay referenced until wait()."""
        self.ctx.sync_stream()
        ts, ptrs = plan._ptrs(xs)
        if y is None:
            res, yt = plan._out(ts, common_kind(xs))
        else:
            res, yt = y, (raw(y))
        t = ctypes.c_int64()
        _lib.check(self.ctx.lib.ome_server_submit(self.h, plan.h, ptrs,
                                                  ctypes.c_void_p(yt.data_ptr()),
                                                  ctypes.byref(t)), self.ctx.h)
        self._live[t.value] = (plan, ts, res)
        return t.value

    def inject_status(self, code):
        """Failure-handling tests: the next submitted request writes `code` into its slot's
        device status word after its steps, and then fails alone (wait() raises)."""
        _lib.check(self.ctx.lib.ome_server_inject_status(self.h, int(code)), self.ctx.h)

    def poll(self, ticket):
        """0 queued, 1 running, 2 done, < 0 failed."""
        st = ctypes.c_int()
        _lib.check(self.ctx.lib.ome_server_poll(self.h, int(ticket), ctypes.byref(st)), self.ctx.h)
        return st.value

    def wait(self, ticket):
        """Block until the request is done: (output, latency in milliseconds)."""
        lat = ctypes.c_double()
        try:
            _lib.check(self.ctx.lib.ome_server_wait(self.h, int(ticket), ctypes.byref(lat)),
                       self.ctx.h)
        finally:
            entry = self._live.pop(ticket, None)
        return entry[2], lat.value

    def hold(self, on=True):
        """Hold admission (on) or release it (off): requests submitted while held queue and
        are admitted together, in one round, on release.  Release before waiting on them."""
        _lib.check(self.ctx.lib.ome_server_hold(self.h, 1 if on else 0), self.ctx.h)

    def stats(self):
        """Counters of the loop: submitted, completed, sliced (admitted as a sliced twin),
        blocks_reused (lent the block the same plan used last), kept_bytes, batched_launches
        and batched_requests (fused small steps of several requests run in one launch)."""
        st = _lib.ServerStats()
        _lib.check(self.ctx.lib.ome_server_get_stats(self.h, ctypes.byref(st)), self.ctx.h)
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.ome_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def plan_bytes(plan):
    """Device bytes the loop lends to a plan at admission."""
    return int(plan.info().arena_bytes)

