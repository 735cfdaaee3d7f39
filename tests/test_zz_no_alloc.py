"""ome_einsum_noalloc through the C ABI: the same results as ome_einsum (bit for bit: the same
plan and kernels, the temporaries in the library's fixed scratch array instead of the pool),
on two calls in a row with the second on smaller inputs, with no pool allocation; and a
contraction whose temporaries exceed the array fails with OME_ERR_OOM."""
import ctypes

import numpy as np
import pytest

from helpers import assert_close, parse, rand, rng, wide
from oracle import default_backend as db

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    yield


def call(entry, spec, xs, sd, y0=None, sx=True, sy=False):
    """One flat einsum through the named C entry point; returns y as numpy."""
    from omeinsum_mi355x import _lib
    from omeinsum_mi355x.device import dtype_code
    from omeinsum_mi355x.ein_ops import _flat_args, scalar
    ixs, iy = parse(spec)
    ts = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in xs]
    if y0 is None:
        yt = torch.empty([sd[l] for l in iy], dtype=ts[0].dtype, device="cuda")
    else:
        yt = torch.from_numpy(np.ascontiguousarray(y0)).cuda()
    ctx = _lib.context(yt.device.index)
    nix, cixs, ciy, nlabels, sizes = _flat_args(ixs, iy, sd)
    ptrs = (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])
    code = dtype_code(yt.dtype, None)
    s_x, s_y = scalar(sx, code), scalar(sy, code)
    ctx.sync_stream()
    fn = getattr(ctx.lib, entry)
    _lib.check(fn(ctx.h, code, len(ts), ptrs, nix, cixs, ctypes.c_void_p(yt.data_ptr()), len(iy),
                  ciy, nlabels, sizes, ctypes.byref(s_x), ctypes.byref(s_y)), ctx.h)
    torch.cuda.synchronize()
    return yt.cpu().numpy()


# contractions whose plans use temporaries: a permuted operand (K labels in different orders),
# split-K partials (few outputs, long K), the complex four-real-products route
CASES = [("ijk,jil->kl", np.float32, {"i": 96, "j": 80, "k": 48, "l": 40},
          {"i": 40, "j": 24, "k": 16, "l": 8}),
         ("ijk,jkl->il", np.float64, {"i": 64, "j": 64, "k": 64, "l": 64},
          {"i": 24, "j": 32, "k": 16, "l": 20}),
         ("ajb,bkj->bka", np.complex64, {"a": 130, "j": 70, "b": 3, "k": 66},
          {"a": 33, "j": 20, "b": 2, "k": 40})]


@pytest.mark.parametrize("spec,dtype,big,small", CASES)
def test_noalloc_matches_einsum_twice(spec, dtype, big, small):
    from omeinsum_mi355x import _lib
    ctx = _lib.context()
    r = rng(7700)
    ixs, iy = parse(spec)
    for sd in (big, small):  # two calls in a row, the second on smaller inputs
        xs = [rand(r, [sd[l] for l in ix], dtype) for ix in ixs]
        y0 = rand(r, [sd[l] for l in iy], dtype)
        ref_alloc = call("ome_einsum", spec, xs, sd)
        before = ctx.pool_stats()
        got = call("ome_einsum_noalloc", spec, xs, sd)
        got_scaled = call("ome_einsum_noalloc", spec, xs, sd, y0=y0, sx=2, sy=0.5)
        after = ctx.pool_stats()
        assert after.n_allocs == before.n_allocs and after.n_device_mallocs == before.n_device_mallocs
        assert np.array_equal(got, ref_alloc), (spec, sd)
        ref_scaled = call("ome_einsum", spec, xs, sd, y0=y0, sx=2, sy=0.5)
        assert np.array_equal(got_scaled, ref_scaled), (spec, sd)
        k = int(np.prod([sd[l] for l in set(ixs[0]) & set(ixs[1]) - set(iy)]))
        W = wide(dtype)
        ref64 = db.einsum(ixs, iy, [x.astype(W) for x in xs], sd)
        assert_close(got, ref64, dtype, k, spec, depth=k)


def test_noalloc_rejects_temporaries_above_the_array():
    # ijk,jil->kl at i = j = 512, k = l = 64: a permuted Float32 copy of one 64 MiB operand
    # does not fit the 64 MiB array; ome_einsum (pool memory) runs it
    from omeinsum_mi355x import _lib
    sd = {"i": 512, "j": 512, "k": 64, "l": 64}
    r = rng(7701)
    ixs, iy = parse("ijk,jil->kl")
    xs = [rand(r, [sd[l] for l in ix], np.float32) for ix in ixs]
    y0 = np.full((64, 64), 7.0, np.float32)
    with pytest.raises(MemoryError, match="scratch array"):
        call("ome_einsum_noalloc", "ijk,jil->kl", xs, sd, y0=y0)
    got = call("ome_einsum", "ijk,jil->kl", xs, sd)
    ref64 = db.einsum(ixs, iy, [x.astype(np.float64) for x in xs], sd)
    assert_close(got, ref64, np.float32, 512 * 512, depth=512 * 512)
    assert _lib.OME_ERR_OOM == 3
