"""Synthetic inputs of the batched-uneven workload on the device.

``uneven_batch`` builds 4096 instances of ein"ijb,jkb->ikb" (i, j, k uniform in
{8, 16, ..., 256}, b uniform in 1..16, Float32 uniform in [-1, 1)) and 1024 ein"ijkl->ljik"
permutes (dims uniform in [4, 64]).  The shapes come from one PCG64 stream; the values come
from one device generator filling a single buffer, which every operand is a view of (each
view starts on a 256-byte boundary, like a fresh allocation).  Building the inputs therefore
takes a handful of kernel dispatches, not one per operand.
"""
import numpy as np

_ALIGN = 64  # elements of 4 bytes: 256-byte aligned views


def _round_up(x, a=_ALIGN):
    return (x + a - 1) // a * a


def uneven_batch(seed=3000, n_gemm=4096, n_perm=1024, device="cuda"):
    """Returns (gemms, perms, info):
    gemms: list of (Bt, At, Ct) torch views; column-major A[i,j,b], B[j,k,b], C[i,k,b] are
           row-major At (b, j, i), Bt (b, k, j), Ct (b, k, i), so C_t = Bt @ At in torch's
           batch-major convention (the order GroupedGemm takes);
    perms: list of (x, y) views; column-major x[i,j,k,l] -> y[l,j,i,k] is row-major
           x (l, k, j, i) -> y (k, i, j, l), i.e. y = x.permute(1, 3, 2, 0);
    info:  dict of gemm_flop, gemm_bytes, permute_bytes (algorithmic, each operand once)."""
    import torch
    r = np.random.Generator(np.random.PCG64(seed))
    dims8 = np.arange(8, 257, 8)
    gshapes = []
    for _ in range(n_gemm):
        i, j, k = (int(x) for x in r.choice(dims8, 3))
        b = int(r.integers(1, 17))
        gshapes.append((i, j, k, b))
    pshapes = [tuple(int(x) for x in r.integers(4, 65, 4)) for _ in range(n_perm)]
    n_in = sum(_round_up(b * j * i) + _round_up(b * k * j) for i, j, k, b in gshapes) + \
        sum(_round_up(i * j * k * l) for i, j, k, l in pshapes)
    n_out = sum(_round_up(b * k * i) for i, j, k, b in gshapes) + \
        sum(_round_up(i * j * k * l) for i, j, k, l in pshapes)
    g = torch.Generator(device=device).manual_seed(seed + 1)
    src = torch.empty(n_in, device=device).uniform_(-1.0, 1.0, generator=g)
    dst = torch.zeros(n_out, device=device)
    oi = oo = 0
    gemms, perms = [], []
    flop = gbytes = pbytes = 0.0
    for i, j, k, b in gshapes:
        At = src[oi:oi + b * j * i].view(b, j, i)
        oi += _round_up(b * j * i)
        Bt = src[oi:oi + b * k * j].view(b, k, j)
        oi += _round_up(b * k * j)
        Ct = dst[oo:oo + b * k * i].view(b, k, i)
        oo += _round_up(b * k * i)
        gemms.append((Bt, At, Ct))
        flop += 2.0 * i * j * k * b
        gbytes += 4.0 * b * (i * j + j * k + i * k)
    for i, j, k, l in pshapes:
        n = i * j * k * l
        x = src[oi:oi + n].view(l, k, j, i)
        oi += _round_up(n)
        y = dst[oo:oo + n].view(k, i, j, l)
        oo += _round_up(n)
        perms.append((x, y))
        pbytes += 8.0 * n
    return gemms, perms, {"gemm_flop": flop, "gemm_bytes": gbytes, "permute_bytes": pbytes,
                          "buffers": (src, dst)}
