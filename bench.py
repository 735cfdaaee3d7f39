"""Benchmark of the MI355X contraction backend (one JSON line on stdout).

Headline workload: ein"abcd,cdef->abef" with every label of size 64 (two 16.7M-element
Float32 tensors; one 4096^3 GEMM-shaped contraction), evaluated by the MI355X backend with
the 3-term BF16 split on the matrix cores.  A step = one einsum forward over inputs already
resident in HBM (split/pack of both operands + the packed GEMM).  value = algorithmic
TFLOP/s (2*64^6 flop per step).

roofline: the dominant kernel is the packed BF16x3 GEMM.  Its launches are bracketed by HIP
events on the stream they run on (the library's kernel-timing hook) during the timed region;
achieved = 2*64^6 flop per launch / average launch duration, against the FP32-equivalent
peak of the 3-product split (dense BF16 MFMA peak / 3).  traffic = HBM bytes per launch from
the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this kernel, when present.

Also reported in the same line: the same contraction in plain FP32 MFMA and its error vs
FP64; ijk,jkl->il at 64^3; the 100-vertex 3-regular ComplexF32 network contracted to a scalar
by the MI355X backend (graph replay) next to the plain per-step GPU path; the CPU baseline
(the oracle's numpy restatement of OMEinsum's DefaultBackend on this host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|c4]
                       [--no-cpu] [--copy-calibration]
"""
import argparse
import json
import os
import re
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "omeinsum.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_BF16 = 2516.6      # TFLOP/s dense (MI355X datasheet)
PEAK_FP32 = 157.3
PEAK_FP64 = 78.6
PEAK_HBM = 8000.0       # GB/s
C4_FLOPS = 2.0 * 64 ** 6
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "gemm_packed_bf16x3_abcd_cdef_64_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_events(fn, steps, warmup, after_warmup=None):
    """Average milliseconds per call: HIP events on torch's current stream (the library
    launches on that stream) bracketing `steps` calls after `warmup` calls."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    if after_warmup is not None:
        after_warmup()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def read_traffic():
    """HBM bytes per launch of the packed GEMM on the headline workload, from the committed
    rocprofv3 counter summary (line `traffic_bytes_per_launch: N`), or None."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                m = re.match(r"\s*traffic_bytes_per_launch:\s*([0-9.eE+]+)", line)
                if m:
                    return float(m.group(1))
    except OSError:
        pass
    return None


def bench_c4(steps, warmup, precs=("bf16x3", "fp32")):
    import numpy as np
    import torch
    import omeinsum_mi355x as om
    from omeinsum_mi355x import _lib

    g = torch.Generator(device="cuda").manual_seed(4000)
    a = torch.randn(64, 64, 64, 64, device="cuda", generator=g)
    b = torch.randn(64, 64, 64, 64, device="cuda", generator=g)
    code = om.ein("abcd,cdef->abef")
    sd = om.get_size_dict(code.ixs, (a, b))
    y = torch.empty(64, 64, 64, 64, device="cuda")
    ctx = _lib.context()
    out = {}
    for prec in precs:
        om.set_einsum_backend(om.MI355XBackend(precision=prec))
        fn = lambda: om.einsum_inplace(code, (a, b), y, True, False, sd)
        kc = "gemm_packed" if prec == "bf16x3" else "gemm_mfma"

        def clear():
            for k in ("gemm_packed", "gemm_mfma", "pack"):
                ctx.kernel_timing(k)

        ctx.set_option(_lib.OPT_KERNEL_TIMING, 1)
        try:
            ms = time_events(fn, steps, warmup, after_warmup=clear)
            k_ms, k_n = ctx.kernel_timing(kc)
            p_ms, p_n = ctx.kernel_timing("pack")
        finally:
            ctx.set_option(_lib.OPT_KERNEL_TIMING, 0)
        kern_ms = k_ms / max(k_n, 1)
        # error of the result against FP64 on a sample of output rows
        fn()
        torch.cuda.synchronize()
        an = a.cpu().numpy().astype(np.float64).reshape(4096, 4096)
        bn = b.cpu().numpy().astype(np.float64).reshape(4096, 4096)
        rows = np.arange(0, 4096, 64)
        ref = an[rows] @ bn           # row-major: abcd -> (ab),(cd)
        got = y.cpu().numpy().reshape(4096, 4096)[rows].astype(np.float64)
        err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        out[prec] = {"ms": ms, "tflops": C4_FLOPS / ms / 1e9, "rel_err_fp64": err,
                     "gemm_kernel_us": kern_ms * 1e3, "gemm_kernel_launches": k_n,
                     "gemm_kernel_tflops": C4_FLOPS / kern_ms / 1e9 if k_n else None,
                     "pack_kernel_us": (p_ms / p_n * 1e3) if p_n else None}
        log("C4 %s: %.3f ms/step, %.1f TFLOP/s; GEMM kernel %.1f us (%d launches, %.1f TFLOP/s); "
            "rel err vs FP64 %.2e" % (prec, ms, C4_FLOPS / ms / 1e9, kern_ms * 1e3, k_n,
                                      C4_FLOPS / kern_ms / 1e9 if k_n else 0, err))
    om.set_einsum_backend(om.DefaultBackend())
    return out, (a, b)


def bench_c1(steps, warmup):
    import torch
    import omeinsum_mi355x as om
    out = {}
    code = om.ein("ijk,jkl->il")
    for name, dt in (("f64", torch.float64), ("f32", torch.float32), ("i64", torch.int64)):
        if dt == torch.int64:
            a = torch.randint(-8, 9, (64, 64, 64), device="cuda")
            b = torch.randint(-8, 9, (64, 64, 64), device="cuda")
        else:
            a = torch.rand(64, 64, 64, dtype=dt, device="cuda") * 2 - 1
            b = torch.rand(64, 64, 64, dtype=dt, device="cuda") * 2 - 1
        sd = om.get_size_dict(code.ixs, (a, b))
        y = torch.empty(64, 64, dtype=dt, device="cuda")
        ms = time_events(lambda: om.einsum_inplace(code, (a, b), y, True, False, sd), steps * 5, warmup)
        out[name] = {"us": ms * 1e3, "tflops": 2 * 64 ** 4 / ms / 1e9}
    log("C1 ijk,jkl->il 64^3: %s" % {k: round(v["us"], 1) for k, v in out.items()})
    return out


def bench_c2(steps, warmup):
    import numpy as np
    import torch
    import omeinsum_mi355x as om
    from omeinsum_mi355x import contractionorder as co, networks
    from omeinsum_mi355x.tree import plan_for
    from oracle import default_backend as db

    n = 100
    edges = networks.random_regular_graph(n, 3, seed=2000)
    code = networks.network_code(n, edges)
    sd = {e: 2 for e in range(len(edges))}
    xs = networks.gaussian_tensors(code, sd, seed=2001, dtype=np.complex64)
    tree = co.optimize_code(code, sd)
    dxs = [torch.from_numpy(x).cuda() for x in xs]
    # FP64 reference of the same tree (the oracle's per-step restatement, in complex128)
    ref = complex(db.nested_einsum(tree.to_plain(), [x.astype(np.complex128) for x in xs], sd)[()])
    out = {"n_vertices": n, "flops_log2": float(np.log2(8 * co.flop(tree, sd)))}
    y = torch.empty((), dtype=torch.complex64, device="cuda")
    for name, be in (("mi355x_fp32", om.MI355XBackend(precision="fp32")),
                     ("mi355x_bf16x3", om.MI355XBackend(precision="bf16x3")),
                     ("per_step_fp32", om.DefaultBackend())):
        om.set_einsum_backend(be)
        fn = lambda: om.einsum_inplace(tree, dxs, y, True, False, sd)
        ms = time_events(fn, steps, warmup)
        fn()
        got = complex(y.cpu().numpy()[()])
        out[name] = {"ms": ms, "rel_err_fp64": abs(got - ref) / abs(ref)}
    om.set_einsum_backend(om.MI355XBackend())
    plan = plan_for(tree, dxs, sd)
    info = plan.info()
    out["tree"] = {"steps": info.n_steps, "kernels": int(info.n_kernels),
                   "arena_bytes": int(info.arena_bytes), "flops": info.flops}
    v, lg = plan.execute_scalar_normalized(dxs)
    out["normalized_rel_err_fp64"] = abs(v * np.exp(lg) - ref) / abs(ref)
    om.set_einsum_backend(om.DefaultBackend())
    log("C2 100-vertex network: %s" % {k: v for k, v in out.items() if isinstance(v, dict)})
    return out


def copy_calibration(reps=5):
    """A streaming copy of 1 GiB through one elementwise kernel (y = x + 0: each element read
    once, written once), repeated: the reading of the FETCH_SIZE / WRITE_SIZE counters for a
    kernel that moves exactly 1 GiB each way."""
    import torch
    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    log("copy calibration: %d kernels, each reading and writing %d bytes" % (reps, x.numel() * 4))


def blas_threads():
    try:
        from threadpoolctl import threadpool_info
        n = [i["num_threads"] for i in threadpool_info() if i.get("user_api") == "blas"]
        if n:
            return int(max(n))
    except Exception:  # noqa: BLE001 - fall back to the visible core count
        pass
    return os.cpu_count()


def cpu_baseline(ab):
    """The oracle's restatement of OMEinsum's DefaultBackend (analyze_binary + permutes +
    OpenBLAS GEMM through numpy) on this host, on the same 64^4 Float32 inputs.  Bounded
    sample: up to 5 timed repetitions after one warm-up, stopping after ~25 s."""
    from oracle import default_backend as db
    a = ab[0].cpu().numpy()
    b = ab[1].cpu().numpy()
    ixs, iy = [list("abcd"), list("cdef")], list("abef")
    db.einsum(ixs, iy, [a, b])
    ts = []
    t_end = time.time() + 25
    for _ in range(5):
        t0 = time.perf_counter()
        db.einsum(ixs, iy, [a, b])
        ts.append(time.perf_counter() - t0)
        if time.time() > t_end:
            break
    med = statistics.median(ts)
    return {"value": round(C4_FLOPS / med / 1e12, 4), "unit": "TFLOP/s", "cores": blas_threads(),
            "kind": "port",
            "sample": "the full abcd,cdef->abef 64^4 Float32 contraction, oracle DefaultBackend "
                      "restatement (numpy permutes + OpenBLAS sgemm), median of %d runs = %.3f s"
                      % (len(ts), med)}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", choices=("all", "c4"), default="all")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--copy-calibration", action="store_true")
    args = p.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this backend runs on a single MI355X (no multi-GPU "
              "path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs the GPU", file=sys.stderr)
        sys.exit(1)
    import omeinsum_mi355x as om  # noqa: F401

    precs = ("bf16x3",) if args.workload == "c4" else ("bf16x3", "fp32")
    c4, ab = bench_c4(args.steps, args.warmup, precs)
    extra = {"c4": c4}
    if args.workload == "all":
        extra["c1_ijk_jkl_il_64"] = bench_c1(args.steps, args.warmup)
        extra["c2_network_100"] = bench_c2(max(5, args.steps // 2), 2)
    if args.copy_calibration:
        copy_calibration()
    cpu = None if args.no_cpu else cpu_baseline(ab)
    head = c4["bf16x3"]
    peak = PEAK_BF16 / 3
    kt = head["gemm_kernel_tflops"]
    line = {
        "metric": "einsum TFLOP/s (% MFMA peak) vs plain GPU path; p50/p99 TN-request latency",
        "value": round(head["tflops"], 2),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (bf16x3 split products, f32 accumulate)",
        "data": "synthetic (seeded normal inputs)",
        "config": {"workload": "ein\"abcd,cdef->abef\", all labels 64 (two 64^4 Float32 tensors), "
                               "MI355XBackend(precision=bf16x3)", "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(kt, 2) if kt else None,
                     "peak": round(peak, 1), "unit": "TFLOP/s",
                     "frac": round(kt / peak, 4) if kt else None, "traffic": read_traffic(),
                     "kernel": "k_gemm_packed_bf16x3 (avg %.1f us over %d launches)"
                               % (head["gemm_kernel_us"], head["gemm_kernel_launches"]),
                     "peak_note": "FP32-equivalent peak of the 3-product BF16 split = dense BF16 "
                                  "MFMA peak 2516.6 / 3"},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
