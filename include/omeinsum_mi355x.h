/*
 * omeinsum_mi355x.h — C ABI of the MI355X (gfx950) contraction backend for OMEinsum.
 *
 * Every entry point takes plain pointers and sizes; no Julia, Python or torch types cross
 * this boundary.  A Julia binding calls them with `ccall`; the Python host package binds
 * them with ctypes.
 *
 * Conventions shared by all entry points
 *   - Tensors are dense and COLUMN-MAJOR (first label fastest), as Julia arrays are.
 *     A tensor with labels (l0, l1, ..., l{r-1}) has element (i0, ..., i{r-1}) at offset
 *     i0 + i1*d0 + i2*d0*d1 + ...   (d_p = sizes[l_p]).
 *   - Labels are dense int32 ids 0..nlabels-1 assigned by the caller; `sizes[id]` is the
 *     extent of label `id`.  A 0-d tensor (no labels) holds one element.
 *   - All data pointers are DEVICE pointers on the context's device.  Work is enqueued on
 *     the context's stream and is stream-ordered; call ome_synchronize() to wait.
 *   - Scaling semantics everywhere:  y <- sy*y + sx*f(xs).  When sy is the zero of the
 *     element type, y is not read (NaN-safe overwrite).  When sx is zero, y <- sy*y.
 *     This is OMEinsum's @addmul! contract (src/utils.jl:1-37, src/loop_einsum.jl:21-32).
 *   - Element types: OME_F32/F64 (Float32/Float64), OME_C32/C64 (ComplexF32/ComplexF64,
 *     interleaved re,im), OME_I64 (Int64, wrapping two's-complement arithmetic, bit-exact),
 *     OME_TROP_F32/F64 (TropicalNumbers' Tropical{T}: a+b := max(a,b), a*b := a+b,
 *     zero = -Inf, one = 0; the stored value is the tropical number's content),
 *     OME_DD (Double64 of DoubleFloats.jl, the element type of the reference's GPU test
 *     test/roceinsum.jl:56-72: value = hi + lo, stored as interleaved (hi, lo) Float64
 *     pairs; double-word addition and multiplication after Joldes, Muller and Popescu,
 *     ACM TOMS 44(2), 2017, Algorithms 6 and 12),
 *     OME_CTROP_F64 (TropicalNumbers' CountingTropical{Float64,Float64}: a*b := (a.n + b.n,
 *     a.c * b.c), a+b := the larger n with its count, counts added on equal n; zero =
 *     (-Inf, 0), one = (0, 1); a scalar's re is n and im is c),
 *     OME_TROP_BOOL (TropicalNumbers' TropicalAndOr: + is or, * is and; a scalar's i is the
 *     value).  The semiring types have no counterpart among the reference's GPU kernels:
 *     OMEinsum runs them through its generic paths, which these calls replace.
 *     All inputs of one call share the output's element type; the caller promotes first
 *     (OMEinsum promotes in get_output_array, src/loop_einsum.jl:48-61).
 *   - Return value: OME_OK or an ome_status code; the message is in ome_last_error(ctx).
 *     The glue re-throws ArgumentError / DimensionMismatch / ErrorException accordingly
 *     (reference error types: src/interfaces.jl:78-90, src/loop_einsum.jl:24).
 *   - Calls on one context are serialised by an internal mutex.
 */
#ifndef OMEINSUM_MI355X_H
#define OMEINSUM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OME_ABI_VERSION 1
#define OME_MAX_RANK 64     /* labels per tensor */
#define OME_MAX_INPUTS 64   /* tensors per flat EinCode */

typedef enum ome_status {
  OME_OK = 0,
  OME_ERR_INVALID_ARG = 1,     /* -> ArgumentError */
  OME_ERR_DIM_MISMATCH = 2,    /* -> DimensionMismatch */
  OME_ERR_OOM = 3,             /* device memory pool exhausted */
  OME_ERR_LOOPS_FORBIDDEN = 4, /* allow_loops(false) and the code needs the index walk */
  OME_ERR_DEVICE = 5,          /* HIP runtime error */
  OME_ERR_UNSUPPORTED = 6      /* shape/type outside what this build handles */
} ome_status;

typedef enum ome_dtype {
  OME_F32 = 0,
  OME_F64 = 1,
  OME_C32 = 2,
  OME_C64 = 3,
  OME_I64 = 4,
  OME_TROP_F32 = 5,
  OME_TROP_F64 = 6,
  OME_DD = 7,        /* Double64 (DoubleFloats.jl): interleaved (hi, lo) Float64 pairs */
  OME_CTROP_F64 = 8, /* CountingTropical{Float64,Float64}: interleaved (n, c) Float64 pairs */
  OME_TROP_BOOL = 9  /* TropicalAndOr: one byte per element, 0 or 1 */
} ome_dtype;

/* Arithmetic used for Float32 / ComplexF32 products on the matrix cores. */
typedef enum ome_precision {
  OME_PREC_FP32 = 0,   /* v_mfma_f32_*_f32: FP32 operands, FP32 accumulate (default) */
  OME_PREC_BF16X3 = 1, /* a = ah + al (bf16 hi/lo); ah*bh + ah*bl + al*bh, FP32 accumulate */
  OME_PREC_FP64 = 2,   /* operands widened to FP64, v_mfma_f64 */
  OME_PREC_BF16X6 = 3  /* a = a0 + a1 + a2 (three bf16 planes); the six products ai*bj with
                          i + j <= 2, FP32 accumulate: FP32-level error on the bf16 matrix
                          cores (large products; smaller ones run at OME_PREC_FP32) */
} ome_precision;

typedef enum ome_complex_algo {
  OME_CPLX_4M = 0, /* four real products per complex product */
  OME_CPLX_3M = 1  /* Gauss: three real products (ar*br, ai*bi, (ar+ai)*(br+bi)) */
} ome_complex_algo;

typedef enum ome_option {
  OME_OPT_PRECISION = 0,    /* ome_precision for F32/C32 GEMM-shaped steps */
  OME_OPT_COMPLEX_ALGO = 1, /* ome_complex_algo */
  OME_OPT_ALLOW_LOOPS = 2,  /* 1 (default) / 0: mirrors OMEinsum.allow_loops (src/loop_einsum.jl:63-73) */
  OME_OPT_DIRECT_GEMM = 3,  /* 1 (default): strided GEMM reads operands in place when layout allows */
  OME_OPT_USE_GRAPHS = 4,   /* 1 (default): tree plans replay a captured hipGraph */
  OME_OPT_AUTOTUNE = 5,     /* 1 (default): time GEMM tile candidates on first use of a shape class */
  OME_OPT_GEMM_KERNEL = 6,  /* force a GEMM kernel family (-1 = automatic; for tests) */
  OME_OPT_KERNEL_TIMING = 7, /* 1: bracket kernel launches with HIP events (process-wide) */
  OME_OPT_GROUPED_VARIANT = 8, /* grouped GEMM tile variant: -1 autotuned (default), 0 64x64,
                                  1 64x64 with K step 32, 2 128x128, 3 BF16x3 64x64,
                                  4 BF16x3 128x128, 5 one wave per 32x32 tile,
                                  6 BF16x3 one wave per 32x32 tile, 7 BF16x3 64x64 with K
                                  step 64, 8 64x64 persistent and pipelined across tiles,
                                  9 BF16x3 64x64 persistent and pipelined (8 and 9 need
                                  buffer resources and otherwise run as 1 and 3),
                                  10 BF16x6 64x64 (Float32 only; for tests) */
  OME_OPT_LAST_GROUPED_CHOICE = 9, /* read-only: 2*variant + persistent of the last grouped
                                      GEMM launch (the autotuner's pick) */
  OME_OPT_TREE_DEFER_MEMORY = 10,  /* 1: ome_tree_create sizes the plan's device memory but
                                      does not allocate it (first use or the request loop
                                      does); ome_tree_get_info reports the size */
  OME_OPT_GROUPED_BUFFERS = 11,    /* 1 (default): Float32 grouped GEMMs whose problems all fit
                                      32-bit offsets read and write through buffer resources;
                                      0: always through pointers (for tests) */
  OME_OPT_TRACE_RANGES = 12        /* 1: roctx ranges (process-wide) around every tree
                                      evaluation, tree step, graph replay, grouped launch and
                                      request (admission to retirement), for rocprofv3
                                      --marker-trace; OME_ERR_UNSUPPORTED when the roctx
                                      library (librocprofiler-sdk-roctx) cannot be loaded */
} ome_option;

/* Kernel classes of ome_get_kernel_timing. */
typedef enum ome_kernel_class {
  OME_KC_GEMM_PACKED = 0, /* BF16x3 GEMM on packed hi/lo operand tiles */
  OME_KC_PACK = 1,        /* FP32 -> BF16 hi/lo split and tile packing */
  OME_KC_GEMM_MFMA = 2,   /* FP32 / FP64 / BF16x3 MFMA GEMM on strided operands */
  OME_KC_GEMM_VALU = 3,   /* vector-ALU GEMM (Int64, tropical, complex, small shapes) */
  OME_KC_PERMUTE = 4,     /* permutedims / tensorpermute */
  OME_KC_WALK = 5,        /* index walk (loop_einsum, unary reductions) */
  OME_KC_GROUPED = 6,     /* grouped GEMM */
  OME_KC_OTHER = 7        /* split-K reductions and the rest */
} ome_kernel_class;

/* A scalar for sx / sy / alpha / beta.  Float types read re (and im for complex);
 * OME_I64 reads i; tropical types read re as the tropical content (so the tropical one
 * is re = 0 and the tropical zero is re = -Inf). */
typedef struct ome_scalar {
  double re;
  double im;
  int64_t i;
} ome_scalar;

typedef struct ome_ctx ome_ctx;
typedef struct ome_tree ome_tree;
typedef struct ome_server ome_server;

typedef struct ome_device_info {
  char name[64];
  char arch[32];       /* "gfx950" on MI355X */
  int32_t compute_units;
  int32_t xcds;
  int64_t total_mem;
  int64_t lds_per_cu;
  int32_t clock_mhz;
  int32_t device;
} ome_device_info;

typedef struct ome_pool_stats {
  int64_t reserved_bytes;  /* device memory held by the pool */
  int64_t in_use_bytes;
  int64_t peak_in_use_bytes;
  int64_t n_allocs;
  int64_t n_device_mallocs;
} ome_pool_stats;

typedef struct ome_tree_info {
  int32_t n_steps;          /* internal nodes of the tree (after drop_slicedim) */
  int32_t n_leaves;
  int64_t n_slices;         /* product of sliced label sizes (1 if not sliced) */
  double flops;             /* algorithmic flops of one full evaluation (all slices) */
  double bytes;             /* algorithmic bytes (sum of step inputs + outputs) */
  int64_t arena_bytes;      /* device arena one evaluation needs (static liveness plan) */
  int64_t peak_live_elems;  /* peak_memory in elements: live tensors when a step replaces
                               its inputs by its output (OMEinsumContractionOrders' measure) */
  int64_t n_kernels;        /* kernels launched by one evaluation */
} ome_tree_info;

/* ---- context ----------------------------------------------------------------------- */

/* No reference counterpart: the HIP runtime owner.  stream may be NULL (null stream) or
 * a hipStream_t created by the caller (torch.cuda.current_stream().cuda_stream). */
int ome_init(int device, size_t pool_bytes, void* stream, ome_ctx** out);
int ome_shutdown(ome_ctx* ctx);
int ome_set_stream(ome_ctx* ctx, void* stream);
const char* ome_last_error(const ome_ctx* ctx);
int ome_abi_version(void);
/* Mirrors OMEinsum.allow_loops (src/loop_einsum.jl:63-73) and backend options. */
int ome_set_option(ome_ctx* ctx, int key, int64_t value);
int ome_get_option(ome_ctx* ctx, int key, int64_t* value);
int ome_get_device_info(ome_ctx* ctx, ome_device_info* info);
int ome_synchronize(ome_ctx* ctx);
/* No reference counterpart: measurement hook.  While OME_OPT_KERNEL_TIMING is 1, every launch
 * of a kernel class is bracketed by HIP events on the stream it runs on (launches captured into
 * a graph are not timed).  Waits for the recorded launches of kclass, returns their summed
 * duration and their number, and clears them. */
int ome_get_kernel_timing(ome_ctx* ctx, int kclass, double* total_ms, int64_t* launches);

/* ---- memory pool (replaces get_output_array / similar / unsafe_free!:
 *      src/loop_einsum.jl:48-61, ext/AMDGPUExt.jl:27-40,124-126) ------------------------ */
int ome_alloc(ome_ctx* ctx, size_t bytes, void** ptr);
int ome_free(ome_ctx* ctx, void* ptr);
int ome_get_pool_stats(ome_ctx* ctx, ome_pool_stats* stats);
int ome_memcpy_h2d(ome_ctx* ctx, void* dst, const void* src, size_t bytes);
int ome_memcpy_d2h(ome_ctx* ctx, void* dst, const void* src, size_t bytes);
int ome_memcpy_d2d(ome_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- flat einsum -------------------------------------------------------------------- */

/* einsum!(ixs, iy, xs, y, sx, sy, size_dict): replaces src/einsum.jl:43-124 for device
 * arrays — unary (1 input: the unary pipeline, :52-101), binary (2 inputs:
 * analyze_binary + simplifyto + GEMM + post-einsum, :104-211) and n-ary (>= 3 inputs: the
 * index-walk kernel, i.e. loop_einsum!, :47-50).  ixs is the concatenation of the label
 * lists of the nin inputs, nix[k] their lengths. */
int ome_einsum(ome_ctx* ctx, int dtype, int nin, const void* const* xs, const int32_t* nix,
               const int32_t* ixs, void* y, int niy, const int32_t* iy, int nlabels,
               const int64_t* sizes, const ome_scalar* sx, const ome_scalar* sy);

/* No counterpart in the reference: ome_einsum for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, semantics and errors as ome_einsum; the planner's temporaries
 * (permuted operands, split-K partials, packed BF16 operands, real-GEMM operands) come from
 * one device array of OME_NOALLOC_SCRATCH_BYTES per GPU, held by the library until the
 * process ends, instead of the pool, and the call allocates nothing.  A contraction whose
 * temporaries exceed the array fails with OME_ERR_OOM before anything is launched.  Every
 * call on one GPU uses that GPU's array: calls that overlap on a GPU (other contexts or
 * streams of the same device) race; calls in order on one stream do not. */
#define OME_NOALLOC_SCRATCH_BYTES (64ll << 20)
int ome_einsum_noalloc(ome_ctx* ctx, int dtype, int nin, const void* const* xs,
                       const int32_t* nix, const int32_t* ixs, void* y, int niy,
                       const int32_t* iy, int nlabels, const int64_t* sizes,
                       const ome_scalar* sx, const ome_scalar* sy);

/* loop_einsum!(ixs, iy, xs, y, sx, sy, size_dict): src/loop_einsum.jl:21-32 (and its
 * ROCArray override ext/AMDGPUExt.jl:54-86).  Always the index-walk kernel; fails with
 * OME_ERR_LOOPS_FORBIDDEN when allow_loops is off. */
int ome_loop_einsum(ome_ctx* ctx, int dtype, int nin, const void* const* xs,
                    const int32_t* nix, const int32_t* ixs, void* y, int niy,
                    const int32_t* iy, int nlabels, const int64_t* sizes,
                    const ome_scalar* sx, const ome_scalar* sy);

/* @flatten_addmul! y .= sy*y .+ sx*x over n elements (src/utils.jl:1-37); conj_x != 0 reads
 * conj(x) (the conj(dy) of einsum_grad, src/autodiff.jl:24-30).  x == y is allowed. */
int ome_addmul(ome_ctx* ctx, int dtype, int64_t n, const void* x, void* y,
               const ome_scalar* sx, const ome_scalar* sy, int conj_x);

/* tensorpermute!(C, A, perm, sx, sy): src/utils.jl:122-156.  dims are A's extents,
 * perm is 0-based (C dim p = A dim perm[p]). */
int ome_permute(ome_ctx* ctx, int dtype, int ndim, const int64_t* dims, const int32_t* perm,
                const void* a, void* c, const ome_scalar* sx, const ome_scalar* sy);

/* _batched_gemm!(C1, C2, alpha, A, B, beta, C): src/utils.jl:162-177 and
 * ext/AMDGPUExt.jl:110-112.  Column-major A (m x k x batch, or k x m x batch when
 * ta == 'T'), B (k x n x batch, or n x k x batch when tb == 'T'), C m x n x batch. */
int ome_batched_gemm(ome_ctx* ctx, int dtype, char ta, char tb, int64_t m, int64_t n,
                     int64_t k, int64_t batch, const ome_scalar* alpha, const void* a,
                     const void* b, const ome_scalar* beta, void* c);

/* No reference counterpart: many independent batched GEMMs of uneven shapes in one
 * launch (the shallow levels of contraction trees, independent networks).
 * Each problem g: C_g[m,n,l] = alpha*sum_k A_g[m,k,l] B_g[k,n,l] + beta*C_g, column-major,
 * dense, no transposes; dims holds (m,n,k,batch) per problem.  Float32 runs on the FP32
 * matrix cores, other types on the semiring VALU tile.  The tile shape and grid are chosen
 * per shape class by timing the candidates on first use (when beta is zero, autotuning is
 * on and the stream is not being captured) and remembered for the context.  Graph capture:
 * a problem list (shapes and pointers) may be captured once it has run on the context; its
 * device descriptors then stay allocated until ome_shutdown.  A list that has not run fails
 * under capture with OME_ERR_UNSUPPORTED (the same holds for ome_grouped_permute). */
int ome_grouped_gemm(ome_ctx* ctx, int dtype, int ngroups, const int64_t* dims,
                     const void* const* a, const void* const* b, void* const* c,
                     const ome_scalar* alpha, const ome_scalar* beta);

/* No reference counterpart: many tensorpermute! calls (src/utils.jl:122-156) in one launch.
 * Problem g permutes a[g] (ndims[g] dims; extents in dims, concatenated over problems) into
 * c[g]: c dim p = a dim perms[p] (0-based, concatenated), c <- sy*c + sx*permutedims(a). */
int ome_grouped_permute(ome_ctx* ctx, int dtype, int nperm, const int32_t* ndims,
                        const int64_t* dims, const int32_t* perms, const void* const* a,
                        void* const* c, const ome_scalar* sx, const ome_scalar* sy);

/* ---- contraction trees -------------------------------------------------------------- */

/* Plan a NestedEinsum (src/einsequence.jl:161-287) or SlicedEinsum (src/slicing.jl:10-137).
 * desc is the tree in post-order:
 *   desc = [n_leaves, n_nodes, node_0, node_1, ...]       (root is the last node)
 *   node = [nargs, arg_0, ..., arg_{nargs-1}, niy, iy...]
 *   arg  = [ref, nix, ix...]  with ref >= 0 a leaf (tensorindex, 0-based) and
 *          ref < 0 the internal node -(ref+1) (an earlier node).
 * slice_labels (may be NULL when nslice == 0) are the SlicedEinsum labels.  Intermediates
 * live in one device arena laid out by a static liveness plan; evaluation is recorded once
 * into a hipGraph and replayed.  The context's planning options (precision, complex
 * algorithm, allow_loops, direct GEMM, GEMM kernel) are frozen into the plan here; later
 * changes of those options apply to plans created afterwards. */
int ome_tree_create(ome_ctx* ctx, int dtype, const int32_t* desc, int64_t desc_len,
                    int nlabels, const int64_t* sizes, int nslice,
                    const int32_t* slice_labels, ome_tree** out);
int ome_get_tree_info(ome_tree* tree, ome_tree_info* info);
/* einsum!(ne, xs, y, sx, sy, size_dict): src/einsequence.jl:270-287, src/slicing.jl:112-121.
 * For a sliced tree the cross-slice sum runs in FP64 (complex FP64 for C32) and is
 * rounded to the output type once. */
int ome_tree_execute(ome_tree* tree, const void* const* xs, void* y, const ome_scalar* sx,
                     const ome_scalar* sy);
/* No reference counterpart: full-network contraction to a scalar with log-scale
 * normalisation.  Every input tensor is divided by its max-abs, and so is every
 * intermediate; the logs accumulate in FP64.  Float32 / ComplexF32 trees are evaluated in
 * FP64: each input is widened to Float64 / ComplexF64 as it is normalised, and an FP64 twin
 * of the plan (built on first use, then kept with the plan) runs every step, the final
 * scalar included, in FP64.  Writes value (re, im) and log_scale so that
 * result = value * exp(log_scale). */
int ome_tree_execute_scalar_normalized(ome_tree* tree, const void* const* xs,
                                       double* value_re_im, double* log_scale);
/* cost_and_gradient(code, xs, ybar): src/bp.jl:122-130 (cached_einsum :33-48, back_propagate
 * :79-96, einsum_backward_rule src/autodiff.jl:24-33).  ybar has the output's shape; xbars[i]
 * receives d cost / d xs[i] (conj convention of src/autodiff.jl:26-29; real inputs keep the
 * real part).  y receives the forward value.  Sliced trees are evaluated unsliced, as the
 * reference does (src/bp.jl:26-29). */
int ome_tree_gradient(ome_tree* tree, const void* const* xs, const void* ybar, void* y,
                      void* const* xbars);
int ome_tree_destroy(ome_tree* tree);

/* ---- request loop (no reference counterpart) ---------------------------------------------
 * Continuous batching of tensor-network contraction requests: a request is admitted when
 * its tree's arena fits the remaining HBM budget and an execution slot is free; admitted
 * requests run concurrently on separate streams; requests that complete free their slot
 * and arena so waiting requests are admitted as they arrive.  Admission orders a request
 * after the work on the context's stream at that moment (an event recorded there), so that
 * stream must not be under graph capture while requests are queued. */
int ome_server_create(ome_ctx* ctx, size_t hbm_budget, int max_active, ome_server** out);
/* xs and y stay owned by the caller and must stay valid until the request completes. */
int ome_server_submit(ome_server* srv, ome_tree* tree, const void* const* xs, void* y,
                      int64_t* ticket);
/* status: 0 queued, 1 running, 2 done, <0 failed (negated ome_status). */
int ome_server_poll(ome_server* srv, int64_t ticket, int* status);
/* Blocks until done; latency_ms = completion - submission (host wall clock). */
int ome_server_wait(ome_server* srv, int64_t ticket, double* latency_ms);
int ome_server_destroy(ome_server* srv);
/* Per-request memory limit (default: the whole budget).  A request whose plan needs more is
 * sliced at submission (src/slicing.jl:104-137 semantics): the loop admits a twin of its tree
 * with labels sliced one at a time (each time the non-output label carried by the most
 * elements of the step outputs) until the twin's memory fits, built once per plan and limit.
 * Submission fails with OME_ERR_OOM when no slicing fits. */
int ome_server_set_request_limit(ome_server* srv, size_t bytes);
typedef struct ome_server_stats {
  int64_t submitted;
  int64_t completed;
  int64_t sliced;           /* submissions admitted as a sliced twin */
  int64_t blocks_reused;    /* admissions lent the block the same plan used last */
  int64_t kept_bytes;       /* retired blocks the loop keeps for later admissions */
  int64_t batched_launches; /* shared launches of several requests' fused small steps */
  int64_t batched_requests; /* requests whose fused small steps ran in a shared launch */
} ome_server_stats;
int ome_server_get_stats(ome_server* srv, ome_server_stats* stats);
/* Cross-request batching (on by default): the requests admitted together run the fused
 * subtrees of small steps of their plans (unsliced plans, one element type) in one launch
 * shared by all of them, then each its other steps.  Each request computes exactly what it
 * computes alone. */
int ome_server_set_batching(ome_server* srv, int on);
/* Admission hold (no reference counterpart; for callers that submit a burst of requests to
 * start together, and for tests of batching): while on, submitted requests queue without
 * being admitted; turning it off admits every queued request that fits in one round.  A
 * request submitted under hold completes only after the hold is released, so release it
 * before waiting on such a request.  Destroying the server admits and drains the queue. */
int ome_server_hold(ome_server* srv, int on);
/* Request status word (no reference counterpart; for tests of failure handling): every
 * execution slot owns a device-visible status word, cleared when a request is admitted to it
 * and read when the request retires; a non-zero word fails that request alone
 * (-OME_ERR_DEVICE) while the requests beside it complete.  This call marks the next request
 * submitted to the loop (code 0 clears the mark) so that, when it runs, a kernel on its slot
 * stream writes `code` into the word after its steps, as a kernel that detects a fault in its
 * own work would.  Every kernel still runs to its end. */
int ome_server_inject_status(ome_server* srv, int code);

#ifdef __cplusplus
}
#endif
#endif /* OMEINSUM_MI355X_H */
