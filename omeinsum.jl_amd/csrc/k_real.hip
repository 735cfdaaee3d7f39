// Operand preparation for GEMMs that run on the real matrix cores in another arithmetic
// than the element type.
//
//   4M (realified complex):  A' = [Ar | Ai]  (M x 2K),  B' = [[Br, Bi], [-Bi, Br]]  (2K x 2N),
//                            A' B' = [Cr | Ci]: one real GEMM of 4MNK multiply-adds.
//   3M (Gauss):              P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi) as one batched
//                            real GEMM of three products; Cr = P1 - P2, Ci = P3 - P1 - P2.
//   widen:                   Float32 operands copied to Float64, FP64 GEMM, rounded once.
//
// The real operands are dense column-major per batch (A' m-fastest, B' k-fastest), so the
// real GEMM always sees unit strides; `combine` reads the real result and applies the
// caller's epilogue y <- sy*y + sx*C on the original (strided, interleaved) output.
// Element indices are decomposed with 32-bit invariant division: the planner routes
// products here only when every operand has fewer than 2^31 elements.
#include "kernels.h"

namespace ome {

namespace {

template <class T> struct Parts;
template <> struct Parts<float> {
  __device__ static float re(float v) { return v; }
  __device__ static float im(float) { return 0.f; }
};
template <> struct Parts<double> {
  __device__ static double re(double v) { return v; }
  __device__ static double im(double) { return 0.0; }
};
template <> struct Parts<c32> {
  __device__ static float re(c32 v) { return v.x; }
  __device__ static float im(c32 v) { return v.y; }
};
template <> struct Parts<c64> {
  __device__ static double re(c64 v) { return v.x; }
  __device__ static double im(c64 v) { return v.y; }
};

inline int blocks_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 1048576 ? 1048576 : g));
}

}  // namespace

RealShape real_shape(int algo, const GemmDesc& d) {
  RealShape r{};
  if (algo == RA_4M) {
    r.M = d.M; r.N = 2 * d.N; r.K = 2 * d.K; r.B = d.B;
  } else if (algo == RA_3M) {
    r.M = d.M; r.N = d.N; r.K = d.K; r.B = 3 * d.B;
  } else {
    r.M = d.M; r.N = d.N; r.K = d.K; r.B = d.B;
  }
  r.a_elems = r.M * r.K * r.B;
  r.b_elems = r.K * r.N * r.B;
  r.c_elems = r.M * r.N * r.B;
  return r;
}

// A (M x K per batch, strides sam/sak/sab) -> real operand(s), m fastest.
template <class TIn, class R, int ALGO>
__global__ __launch_bounds__(256) void k_prep_a(GemmDesc d, R* __restrict__ out, FDiv dm, FDiv dk) {
  const TIn* __restrict__ A = (const TIn*)d.a;
  const uint32_t total = (uint32_t)(d.M * d.K * d.B);
  const int64_t MK = d.M * d.K;
  for (uint32_t e = blockIdx.x * 256u + threadIdx.x; e < total; e += gridDim.x * 256u) {
    const uint32_t q = fdiv_q(e, dm), m = e - q * dm.d;
    const uint32_t b = fdiv_q(q, dk), k = q - b * dk.d;
    const TIn v = A[b * d.sab + m * d.sam + k * d.sak];
    const R re = (R)Parts<TIn>::re(v), im = (R)Parts<TIn>::im(v);
    if (ALGO == RA_4M) {
      R* o = out + (int64_t)b * 2 * MK + m;
      o[d.M * (int64_t)k] = re;
      o[d.M * (int64_t)(d.K + k)] = im;
    } else if (ALGO == RA_3M) {
      R* o = out + (int64_t)b * 3 * MK + m + d.M * (int64_t)k;
      o[0] = re;
      o[MK] = im;
      o[2 * MK] = re + im;
    } else {
      out[(int64_t)b * MK + m + d.M * (int64_t)k] = re;
    }
  }
}

// B (K x N per batch, strides sbk/sbn/sbb) -> real operand(s), k fastest.
template <class TIn, class R, int ALGO>
__global__ __launch_bounds__(256) void k_prep_b(GemmDesc d, R* __restrict__ out, FDiv dk, FDiv dn) {
  const TIn* __restrict__ Bm = (const TIn*)d.b;
  const uint32_t total = (uint32_t)(d.K * d.N * d.B);
  const int64_t KN = d.K * d.N;
  for (uint32_t e = blockIdx.x * 256u + threadIdx.x; e < total; e += gridDim.x * 256u) {
    const uint32_t q = fdiv_q(e, dk), k = e - q * dk.d;
    const uint32_t b = fdiv_q(q, dn), n = q - b * dn.d;
    const TIn v = Bm[b * d.sbb + k * d.sbk + n * d.sbn];
    const R re = (R)Parts<TIn>::re(v), im = (R)Parts<TIn>::im(v);
    if (ALGO == RA_4M) {
      // B' is 2K x 2N: column n = [Br; -Bi], column N + n = [Bi; Br]
      R* o = out + (int64_t)b * 4 * KN;
      const int64_t K2 = 2 * d.K;
      o[k + K2 * n] = re;
      o[d.K + k + K2 * n] = -im;
      o[k + K2 * (d.N + n)] = im;
      o[d.K + k + K2 * (d.N + n)] = re;
    } else if (ALGO == RA_3M) {
      R* o = out + (int64_t)b * 3 * KN + k + d.K * (int64_t)n;
      o[0] = re;
      o[KN] = im;
      o[2 * KN] = re + im;
    } else {
      out[(int64_t)b * KN + k + d.K * (int64_t)n] = re;
    }
  }
}

template <class O, class R> struct Make;
template <class R> struct Make<OpsF32, R> {
  __device__ static float mk(R re, R) { return (float)re; }
};
template <class R> struct Make<OpsF64, R> {
  __device__ static double mk(R re, R) { return (double)re; }
};
template <class R> struct Make<OpsC32, R> {
  __device__ static c32 mk(R re, R im) { return {(float)re, (float)im}; }
};
template <class R> struct Make<OpsC64, R> {
  __device__ static c64 mk(R re, R im) { return {(double)re, (double)im}; }
};

// Real result -> y (M x N per batch, strides sym/syn/syb) with the epilogue.
template <class O, class R, int ALGO>
__global__ __launch_bounds__(256) void k_combine(GemmDesc d, const R* __restrict__ C, FDiv dm, FDiv dn) {
  using T = typename O::T;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  const uint32_t total = (uint32_t)(d.M * d.N * d.B);
  const int64_t MN = d.M * d.N;
  for (uint32_t e = blockIdx.x * 256u + threadIdx.x; e < total; e += gridDim.x * 256u) {
    const uint32_t q = fdiv_q(e, dm), m = e - q * dm.d;
    const uint32_t b = fdiv_q(q, dn), n = q - b * dn.d;
    R re, im = 0;
    if (ALGO == RA_4M) {
      const R* c = C + (int64_t)b * 2 * MN + m;
      re = c[d.M * (int64_t)n];
      im = c[d.M * (int64_t)(d.N + n)];
    } else if (ALGO == RA_3M) {
      const R* c = C + (int64_t)b * 3 * MN + m + d.M * (int64_t)n;
      const R p1 = c[0], p2 = c[MN], p3 = c[2 * MN];
      re = p1 - p2;
      im = p3 - p1 - p2;
    } else {
      re = C[(int64_t)b * MN + m + d.M * (int64_t)n];
    }
    T* yp = (T*)d.y + b * d.syb + m * d.sym + n * d.syn;
    *yp = epilogue<O>(d.mode, sx, sy, Make<O, R>::mk(re, im), yp);
  }
}

// dt: the element type (F32, C32, C64); rdouble: the real GEMM's type is Float64.
#define OME_REAL_DISPATCH(dt, rdouble, algo, BODY)                                        \
  do {                                                                                    \
    if (dt == F32 && rdouble) {                                                           \
      using TIn = float; using R = double; using O = OpsF32;                              \
      constexpr int AL = RA_WIDEN; BODY;                                                  \
    } else if (dt == C32 && !rdouble) {                                                   \
      using TIn = c32; using R = float; using O = OpsC32;                                 \
      if (algo == RA_3M) { constexpr int AL = RA_3M; BODY; }                              \
      else { constexpr int AL = RA_4M; BODY; }                                            \
    } else if (dt == C32) {                                                               \
      using TIn = c32; using R = double; using O = OpsC32;                                \
      if (algo == RA_3M) { constexpr int AL = RA_3M; BODY; }                              \
      else { constexpr int AL = RA_4M; BODY; }                                            \
    } else if (dt == C64) {                                                               \
      using TIn = c64; using R = double; using O = OpsC64;                                \
      if (algo == RA_3M) { constexpr int AL = RA_3M; BODY; }                              \
      else { constexpr int AL = RA_4M; BODY; }                                            \
    }                                                                                     \
  } while (0)

void launch_prep_a(int dt, bool rdouble, int algo, const GemmDesc& d, void* out, hipStream_t st) {
  const int64_t n = d.M * d.K * d.B;
  if (n <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_PACK, st);
  const FDiv dm = make_fdiv((uint32_t)d.M), dk = make_fdiv((uint32_t)d.K);
  OME_REAL_DISPATCH(dt, rdouble, algo, {
    (void)sizeof(O);
    hipLaunchKernelGGL((k_prep_a<TIn, R, AL>), dim3(blocks_for(n)), dim3(256), 0, st, d, (R*)out, dm, dk);
  });
}

void launch_prep_b(int dt, bool rdouble, int algo, const GemmDesc& d, void* out, hipStream_t st) {
  const int64_t n = d.K * d.N * d.B;
  if (n <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_PACK, st);
  const FDiv dk = make_fdiv((uint32_t)d.K), dn = make_fdiv((uint32_t)d.N);
  OME_REAL_DISPATCH(dt, rdouble, algo, {
    (void)sizeof(O);
    hipLaunchKernelGGL((k_prep_b<TIn, R, AL>), dim3(blocks_for(n)), dim3(256), 0, st, d, (R*)out, dk, dn);
  });
}

void launch_combine(int dt, bool rdouble, int algo, const GemmDesc& dd, const void* c,
                    hipStream_t st) {
  const int64_t n = dd.M * dd.N * dd.B;
  if (n <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_OTHER, st);
  GemmDesc d = dd;
  const FDiv dm = make_fdiv((uint32_t)d.M), dn = make_fdiv((uint32_t)d.N);
  OME_REAL_DISPATCH(dt, rdouble, algo, {
    (void)sizeof(TIn);
    d.mode = epi_mode<O>(d.sx, d.sy);
    hipLaunchKernelGGL((k_combine<O, R, AL>), dim3(blocks_for(n)), dim3(256), 0, st, d, (const R*)c, dm, dn);
  });
}

#undef OME_REAL_DISPATCH

}  // namespace ome
