// Grouped launches: many independent problems of uneven shapes in one persistent launch.
//
// Grouped GEMM: problem g computes C_g[:,:,l] = alpha*A_g[:,:,l]*B_g[:,:,l] + beta*C_g[:,:,l]
// (column-major, dense, no transposes; A m x k, B k x n).  Every problem is cut into output
// tiles; the tiles of all problems form one list, ordered by problem and cut into one range
// per XCD (the host deals the problems, longest K first, to the least loaded XCD, so each
// range holds whole problems of every cost).  The workgroups of XCD c walk its range
// round-robin and find the problem of a tile in a tile -> problem table (one load, not a
// search); a problem's operand panels are then read into that XCD's L2 only.
//
//   Float32:          MFMA v_mfma_f32_32x32x2_f32, workgroup tile TM x TN (4 waves, 2x2),
//                     K step BK staged through LDS with a register prefetch of the next step;
//                     at BF16x3 precision the same tiles split into bf16 hi/lo while staged
//                     and run on v_mfma_f32_32x32x16_bf16.
//   other types:      semiring VALU tile 64 x 64 (4 x 4 outputs per thread), as in the
//                     strided GEMM (exact Int64, max-plus, complex).
//
// Grouped permute: problem g writes y_g = sx*permutedims(x_g, perm_g) + sy*y_g.  Each problem
// is a list of 64 x 64 LDS-transposed tiles (when the input's contiguous dim moves) or of
// 1024-element runs along the output's fastest dim; the blocks of all problems form one grid.
#include <type_traits>

#include "kernels.h"

namespace ome {

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Loads and stores through global-address-space pointers: the problem descriptors hold generic
// pointers, and generic (flat) accesses would also count against the LDS wait counter.
typedef float gg_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gg_load4(const float* p) {
  const gg_f4 v = *(const __attribute__((address_space(1))) gg_f4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float gg_load(const float* p) {
  return *(const __attribute__((address_space(1))) float*)p;
}
__device__ __forceinline__ float* gg_global(float* p) {
  return (float*)(__attribute__((address_space(1))) float*)p;
}

// One K-step of an operand tile as float4 groups q (A: 4 rows m..m+3 at one k; B: 4 k at one
// n).  Every lane loads: elements outside the problem read the operand's first element and
// their bit (4q + i) is left clear in the returned mask; gg_stage_zero applies the zeros when
// the step is staged, so no wait follows the loads.
template <int NQ>
__device__ __forceinline__ uint32_t gg_load_a(const float* A, int64_t M, int64_t K, bool vec,
                                              int64_t m0, int64_t k0, int TM, float4 (&ra)[NQ]) {
  uint32_t mask = 0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int e = threadIdx.x + 256 * q;
    const int64_t m = m0 + (e % (TM / 4)) * 4, k = k0 + e / (TM / 4);
    if (vec) {  // M % 4 == 0: a float4 is wholly inside or wholly outside
      const bool ok = m < M && k < K;
      ra[q] = gg_load4(A + (ok ? m + k * M : 0));
      mask |= ok ? 0xFu << (4 * q) : 0u;
    } else {
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool ok = m + i < M && k < K;
        v[i] = gg_load(A + (ok ? m + i + k * M : 0));
        mask |= ok ? 1u << (4 * q + i) : 0u;
      }
      ra[q] = make_float4(v[0], v[1], v[2], v[3]);
    }
  }
  return mask;
}
template <int NQ>
__device__ __forceinline__ uint32_t gg_load_b(const float* B, int64_t K, int64_t N, bool vec,
                                              int64_t n0, int64_t k0, int BK, float4 (&rb)[NQ]) {
  uint32_t mask = 0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int e = threadIdx.x + 256 * q;
    const int64_t k = k0 + (e % (BK / 4)) * 4, n = n0 + e / (BK / 4);
    if (vec) {  // K % 4 == 0
      const bool ok = k < K && n < N;
      rb[q] = gg_load4(B + (ok ? k + n * K : 0));
      mask |= ok ? 0xFu << (4 * q) : 0u;
    } else {
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool ok = k + i < K && n < N;
        v[i] = gg_load(B + (ok ? k + i + n * K : 0));
        mask |= ok ? 1u << (4 * q + i) : 0u;
      }
      rb[q] = make_float4(v[0], v[1], v[2], v[3]);
    }
  }
  return mask;
}
// Buffer resources over one operand slice (raw: no stride, range-checked against `bytes`).
// Offsets at or past GG_OOB lie outside every slice the host admits (< 2^31 bytes).
typedef uint32_t gg_u4 __attribute__((ext_vector_type(4)));
constexpr uint32_t GG_OOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t gg_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}

// A: EPT consecutive k (from k) of row m through rs; a k past K lands past the slice (zero),
// and a row past M starts past it and stays there.
template <int EPT>
__device__ __forceinline__ void gg_buf_load_a(__amdgpu_buffer_rsrc_t rs, bool mok, int64_t m,
                                              int64_t k, int64_t M, float (&ra)[EPT]) {
  uint32_t off = mok ? (uint32_t)(m + k * M) * 4u : GG_OOB;
  const uint32_t inc = mok ? (uint32_t)M * 4u : 0u;
#pragma unroll
  for (int e = 0; e < EPT; ++e, off += inc)
    ra[e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
}
// B: the float4 groups of gg_load_b (4 k at one n) through rs; a column past N lands past the
// slice, and a k past K is sent past it explicitly (it would reach the next column).
template <int NQ, int BK>
__device__ __forceinline__ void gg_buf_load_b(__amdgpu_buffer_rsrc_t rs, bool vec, int64_t K,
                                              int64_t n0, int64_t k0, float4 (&rb)[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int e = threadIdx.x + 256 * q;
    const int64_t k = k0 + (e % (BK / 4)) * 4, n = n0 + e / (BK / 4);
    if (vec) {  // K % 4 == 0: a float4 is wholly inside or wholly outside
      const uint32_t o = k < K ? (uint32_t)(k + n * K) * 4u : GG_OOB;
      const gg_u4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)o, 0, 0);
      rb[q] = make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                          __uint_as_float(v.w));
    } else {
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t o = k + i < K ? (uint32_t)(k + i + n * K) * 4u : GG_OOB;
        v[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)o, 0, 0));
      }
      rb[q] = make_float4(v[0], v[1], v[2], v[3]);
    }
  }
}
// C: the 32x32 MFMA blocks of a wave (lane: row m, columns 8 (q / 4) + 4 (lane / 32) + q % 4)
// through a resource over the C slice; a column past N lands past it (the store is dropped),
// a row past M is sent past it explicitly.
template <int NB, int MB>
__device__ __forceinline__ void gg_buf_store_c(float* C, int64_t M, int64_t N, int64_t m0,
                                               int64_t n0, int lane, int mode, float alpha,
                                               float beta, const floatx16 (&acc)[NB][MB]) {
  const __amdgpu_buffer_rsrc_t rsc = gg_rsrc(C, (uint32_t)(M * N * 4));
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int64_t m = m0 + mb * 32 + (lane & 31);
      const int64_t nq = n0 + nb * 32 + 4 * (lane >> 5);
      const uint32_t base = m < M ? (uint32_t)(m + nq * M) * 4u : GG_OOB;
      const uint32_t inc = m < M ? (uint32_t)M * 4u : 0u;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const uint32_t o = base + (uint32_t)(8 * (q >> 2) + (q & 3)) * inc;
        const float v = acc[nb][mb][q];
        float y = 0.f;
        if (mode != EPI_OVERWRITE)
          y = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsc, (int)o, 0, 0));
        const float r = mode == EPI_OVERWRITE ? alpha * v
                        : mode == EPI_SCALE_ONLY ? beta * y : beta * y + alpha * v;
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(r), rsc, (int)o, 0, 0);
      }
    }
}

template <int NQ>
__device__ __forceinline__ void gg_stage_zero(float4 (&r)[NQ], uint32_t mask) {
  if (mask == (NQ == 8 ? 0xFFFFFFFFu : (1u << (4 * NQ)) - 1u)) return;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    r[q].x = (mask >> (4 * q)) & 1u ? r[q].x : 0.f;
    r[q].y = (mask >> (4 * q + 1)) & 1u ? r[q].y : 0.f;
    r[q].z = (mask >> (4 * q + 2)) & 1u ? r[q].z : 0.f;
    r[q].w = (mask >> (4 * q + 3)) & 1u ? r[q].w : 0.f;
  }
}

// ---- grouped GEMM, Float32 on the matrix cores -------------------------------------------

// BUF: operands and C through buffer resources, A staged as rows of consecutive k (as the
// BF16x3 kernel below does); otherwise masked loads through pointers.
template <int TM, int TN, int BK, bool BUF>
__global__ __launch_bounds__(256) void k_grouped_f32(GGDesc g, const GGProb* __restrict__ probs,
                                                     const int64_t* __restrict__ first,
                                                     const int32_t* __restrict__ map) {
  constexpr int LDK = BK + 4;
  constexpr int MB = TM / 64, NB = TN / 64;       // 32x32 blocks per wave along m / n
  constexpr int NA = TM * BK / 1024;              // float4 loads per thread, A tile
  constexpr int NBL = TN * BK / 1024;             // float4 loads per thread, B tile
  static_assert(NA >= 1 && NBL >= 1, "tile too small for 256 threads");
  __shared__ __attribute__((aligned(16))) float sA[TM][LDK];
  __shared__ __attribute__((aligned(16))) float sB[TN][LDK];
  const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = (w & 1) * (TM / 2), wn = (w >> 1) * (TN / 2);
  const float alpha = (float)g.alpha.re, beta = (float)g.beta.re;
  const int xc = blockIdx.x % GG_XCDS;
  const int64_t xstep = gridDim.x / GG_XCDS;
  for (int64_t tile = g.xcd_first[xc] + blockIdx.x / GG_XCDS; tile < g.xcd_first[xc + 1];
       tile += xstep) {
    // the tables are restrict arguments: the kernel's stores cannot clobber them, so they are
    // read with scalar loads and every per-problem branch below is uniform
    const int p = map[tile];
    const GGProb pr = probs[p];
    int64_t batch, m0, n0;
    if constexpr (BUF) {  // the problem's tile count fits 32 bits: invariant division
      const uint32_t r = (uint32_t)(tile - first[p]);
      const uint32_t b = fdiv_q(r, pr.per_div), rr = r - b * pr.per_div.d;
      const uint32_t tn = fdiv_q(rr, pr.tm_div);
      batch = b;
      m0 = (int64_t)(rr - tn * pr.tm_div.d) * TM;
      n0 = (int64_t)tn * TN;
    } else {
      const int64_t r = tile - first[p];
      const int64_t per = (int64_t)pr.tiles_m * pr.tiles_n;
      batch = r / per;
      const int64_t rr = r - batch * per;
      m0 = (rr % pr.tiles_m) * TM;
      n0 = (rr / pr.tiles_m) * TN;
    }
    const int64_t M = pr.m, N = pr.n, K = pr.k;
    const float* __restrict__ A = (const float*)pr.a + batch * M * K;
    const float* __restrict__ B = (const float*)pr.b + batch * K * N;
    floatx16 acc[NB][MB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < MB; ++j)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.f;
    // A tile (m fastest in memory): load q covers m4 = (e % (TM/4))*4, k = e / (TM/4);
    // with BUF, thread t holds EPT consecutive k of row am instead
    // B tile (k fastest in memory): load q covers k4 = (e % (BK/4))*4, n = e / (BK/4)
    constexpr int EPT = TM * BK / 256;
    static_assert(EPT % 4 == 0 && BK % EPT == 0, "A row staging shape");
    const int am = t % TM, ak = (t / TM) * EPT;
    float4 ra[NA], rb[NBL];
    float rr_a[EPT];
    uint32_t mka = 0, mkb = 0;
    __amdgpu_buffer_rsrc_t rsa, rsb;
    if constexpr (BUF) {
      rsa = gg_rsrc(A, (uint32_t)(M * K * 4));
      rsb = gg_rsrc(B, (uint32_t)(K * N * 4));
    }
    const bool mok = m0 + am < M;
    auto load = [&](int64_t k0) __attribute__((always_inline)) {
      if constexpr (BUF) {
        gg_buf_load_a<EPT>(rsa, mok, m0 + am, k0 + ak, M, rr_a);
        gg_buf_load_b<NBL, BK>(rsb, pr.vec_b, K, n0, k0, rb);
      } else {
        mka = gg_load_a<NA>(A, M, K, pr.vec_a, m0, k0, TM, ra);
        mkb = gg_load_b<NBL>(B, K, N, pr.vec_b, n0, k0, BK, rb);
      }
    };
    // an empty sum (K == 0) loads nothing: its operands may hold no element at all
    if (K > 0) load(0);
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
      if constexpr (BUF) {
#pragma unroll
        for (int c = 0; c < EPT / 4; ++c)
          *(float4*)&sA[am][ak + 4 * c] =
              make_float4(rr_a[4 * c], rr_a[4 * c + 1], rr_a[4 * c + 2], rr_a[4 * c + 3]);
      } else {
        gg_stage_zero<NA>(ra, mka);
        gg_stage_zero<NBL>(rb, mkb);
#pragma unroll
        for (int q = 0; q < NA; ++q) {
          const int e = t + 256 * q;
          const int m = (e % (TM / 4)) * 4, k = e / (TM / 4);
          sA[m][k] = ra[q].x;
          sA[m + 1][k] = ra[q].y;
          sA[m + 2][k] = ra[q].z;
          sA[m + 3][k] = ra[q].w;
        }
      }
#pragma unroll
      for (int q = 0; q < NBL; ++q) {
        const int e = t + 256 * q;
        const int k = (e % (BK / 4)) * 4, n = e / (BK / 4);
        *(float4*)&sB[n][k] = rb[q];
      }
      __syncthreads();
      if (k0 + BK < K) load(k0 + BK);
      // lanes 0-31 take k = kb + s, lanes 32-63 k = kb + 4 + s in the s-th MFMA of a group;
      // a wave whose whole block lies outside the problem skips its MFMAs, and the 8-deep
      // groups past K (the zero padding of the last step) are skipped
      if (m0 + wm < M && n0 + wn < N) {
#pragma unroll
        for (int kb = 0; kb < BK; kb += 8) {
          if (k0 + kb >= K) break;
          const int kk = kb + 4 * (lane >> 5);
          float4 af[NB], bf[MB];
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) af[nb] = *(const float4*)&sB[wn + nb * 32 + (lane & 31)][kk];
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) bf[mb] = *(const float4*)&sA[wm + mb * 32 + (lane & 31)][kk];
#pragma unroll
          for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
              acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].x, bf[mb].x, acc[nb][mb], 0, 0, 0);
              acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].y, bf[mb].y, acc[nb][mb], 0, 0, 0);
              acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].z, bf[mb].z, acc[nb][mb], 0, 0, 0);
              acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].w, bf[mb].w, acc[nb][mb], 0, 0, 0);
            }
        }
      }
      __syncthreads();
    }
    // lane holds m = m0 + wm + mb*32 + (lane & 31), n = n0 + wn + nb*32 + 8*(q/4) + 4*(lane/32) + q%4
    float* __restrict__ C = (float*)pr.c + batch * M * N;
    if constexpr (BUF) {
      gg_buf_store_c<NB, MB>(C, M, N, m0 + wm, n0 + wn, lane, g.mode, alpha, beta, acc);
    } else {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int64_t m = m0 + wm + mb * 32 + (lane & 31);
            const int64_t n = n0 + wn + nb * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            if (m < M && n < N) {
              float* yp = gg_global(C + m + n * M);
              *yp = epilogue<OpsF32>(g.mode, alpha, beta, acc[nb][mb][q], yp);
            }
          }
    }
  }
}

// ---- grouped GEMM, Float32 as BF16x3 on the matrix cores ---------------------------------
// Operands are split while they are staged: x = hi + lo (bf16, round to nearest even), and
// each 16-deep K sub-step runs al*bh + ah*bl + ah*bh on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation.  K step 32; LDS rows hold 32 bf16 of one row (+ 8 pad: 80-byte rows).

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t gg_f2bf(float x) {  // round to nearest even, in hardware
  const __bf16 h = (__bf16)x;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ void gg_split(float x, uint16_t& hi, uint16_t& lo) {
  hi = gg_f2bf(x);
  lo = gg_f2bf(x - __uint_as_float((uint32_t)hi << 16));
}

// 64 x 64 tiles at five waves per SIMD with buffer operands (at most 102 VGPRs, 94 used), four
// otherwise (121 used): the many short tiles are latency-bound.  128 x 128 tiles keep the
// compiler's choice.
// BUF: every problem's A, B and C batch slices span less than 2^31 bytes and M, N < 2^24 (the
// host checks): operands are read through buffer resources with 32-bit offsets, whose range
// check returns zero past the slice, and C is written through one, whose range check drops
// the stores past it; only a row past M (A, C) or a k past K (B) needs an explicit check.
// BK: k per step (32, or 64 for half the round trips of a short tile at the cost of LDS).
// NP: bf16 planes per operand: 2 (BF16x3: three products) or 3 (BF16x6: hi, mid, lo and the
// six products a_p*b_q with p + q <= 2, smallest first; FP32-level error).
template <int TM, int TN, bool BUF, int BK, int NP>
__device__ __forceinline__ void grouped_split_body(const GGDesc& g, const GGProb* __restrict__ probs,
                                                   const int64_t* __restrict__ first,
                                                   const int32_t* __restrict__ map) {
  constexpr int LDK = BK + 8;  // row pitch 80 or 144 bytes: 8 rows hit 8 distinct bank groups
  constexpr int MB = TM / 64, NB = TN / 64;
  constexpr int NBL = TN * BK / 1024;
  // A: thread t loads EPT consecutive k of row am (lanes along m: each load is one coalesced
  // run of the column) and stores their hi / lo halves as 16-byte rows of sA, whose 80-byte
  // row pitch puts eight consecutive rows in distinct bank groups
  constexpr int EPT = TM * BK / 256, KQ = BK / EPT;
  static_assert(TM * KQ == 256 && EPT % 8 == 0 && EPT <= 32, "A staging shape");
  __shared__ __attribute__((aligned(16))) uint16_t sA[NP][TM][LDK];  // [plane][m][k]
  __shared__ __attribute__((aligned(16))) uint16_t sB[NP][TN][LDK];  // [plane][n][k]
  const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = (w & 1) * (TM / 2), wn = (w >> 1) * (TN / 2);
  const float alpha = (float)g.alpha.re, beta = (float)g.beta.re;
  const int xc = blockIdx.x % GG_XCDS;
  const int64_t xstep = gridDim.x / GG_XCDS;
  for (int64_t tile = g.xcd_first[xc] + blockIdx.x / GG_XCDS; tile < g.xcd_first[xc + 1];
       tile += xstep) {
    // the tables are restrict arguments: the kernel's stores cannot clobber them, so they are
    // read with scalar loads and every per-problem branch below is uniform
    const int p = map[tile];
    const GGProb pr = probs[p];
    int64_t batch, m0, n0;
    if constexpr (BUF) {  // the problem's tile count fits 32 bits: invariant division
      const uint32_t r = (uint32_t)(tile - first[p]);
      const uint32_t b = fdiv_q(r, pr.per_div), rr = r - b * pr.per_div.d;
      const uint32_t tn = fdiv_q(rr, pr.tm_div);
      batch = b;
      m0 = (int64_t)(rr - tn * pr.tm_div.d) * TM;
      n0 = (int64_t)tn * TN;
    } else {
      const int64_t r = tile - first[p];
      const int64_t per = (int64_t)pr.tiles_m * pr.tiles_n;
      batch = r / per;
      const int64_t rr = r - batch * per;
      m0 = (rr % pr.tiles_m) * TM;
      n0 = (rr / pr.tiles_m) * TN;
    }
    const int64_t M = pr.m, N = pr.n, K = pr.k;
    const float* __restrict__ A = (const float*)pr.a + batch * M * K;
    const float* __restrict__ B = (const float*)pr.b + batch * K * N;
    floatx16 acc[NB][MB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < MB; ++j)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.f;
    const int am = t % TM, ak = (t / TM) * EPT;
    float ra[EPT];
    float4 rb[NBL];
    uint32_t mka = 0, mkb = 0;
    __amdgpu_buffer_rsrc_t rsa, rsb;
    if constexpr (BUF) {
      rsa = gg_rsrc(A, (uint32_t)(M * K * 4));
      rsb = gg_rsrc(B, (uint32_t)(K * N * 4));
    }
    const bool mok = m0 + am < M;
    auto load = [&](int64_t k0) __attribute__((always_inline)) {
      if constexpr (BUF) {
        gg_buf_load_a<EPT>(rsa, mok, m0 + am, k0 + ak, M, ra);
        gg_buf_load_b<NBL, BK>(rsb, pr.vec_b, K, n0, k0, rb);
      } else {
        // out-of-range elements read A's first element; their mask bit stays clear
        mka = 0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int64_t k = k0 + ak + e;
          const bool ok = mok && k < K;
          ra[e] = gg_load(A + (ok ? m0 + am + k * M : 0));
          mka |= ok ? 1u << e : 0u;
        }
        mkb = gg_load_b<NBL>(B, K, N, pr.vec_b, n0, k0, BK, rb);
      }
    };
    // an empty sum (K == 0) loads nothing: its operands may hold no element at all
    if (K > 0) load(0);
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
      if constexpr (!BUF) {
        if (mka != (EPT == 32 ? 0xFFFFFFFFu : (1u << EPT) - 1u)) {
#pragma unroll
          for (int e = 0; e < EPT; ++e) ra[e] = (mka >> e) & 1u ? ra[e] : 0.f;
        }
        gg_stage_zero<NBL>(rb, mkb);
      }
#pragma unroll
      for (int c = 0; c < EPT / 8; ++c) {
        uint4 p0, p1, p2;
        if constexpr (NP == 2) {
          split2_bf16(ra[8 * c + 0], ra[8 * c + 1], p0.x, p1.x);
          split2_bf16(ra[8 * c + 2], ra[8 * c + 3], p0.y, p1.y);
          split2_bf16(ra[8 * c + 4], ra[8 * c + 5], p0.z, p1.z);
          split2_bf16(ra[8 * c + 6], ra[8 * c + 7], p0.w, p1.w);
        } else {
          split3_bf16(ra[8 * c + 0], ra[8 * c + 1], p0.x, p1.x, p2.x);
          split3_bf16(ra[8 * c + 2], ra[8 * c + 3], p0.y, p1.y, p2.y);
          split3_bf16(ra[8 * c + 4], ra[8 * c + 5], p0.z, p1.z, p2.z);
          split3_bf16(ra[8 * c + 6], ra[8 * c + 7], p0.w, p1.w, p2.w);
          *(uint4*)&sA[NP - 1][am][ak + 8 * c] = p2;
        }
        *(uint4*)&sA[0][am][ak + 8 * c] = p0;
        *(uint4*)&sA[1][am][ak + 8 * c] = p1;
      }
#pragma unroll
      for (int q = 0; q < NBL; ++q) {
        const int e = t + 256 * q;
        const int k = (e % (BK / 4)) * 4, n = e / (BK / 4);
        uint2 p0, p1, p2;
        if constexpr (NP == 2) {
          split2_bf16(rb[q].x, rb[q].y, p0.x, p1.x);
          split2_bf16(rb[q].z, rb[q].w, p0.y, p1.y);
        } else {
          split3_bf16(rb[q].x, rb[q].y, p0.x, p1.x, p2.x);
          split3_bf16(rb[q].z, rb[q].w, p0.y, p1.y, p2.y);
          *(uint2*)&sB[NP - 1][n][k] = p2;
        }
        *(uint2*)&sB[0][n][k] = p0;
        *(uint2*)&sB[1][n][k] = p1;
      }
      __syncthreads();
      if (k0 + BK < K) load(k0 + BK);
      if (m0 + wm < M && n0 + wn < N) {
#pragma unroll
        for (int s2 = 0; s2 < BK / 16; ++s2) {
          if (k0 + 16 * s2 >= K) break;  // a 16-deep sub-step wholly in the padding
          // lane l holds row l % 32, k = 16*s2 + 8*(l / 32) + 0..7; fn: plane p of the B
          // tile (rows n, the MFMA's first operand), fm: plane p of the A tile (rows m)
          const int kk = 16 * s2 + 8 * (lane >> 5);
          bf16x8_t fn[NP][NB], fm[NP][MB];
#pragma unroll
          for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
              fn[p][nb] = *(const bf16x8_t*)&sB[p][wn + nb * 32 + (lane & 31)][kk];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
              fm[p][mb] = *(const bf16x8_t*)&sA[p][wm + mb * 32 + (lane & 31)][kk];
          }
          // products a_p * b_q, smallest first: (1,0) (0,1) (0,0) for two planes;
          // (2,0) (1,1) (0,2) (1,0) (0,1) (0,0) for three
          constexpr int NQ = NP == 2 ? 3 : 6;
          constexpr int PN2[3] = {1, 0, 0}, PM2[3] = {0, 1, 0};
          constexpr int PN3[6] = {2, 1, 0, 1, 0, 0}, PM3[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
              for (int mb = 0; mb < MB; ++mb)
                acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    fn[NP == 2 ? PN2[q % 3] : PN3[q]][nb], fm[NP == 2 ? PM2[q % 3] : PM3[q]][mb],
                    acc[nb][mb], 0, 0, 0);
        }
      }
      __syncthreads();
    }
    float* __restrict__ C = (float*)pr.c + batch * M * N;
    if constexpr (BUF) {
      gg_buf_store_c<NB, MB>(C, M, N, m0 + wm, n0 + wn, lane, g.mode, alpha, beta, acc);
    } else {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int64_t m = m0 + wm + mb * 32 + (lane & 31);
            const int64_t n = n0 + wn + nb * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            if (m < M && n < N) {
              float* yp = gg_global(C + m + n * M);
              *yp = epilogue<OpsF32>(g.mode, alpha, beta, acc[nb][mb][q], yp);
            }
          }
    }
  }
}

template <int TM, int TN, bool BUF, int BK = 32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TM == 64 && BK == 32 ? (BUF ? 6 : 4) : 1, TM == 64 && BK == 32 ? (BUF ? 6 : 4) : 8))) void k_grouped_bf16x3(GGDesc g, const GGProb* __restrict__ probs,
                                                        const int64_t* __restrict__ first,
                                                        const int32_t* __restrict__ map) {
  grouped_split_body<TM, TN, BUF, BK, 2>(g, probs, first, map);
}

// The six-product form, 64 x 64 tiles, K step 32 (30 KiB of LDS).
template <bool BUF>
__global__ __launch_bounds__(256) void k_grouped_bf16x6(GGDesc g, const GGProb* __restrict__ probs,
                                                        const int64_t* __restrict__ first,
                                                        const int32_t* __restrict__ map) {
  grouped_split_body<64, 64, BUF, 32, 3>(g, probs, first, map);
}

// ---- grouped GEMM, persistent and pipelined across tiles ----------------------------------
// gg_run_per_cu(X3) workgroups per CU walk their XCD's tile range round-robin (workgroup j of an
// XCD takes its tiles j, j + W, j + 2W, ...), so the tiles in flight on an XCD stay neighbours
// and their problems' operand panels share its L2.  The K steps of a workgroup's tiles form
// one sequence with one step in flight: the first step of the next tile is fetched as soon as
// the current tile's C block is written, while the other workgroups of the CU run.  The next
// tile's problem index is read a tile ahead, so a tile boundary adds one scalar read (the
// problem's descriptor, an L2 hit) instead of a workgroup launch, a chain of descriptor
// reads and a first operand round trip.  64 x 64 tiles, 32-deep K
// steps, operands and C through buffer resources (the host admits only lists that fit 32-bit
// offsets); X3: BF16x3 products, otherwise FP32 MFMA.
template <bool X3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gg_run_per_cu(X3), gg_run_per_cu(X3)))) void k_grouped_run(GGDesc g, const GGProb* __restrict__ probs,
                                                     const int32_t* __restrict__ map) {
  constexpr int TM = 64, TN = 64, BK = 32;
  constexpr int EPT = TM * BK / 256, NBL = TN * BK / 1024;  // 8 k of one A row; 2 float4 of B
  constexpr int LDK = X3 ? BK + 8 : BK + 4;
  using LT = std::conditional_t<X3, uint16_t, float>;
  __shared__ __attribute__((aligned(16))) LT sA[X3 ? 2 : 1][TM][LDK];  // [hi/lo][m][k]
  __shared__ __attribute__((aligned(16))) LT sB[X3 ? 2 : 1][TN][LDK];  // [hi/lo][n][k]
  const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = (w & 1) * (TM / 2), wn = (w >> 1) * (TN / 2);
  const int am = t % TM, ak = (t / TM) * EPT;
  const float alpha = (float)g.alpha.re, beta = (float)g.beta.re;
  const int xc = blockIdx.x % GG_XCDS;
  const int64_t W = gridDim.x / GG_XCDS, xend = g.xcd_first[xc + 1];
  int64_t tile = g.xcd_first[xc] + blockIdx.x / GG_XCDS;
  if (tile >= xend) return;
  // load side: the batch slices, sizes and tile corner whose K step lk0 is being fetched
  // (sizes below 2^22 and slices below 2^31 bytes: 32-bit arithmetic)
  const float *lA, *lB;
  float* lC;
  uint32_t lM, lN, lK, lm0, ln0, lvec;
  auto place = [&](const GGProb& pr) __attribute__((always_inline)) {
    const uint32_t r = (uint32_t)(tile - pr.first_tile);
    const uint32_t b = fdiv_q(r, pr.per_div), q = r - b * pr.per_div.d;
    const uint32_t tn = fdiv_q(q, pr.tm_div);
    lM = (uint32_t)pr.m;
    lN = (uint32_t)pr.n;
    lK = (uint32_t)pr.k;
    lm0 = (q - tn * pr.tm_div.d) * TM;
    ln0 = tn * TN;
    lvec = (uint32_t)pr.vec_b;
    lA = (const float*)pr.a + (int64_t)b * lM * lK;
    lB = (const float*)pr.b + (int64_t)b * lK * lN;
    lC = (float*)pr.c + (int64_t)b * lM * lN;
  };
  place(probs[map[tile]]);
  int64_t ntile = tile + W;  // the next tile and its problem (read a tile ahead)
  int np = ntile < xend ? map[ntile] : 0;
  uint32_t lk0 = 0;
  float ra[EPT];
  float4 rb[NBL];
  auto load = [&]() __attribute__((always_inline)) {
    gg_buf_load_a<EPT>(gg_rsrc(lA, lM * lK * 4u), lm0 + am < lM, lm0 + am, lk0 + ak, lM, ra);
    gg_buf_load_b<NBL, BK>(gg_rsrc(lB, lK * lN * 4u), lvec != 0, lK, ln0, lk0, rb);
  };
  load();
  // compute side: the tile whose step was fetched last (one step behind the load side)
  float* cC;
  uint32_t cM, cN, cm0, cn0;
  auto take = [&]() __attribute__((always_inline)) {
    cM = lM;
    cN = lN;
    cm0 = lm0;
    cn0 = ln0;
    cC = lC;
  };
  take();
  floatx16 acc;
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.f;
  for (;;) {
    const bool last = lk0 + BK >= lK;  // the step in registers ends its tile
    const uint32_t kv = lK > lk0 ? lK - lk0 : 0;  // k of this step inside the problem
    const bool more = !last || ntile < xend;
    // the step in registers -> LDS (BF16x3: split into hi and lo halves)
    if constexpr (X3) {
      uint4 ph, pl;
      split2_bf16(ra[0], ra[1], ph.x, pl.x);
      split2_bf16(ra[2], ra[3], ph.y, pl.y);
      split2_bf16(ra[4], ra[5], ph.z, pl.z);
      split2_bf16(ra[6], ra[7], ph.w, pl.w);
      *(uint4*)&sA[0][am][ak] = ph;
      *(uint4*)&sA[1][am][ak] = pl;
#pragma unroll
      for (int q = 0; q < NBL; ++q) {
        const int e = t + 256 * q;
        const int k = (e % (BK / 4)) * 4, n = e / (BK / 4);
        uint2 bh, bl;
        split2_bf16(rb[q].x, rb[q].y, bh.x, bl.x);
        split2_bf16(rb[q].z, rb[q].w, bh.y, bl.y);
        *(uint2*)&sB[0][n][k] = bh;
        *(uint2*)&sB[1][n][k] = bl;
      }
    } else {
      *(float4*)&sA[0][am][ak] = make_float4(ra[0], ra[1], ra[2], ra[3]);
      *(float4*)&sA[0][am][ak + 4] = make_float4(ra[4], ra[5], ra[6], ra[7]);
#pragma unroll
      for (int q = 0; q < NBL; ++q) {
        const int e = t + 256 * q;
        *(float4*)&sB[0][e / (BK / 4)][(e % (BK / 4)) * 4] = rb[q];
      }
    }
    __syncthreads();
    if (!last) {  // the next step of this tile
      lk0 += BK;
      load();
    }
    // a wave whose 32 x 32 block lies outside the problem skips its MFMAs
    if (cm0 + wm < cM && cn0 + wn < cN) {
      if constexpr (X3) {
#pragma unroll
        for (int s2 = 0; s2 < BK / 16; ++s2) {
          if (16u * s2 >= kv) break;
          const int kk = 16 * s2 + 8 * (lane >> 5);
          const bf16x8_t ah = *(const bf16x8_t*)&sB[0][wn + (lane & 31)][kk];
          const bf16x8_t al = *(const bf16x8_t*)&sB[1][wn + (lane & 31)][kk];
          const bf16x8_t bh = *(const bf16x8_t*)&sA[0][wm + (lane & 31)][kk];
          const bf16x8_t bl = *(const bf16x8_t*)&sA[1][wm + (lane & 31)][kk];
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int kb = 0; kb < BK; kb += 8) {
          if ((uint32_t)kb >= kv) break;
          const int kk = kb + 4 * (lane >> 5);
          const float4 af = *(const float4*)&sB[0][wn + (lane & 31)][kk];
          const float4 bf = *(const float4*)&sA[0][wm + (lane & 31)][kk];
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(af.x, bf.x, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(af.y, bf.y, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(af.z, bf.z, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(af.w, bf.w, acc, 0, 0, 0);
        }
      }
    }
    if (last) {
      // the C block goes out before the next tile's first step is fetched: stores and loads
      // retire in issue order, so the wait for that step then waits for no store
      const floatx16 (&a1)[1][1] = *(const floatx16(*)[1][1])&acc;
      gg_buf_store_c<1, 1>(cC, cM, cN, (int64_t)(cm0 + wm), (int64_t)(cn0 + wn), lane, g.mode, alpha,
                           beta, a1);
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] = 0.f;
      if (more) {
        tile = ntile;
        place(probs[np]);
        lk0 = 0;
        load();
        take();
        ntile = tile + W;
        if (ntile < xend) np = map[ntile];
      }
    }
    __syncthreads();
    if (!more) break;
  }
}

// ---- grouped GEMM, one wave per 32 x 32 tile ----------------------------------------------
// No LDS and no barriers: every wave owns a tile and loads its MFMA fragments straight from
// global memory (A along m: 32 lanes read 128 contiguous bytes per k; B along k: 16 bytes per
// lane), D K-steps ahead in registers, so many independent loads are in flight per SIMD.
// Lane l holds fragment row l % 32 (n for the B tile, m for the A tile); lanes 0-31 take
// k = k0 + s and lanes 32-63 k = k0 + 4 + s in the s-th FP32 MFMA of a K-step of 8.

template <int D>
__global__ __launch_bounds__(256) void k_grouped_f32_wave(GGDesc g) {
  const int lane = threadIdx.x & 63;
  const int xc = blockIdx.x % GG_XCDS;
  const int64_t nw = (int64_t)(gridDim.x / GG_XCDS) * 4;
  const float alpha = (float)g.alpha.re, beta = (float)g.beta.re;
  for (int64_t tile = g.xcd_first[xc] + (int64_t)(blockIdx.x / GG_XCDS) * 4 + (threadIdx.x >> 6);
       tile < g.xcd_first[xc + 1]; tile += nw) {
    const int p = g.map[tile];
    const GGProb pr = g.probs[p];
    const int64_t r = tile - g.first[p];
    const int64_t per = (int64_t)pr.tiles_m * pr.tiles_n;
    const int64_t batch = r / per;
    const int64_t rr = r - batch * per;
    const int64_t m0 = (rr % pr.tiles_m) * 32, n0 = (rr / pr.tiles_m) * 32;
    const int64_t M = pr.m, N = pr.n, K = pr.k;
    const float* __restrict__ A = (const float*)pr.a + batch * M * K;
    const float* __restrict__ B = (const float*)pr.b + batch * K * N;
    const int64_t m = m0 + (lane & 31), n = n0 + (lane & 31);
    const int kq = 4 * (lane >> 5);
    const bool mok = m < M, nok = n < N;
    const float* __restrict__ ap = A + (mok ? m : 0);
    const float* __restrict__ bp = B + (nok ? n : 0) * K;
    float4 fa[D], fb[D];
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // rows outside the tile read row 0 (clamped pointers) and are zeroed after the load
    auto load = [&](int d, int64_t k0) __attribute__((always_inline)) {
      const int64_t k = k0 + kq;
      if (k + 4 <= K) {
        const float* bk = bp + k;
        const float4 v = pr.vec_b ? *(const float4*)bk : make_float4(bk[0], bk[1], bk[2], bk[3]);
        const float* ak = ap + k * M;
        const float4 u = make_float4(ak[0], ak[M], ak[2 * M], ak[3 * M]);
        fa[d] = nok ? v : z4;
        fb[d] = mok ? u : z4;
      } else {
        fa[d].x = (nok && k < K) ? bp[k] : 0.f;
        fa[d].y = (nok && k + 1 < K) ? bp[k + 1] : 0.f;
        fa[d].z = (nok && k + 2 < K) ? bp[k + 2] : 0.f;
        fa[d].w = (nok && k + 3 < K) ? bp[k + 3] : 0.f;
        fb[d].x = (mok && k < K) ? ap[k * M] : 0.f;
        fb[d].y = (mok && k + 1 < K) ? ap[(k + 1) * M] : 0.f;
        fb[d].z = (mok && k + 2 < K) ? ap[(k + 2) * M] : 0.f;
        fb[d].w = (mok && k + 3 < K) ? ap[(k + 3) * M] : 0.f;
      }
    };
    floatx16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) load(d, 8 * d);
    for (int64_t k0 = 0; k0 < K; k0 += 8 * D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (k0 + 8 * d < K) {
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[d].x, fb[d].x, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[d].y, fb[d].y, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[d].z, fb[d].z, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[d].w, fb[d].w, acc, 0, 0, 0);
          if (k0 + 8 * (d + D) < K) load(d, k0 + 8 * (d + D));
        }
      }
    }
    float* __restrict__ C = (float*)pr.c + batch * M * N;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int64_t nn = n0 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
      if (mok && nn < N) {
        float* yp = C + m + nn * M;
        *yp = epilogue<OpsF32>(g.mode, alpha, beta, acc[q], yp);
      }
    }
  }
}

// BF16x3 form: K-steps of 16, lane l holds k = k0 + 8*(l/32) + 0..7 of its row, split into
// bf16 hi/lo in registers; three v_mfma_f32_32x32x16_bf16 per step.
typedef __bf16 bf16x8w_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8w_t gg_pack8(const uint16_t (&h)[8]) {
  uint32_t u[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) u[i] = (uint32_t)h[2 * i] | ((uint32_t)h[2 * i + 1] << 16);
  bf16x8w_t v;
  __builtin_memcpy(&v, u, 16);
  return v;
}

template <int D>
__global__ __launch_bounds__(256) void k_grouped_bf16x3_wave(GGDesc g) {
  const int lane = threadIdx.x & 63;
  const int xc = blockIdx.x % GG_XCDS;
  const int64_t nw = (int64_t)(gridDim.x / GG_XCDS) * 4;
  const float alpha = (float)g.alpha.re, beta = (float)g.beta.re;
  for (int64_t tile = g.xcd_first[xc] + (int64_t)(blockIdx.x / GG_XCDS) * 4 + (threadIdx.x >> 6);
       tile < g.xcd_first[xc + 1]; tile += nw) {
    const int p = g.map[tile];
    const GGProb pr = g.probs[p];
    const int64_t r = tile - g.first[p];
    const int64_t per = (int64_t)pr.tiles_m * pr.tiles_n;
    const int64_t batch = r / per;
    const int64_t rr = r - batch * per;
    const int64_t m0 = (rr % pr.tiles_m) * 32, n0 = (rr / pr.tiles_m) * 32;
    const int64_t M = pr.m, N = pr.n, K = pr.k;
    const float* __restrict__ A = (const float*)pr.a + batch * M * K;
    const float* __restrict__ B = (const float*)pr.b + batch * K * N;
    const int64_t m = m0 + (lane & 31), n = n0 + (lane & 31);
    const int kq = 8 * (lane >> 5);
    const bool mok = m < M, nok = n < N;
    const float* __restrict__ ap = A + (mok ? m : 0);
    const float* __restrict__ bp = B + (nok ? n : 0) * K;
    float fa[D][8], fb[D][8];
    auto load = [&](int d, int64_t k0) __attribute__((always_inline)) {
      const int64_t k = k0 + kq;
      if (k + 8 <= K) {
        const float* bk = bp + k;
        if (pr.vec_b) {
          const float4 u = *(const float4*)bk, v = *(const float4*)(bk + 4);
          fa[d][0] = u.x; fa[d][1] = u.y; fa[d][2] = u.z; fa[d][3] = u.w;
          fa[d][4] = v.x; fa[d][5] = v.y; fa[d][6] = v.z; fa[d][7] = v.w;
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) fa[d][i] = bk[i];
        }
        const float* ak = ap + k * M;
#pragma unroll
        for (int i = 0; i < 8; ++i) fb[d][i] = ak[i * M];
        if (!nok)
#pragma unroll
          for (int i = 0; i < 8; ++i) fa[d][i] = 0.f;
        if (!mok)
#pragma unroll
          for (int i = 0; i < 8; ++i) fb[d][i] = 0.f;
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) fa[d][i] = (nok && k + i < K) ? bp[k + i] : 0.f;
#pragma unroll
        for (int i = 0; i < 8; ++i) fb[d][i] = (mok && k + i < K) ? ap[(k + i) * M] : 0.f;
      }
    };
    floatx16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) load(d, 16 * d);
    for (int64_t k0 = 0; k0 < K; k0 += 16 * D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (k0 + 16 * d < K) {
          uint16_t ah[8], al[8], bh[8], bl[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            gg_split(fa[d][i], ah[i], al[i]);
            gg_split(fb[d][i], bh[i], bl[i]);
          }
          const bf16x8w_t vah = gg_pack8(ah), val = gg_pack8(al), vbh = gg_pack8(bh), vbl = gg_pack8(bl);
          if (k0 + 16 * (d + D) < K) load(d, k0 + 16 * (d + D));
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(val, vbh, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(vah, vbl, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(vah, vbh, acc, 0, 0, 0);
        }
      }
    }
    float* __restrict__ C = (float*)pr.c + batch * M * N;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int64_t nn = n0 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
      if (mok && nn < N) {
        float* yp = C + m + nn * M;
        *yp = epilogue<OpsF32>(g.mode, alpha, beta, acc[q], yp);
      }
    }
  }
}

// ---- grouped GEMM, semiring VALU tile (any element type) --------------------------------

template <class O>
__global__ __launch_bounds__(256) void k_grouped_valu(GGDesc g, const GGProb* __restrict__ probs,
                                                      const int64_t* __restrict__ first,
                                                      const int32_t* __restrict__ map) {
  using T = typename O::T;
  constexpr int BM = 64, BN = 64, BK = 16;
  __shared__ T As[BK][BM + 1];
  __shared__ T Bs[BK][BN + 1];
  const int t = threadIdx.x, tm = t & 15, tn = t >> 4;
  const T alpha = O::from(g.alpha), beta = O::from(g.beta);
  const int xc = blockIdx.x % GG_XCDS;
  const int64_t xstep = gridDim.x / GG_XCDS;
  for (int64_t tile = g.xcd_first[xc] + blockIdx.x / GG_XCDS; tile < g.xcd_first[xc + 1];
       tile += xstep) {
    const int p = map[tile];
    const GGProb pr = probs[p];
    const int64_t r = tile - first[p];
    const int64_t per = (int64_t)pr.tiles_m * pr.tiles_n;
    const int64_t batch = r / per;
    const int64_t rr = r - batch * per;
    const int64_t m0 = (rr % pr.tiles_m) * BM, n0 = (rr / pr.tiles_m) * BN;
    const int64_t M = pr.m, N = pr.n, K = pr.k;
    const T* __restrict__ A = (const T*)pr.a + batch * M * K;
    const T* __restrict__ B = (const T*)pr.b + batch * K * N;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = O::zero();
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
      // the step's eight loads are issued before any is staged (lanes outside the problem load
      // its first element and stage a zero)
      T av[4], bv[4];
      uint32_t ok = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = t + 256 * q;
        const int64_t gm = m0 + (e & 63), gk = k0 + (e >> 6);  // A: m fastest
        const bool oka = gm < M && gk < K;
        av[q] = A[oka ? gm + gk * M : 0];
        const int64_t gn = n0 + (e >> 4), gkb = k0 + (e & 15);  // B: k fastest
        const bool okb = gn < N && gkb < K;
        bv[q] = B[okb ? gkb + gn * K : 0];
        ok |= (oka ? 1u : 0u) << q | (okb ? 16u : 0u) << q;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = t + 256 * q;
        As[e >> 6][e & 63] = (ok >> q) & 1u ? av[q] : O::zero();
        Bs[e & 15][e >> 4] = (ok >> (4 + q)) & 1u ? bv[q] : O::zero();
      }
      __syncthreads();
      const int kn = (int)((K - k0) < BK ? (K - k0) : BK);
      for (int kk = 0; kk < kn; ++kk) {
        T a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[kk][tm + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tn + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = O::fma(a[i], b[j], acc[i][j]);
      }
      __syncthreads();
    }
    T* __restrict__ C = (T*)pr.c + batch * M * N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t m = m0 + tm + 16 * i, n = n0 + tn + 16 * j;
        if (m < M && n < N) {
          T* yp = C + m + n * M;
          *yp = epilogue<O>(g.mode, alpha, beta, acc[i][j], yp);
        }
      }
  }
}

void gg_tile_shape(int dt, int variant, int& tm, int& tn) {
  if (dt != F32) { tm = tn = 64; return; }
  if (variant == GG_F32_W32 || variant == GG_BF16X3_W32) { tm = tn = 32; return; }
  if (variant == GG_F32_128 || variant == GG_BF16X3_128) { tm = tn = 128; return; }
  tm = tn = 64;
}

void launch_grouped_gemm(int dt, int variant, const GGDesc& g, int grid, hipStream_t st) {
  if (g.total_tiles <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_GROUPED, st);
  if (dt == F32) {
    // run variants: the host launches them only when every problem fits 32-bit offsets
    if (variant == GG_BF16X3_RUN)
      hipLaunchKernelGGL((k_grouped_run<true>), dim3(grid), dim3(256), 0, st, g, g.probs, g.map);
    else if (variant == GG_F32_RUN)
      hipLaunchKernelGGL((k_grouped_run<false>), dim3(grid), dim3(256), 0, st, g, g.probs, g.map);
    else if (variant == GG_F32_W32)
      hipLaunchKernelGGL((k_grouped_f32_wave<4>), dim3(grid), dim3(256), 0, st, g);
    else if (variant == GG_BF16X3_W32)
      hipLaunchKernelGGL((k_grouped_bf16x3_wave<2>), dim3(grid), dim3(256), 0, st, g);
    else if (variant == GG_BF16X6_64 && g.offsets32)
      hipLaunchKernelGGL((k_grouped_bf16x6<true>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_BF16X6_64)
      hipLaunchKernelGGL((k_grouped_bf16x6<false>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_BF16X3_64_BK64 && g.offsets32)
      hipLaunchKernelGGL((k_grouped_bf16x3<64, 64, true, 64>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_BF16X3_64_BK64)
      hipLaunchKernelGGL((k_grouped_bf16x3<64, 64, false, 64>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_BF16X3_64 && g.offsets32)
      hipLaunchKernelGGL((k_grouped_bf16x3<64, 64, true>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_BF16X3_64)
      hipLaunchKernelGGL((k_grouped_bf16x3<64, 64, false>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_BF16X3_128 && g.offsets32)
      hipLaunchKernelGGL((k_grouped_bf16x3<128, 128, true>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_BF16X3_128)
      hipLaunchKernelGGL((k_grouped_bf16x3<128, 128, false>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_F32_128 && g.offsets32)
      hipLaunchKernelGGL((k_grouped_f32<128, 128, 16, true>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_F32_128)
      hipLaunchKernelGGL((k_grouped_f32<128, 128, 16, false>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_F32_64_BK32 && g.offsets32)
      hipLaunchKernelGGL((k_grouped_f32<64, 64, 32, true>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (variant == GG_F32_64_BK32)
      hipLaunchKernelGGL((k_grouped_f32<64, 64, 32, false>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else if (g.offsets32)
      hipLaunchKernelGGL((k_grouped_f32<64, 64, 16, true>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    else
      hipLaunchKernelGGL((k_grouped_f32<64, 64, 16, false>), dim3(grid), dim3(256), 0, st, g, g.probs, g.first, g.map);
    return;
  }
  OME_DISPATCH(dt, O, {
    GGDesc gg = g;
    gg.mode = epi_mode<O>(g.alpha, g.beta);
    hipLaunchKernelGGL(k_grouped_valu<O>, dim3(grid), dim3(256), 0, st, gg, g.probs, g.first, g.map);
  });
}

// ---- grouped permute ----------------------------------------------------------------------

// SHARED: the instantiation for shared-dim problems (tiled == 2, launched over their block
// range); otherwise the one for the two-group tile (tiled == 1) and the plain walk.  Each
// branch gets its own register allocation and LDS footprint.
template <class O, bool SHARED>
__global__ __launch_bounds__(256) void k_grouped_permute(GPDesc g) {
  using T = typename O::T;
  // two-group tile: tla rows of tlb + 1 (tla <= 256, tla * tlb <= 4096); shared-dim tile: tla
  // rows of LDA = tsc * tlb + (1 or 2) with tla * tlb * tsc <= GP_SCAP_BYTES / sizeof(T)
  constexpr int SCAP = GP_SCAP_BYTES / (int)sizeof(T);
  constexpr int TILE = SHARED ? SCAP + 512 : 4096 + 256;
  __shared__ T tile[TILE];
  __shared__ int64_t soff_a[256], soff_b[SHARED ? 256 : 64];
  const T sx = O::from(g.sx), sy = O::from(g.sy);
  for (int64_t blk = g.blk_begin + blockIdx.x; blk < g.blk_end; blk += gridDim.x) {
    const int p = g.map[blk];
    const GPProb& pr = g.probs[p];
    const T* __restrict__ x = (const T*)pr.x;
    T* __restrict__ y = (T*)pr.y;
    int64_t b = blk - g.first[p];
    if constexpr (SHARED) {
      const uint32_t sk = (uint32_t)(b % pr.ts_tiles);
      uint32_t rest = (uint32_t)(b / pr.ts_tiles);
      int64_t ibase = 0, obase = 0;
      for (int q = 0; q < pr.nr; ++q) {
        const uint32_t qq = fdiv_q(rest, pr.rdv[q]);
        const uint32_t v = rest - qq * pr.rdv[q].d;
        rest = qq;
        ibase += (int64_t)v * pr.r_is[q];
        obase += (int64_t)v * pr.r_os[q];
      }
      const int tla = (int)pr.pa, tlb = (int)pr.pb, tsc = pr.tsc;
      const int64_t s0 = (int64_t)sk * tsc;
      const uint32_t nsv = (uint32_t)((pr.s_ext - s0 < tsc) ? pr.s_ext - s0 : tsc);
      ibase += s0 * pr.s_is;
      obase += s0 * pr.s_os;
      // input offset of (a, s, b) = b + s * s_is + fa(a); output offset = a + s * s_os + fb(b)
      for (int u = threadIdx.x; u < tla; u += 256) {
        uint32_t ra = (uint32_t)u;
        int64_t off = 0;
        for (int q = 0; q < pr.no; ++q) {
          const uint32_t qq = fdiv_q(ra, pr.odv[q]);
          off += (int64_t)(ra - qq * pr.odv[q].d) * pr.o_is[q];
          ra = qq;
        }
        soff_a[u] = off;
      }
      for (int u = threadIdx.x; u < tlb; u += 256) {
        uint32_t rb = (uint32_t)u;
        int64_t off = 0;
        for (int q = 0; q < pr.ni; ++q) {
          const uint32_t qq = fdiv_q(rb, pr.idv[q]);
          off += (int64_t)(rb - qq * pr.idv[q].d) * pr.i_os[q];
          rb = qq;
        }
        soff_b[u] = off;
      }
      __syncthreads();
      const uint32_t st = (uint32_t)(tsc * tlb);
      const uint32_t lda = st + 1 + ((st + 1) % 2 == 0 ? 1 : 0);  // odd row pitch
      const uint32_t cnt = (uint32_t)tla * st;                     // <= SCAP
      const T* __restrict__ xb = x + ibase;
      T* __restrict__ yb = y + obase;
      // read: (b, s) runs contiguous in x across lanes, a slowest.  S follows I in the input
      // (s_is == tlb), so element e = (a, r) with r = s * tlb + b over the run of one a sits
      // at x offset r + fa(a), and at LDS index a * lda + r: one division per element
      constexpr int NBT = sizeof(T) >= 16 ? 8 : 16;
      const uint32_t rvalid = nsv * (uint32_t)tlb;  // r past the chunk's last s is not read
      for (uint32_t h = 0; h < cnt; h += 256u * NBT) {
        T v[NBT];
        uint32_t li[NBT];
#pragma unroll
        for (int j = 0; j < NBT; ++j) {
          const uint32_t e = h + threadIdx.x + 256u * j;
          const uint32_t al = fdiv_q(e, pr.st_div), r = e - al * st;
          li[j] = (e < cnt) ? al * lda + r : 0xFFFFFFFFu;
          v[j] = O::zero();
          if (e < cnt && r < rvalid) v[j] = xb[r + soff_a[al]];
        }
#pragma unroll
        for (int j = 0; j < NBT; ++j)
          if (li[j] != 0xFFFFFFFFu) tile[li[j]] = v[j];
      }
      __syncthreads();
      // write: (a, s) runs contiguous in y across lanes, b slowest; 8 elements per thread
      // gathered from LDS before their stores
      for (uint32_t h = 0; h < cnt; h += 256u * 8) {
        T v[8];
        int64_t yo[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const uint32_t e = h + threadIdx.x + 256u * j;
          const uint32_t t1 = fdiv_q(e, pr.tla_div), al = e - t1 * (uint32_t)tla;
          const uint32_t bl = fdiv_q(t1, pr.tsc_div), sl = t1 - bl * (uint32_t)tsc;
          yo[j] = -1;
          if (e < cnt && sl < nsv) {
            v[j] = tile[al * lda + sl * (uint32_t)tlb + bl];
            yo[j] = al + (int64_t)sl * pr.s_os + soff_b[bl];
          }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (yo[j] >= 0) {
            T* yp = yb + yo[j];
            *yp = epilogue<O>(g.mode, sx, sy, v[j], yp);
          }
      }
      __syncthreads();
    } else if (pr.tiled) {
      const int64_t tbk = b % pr.tb;  // b-tiles fastest: consecutive blocks read neighbouring x
      b /= pr.tb;
      const int64_t tak = b % pr.ta;
      uint32_t rest = (uint32_t)(b / pr.ta);
      int64_t ibase = 0, obase = 0;
      for (int q = 0; q < pr.nr; ++q) {
        const uint32_t qq = fdiv_q(rest, pr.rdv[q]);
        const uint32_t v = rest - qq * pr.rdv[q].d;
        rest = qq;
        ibase += (int64_t)v * pr.r_is[q];
        obase += (int64_t)v * pr.r_os[q];
      }
      const int tla = pr.tla, tlb = pr.tlb, ld = tlb + 1;
      const uint32_t cnt = (uint32_t)(tla * tlb);  // <= 4096: at most 16 per thread
      // the input offset of element (a, b) is b + fa(a) and its output offset a + fb(b):
      // tabulate fa over the tile's a values and fb over its b values once
      for (int u = threadIdx.x; u < tla; u += 256) {
        uint32_t ra = (uint32_t)(tak * tla + u);
        int64_t off = 0;
        for (int q = 0; q < pr.no; ++q) {
          const uint32_t qq = fdiv_q(ra, pr.odv[q]);
          off += (int64_t)(ra - qq * pr.odv[q].d) * pr.o_is[q];
          ra = qq;
        }
        soff_a[u] = off;
      }
      for (int u = threadIdx.x; u < tlb; u += 256) {
        uint32_t rb = (uint32_t)(tbk * tlb + u);
        int64_t off = 0;
        for (int q = 0; q < pr.ni; ++q) {
          const uint32_t qq = fdiv_q(rb, pr.idv[q]);
          off += (int64_t)(rb - qq * pr.idv[q].d) * pr.i_os[q];
          rb = qq;
        }
        soff_b[u] = off;
      }
      __syncthreads();
      const int64_t na = pr.pa - tak * tla, nb = pr.pb - tbk * tlb;  // valid extents
      const T* __restrict__ xb = x + ibase + tbk * tlb;
      T* __restrict__ yb = y + obase + tak * tla;
      // read: b (contiguous in x) fastest across lanes; the loads of a batch are all issued
      // before any LDS store (one batch of 16, or two of 8 for 16-byte elements)
      constexpr int NBT = sizeof(T) >= 16 ? 8 : 16;
#pragma unroll
      for (int h = 0; h < 16; h += NBT) {
        T v[NBT];
#pragma unroll
        for (int j = 0; j < NBT; ++j) {
          const uint32_t e = threadIdx.x + 256u * (h + j);
          const uint32_t al = fdiv_q(e, pr.tlb_div), bl = e - al * (uint32_t)tlb;
          if (e < cnt && al < na && bl < nb) v[j] = xb[bl + soff_a[al]];
        }
#pragma unroll
        for (int j = 0; j < NBT; ++j) {
          const uint32_t e = threadIdx.x + 256u * (h + j);
          const uint32_t al = fdiv_q(e, pr.tlb_div), bl = e - al * (uint32_t)tlb;
          if (e < cnt) tile[al * ld + bl] = v[j];
        }
      }
      __syncthreads();
      // write: a (contiguous in y) fastest across lanes
#pragma unroll 4
      for (int j = 0; j < 16; ++j) {
        const uint32_t e = threadIdx.x + 256u * j;
        if (e >= cnt) break;
        const uint32_t bl = fdiv_q(e, pr.tla_div), al = e - bl * (uint32_t)tla;
        if (al < na && bl < nb) {
          T* yp = yb + al + soff_b[bl];
          *yp = epilogue<O>(g.mode, sx, sy, tile[al * ld + bl], yp);
        }
      }
      __syncthreads();
    } else {
      // 1024 consecutive outputs along the flattened output
      const int64_t base = b * 1024;
      T v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t o = base + threadIdx.x + 256 * j;
        if (o < pr.n) {
          uint32_t rem = (uint32_t)o;
          int64_t off = 0;
          for (int q = 0; q < pr.rank; ++q) {
            const uint32_t qq = fdiv_q(rem, pr.odiv[q]);
            off += (int64_t)(rem - qq * pr.odiv[q].d) * pr.istride[q];
            rem = qq;
          }
          v[j] = x[off];
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t o = base + threadIdx.x + 256 * j;
        if (o < pr.n) y[o] = epilogue<O>(g.mode, sx, sy, v[j], y + o);
      }
    }
  }
}

void launch_grouped_permute(int dt, const GPDesc& g, int64_t split, hipStream_t st) {
  if (g.total_blocks <= 0) return;
  OME_DRY_RETURN((split > 0 ? 1 : 0) + (g.total_blocks > split ? 1 : 0));
  TimedScope ts(KC_PERMUTE, st);
  OME_DISPATCH(dt, O, {
    GPDesc gg = g;
    gg.mode = epi_mode<O>(g.sx, g.sy);
    if (split > 0) {
      gg.blk_begin = 0;
      gg.blk_end = split;
      hipLaunchKernelGGL((k_grouped_permute<O, true>), dim3((uint32_t)std::min<int64_t>(split, 1 << 30)),
                         dim3(256), 0, st, gg);
    }
    if (g.total_blocks > split) {
      gg.blk_begin = split;
      gg.blk_end = g.total_blocks;
      hipLaunchKernelGGL((k_grouped_permute<O, false>),
                         dim3((uint32_t)std::min<int64_t>(g.total_blocks - split, 1 << 30)), dim3(256), 0,
                         st, gg);
    }
  });
}

}  // namespace ome
