// Matrix-core GEMMs for gfx950 (CDNA4, 64-wide waves):
//   y[m,n,b] = epilogue(sx, sy, sum_k a[m,k,b] b[k,n,b], y[m,n,b])
// Operands are read in place with their strides: A must have unit stride along m or k,
// B along k or n (the stride-aware direct form; other layouts are permuted first).
//
// Tile 128x128 per 256-thread workgroup (4 waves, 2x2, 64x64 each).  Global tiles are
// prefetched into registers while the matrix cores work on the LDS copy of the previous
// K-chunk.  In LDS both operands are stored K-contiguous ([row][k]) so every lane fetches
// its MFMA fragment with one 64/128-bit read.
//
// The MFMA "A" operand is the B tile (rows n) and the "B" operand is the A tile (cols m),
// so the accumulator's lane index runs along m: stores to a column-major y are coalesced.
//
// Precisions:
//   FP32    v_mfma_f32_32x32x2_f32   (FP32 in, FP32 accumulate)
//   BF16X3  v_mfma_f32_32x32x16_bf16 on x = hi + lo splits: ah*bh + ah*bl + al*bh
//   FP64    v_mfma_f64_16x16x4_f64
#include "kernels.h"

namespace ome {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = f2bf(x);
  lo = f2bf(x - bf2f(hi));
}

// Map the linear workgroup id to (tile_m, tile_n, batch, ksplit).  Workgroups are
// dispatched round-robin over the 8 XCDs; consecutive ids are regrouped so that each XCD
// receives a contiguous range of tiles, and tiles are ordered in groups of 8 tile-rows so
// neighbouring workgroups share A and B panels in that XCD's L2.
struct TileIdx {
  int64_t tm, tn, batch, ks;
};
__device__ __forceinline__ TileIdx map_tile(int64_t bid, int64_t nblocks, int64_t tiles_m,
                                            int64_t tiles_n, int64_t nbatch) {
  if ((nblocks & 7) == 0) {
    const int64_t per = nblocks >> 3;
    bid = (bid & 7) * per + (bid >> 3);
  }
  const int64_t per_mat = tiles_m * tiles_n;
  TileIdx t;
  int64_t mat = bid / per_mat;
  int64_t r = bid - mat * per_mat;
  t.batch = mat % nbatch;
  t.ks = mat / nbatch;
  constexpr int64_t G = 8;
  const int64_t group = r / (G * tiles_n);
  const int64_t first_m = group * G;
  int64_t gsz = tiles_m - first_m;
  if (gsz > G) gsz = G;
  const int64_t in = r - group * G * tiles_n;
  t.tm = first_m + in % gsz;
  t.tn = in / gsz;
  return t;
}

// Write a wave's 64x64 block of 32x32 accumulators: lane holds m = mbase + mb*32 + (lane&31)
// and n = nbase + nb*32 + 8*(r/4) + 4*(lane/32) + r%4.  Split-K partials go to the workspace.
template <class O>
__device__ __forceinline__ void store_acc32(const GemmDesc& d, floatx16 (&acc)[2][2],
                                            int64_t mbase, int64_t nbase, int64_t batch,
                                            int64_t ks, int lane) {
  const float sx = O::from(d.sx), sy = O::from(d.sy);
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t m = mbase + mb * 32 + (lane & 31);
        const int64_t n = nbase + nb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m < d.M && n < d.N) {
          if (d.ksplit > 1) {
            ((float*)d.work)[((ks * d.B + batch) * d.M) * d.N + m + n * d.M] = acc[nb][mb][r];
          } else {
            float* yp = (float*)d.y + batch * d.syb + m * d.sym + n * d.syn;
            *yp = epilogue<O>(d.mode, sx, sy, acc[nb][mb][r], yp);
          }
        }
      }
}

// ------------------------------------------------------------------------------------------
// BF16X3 and FP32 share the staging scheme for FP32 sources.

template <bool KMAJ, int BKK>
__device__ __forceinline__ void load_f32_tile(const float* __restrict__ P, int64_t r0, int64_t k0,
                                              int64_t R, int64_t K, int64_t sr, int64_t sk,
                                              bool vec, float* __restrict__ v) {
  // 128 rows x BKK k; each thread loads BKK*128/256 = BKK/2 values.
  // KMAJ (sk == 1): row = t/2, k = (t&1)*(BKK/2) + i        (i < BKK/2)
  // !KMAJ (sr == 1): rows r4..r4+3 with r4 = (t&31)*4, k = (t>>5)*(BKK/8) + j  (j < BKK/8)
  const int t = threadIdx.x;
  if constexpr (KMAJ) {
    constexpr int NV = BKK / 2;
    const int64_t r = r0 + (t >> 1);
    const int64_t kb = k0 + (t & 1) * NV;
    const float* p = P + r * sr + kb;
    if (vec && r < R && kb + NV <= K) {
#pragma unroll
      for (int i = 0; i < NV; i += 4) {
        float4 q = *(const float4*)(p + i);
        v[i] = q.x; v[i + 1] = q.y; v[i + 2] = q.z; v[i + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NV; ++i) v[i] = (r < R && kb + i < K) ? p[i] : 0.f;
    }
  } else {
    constexpr int NK = BKK / 8;
    const int64_t r4 = r0 + (t & 31) * 4;
    const int64_t kb = k0 + (t >> 5) * NK;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      const float* p = P + r4 + (kb + j) * sk;
      if (vec && r4 + 4 <= R && kb + j < K) {
        float4 q = *(const float4*)p;
        v[4 * j] = q.x; v[4 * j + 1] = q.y; v[4 * j + 2] = q.z; v[4 * j + 3] = q.w;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) v[4 * j + i] = (r4 + i < R && kb + j < K) ? p[i] : 0.f;
      }
    }
  }
}

// ---- BF16X3 --------------------------------------------------------------------------------
constexpr int B3_BK = 32;
constexpr int B3_LDK = B3_BK + 8;  // bf16 per LDS row (80 bytes, keeps 16-byte alignment)

template <bool KMAJ>
__device__ __forceinline__ void store_bf16x3(uint16_t (*__restrict__ hi)[B3_LDK],
                                             uint16_t (*__restrict__ lo)[B3_LDK],
                                             const float* __restrict__ v) {
  const int t = threadIdx.x;
  if constexpr (KMAJ) {
    const int r = t >> 1, kb = (t & 1) * 16;
    uint16_t h[16], l[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) split_bf16(v[i], h[i], l[i]);
    uint4* ph = (uint4*)&hi[r][kb];
    uint4* pl = (uint4*)&lo[r][kb];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      uint4 a, b;
      a.x = h[8 * q] | ((uint32_t)h[8 * q + 1] << 16); a.y = h[8 * q + 2] | ((uint32_t)h[8 * q + 3] << 16);
      a.z = h[8 * q + 4] | ((uint32_t)h[8 * q + 5] << 16); a.w = h[8 * q + 6] | ((uint32_t)h[8 * q + 7] << 16);
      b.x = l[8 * q] | ((uint32_t)l[8 * q + 1] << 16); b.y = l[8 * q + 2] | ((uint32_t)l[8 * q + 3] << 16);
      b.z = l[8 * q + 4] | ((uint32_t)l[8 * q + 5] << 16); b.w = l[8 * q + 6] | ((uint32_t)l[8 * q + 7] << 16);
      ph[q] = a;
      pl[q] = b;
    }
  } else {
    const int r4 = (t & 31) * 4, kb = (t >> 5) * 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      uint16_t h[4], l[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) split_bf16(v[4 * j + i], h[j], l[j]);
      uint2 a, b;
      a.x = h[0] | ((uint32_t)h[1] << 16); a.y = h[2] | ((uint32_t)h[3] << 16);
      b.x = l[0] | ((uint32_t)l[1] << 16); b.y = l[2] | ((uint32_t)l[3] << 16);
      *(uint2*)&hi[r4 + i][kb] = a;
      *(uint2*)&lo[r4 + i][kb] = b;
    }
  }
}

template <bool AK, bool BK>
__global__ __launch_bounds__(256) void k_mfma_bf16x3(GemmDesc d, int64_t tiles_m, int64_t tiles_n,
                                                     int64_t kchunk, int vec_a, int vec_b) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][128][B3_LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][128][B3_LDK];
  const TileIdx ti = map_tile(blockIdx.x, gridDim.x, tiles_m, tiles_n, d.B);
  const int64_t m0 = ti.tm * 128, n0 = ti.tn * 128;
  const int64_t kbeg = ti.ks * kchunk;
  const int64_t kend = (kbeg + kchunk < d.K) ? kbeg + kchunk : d.K;
  const float* A = (const float*)d.a + ti.batch * d.sab;
  const float* B = (const float*)d.b + ti.batch * d.sbb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  float ra[16], rb[16];
  // A rows are m (stride sam), k stride sak; B rows are n (stride sbn), k stride sbk
  load_f32_tile<AK, B3_BK>(A, m0, kbeg, d.M, kend, d.sam, d.sak, vec_a, ra);
  load_f32_tile<BK, B3_BK>(B, n0, kbeg, d.N, kend, d.sbn, d.sbk, vec_b, rb);
  for (int64_t k0 = kbeg; k0 < kend; k0 += B3_BK) {
    store_bf16x3<AK>(sA[0], sA[1], ra);
    store_bf16x3<BK>(sB[0], sB[1], rb);
    __syncthreads();
    if (k0 + B3_BK < kend) {
      load_f32_tile<AK, B3_BK>(A, m0, k0 + B3_BK, d.M, kend, d.sam, d.sak, vec_a, ra);
      load_f32_tile<BK, B3_BK>(B, n0, k0 + B3_BK, d.N, kend, d.sbn, d.sbk, vec_b, rb);
    }
#pragma unroll
    for (int ks = 0; ks < B3_BK; ks += 16) {
      const int kk = ks + 8 * (lane >> 5);
      bf16x8_t ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        ah[nb] = *(const bf16x8_t*)&sB[0][wn + nb * 32 + (lane & 31)][kk];
        al[nb] = *(const bf16x8_t*)&sB[1][wn + nb * 32 + (lane & 31)][kk];
      }
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        bh[mb] = *(const bf16x8_t*)&sA[0][wm + mb * 32 + (lane & 31)][kk];
        bl[mb] = *(const bf16x8_t*)&sA[1][wm + mb * 32 + (lane & 31)][kk];
      }
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[nb], bh[mb], acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[nb], bl[mb], acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[nb], bh[mb], acc[nb][mb], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  store_acc32<OpsF32>(d, acc, m0 + wm, n0 + wn, ti.batch, ti.ks, lane);
}

// ---- FP32 ------------------------------------------------------------------------------------
// Two-level accumulation for long K: an MFMA accumulator sums its products in order, so its
// FP32 rounding error grows as sqrt(depth) (about u*sqrt(d/6) relative, 1.6e-6 at d = 4096).
// With FOLD, every FOLD_DEPTH real products the accumulators are added into running sums
// and restart from zero, so the error is about u*(sqrt(FOLD_DEPTH/6) + sqrt(K/FOLD_DEPTH/6)):
// 6e-7 at K = 2^16 instead of 6e-6.  The adds are 64 VALU operations per lane per 256-deep
// stretch of MFMA work.
constexpr int FOLD_DEPTH = 256;

// init: tot = 0.  Otherwise tot += acc and acc = 0; last: acc = tot + acc (the final sum).
__device__ __forceinline__ void fold_acc(floatx16 (&tot)[2][2], floatx16 (&acc)[2][2], bool init,
                                         bool last = false) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        if (init) {
          tot[i][j][r] = 0.f;
        } else if (last) {
          acc[i][j][r] = tot[i][j][r] + acc[i][j][r];
        } else {
          tot[i][j][r] += acc[i][j][r];
          acc[i][j][r] = 0.f;
        }
      }
}

constexpr int F32_BK = 32;
constexpr int F32_LDK = F32_BK + 4;

// Stage one thread's F32_BK/2 values (the layout of load_f32_tile) into a 128 x F32_BK tile.
template <bool KMAJ>
__device__ __forceinline__ void store_f32(float (*__restrict__ s)[F32_LDK], const float* __restrict__ v) {
  const int t = threadIdx.x;
  if constexpr (KMAJ) {
    constexpr int NV = F32_BK / 2;
    const int r = t >> 1, kb = (t & 1) * NV;
#pragma unroll
    for (int i = 0; i < NV; i += 4)
      *(float4*)&s[r][kb + i] = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
  } else {
    constexpr int NK = F32_BK / 8;
    const int r4 = (t & 31) * 4, kb = (t >> 5) * NK;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if constexpr (NK == 2) {
        *(float2*)&s[r4 + i][kb] = make_float2(v[i], v[4 + i]);
      } else {
#pragma unroll
        for (int j = 0; j < NK; j += 4)
          *(float4*)&s[r4 + i][kb + j] =
              make_float4(v[4 * j + i], v[4 * (j + 1) + i], v[4 * (j + 2) + i], v[4 * (j + 3) + i]);
      }
    }
  }
}

template <bool AK, bool BK, bool FOLD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FOLD ? 2 : 1, FOLD ? 2 : 8))) void k_mfma_f32(GemmDesc d, int64_t tiles_m, int64_t tiles_n,
                                                  int64_t kchunk, int vec_a, int vec_b) {
  __shared__ __attribute__((aligned(16))) float sA[128][F32_LDK];
  __shared__ __attribute__((aligned(16))) float sB[128][F32_LDK];
  const TileIdx ti = map_tile(blockIdx.x, gridDim.x, tiles_m, tiles_n, d.B);
  const int64_t m0 = ti.tm * 128, n0 = ti.tn * 128;
  const int64_t kbeg = ti.ks * kchunk;
  const int64_t kend = (kbeg + kchunk < d.K) ? kbeg + kchunk : d.K;
  const float* A = (const float*)d.a + ti.batch * d.sab;
  const float* B = (const float*)d.b + ti.batch * d.sbb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  floatx16 tot[2][2];  // FOLD: running sums the MFMA accumulators fold into (see fold_acc)
  if constexpr (FOLD) fold_acc(tot, acc, true);
  int since_fold = 0;
  float ra[F32_BK / 2], rb[F32_BK / 2];
  load_f32_tile<AK, F32_BK>(A, m0, kbeg, d.M, kend, d.sam, d.sak, vec_a, ra);
  load_f32_tile<BK, F32_BK>(B, n0, kbeg, d.N, kend, d.sbn, d.sbk, vec_b, rb);
  for (int64_t k0 = kbeg; k0 < kend; k0 += F32_BK) {
    store_f32<AK>(sA, ra);
    store_f32<BK>(sB, rb);
    __syncthreads();
    if (k0 + F32_BK < kend) {
      load_f32_tile<AK, F32_BK>(A, m0, k0 + F32_BK, d.M, kend, d.sam, d.sak, vec_a, ra);
      load_f32_tile<BK, F32_BK>(B, n0, k0 + F32_BK, d.N, kend, d.sbn, d.sbk, vec_b, rb);
    }
    // lanes 0-31 hold k = kb+s, lanes 32-63 hold k = kb+4+s for the s-th MFMA (s < 4)
#pragma unroll
    for (int kb = 0; kb < F32_BK; kb += 8) {
      const int kk = kb + 4 * (lane >> 5);
      float4 af[2], bf[2];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) af[nb] = *(const float4*)&sB[wn + nb * 32 + (lane & 31)][kk];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) bf[mb] = *(const float4*)&sA[wm + mb * 32 + (lane & 31)][kk];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].x, bf[mb].x, acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].y, bf[mb].y, acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].z, bf[mb].z, acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].w, bf[mb].w, acc[nb][mb], 0, 0, 0);
        }
    }
    if constexpr (FOLD) {
      if (++since_fold == FOLD_DEPTH / F32_BK) {
        since_fold = 0;
        fold_acc(tot, acc, false);
      }
    }
    __syncthreads();
  }
  if constexpr (FOLD) fold_acc(tot, acc, false, true);
  store_acc32<OpsF32>(d, acc, m0 + wm, n0 + wn, ti.batch, ti.ks, lane);
}

// ---- ComplexF32 on the FP32 matrix cores, realified while staged ---------------------------
// Output tile 128 m x 64 n (complex), K-steps of 8 complex k (16 real).  The LDS tiles hold the
// real operands of the four-real-products form, built from the complex operands as they are
// staged (no operand preparation in HBM, no combine pass):
//   A'[m][k] = Ar(m, k), A'[m][8 + k] = Ai(m, k);
//   B' rows n' run over [re n 0..31 | im n 0..31 | re n 32..63 | im n 32..63]:
//   B'[re n][k] = Br, B'[im n][k] = Bi, B'[re n][8 + k] = -Bi, B'[im n][8 + k] = Br,
// so C' = A' B' holds Cr in the re rows and Ci in the im rows, and the two 32-wide blocks of a
// wave are the real and imaginary parts of the same 32 x 32 outputs.
constexpr int C32_BK = 8;
constexpr int C32_LDK = 2 * C32_BK + 4;  // real k per LDS row: [re | im] + 4 pad

// A tile: 128 rows x 8 k, 4 complex per thread.
//   KMAJ (sk == 1): row t / 2, k = (t % 2) * 4 + i;  !KMAJ (sr == 1): row t % 128, k = (t / 128) * 4 + i
template <bool KMAJ>
__device__ __forceinline__ void load_c32_a(const c32* __restrict__ P, int64_t r0, int64_t k0,
                                           int64_t R, int64_t K, int64_t sr, int64_t sk,
                                           c32 (&v)[4]) {
  const int t = threadIdx.x;
  const int64_t r = r0 + (KMAJ ? (t >> 1) : (t & 127));
  const int64_t kb = k0 + (KMAJ ? (t & 1) * 4 : (t >> 7) * 4);
  // every lane loads (outside the tile: the first element), the zeros are applied when the
  // values are staged, so no wait follows the loads here
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const bool ok = r < R && kb + i < K;
    v[i] = P[ok ? r * sr + (kb + i) * sk : 0];
  }
}

// B tile: 64 rows (n) x 8 k, 2 complex per thread.
//   KMAJ (sk == 1): row t / 4, k = (t % 4) * 2 + i;  !KMAJ (sn == 1): row t % 64, k = (t / 64) * 2 + i
template <bool KMAJ>
__device__ __forceinline__ void load_c32_b(const c32* __restrict__ P, int64_t r0, int64_t k0,
                                           int64_t R, int64_t K, int64_t sr, int64_t sk,
                                           c32 (&v)[2]) {
  const int t = threadIdx.x;
  const int64_t r = r0 + (KMAJ ? (t >> 2) : (t & 63));
  const int64_t kb = k0 + (KMAJ ? (t & 3) * 2 : (t >> 6) * 2);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const bool ok = r < R && kb + i < K;
    v[i] = P[ok ? r * sr + (kb + i) * sk : 0];
  }
}

template <bool AK, bool BK, bool FOLD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FOLD ? 2 : 1, FOLD ? 2 : 8))) void k_mfma_c32(GemmDesc d, int64_t tiles_m, int64_t tiles_n,
                                                  int64_t kchunk) {
  __shared__ __attribute__((aligned(16))) float sA[128][C32_LDK];
  __shared__ __attribute__((aligned(16))) float sB[128][C32_LDK];
  const TileIdx ti = map_tile(blockIdx.x, gridDim.x, tiles_m, tiles_n, d.B);
  const int64_t m0 = ti.tm * 128, n0 = ti.tn * 64;
  const int64_t kbeg = ti.ks * kchunk;
  const int64_t kend = (kbeg + kchunk < d.K) ? kbeg + kchunk : d.K;
  const c32* A = (const c32*)d.a + ti.batch * d.sab;
  const c32* B = (const c32*)d.b + ti.batch * d.sbb;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  floatx16 tot[2][2];  // FOLD: running sums the MFMA accumulators fold into (see fold_acc)
  if constexpr (FOLD) fold_acc(tot, acc, true);
  int since_fold = 0;
  c32 ra[4], rb[2];
  load_c32_a<AK>(A, m0, kbeg, d.M, kend, d.sam, d.sak, ra);
  load_c32_b<BK>(B, n0, kbeg, d.N, kend, d.sbn, d.sbk, rb);
  for (int64_t k0 = kbeg; k0 < kend; k0 += C32_BK) {
    {
      // elements outside the operand (loaded from its first element) stage as zeros
      const int r = AK ? (t >> 1) : (t & 127), kb = AK ? (t & 1) * 4 : (t >> 7) * 4;
      const bool rok = m0 + r < d.M;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool ok = rok && k0 + kb + i < kend;
        sA[r][kb + i] = ok ? ra[i].x : 0.f;
        sA[r][C32_BK + kb + i] = ok ? ra[i].y : 0.f;
      }
    }
    {
      const int n = BK ? (t >> 2) : (t & 63), kb = BK ? (t & 3) * 2 : (t >> 6) * 2;
      const int nre = (n & 31) + ((n >> 5) << 6), nim = nre + 32;
      const bool nok = n0 + n < d.N;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const bool ok = nok && k0 + kb + i < kend;
        const float br = ok ? rb[i].x : 0.f, bi = ok ? rb[i].y : 0.f;
        sB[nre][kb + i] = br;
        sB[nim][kb + i] = bi;
        sB[nre][C32_BK + kb + i] = -bi;
        sB[nim][C32_BK + kb + i] = br;
      }
    }
    __syncthreads();
    if (k0 + C32_BK < kend) {
      load_c32_a<AK>(A, m0, k0 + C32_BK, d.M, kend, d.sam, d.sak, ra);
      load_c32_b<BK>(B, n0, k0 + C32_BK, d.N, kend, d.sbn, d.sbk, rb);
    }
    // lanes 0-31 hold k' = kb+s, lanes 32-63 hold k' = kb+4+s for the s-th MFMA (s < 4)
#pragma unroll
    for (int kb = 0; kb < 2 * C32_BK; kb += 8) {
      const int kk = kb + 4 * (lane >> 5);
      float4 af[2], bf[2];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) af[nb] = *(const float4*)&sB[wn + nb * 32 + (lane & 31)][kk];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) bf[mb] = *(const float4*)&sA[wm + mb * 32 + (lane & 31)][kk];
#pragma unroll
      for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].x, bf[mb].x, acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].y, bf[mb].y, acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].z, bf[mb].z, acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[nb].w, bf[mb].w, acc[nb][mb], 0, 0, 0);
        }
    }
    if constexpr (FOLD) {
      if (++since_fold == FOLD_DEPTH / (2 * C32_BK)) {
        since_fold = 0;
        fold_acc(tot, acc, false);
      }
    }
    __syncthreads();
  }
  if constexpr (FOLD) fold_acc(tot, acc, false, true);
  // lane holds m = m0 + wm + mb*32 + (lane & 31) and complex n = n0 + wn/2 + 8*(r/4) +
  // 4*(lane/32) + r%4, with the real part in block nb = 0 and the imaginary part in nb = 1
  const c32 sx = OpsC32::from(d.sx), sy = OpsC32::from(d.sy);
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t m = m0 + wm + mb * 32 + (lane & 31);
      const int64_t n = n0 + (wn >> 1) + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (m < d.M && n < d.N) {
        const c32 v{acc[0][mb][r], acc[1][mb][r]};
        if (d.ksplit > 1) {
          ((c32*)d.work)[((ti.ks * d.B + ti.batch) * d.M) * d.N + m + n * d.M] = v;
        } else {
          c32* yp = (c32*)d.y + ti.batch * d.syb + m * d.sym + n * d.syn;
          *yp = epilogue<OpsC32>(d.mode, sx, sy, v, yp);
        }
      }
    }
}

// ---- FP64 ------------------------------------------------------------------------------------
constexpr int F64_BK = 32;
constexpr int F64_LDK = F64_BK + 2;

template <bool KMAJ>
__device__ __forceinline__ void load_f64_tile(const double* __restrict__ P, int64_t r0, int64_t k0,
                                              int64_t R, int64_t K, int64_t sr, int64_t sk,
                                              bool vec, double* __restrict__ v) {
  // 128 rows x F64_BK k, F64_BK / 2 values per thread.
  // KMAJ: row = t/2, k = (t&1)*(F64_BK/2) + i.
  // !KMAJ: rows r2, r2+1 with r2 = (t&63)*2, k = (t>>6)*(F64_BK/4) + j.
  const int t = threadIdx.x;
  if constexpr (KMAJ) {
    constexpr int NV = F64_BK / 2;
    const int64_t r = r0 + (t >> 1);
    const int64_t kb = k0 + (t & 1) * NV;
    const double* p = P + r * sr + kb;
    if (vec && r < R && kb + NV <= K) {
#pragma unroll
      for (int i = 0; i < NV; i += 2) {
        double2 q = *(const double2*)(p + i);
        v[i] = q.x; v[i + 1] = q.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NV; ++i) v[i] = (r < R && kb + i < K) ? p[i] : 0.0;
    }
  } else {
    constexpr int NK = F64_BK / 4;
    const int64_t r2 = r0 + (t & 63) * 2;
    const int64_t kb = k0 + (t >> 6) * NK;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      const double* p = P + r2 + (kb + j) * sk;
      if (vec && r2 + 2 <= R && kb + j < K) {
        double2 q = *(const double2*)p;
        v[2 * j] = q.x; v[2 * j + 1] = q.y;
      } else {
        v[2 * j] = (r2 < R && kb + j < K) ? p[0] : 0.0;
        v[2 * j + 1] = (r2 + 1 < R && kb + j < K) ? p[1] : 0.0;
      }
    }
  }
}

template <bool KMAJ>
__device__ __forceinline__ void store_f64(double (*__restrict__ s)[F64_LDK], const double* __restrict__ v) {
  const int t = threadIdx.x;
  if constexpr (KMAJ) {
    constexpr int NV = F64_BK / 2;
    const int r = t >> 1, kb = (t & 1) * NV;
#pragma unroll
    for (int i = 0; i < NV; i += 2) *(double2*)&s[r][kb + i] = make_double2(v[i], v[i + 1]);
  } else {
    constexpr int NK = F64_BK / 4;
    const int r2 = (t & 63) * 2, kb = (t >> 6) * NK;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < NK; j += 2)
        *(double2*)&s[r2 + i][kb + j] = make_double2(v[2 * j + i], v[2 * (j + 1) + i]);
  }
}

// Two waves per SIMD: the 128 FP64 accumulator registers per lane plus operands fit 256
// registers, and a second resident workgroup covers the other's barriers and loads (left to
// itself the compiler takes 338 registers and one wave per SIMD, 18% slower on 4096^3).
template <bool AK, bool BK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mfma_f64(GemmDesc d, int64_t tiles_m, int64_t tiles_n,
                                                  int64_t kchunk, int vec_a, int vec_b) {
  __shared__ __attribute__((aligned(16))) double sA[128][F64_LDK];
  __shared__ __attribute__((aligned(16))) double sB[128][F64_LDK];
  const TileIdx ti = map_tile(blockIdx.x, gridDim.x, tiles_m, tiles_n, d.B);
  const int64_t m0 = ti.tm * 128, n0 = ti.tn * 128;
  const int64_t kbeg = ti.ks * kchunk;
  const int64_t kend = (kbeg + kchunk < d.K) ? kbeg + kchunk : d.K;
  const double* A = (const double*)d.a + ti.batch * d.sab;
  const double* B = (const double*)d.b + ti.batch * d.sbb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
  doublex4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
  double ra[F64_BK / 2], rb[F64_BK / 2];
  load_f64_tile<AK>(A, m0, kbeg, d.M, kend, d.sam, d.sak, vec_a, ra);
  load_f64_tile<BK>(B, n0, kbeg, d.N, kend, d.sbn, d.sbk, vec_b, rb);
  for (int64_t k0 = kbeg; k0 < kend; k0 += F64_BK) {
    store_f64<AK>(sA, ra);
    store_f64<BK>(sB, rb);
    __syncthreads();
    if (k0 + F64_BK < kend) {
      load_f64_tile<AK>(A, m0, k0 + F64_BK, d.M, kend, d.sam, d.sak, vec_a, ra);
      load_f64_tile<BK>(B, n0, k0 + F64_BK, d.N, kend, d.sbn, d.sbk, vec_b, rb);
    }
    // 16x16x4: lane group g = lane/16 holds k = kb + 2g + s for the s-th MFMA (s < 2)
#pragma unroll
    for (int kb = 0; kb < F64_BK; kb += 8) {
      const int kk = kb + 2 * (lane >> 4);
      double2 af[4], bf[4];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) af[nb] = *(const double2*)&sB[wn + nb * 16 + (lane & 15)][kk];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) bf[mb] = *(const double2*)&sA[wm + mb * 16 + (lane & 15)][kk];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          acc[nb][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[nb].x, bf[mb].x, acc[nb][mb], 0, 0, 0);
          acc[nb][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[nb].y, bf[mb].y, acc[nb][mb], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // epilogue: v_mfma_f64_16x16x4 leaves D[row][col] with col = lane & 15 and
  // row = (lane >> 4) + 4*r, so the lane holds m = mbase + (lane & 15), n = nbase + (lane>>4) + 4r
  const OpsF64::T sx = OpsF64::from(d.sx), sy = OpsF64::from(d.sy);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm + mb * 16 + (lane & 15);
        const int64_t n = n0 + wn + nb * 16 + (lane >> 4) + 4 * r;
        if (m < d.M && n < d.N) {
          if (d.ksplit > 1) {
            ((double*)d.work)[((ti.ks * d.B + ti.batch) * d.M) * d.N + m + n * d.M] = acc[nb][mb][r];
          } else {
            double* yp = (double*)d.y + ti.batch * d.syb + m * d.sym + n * d.syn;
            *yp = epilogue<OpsF64>(d.mode, sx, sy, acc[nb][mb][r], yp);
          }
        }
      }
}

// ------------------------------------------------------------------------------------------

template <class O>
__global__ __launch_bounds__(256) void k_splitk_reduce(GemmDesc d) {
  using T = typename O::T;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  const int64_t per = d.M * d.N, total = per * d.B;
  const T* W = (const T*)d.work;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t batch = e / per, r = e - batch * per;
    const int64_t n = r / d.M, m = r - n * d.M;
    // partials summed in split order in the accumulation arithmetic (FP64 for Float32 and
    // ComplexF32), loaded 16 at a time so their latencies overlap
    using Wd = AccOps<O>;
    using A = typename Wd::A;
    typename A::T acc = A::zero();
    int s = 0;
    for (; s + 16 <= d.ksplit; s += 16) {
      T v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = W[(int64_t)(s + q) * total + e];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc = A::add(acc, Wd::widen(v[q]));
    }
    for (; s < d.ksplit; ++s) acc = A::add(acc, Wd::widen(W[(int64_t)s * total + e]));
    T* yp = (T*)d.y + batch * d.syb + m * d.sym + n * d.syn;
    *yp = epilogue<O>(d.mode, sx, sy, Wd::narrow(acc), yp);
  }
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

template <bool FOLD>
static void launch_f32(dim3 grid, dim3 block, hipStream_t st, const GemmDesc& d, int64_t tiles_m,
                       int64_t tiles_n, int64_t kchunk, int va, int vb, bool ak, bool bk) {
  if (ak && bk) hipLaunchKernelGGL((k_mfma_f32<true, true, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);
  else if (ak) hipLaunchKernelGGL((k_mfma_f32<true, false, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);
  else if (bk) hipLaunchKernelGGL((k_mfma_f32<false, true, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);
  else hipLaunchKernelGGL((k_mfma_f32<false, false, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);
}

template <bool FOLD>
static void launch_c32(dim3 grid, dim3 block, hipStream_t st, const GemmDesc& d, int64_t tiles_m,
                       int64_t tiles_n, int64_t kchunk, bool ak, bool bk) {
  if (ak && bk) hipLaunchKernelGGL((k_mfma_c32<true, true, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk);
  else if (ak) hipLaunchKernelGGL((k_mfma_c32<true, false, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk);
  else if (bk) hipLaunchKernelGGL((k_mfma_c32<false, true, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk);
  else hipLaunchKernelGGL((k_mfma_c32<false, false, FOLD>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk);
}

int mfma_plan_ksplit(const GemmDesc& d, int requested) {
  if (requested > 0) return requested;
  const int64_t tiles = ((d.M + 127) / 128) * ((d.N + 127) / 128) * d.B;
  int ks = 1;
  // few output tiles and a long K: split K (chunks of at least 64) so that ~512 workgroups run
  while (tiles * ks < 512 && d.K / (ks * 2) >= 64) ks *= 2;
  return ks;
}

int64_t mfma_workspace(int dt, const GemmDesc& d, int ksplit) {
  if (ksplit <= 1) return 0;
  return (int64_t)ksplit * d.B * d.M * d.N * (dt == F64 ? 8 : 4);
}

void launch_gemm_mfma(int prec, GemmDesc d, int ksplit, hipStream_t st) {
  if (d.M <= 0 || d.N <= 0 || d.B <= 0) return;
  OME_DRY_RETURN(ksplit > 1 ? 2 : 1);
  const bool f64 = (prec == PREC_FP64);
  d.mode = f64 ? epi_mode<OpsF64>(d.sx, d.sy) : epi_mode<OpsF32>(d.sx, d.sy);
  if (ksplit < 1) ksplit = 1;
  const int kstep = (prec == PREC_BF16X3) ? B3_BK : f64 ? F64_BK : F32_BK;
  int64_t kchunk = (d.K + ksplit - 1) / ksplit;
  kchunk = ((kchunk + kstep - 1) / kstep) * kstep;
  if (kchunk < kstep) kchunk = kstep;
  ksplit = (int)((d.K + kchunk - 1) / kchunk);
  if (ksplit < 1) ksplit = 1;
  d.ksplit = ksplit;
  const int64_t tiles_m = (d.M + 127) / 128, tiles_n = (d.N + 127) / 128;
  const int64_t blocks = tiles_m * tiles_n * d.B * ksplit;
  const bool ak = (d.sak == 1) || (d.M == 1 && d.sam != 1);
  const bool bk = (d.sbk == 1) || (d.N == 1 && d.sbn != 1);
  const int esz = f64 ? 8 : 4;
  const int vw = f64 ? 2 : 4;  // elements per 16-byte vector
  // vector loads need 16-byte aligned rows: base pointer, batch stride and leading stride
  auto vec_ok = [&](const void* p, int64_t lead, int64_t sbatch) {
    return aligned16(p) && (lead % vw == 0) && (d.B == 1 || sbatch % vw == 0);
  };
  (void)esz;
  const int va = vec_ok(d.a, ak ? d.sam : d.sak, d.sab) ? 1 : 0;
  const int vb = vec_ok(d.b, bk ? d.sbn : d.sbk, d.sbb) ? 1 : 0;
  dim3 grid((uint32_t)blocks), block(256);
#define OME_MFMA_LAUNCH(KER)                                                                   \
  if (ak && bk) hipLaunchKernelGGL((KER<true, true>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);   \
  else if (ak) hipLaunchKernelGGL((KER<true, false>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);  \
  else if (bk) hipLaunchKernelGGL((KER<false, true>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);  \
  else hipLaunchKernelGGL((KER<false, false>), grid, block, 0, st, d, tiles_m, tiles_n, kchunk, va, vb);
  {
  TimedScope ts(KC_GEMM_MFMA, st);
  if (prec == PREC_BF16X3) {
    OME_MFMA_LAUNCH(k_mfma_bf16x3)
  } else if (prec == PREC_FP32) {
    if (kchunk > FOLD_DEPTH) launch_f32<true>(grid, block, st, d, tiles_m, tiles_n, kchunk, va, vb, ak, bk);
    else launch_f32<false>(grid, block, st, d, tiles_m, tiles_n, kchunk, va, vb, ak, bk);
  } else {
    OME_MFMA_LAUNCH(k_mfma_f64)
  }
  }
#undef OME_MFMA_LAUNCH
  if (ksplit > 1) {
    TimedScope ts(KC_OTHER, st);
    const int64_t outputs = d.M * d.N * d.B;
    int g = (int)((outputs + 255) / 256);
    if (g > 65536) g = 65536;
    if (f64) hipLaunchKernelGGL(k_splitk_reduce<OpsF64>, dim3(g), dim3(256), 0, st, d);
    else hipLaunchKernelGGL(k_splitk_reduce<OpsF32>, dim3(g), dim3(256), 0, st, d);
  }
}

int mfma_c32_plan_ksplit(const GemmDesc& d) {
  const int64_t tiles = ((d.M + 127) / 128) * ((d.N + 63) / 64) * d.B;
  int ks = 1;
  // as for the real kernels: chunks of at least 32 complex k (64 real) until ~512 workgroups
  while (tiles * ks < 512 && d.K / (ks * 2) >= 32) ks *= 2;
  return ks;
}

int64_t mfma_c32_workspace(const GemmDesc& d, int ksplit) {
  if (ksplit <= 1) return 0;
  return (int64_t)ksplit * d.B * d.M * d.N * (int64_t)sizeof(c32);
}

void launch_gemm_mfma_c32(GemmDesc d, int ksplit, hipStream_t st) {
  if (d.M <= 0 || d.N <= 0 || d.B <= 0) return;
  OME_DRY_RETURN(ksplit > 1 ? 2 : 1);
  d.mode = epi_mode<OpsC32>(d.sx, d.sy);
  if (ksplit < 1) ksplit = 1;
  int64_t kchunk = (d.K + ksplit - 1) / ksplit;
  kchunk = ((kchunk + C32_BK - 1) / C32_BK) * C32_BK;
  if (kchunk < C32_BK) kchunk = C32_BK;
  ksplit = (int)((d.K + kchunk - 1) / kchunk);
  if (ksplit < 1) ksplit = 1;
  d.ksplit = ksplit;
  const int64_t tiles_m = (d.M + 127) / 128, tiles_n = (d.N + 63) / 64;
  const int64_t blocks = tiles_m * tiles_n * d.B * ksplit;
  const bool ak = (d.sak == 1) || (d.M == 1 && d.sam != 1);
  const bool bk = (d.sbk == 1) || (d.N == 1 && d.sbn != 1);
  const dim3 grid((uint32_t)blocks), block(256);
  {
    TimedScope ts(KC_GEMM_MFMA, st);
    // a complex k is two real products of the realified sum
    if (2 * kchunk > FOLD_DEPTH) launch_c32<true>(grid, block, st, d, tiles_m, tiles_n, kchunk, ak, bk);
    else launch_c32<false>(grid, block, st, d, tiles_m, tiles_n, kchunk, ak, bk);
  }
  if (ksplit > 1) {
    TimedScope ts(KC_OTHER, st);
    const int64_t outputs = d.M * d.N * d.B;
    int g = (int)((outputs + 255) / 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_splitk_reduce<OpsC32>, dim3(g), dim3(256), 0, st, d);
  }
}

}  // namespace ome
