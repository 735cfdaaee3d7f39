// Packed split-precision GEMM for large Float32 / ComplexF32 contractions.
//
// 1. pack: each operand, read once with arbitrary (row, k, batch) strides, is split into
//    BF16 hi/lo planes (x = hi + lo, |x - hi - lo| <= 2^-18 |x|) and written tile-major:
//      [batch][row tile (128)][k tile (32)][plane (2)][128 rows][32 k]  (16 KiB per tile)
//    zero-padded past the edges.  The pack is where any permutation of the operand's labels
//    is absorbed (the fused permute+reshape of the binary rule), and ComplexF32 operands are
//    realified here: A' = [Ar | Ai] (M x 2K), B' = [Br Bi ; -Bi Br] (2K x 2N), so that one real
//    GEMM gives C' = [Cr | Ci] (the four-real-products form).
// 2. GEMM: 128x128 tiles, 4 waves of 64x64, K-steps of 32 with a register prefetch of the next
//    step (plain 16-byte loads, no conversion in the loop); per 16-k sub-step every wave issues
//    3 v_mfma_f32_32x32x16_bf16 per 32x32 block: ah*bh + ah*bl + al*bh, FP32 accumulate.
#include "kernels.h"

namespace ome {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int PK_ROWS = 128;
constexpr int PK_K = 32;
constexpr int PK_TILE = 2 * PK_ROWS * PK_K;  // bf16 elements per packed tile
constexpr int PK_LDK = PK_K + 8;             // LDS row stride (bf16): 80 bytes

__device__ __forceinline__ uint16_t f2bf_rn(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// ---- pack ----------------------------------------------------------------------------------
// One workgroup per packed tile.  The 128x32 source tile is staged through LDS so that the
// global reads follow whichever of (row, k) has unit stride; then every thread emits 16-byte
// chunks of the hi and lo planes.
//
// mode: 0 real operand; 1 complex "A" operand (row i, k' < K: re, k' >= K: im);
//       2 complex "B" operand, realified [Br Bi; -Bi Br] with rows n' in [0, 2N), k' in [0, 2K)
struct PackDesc {
  const void* src;
  uint16_t* dst;
  int64_t R, K;       // logical rows / k of the packed operand (after realification)
  int64_t sr, sk, sb; // strides (in complex elements for complex sources)
  int64_t R0, K0;     // complex source extents (rows, k) for modes 1 and 2
  int64_t tiles_r, tiles_k;
  int mode;
  int rfast;          // 1: unit stride along rows (read rows-fastest)
};

__device__ __forceinline__ float pack_fetch(const PackDesc& d, const float* __restrict__ base,
                                            int64_t r, int64_t k) {
  if (r >= d.R || k >= d.K) return 0.f;
  if (d.mode == 0) return base[r * d.sr + k * d.sk];
  if (d.mode == 1) {  // A' = [Ar | Ai]
    const bool im = k >= d.K0;
    const int64_t kk = im ? k - d.K0 : k;
    return base[2 * (r * d.sr + kk * d.sk) + (im ? 1 : 0)];
  }
  // B' (2K x 2N) stored as rows n' (2N), k' (2K): B'[k'][n'] with
  //   k' <  K, n' <  N:  Br[k][n]     k' <  K, n' >= N:  Bi[k][n]
  //   k' >= K, n' <  N: -Bi[k][n]     k' >= K, n' >= N:  Br[k][n]
  const bool kim = k >= d.K0, nim = r >= d.R0;
  const int64_t kk = kim ? k - d.K0 : k, nn = nim ? r - d.R0 : r;
  const float* p = base + 2 * (nn * d.sr + kk * d.sk);
  if (!kim) return nim ? p[1] : p[0];
  return nim ? p[0] : -p[1];
}

__global__ __launch_bounds__(256) void k_pack_bf16x3(PackDesc d) {
  __shared__ float tile[PK_ROWS][PK_K + 1];
  int64_t bid = blockIdx.x;
  const int64_t tk = bid % d.tiles_k; bid /= d.tiles_k;
  const int64_t tr = bid % d.tiles_r; bid /= d.tiles_r;
  const int64_t batch = bid;
  const float* base = (const float*)d.src + batch * d.sb * (d.mode ? 2 : 1);
  const int64_t r0 = tr * PK_ROWS, k0 = tk * PK_K;
  const int t = threadIdx.x;
  if (d.rfast) {
    // lanes along rows
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      const int r = t & 127, k = (t >> 7) + 2 * i;
      tile[r][k] = pack_fetch(d, base, r0 + r, k0 + k);
    }
  } else {
    // lanes along k
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
      const int k = t & 31, r = (t >> 5) + 8 * i;
      tile[r][k] = pack_fetch(d, base, r0 + r, k0 + k);
    }
  }
  __syncthreads();
  uint16_t* out = d.dst + ((batch * d.tiles_r + tr) * d.tiles_k + tk) * (int64_t)PK_TILE;
  // 512 chunks of 8 bf16 per plane; thread t emits chunks t and t + 256 of both planes
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = t + 256 * q;
    const int r = c >> 2, kq = (c & 3) * 8;
    uint16_t h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float x = tile[r][kq + e];
      h[e] = f2bf_rn(x);
      l[e] = f2bf_rn(x - __uint_as_float((uint32_t)h[e] << 16));
    }
    uint4 vh, vl;
    vh.x = h[0] | ((uint32_t)h[1] << 16); vh.y = h[2] | ((uint32_t)h[3] << 16);
    vh.z = h[4] | ((uint32_t)h[5] << 16); vh.w = h[6] | ((uint32_t)h[7] << 16);
    vl.x = l[0] | ((uint32_t)l[1] << 16); vl.y = l[2] | ((uint32_t)l[3] << 16);
    vl.z = l[4] | ((uint32_t)l[5] << 16); vl.w = l[6] | ((uint32_t)l[7] << 16);
    *(uint4*)(out + c * 8) = vh;
    *(uint4*)(out + PK_ROWS * PK_K + c * 8) = vl;
  }
}

// ---- GEMM on packed operands ---------------------------------------------------------------
struct PackedGemmDesc {
  const uint16_t* a;   // packed A' (rows m)
  const uint16_t* b;   // packed B' (rows n)
  void* y;
  int64_t M, N, K;     // logical GEMM (after realification: N = 2N0, K = 2K0 for complex)
  int64_t B;
  int64_t tiles_m, tiles_n, tiles_k;
  int64_t sym, syn, syb;  // output strides (complex elements when cplx)
  int64_t N0;             // complex: columns n' >= N0 are imaginary parts
  Scal sx, sy;
  int mode;
  int cplx;
  int ksplit;
  int64_t kt_per_split;
  float* work;            // split-K partials [ksplit][B][M][N]
};

__device__ __forceinline__ void pk_map(int64_t bid, int64_t nblocks, const PackedGemmDesc& d,
                                       int64_t& tm, int64_t& tn, int64_t& batch, int64_t& ks) {
  if ((nblocks & 7) == 0) {
    const int64_t per = nblocks >> 3;
    bid = (bid & 7) * per + (bid >> 3);
  }
  const int64_t per_mat = d.tiles_m * d.tiles_n;
  const int64_t mat = bid / per_mat;
  const int64_t r = bid - mat * per_mat;
  batch = mat % d.B;
  ks = mat / d.B;
  constexpr int64_t G = 8;
  const int64_t group = r / (G * d.tiles_n);
  const int64_t first = group * G;
  int64_t gsz = d.tiles_m - first;
  if (gsz > G) gsz = G;
  const int64_t in = r - group * G * d.tiles_n;
  tm = first + in % gsz;
  tn = in / gsz;
}

__global__ __launch_bounds__(256, 2) void k_gemm_packed_bf16x3(PackedGemmDesc d) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][PK_ROWS][PK_LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][PK_ROWS][PK_LDK];
  int64_t tm, tn, batch, ks;
  pk_map(blockIdx.x, gridDim.x, d, tm, tn, batch, ks);
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
  const int64_t kt0 = ks * d.kt_per_split;
  int64_t kt1 = kt0 + d.kt_per_split;
  if (kt1 > d.tiles_k) kt1 = d.tiles_k;
  const uint4* pa = (const uint4*)(d.a + ((batch * d.tiles_m + tm) * d.tiles_k) * (int64_t)PK_TILE);
  const uint4* pb = (const uint4*)(d.b + ((batch * d.tiles_n + tn) * d.tiles_k) * (int64_t)PK_TILE);
  // chunk c (16 bytes) of a tile: plane c/512, row (c%512)/4, k (c%4)*8
  uint4 ra[4], rb[4];
  auto gload = [&](int64_t kt) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra[q] = pa[kt * (PK_TILE / 8) + t + 256 * q];
      rb[q] = pb[kt * (PK_TILE / 8) + t + 256 * q];
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = t + 256 * q;
      const int p = c >> 9, r = (c & 511) >> 2, kq = (c & 3) * 8;
      *(uint4*)&sA[p][r][kq] = ra[q];
      *(uint4*)&sB[p][r][kq] = rb[q];
    }
  };
  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  if (kt0 < kt1) gload(kt0);
  for (int64_t kt = kt0; kt < kt1; ++kt) {
    lstore();
    __syncthreads();
    if (kt + 1 < kt1) gload(kt + 1);
#pragma unroll
    for (int ksub = 0; ksub < PK_K; ksub += 16) {
      const int kk = ksub + 8 * (lane >> 5);
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
  // epilogue: lane holds m = mbase + (lane & 31), n = nbase + 8*(r/4) + 4*(lane/32) + r%4
  const int64_t m0 = tm * PK_ROWS + wm, n0 = tn * PK_ROWS + wn;
  if (d.work) {
    float* W = d.work + ((ks * d.B + batch) * d.M) * d.N;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t m = m0 + mb * 32 + (lane & 31);
          const int64_t n = n0 + nb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          if (m < d.M && n < d.N) W[m + n * d.M] = acc[nb][mb][r];
        }
    return;
  }
  if (!d.cplx) {
    const float sx = (float)d.sx.re, sy = (float)d.sy.re;
    float* Y = (float*)d.y + batch * d.syb;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t m = m0 + mb * 32 + (lane & 31);
          const int64_t n = n0 + nb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          if (m < d.M && n < d.N) {
            float* yp = Y + m * d.sym + n * d.syn;
            *yp = epilogue<OpsF32>(d.mode, sx, sy, acc[nb][mb][r], yp);
          }
        }
    return;
  }
  // complex with a plain overwrite (sx = 1, sy = 0): column n' < N0 is the real part of
  // column n', n' >= N0 the imaginary part of column n' - N0, stored into their interleaved
  // slots.  Any other epilogue goes through the work buffer and k_packed_reduce.
  float* Y = (float*)d.y + 2 * batch * d.syb;
#pragma unroll
  for (int nb = 0; nb < 2; ++nb)
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t m = m0 + mb * 32 + (lane & 31);
        const int64_t n = n0 + nb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m < d.M && n < d.N) {
          const bool im = n >= d.N0;
          const int64_t nn = im ? n - d.N0 : n;
          Y[2 * (m * d.sym + nn * d.syn) + (im ? 1 : 0)] = acc[nb][mb][r];
        }
      }
}

// Split-K reduction for the packed GEMM (real, or complex with column halves re / im).
__global__ __launch_bounds__(256) void k_packed_reduce(PackedGemmDesc d) {
  const int64_t per = d.M * d.N, total = per * d.B;
  const int64_t outs = d.cplx ? d.M * d.N0 * d.B : total;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < outs;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!d.cplx) {
      const int64_t batch = e / per, r = e - batch * per;
      const int64_t n = r / d.M, m = r - n * d.M;
      float acc = 0.f;
      for (int s = 0; s < d.ksplit; ++s) acc += d.work[s * total + e];
      float* yp = (float*)d.y + batch * d.syb + m * d.sym + n * d.syn;
      *yp = epilogue<OpsF32>(d.mode, (float)d.sx.re, (float)d.sy.re, acc, yp);
    } else {
      const int64_t per0 = d.M * d.N0;
      const int64_t batch = e / per0, r = e - batch * per0;
      const int64_t n = r / d.M, m = r - n * d.M;
      c32 acc = {0.f, 0.f};
      for (int s = 0; s < d.ksplit; ++s) {
        acc.x += d.work[s * total + batch * per + m + n * d.M];
        acc.y += d.work[s * total + batch * per + m + (n + d.N0) * d.M];
      }
      c32* yp = (c32*)d.y + batch * d.syb + m * d.sym + n * d.syn;
      *yp = epilogue<OpsC32>(d.mode, OpsC32::from(d.sx), OpsC32::from(d.sy), acc, yp);
    }
  }
}

// ------------------------------------------------------------------------------------------

bool packed_plain_overwrite(const GemmDesc& g) {
  return g.sx.re == 1.0 && g.sx.im == 0.0 && g.sy.re == 0.0 && g.sy.im == 0.0;
}

int64_t packed_bytes(int64_t rows, int64_t k, int64_t batch) {
  const int64_t tr = (rows + PK_ROWS - 1) / PK_ROWS, tk = (k + PK_K - 1) / PK_K;
  return batch * tr * tk * (int64_t)PK_TILE * 2;
}

static void launch_pack(const void* src, uint16_t* dst, int64_t R, int64_t K, int64_t B,
                        int64_t sr, int64_t sk, int64_t sb, int mode, int64_t R0, int64_t K0,
                        hipStream_t st) {
  PackDesc p{};
  p.src = src;
  p.dst = dst;
  p.R = R;
  p.K = K;
  p.sr = sr;
  p.sk = sk;
  p.sb = sb;
  p.R0 = R0;
  p.K0 = K0;
  p.mode = mode;
  p.tiles_r = (R + PK_ROWS - 1) / PK_ROWS;
  p.tiles_k = (K + PK_K - 1) / PK_K;
  p.rfast = (sr == 1 && sk != 1) ? 1 : 0;
  const int64_t blocks = p.tiles_r * p.tiles_k * B;
  hipLaunchKernelGGL(k_pack_bf16x3, dim3((uint32_t)blocks), dim3(256), 0, st, p);
}

// Real F32 or complex C32 GEMM on packed BF16x3 operands.  pa/pb: workspace of
// packed_bytes(...) each; work: split-K partials (ksplit*B*M'*N' floats) or null.
void launch_gemm_packed(int dt, const GemmDesc& g, void* pa, void* pb, float* work, int ksplit,
                        hipStream_t st) {
  OME_DRY_RETURN(ksplit > 1 ? 4 : 3);
  const bool cplx = (dt == C32);
  PackedGemmDesc d{};
  const int64_t M = g.M, N = cplx ? 2 * g.N : g.N, K = cplx ? 2 * g.K : g.K;
  // A' rows m (stride sam), k (stride sak); B' rows n (stride sbn), k (stride sbk)
  launch_pack(g.a, (uint16_t*)pa, M, K, g.B, g.sam, g.sak, g.sab, cplx ? 1 : 0, g.M, g.K, st);
  launch_pack(g.b, (uint16_t*)pb, N, K, g.B, g.sbn, g.sbk, g.sbb, cplx ? 2 : 0, g.N, g.K, st);
  d.a = (const uint16_t*)pa;
  d.b = (const uint16_t*)pb;
  d.y = g.y;
  d.M = M;
  d.N = N;
  d.K = K;
  d.B = g.B;
  d.tiles_m = (M + PK_ROWS - 1) / PK_ROWS;
  d.tiles_n = (N + PK_ROWS - 1) / PK_ROWS;
  d.tiles_k = (K + PK_K - 1) / PK_K;
  d.sym = g.sym;
  d.syn = g.syn;
  d.syb = g.syb;
  d.N0 = g.N;
  d.sx = g.sx;
  d.sy = g.sy;
  d.mode = cplx ? epi_mode<OpsC32>(g.sx, g.sy) : epi_mode<OpsF32>(g.sx, g.sy);
  d.cplx = cplx ? 1 : 0;
  if (ksplit < 1) ksplit = 1;
  int64_t per = (d.tiles_k + ksplit - 1) / ksplit;
  ksplit = (int)((d.tiles_k + per - 1) / per);
  d.kt_per_split = per;
  d.ksplit = ksplit;
  // partial products go through the work buffer when K is split or when a complex result
  // needs an epilogue other than a plain overwrite
  const bool via_work = ksplit > 1 || (cplx && !packed_plain_overwrite(g));
  d.work = via_work ? work : nullptr;
  const int64_t blocks = d.tiles_m * d.tiles_n * d.B * ksplit;
  hipLaunchKernelGGL(k_gemm_packed_bf16x3, dim3((uint32_t)blocks), dim3(256), 0, st, d);
  if (via_work) {
    const int64_t outs = cplx ? g.M * g.N * g.B : M * N * g.B;
    int64_t gr = (outs + 255) / 256;
    if (gr > 65536) gr = 65536;
    hipLaunchKernelGGL(k_packed_reduce, dim3((uint32_t)gr), dim3(256), 0, st, d);
  }
}

}  // namespace ome
