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
#include <cstdlib>
#include <cstring>

#include "kernels.h"

namespace ome {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int PK_K = 32;          // k per packed tile
// elements of one packed (row tile, k tile) block: np bf16 planes of rows x 32
__host__ __device__ constexpr int pk_tile_elems(int rows, int np = 2) { return np * rows * PK_K; }

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

// The element (r, k) of the packed operand as a float address (the operand's first float
// when (r, k) lies outside it) and whether its value is negated or replaced by zero: the
// general pack paths issue a batch of these loads before any of them is used.
__device__ __forceinline__ const float* pack_addr(const PackDesc& d, const float* __restrict__ base,
                                                  int64_t r, int64_t k, bool& neg, bool& ok) {
  ok = r < d.R && k < d.K;
  neg = false;
  if (!ok) return base;
  if (d.mode == 0) return base + r * d.sr + k * d.sk;
  if (d.mode == 1) {  // A' = [Ar | Ai]
    const bool im = k >= d.K0;
    const int64_t kk = im ? k - d.K0 : k;
    return base + 2 * (r * d.sr + kk * d.sk) + (im ? 1 : 0);
  }
  // B' (2K x 2N) stored as rows n' (2N), k' (2K): B'[k'][n'] with
  //   k' <  K, n' <  N:  Br[k][n]     k' <  K, n' >= N:  Bi[k][n]
  //   k' >= K, n' <  N: -Bi[k][n]     k' >= K, n' >= N:  Br[k][n]
  const bool kim = k >= d.K0, nim = r >= d.R0;
  const int64_t kk = kim ? k - d.K0 : k, nn = nim ? r - d.R0 : r;
  neg = kim && !nim;
  return base + 2 * (nn * d.sr + kk * d.sk) + (kim != nim ? 1 : 0);
}

// Swizzle of the k-major staging tile: float4 group g of rows at k is stored at group
// g ^ pk_swz(k).  The four emitter lanes of one row read k 8 apart (same bank without the
// swizzle); xor-ing bits 2-3 of the group with k / 8 moves them 16 banks apart, so the 64
// lanes of a wave reach every bank equally.  Within a row of groups the xor is a permutation,
// so the float4 stores stay contiguous.
__device__ __forceinline__ int pk_swz(int k) { return ((k >> 3) & 3) << 2; }
static_assert(PK_K == 32, "the staging tile holds 32 k per row");

// TR rows per packed tile (128 or 256); NP bf16 planes (2: hi, lo; 3: hi, mid, lo).
template <int TR, int NP>
__device__ __forceinline__ void pack_tile(const PackDesc& d, int64_t bid, float (*tile)[PK_K + 1]) {
  const int64_t tk = bid % d.tiles_k; bid /= d.tiles_k;
  const int64_t tr = bid % d.tiles_r; bid /= d.tiles_r;
  const int64_t batch = bid;
  const float* base = (const float*)d.src + batch * d.sb * (d.mode ? 2 : 1);
  const int64_t r0 = tr * TR, k0 = tk * PK_K;
  const int t = threadIdx.x;
  // fast path: real operand, a unit-stride dim, a full tile and 16-byte aligned runs:
  // TR*32/1024 float4 loads per thread, all issued before the LDS stores
  constexpr int NV = TR * PK_K / 1024;
  const bool full = d.mode == 0 && r0 + TR <= d.R && k0 + PK_K <= d.K;
  // the row-fast fast path stages the tile k-major, tt[k][row] (same LDS), with the float4
  // groups of rows swizzled by bit 3 of k: its float4 stores and the emitter's reads along k
  // are then free of bank conflicts (a row-major tile took 8-way conflicted scalar stores)
  float(*tt)[TR] = (float(*)[TR])&tile[0][0];
  const bool kmaj = full && d.rfast && d.sr == 1 && (d.sk % 4) == 0 && (((uintptr_t)base) & 15) == 0;
  if (kmaj) {
    float4 v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = t + 256 * i;
      const int r4 = (e % (TR / 4)) * 4, k = e / (TR / 4);
      v[i] = *(const float4*)(base + (r0 + r4) + (k0 + k) * d.sk);
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = t + 256 * i;
      const int g = e % (TR / 4), k = e / (TR / 4);
      *(float4*)&tt[k][4 * (g ^ pk_swz(k))] = v[i];
    }
  } else if (full && d.sk == 1 && (d.sr % 4) == 0 && (((uintptr_t)base) & 15) == 0) {
    float4 v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = t + 256 * i;
      const int k4 = (e % (PK_K / 4)) * 4, r = e / (PK_K / 4);
      v[i] = *(const float4*)(base + (r0 + r) * d.sr + k0 + k4);
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = t + 256 * i;
      const int k4 = (e % (PK_K / 4)) * 4, r = e / (PK_K / 4);
      tile[r][k4] = v[i].x;
      tile[r][k4 + 1] = v[i].y;
      tile[r][k4 + 2] = v[i].z;
      tile[r][k4 + 3] = v[i].w;
    }
  } else {
    // lanes along rows (rfast) or along k; batches of 8 loads issued before their LDS stores
#pragma unroll
    for (int i0 = 0; i0 < TR / 8; i0 += 8) {
      float v[8];
      uint32_t ok = 0, neg = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int i = i0 + j;
        const int r = d.rfast ? (t + 256 * i) % TR : (t >> 5) + 8 * i;
        const int k = d.rfast ? (t + 256 * i) / TR : (t & 31);
        bool n_, o_;
        v[j] = *pack_addr(d, base, r0 + r, k0 + k, n_, o_);
        ok |= o_ ? 1u << j : 0u;
        neg |= n_ ? 1u << j : 0u;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int i = i0 + j;
        const int r = d.rfast ? (t + 256 * i) % TR : (t >> 5) + 8 * i;
        const int k = d.rfast ? (t + 256 * i) / TR : (t & 31);
        const float x = (neg >> j) & 1u ? -v[j] : v[j];
        tile[r][k] = (ok >> j) & 1u ? x : 0.f;
      }
    }
  }
  __syncthreads();
  uint16_t* out = d.dst + ((batch * d.tiles_r + tr) * d.tiles_k + tk) * (int64_t)pk_tile_elems(TR, NP);
  // TR*4 chunks of 8 bf16 per plane; thread t emits chunks t + 256q of every plane
#pragma unroll
  for (int q = 0; q < TR / 64; ++q) {
    const int c = t + 256 * q;
    const int r = c >> 2, kq = (c & 3) * 8;
    float x[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = kq + e;
      x[e] = kmaj ? tt[k][4 * ((r >> 2) ^ pk_swz(k)) + (r & 3)] : tile[r][k];
    }
    // row r's 16-byte chunk j lives at slot j ^ ((r >> 2) & 3): conflict-free b128 reads
    const int slot = r * 4 + ((c & 3) ^ ((r >> 2) & 3));
    if constexpr (NP == 2) {
      uint4 vh, vl;
      split2_bf16(x[0], x[1], vh.x, vl.x);
      split2_bf16(x[2], x[3], vh.y, vl.y);
      split2_bf16(x[4], x[5], vh.z, vl.z);
      split2_bf16(x[6], x[7], vh.w, vl.w);
      *(uint4*)(out + slot * 8) = vh;
      *(uint4*)(out + TR * PK_K + slot * 8) = vl;
    } else {
      uint4 vh, vm, vl;
      split3_bf16(x[0], x[1], vh.x, vm.x, vl.x);
      split3_bf16(x[2], x[3], vh.y, vm.y, vl.y);
      split3_bf16(x[4], x[5], vh.z, vm.z, vl.z);
      split3_bf16(x[6], x[7], vh.w, vm.w, vl.w);
      *(uint4*)(out + slot * 8) = vh;
      *(uint4*)(out + TR * PK_K + slot * 8) = vm;
      *(uint4*)(out + 2 * TR * PK_K + slot * 8) = vl;
    }
  }
}

template <int TR>
__global__ __launch_bounds__(256) void k_pack_bf16x3(PackDesc d) {
  __shared__ float tile[TR][PK_K + 1];
  pack_tile<TR, 2>(d, blockIdx.x, tile);
}

// Both operands of one GEMM in one launch: blocks [0, na) pack a, the rest pack b.
template <int TR>
__global__ __launch_bounds__(256) void k_pack2_bf16x3(PackDesc a, PackDesc b, int64_t na) {
  __shared__ float tile[TR][PK_K + 1];
  if ((int64_t)blockIdx.x < na) pack_tile<TR, 2>(a, blockIdx.x, tile);
  else pack_tile<TR, 2>(b, (int64_t)blockIdx.x - na, tile);
}

// Three-plane pack of both operands in one launch: a in TRA-row tiles, b in TRB-row tiles.
template <int TRA, int TRB>
__global__ __launch_bounds__(256) void k_pack2_bf16x6(PackDesc a, PackDesc b, int64_t na) {
  __shared__ float tile[TRA > TRB ? TRA : TRB][PK_K + 1];
  if ((int64_t)blockIdx.x < na) pack_tile<TRA, 3>(a, blockIdx.x, tile);
  else pack_tile<TRB, 3>(b, (int64_t)blockIdx.x - na, tile);
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

// Workgroup tile TM x TN (4 waves, 2 x 2, each (TM/2) x (TN/2) = MB x NB blocks of 32x32).
// Per K-step of 32:
//   1. wait for this step's tiles (copied global -> LDS by global_load_lds_dwordx4, one 1 KiB
//      block per instruction), barrier;
//   2. every wave reads all of its fragments for the step (both 16-k halves, hi and lo planes);
//   3. barrier, then issue the copies of the next step into the same LDS buffer;
//   4. run the step's MFMAs from registers while those copies are in flight.
// s_waitcnt immediate that waits until at most n vector-memory operations are outstanding
// (vmcnt bits [3:0] and [15:14]; expcnt and lgkmcnt left at their no-wait maxima).
constexpr int waitcnt_vm(int n) { return (n & 0xF) | (((n >> 4) & 3) << 14) | (0x7 << 4) | (0xF << 8); }

// Workgroup barrier without the memory fence of __syncthreads (which would wait for every
// outstanding copy).  Callers wait for what they need first; the empty asm statements keep
// the compiler from moving LDS accesses across the barrier.
__device__ __forceinline__ void raw_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
constexpr int WAIT_LGKM0 = 0xC07F;  // lgkmcnt(0): this wave's LDS reads have completed

// Scheduling groups for a region holding NM MFMAs, NV vector-memory loads (the LDS copies) and
// ND LDS reads: one MFMA ahead of each load, then one ahead of each read, then the remaining
// MFMAs (memory operations beyond NM follow in one group).  Masks: 0x008 MFMA, 0x020 VMEM
// read, 0x100 DS read.
template <int NM, int NV, int ND>
__device__ __forceinline__ void sched_interleave() {
  constexpr int PV = NV < NM ? NV : NM;
  constexpr int PD = ND < NM - PV ? ND : NM - PV;
#pragma unroll
  for (int i = 0; i < PV; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
  }
#pragma unroll
  for (int i = 0; i < PD; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
  }
  if constexpr (NV > PV) __builtin_amdgcn_sched_group_barrier(0x020, NV - PV, 0);
  if constexpr (ND > PD) __builtin_amdgcn_sched_group_barrier(0x100, ND - PD, 0);
  if constexpr (NM > PV + PD) __builtin_amdgcn_sched_group_barrier(0x008, NM - PV - PD, 0);
}

template <int TM, int TN, int STAGES>
__global__ __launch_bounds__(256) void k_gemm_packed_bf16x3(PackedGemmDesc d) {
  constexpr int MB = TM / 64, NB = TN / 64;
  constexpr int A_BYTES = TM * PK_K * 2 * 2, B_BYTES = TN * PK_K * 2 * 2;
  constexpr int STAGE = A_BYTES + B_BYTES;
  constexpr int CHUNKS = STAGE / 1024;  // 1 KiB copies per K-step
  constexpr int CPW = CHUNKS / 4;       // copies per wave per K-step
  static_assert(CHUNKS % 4 == 0, "copies must split evenly over 4 waves");
  __shared__ __attribute__((aligned(1024))) uint8_t lds[STAGES * STAGE];
  int64_t tm, tn, batch, ks;
  pk_map(blockIdx.x, gridDim.x, d, tm, tn, batch, ks);
  const int t = threadIdx.x;
  // the wave index is uniform across the wave: say so, the LDS-copy destination must be
  const int lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = (w & 1) * (TM / 2), wn = (w >> 1) * (TN / 2);
  const int64_t kt0 = ks * d.kt_per_split;
  int64_t kt1 = kt0 + d.kt_per_split;
  if (kt1 > d.tiles_k) kt1 = d.tiles_k;
  const uint8_t* ga = (const uint8_t*)(d.a + ((batch * d.tiles_m + tm) * d.tiles_k) * (int64_t)pk_tile_elems(TM));
  const uint8_t* gb = (const uint8_t*)(d.b + ((batch * d.tiles_n + tn) * d.tiles_k) * (int64_t)pk_tile_elems(TN));
  // wave w copies blocks w*CPW .. +CPW of [A tile | B tile] of step kt into stage kt % STAGES
  auto issue = [&](int64_t kt) __attribute__((always_inline)) {
    const int sbase = (int)(kt % STAGES) * STAGE;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
      const int c = w * CPW + i;
      const uint8_t* src = (c * 1024 < A_BYTES) ? ga + kt * A_BYTES + c * 1024
                                                 : gb + kt * B_BYTES + (c * 1024 - A_BYTES);
      const int loff = __builtin_amdgcn_readfirstlane(sbase + c * 1024);  // M0: wave-uniform
      __builtin_amdgcn_global_load_lds((const void*)(src + lane * 16),
                                       (__attribute__((address_space(3))) void*)(lds + loff),
                                       16, 0, 0);
    }
  };
  // fragment of row r (within its tile), 16-byte chunk j, plane p
  auto frag = [&](int base, int rows, int p, int r, int j) __attribute__((always_inline)) {
    const int off = base + p * rows * 64 + r * 64 + 16 * (j ^ ((r >> 2) & 3));
    return *(const bf16x8_t*)(lds + off);
  };
  floatx16 acc[NB][MB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < MB; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  // prologue: the first copies in flight (STAGES - 1 steps; one step when STAGES == 1)
  constexpr int AHEAD = STAGES > 1 ? STAGES - 1 : 1;
#pragma unroll
  for (int s = 0; s < AHEAD; ++s)
    if (kt0 + s < kt1) issue(kt0 + s);
  const int jl = lane >> 5;  // k half of the 16-k MFMA operand (0: k 0-7, 1: k 8-15)
  for (int64_t kt = kt0; kt < kt1; ++kt) {
    // with STAGES > 1 the copy of step kt+STAGES-1 goes out now, into the buffer of step
    // kt-1 whose fragment reads all waves finished before the last barrier
    if (STAGES > 1 && kt + STAGES - 1 < kt1) issue(kt + STAGES - 1);
    // wait for this wave's copies of step kt; copies of later steps may stay in flight
    const int64_t later = STAGES > 1 ? ((kt1 - 1 - kt) < STAGES - 1 ? (kt1 - 1 - kt) : STAGES - 1) : 0;
    if (later >= 2) __builtin_amdgcn_s_waitcnt(waitcnt_vm(2 * CPW));
    else if (later == 1) __builtin_amdgcn_s_waitcnt(waitcnt_vm(CPW));
    else __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    raw_barrier();  // every wave's copies of step kt are in LDS
    const int sbase = (int)(kt % STAGES) * STAGE;
    bf16x8_t ah[2][NB], al[2][NB], bh[2][MB], bl[2][MB];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int r = wn + nb * 32 + (lane & 31);
        ah[h][nb] = frag(sbase + A_BYTES, TN, 0, r, 2 * h + jl);
        al[h][nb] = frag(sbase + A_BYTES, TN, 1, r, 2 * h + jl);
      }
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const int r = wm + mb * 32 + (lane & 31);
        bh[h][mb] = frag(sbase, TM, 0, r, 2 * h + jl);
        bl[h][mb] = frag(sbase, TM, 1, r, 2 * h + jl);
      }
    }
    __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
    raw_barrier();  // every wave has its fragments of step kt: the buffer may be refilled
    if (STAGES == 1 && kt + 1 < kt1) issue(kt + 1);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[h][nb], bh[h][mb], acc[nb][mb], 0, 0, 0);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[h][nb], bl[h][mb], acc[nb][mb], 0, 0, 0);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[h][nb], bh[h][mb], acc[nb][mb], 0, 0, 0);
    }
  }
  // epilogue: lane holds m = mbase + (lane & 31), n = nbase + 8*(r/4) + 4*(lane/32) + r%4
  const int64_t m0 = tm * TM + wm, n0 = tn * TN + wn;
  if (d.work) {
    float* W = d.work + ((ks * d.B + batch) * d.M) * d.N;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
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
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
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
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
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

// Software-pipelined variant (two LDS stages, fragments double-buffered in registers).
// Per K-step kt (two 16-k sub-steps s0, s1), with the fragments of s0(kt) already loaded:
//   read s1(kt) | MFMAs of s0(kt)
//   wait: copy of kt+1 landed, own reads done; barrier
//   issue the copy of kt+2 into the stage kt just vacated
//   read s0(kt+1) | MFMAs of s1(kt)
// so LDS reads are always in flight under MFMAs and there is one barrier per K-step.
// WM x WN waves (64 * WM * WN threads); each wave owns a (TM/WM) x (TN/WN) block of 32x32
// accumulators.
template <int TM, int TN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_packed_bf16x3_pipe(PackedGemmDesc d) {
  constexpr int NW = WM * WN;
  constexpr int MB = TM / WM / 32, NB = TN / WN / 32;
  constexpr int A_BYTES = TM * PK_K * 2 * 2, B_BYTES = TN * PK_K * 2 * 2;
  constexpr int STAGE = A_BYTES + B_BYTES;
  constexpr int CHUNKS = STAGE / 1024;
  constexpr int CPW = CHUNKS / NW;
  static_assert(CHUNKS % NW == 0, "copies must split evenly over the waves");
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * STAGE];
  int64_t tm, tn, batch, ks;
  pk_map(blockIdx.x, gridDim.x, d, tm, tn, batch, ks);
  const int t = threadIdx.x;
  const int lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = (w % WM) * (TM / WM), wn = (w / WM) * (TN / WN);
  const int64_t kt0 = ks * d.kt_per_split;
  int64_t kt1 = kt0 + d.kt_per_split;
  if (kt1 > d.tiles_k) kt1 = d.tiles_k;
  const uint8_t* ga = (const uint8_t*)(d.a + ((batch * d.tiles_m + tm) * d.tiles_k) * (int64_t)pk_tile_elems(TM));
  const uint8_t* gb = (const uint8_t*)(d.b + ((batch * d.tiles_n + tn) * d.tiles_k) * (int64_t)pk_tile_elems(TN));
  // each wave copies CPW consecutive 1 KiB blocks of one operand: the first A_BYTES / 1024
  // blocks of a stage are A's, so the source and its per-step stride are fixed per wave
  constexpr int A_CH = A_BYTES / 1024;
  static_assert(A_CH % CPW == 0, "no wave may straddle the A / B boundary");
  const int c0 = w * CPW;
  const bool wave_a = c0 < A_CH;
  const uint8_t* gsrc = (wave_a ? ga + c0 * 1024 : gb + (c0 - A_CH) * 1024) + lane * 16;
  const int64_t gstep = wave_a ? A_BYTES : B_BYTES;
  auto issue = [&](int64_t kt) __attribute__((always_inline)) {
    const int sbase = (int)(kt & 1) * STAGE + c0 * 1024;
    const uint8_t* src = gsrc + kt * gstep;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
      const int loff = __builtin_amdgcn_readfirstlane(sbase + i * 1024);
      __builtin_amdgcn_global_load_lds((const void*)(src + i * 1024),
                                       (__attribute__((address_space(3))) void*)(lds + loff),
                                       16, 0, 0);
    }
  };
  const int jl = lane >> 5;
  // fragment reads of sub-step h of the step held in stage sb
  auto rd = [&](int sb, int h, bf16x8_t (&ah)[NB], bf16x8_t (&al)[NB], bf16x8_t (&bh)[MB],
                bf16x8_t (&bl)[MB]) __attribute__((always_inline)) {
    const int base = sb * STAGE;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int r = wn + nb * 32 + (lane & 31);
      const int off = base + A_BYTES + r * 64 + 16 * ((2 * h + jl) ^ ((r >> 2) & 3));
      ah[nb] = *(const bf16x8_t*)(lds + off);
      al[nb] = *(const bf16x8_t*)(lds + off + TN * 64);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int r = wm + mb * 32 + (lane & 31);
      const int off = base + r * 64 + 16 * ((2 * h + jl) ^ ((r >> 2) & 3));
      bh[mb] = *(const bf16x8_t*)(lds + off);
      bl[mb] = *(const bf16x8_t*)(lds + off + TM * 64);
    }
  };
  floatx16 acc[NB][MB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < MB; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  auto mm = [&](const bf16x8_t (&ah)[NB], const bf16x8_t (&al)[NB], const bf16x8_t (&bh)[MB],
                const bf16x8_t (&bl)[MB]) __attribute__((always_inline)) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[nb], bh[mb], acc[nb][mb], 0, 0, 0);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[nb], bl[mb], acc[nb][mb], 0, 0, 0);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[nb], bh[mb], acc[nb][mb], 0, 0, 0);
  };
  bf16x8_t ah0[NB], al0[NB], bh0[MB], bl0[MB], ah1[NB], al1[NB], bh1[MB], bl1[MB];
  if (kt0 < kt1) {
    issue(kt0);
    if (kt0 + 1 < kt1) issue(kt0 + 1);
    if (kt0 + 1 < kt1) __builtin_amdgcn_s_waitcnt(waitcnt_vm(CPW));
    else __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    raw_barrier();
    rd((int)(kt0 & 1), 0, ah0, al0, bh0, bl0);  // a split-K slice may start on an odd step
  }
  // sched_barrier(0) fences the instruction scheduler so that each group of LDS reads stays
  // ahead of the MFMAs it overlaps, and the MFMAs stay ahead of the next wait.
  // Steady state (a copy to issue every step): after the barrier the copies of kt+2 and the
  // reads of s0(kt+1) are interleaved one per MFMA of s1(kt), so the matrix pipe keeps running
  // while both waves of a SIMD issue their memory operations.
  constexpr int N_MFMA = 3 * NB * MB, N_RD = 2 * (NB + MB);
  int64_t kt = kt0;
  for (; kt + 2 < kt1; ++kt) {
    const int sb = (int)(kt & 1);
    // the fragments of sub-step 0 were read under the previous MFMAs; retire them here so
    // that the reads below are not waited on before the MFMAs that do not need them
    __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
    __builtin_amdgcn_sched_barrier(0);
    rd(sb, 1, ah1, al1, bh1, bl1);
    __builtin_amdgcn_sched_barrier(0);
    mm(ah0, al0, bh0, bl0);
    __builtin_amdgcn_sched_barrier(0);
    // copy of kt+1 landed (the copy of kt+2 is not issued yet), own reads of stage sb done
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0) & WAIT_LGKM0);
    raw_barrier();
    issue(kt + 2);
    rd(sb ^ 1, 0, ah0, al0, bh0, bl0);
    mm(ah1, al1, bh1, bl1);
    sched_interleave<N_MFMA, CPW, N_RD>();
    __builtin_amdgcn_sched_barrier(0);
  }
  // the last two steps: nothing left to copy
  for (; kt < kt1; ++kt) {
    const int sb = (int)(kt & 1);
    __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
    __builtin_amdgcn_sched_barrier(0);
    rd(sb, 1, ah1, al1, bh1, bl1);
    __builtin_amdgcn_sched_barrier(0);
    mm(ah0, al0, bh0, bl0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0) & WAIT_LGKM0);
    raw_barrier();
    if (kt + 1 < kt1) rd(sb ^ 1, 0, ah0, al0, bh0, bl0);
    __builtin_amdgcn_sched_barrier(0);
    mm(ah1, al1, bh1, bl1);
    __builtin_amdgcn_sched_barrier(0);
  }
  const int64_t m0 = tm * TM + wm, n0 = tn * TN + wn;
  if (d.work) {
    float* W = d.work + ((ks * d.B + batch) * d.M) * d.N;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
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
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
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
  float* Y = (float*)d.y + 2 * batch * d.syb;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
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

// Six-product form on three planes (BF16x6): the same pipeline as the kernel above, with
// the packed tiles holding planes 0 (hi), 1 (mid), 2 (lo) and every fragment set read for
// all three.  Per 32 x 32 block and 16-k sub-step it issues the six products a_p * b_q with
// p + q <= 2, smallest first: (2,0) (1,1) (0,2) (1,0) (0,1) (0,0).  The dropped terms are
// below 2^-23 |a||b| per product, the FP32 rounding scale.
// Tiles TM x TN = 256 x 128: a stage holds (256 + 128) x 32 k x 3 planes (72 KiB), two stages
// fit the LDS; WM x WN = 4 x 2 waves of 64 x 64.  A stage's 72 1-KiB copies split 9 per wave,
// so one wave's copies straddle the A / B boundary: the source is chosen per copy.
template <int TM, int TN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_packed_bf16x6_pipe(PackedGemmDesc d) {
  constexpr int NP = 3;
  constexpr int NW = WM * WN;
  constexpr int MB = TM / WM / 32, NB = TN / WN / 32;
  constexpr int A_PLANE = TM * PK_K * 2, B_PLANE = TN * PK_K * 2;
  constexpr int A_BYTES = NP * A_PLANE, B_BYTES = NP * B_PLANE;
  constexpr int STAGE = A_BYTES + B_BYTES;
  constexpr int CHUNKS = STAGE / 1024;
  constexpr int CPW = CHUNKS / NW;
  constexpr int A_CH = A_BYTES / 1024;
  static_assert(CHUNKS % NW == 0, "copies must split evenly over the waves");
  static_assert(2 * STAGE <= 160 * 1024, "two stages must fit the LDS");
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * STAGE];
  int64_t tm, tn, batch, ks;
  pk_map(blockIdx.x, gridDim.x, d, tm, tn, batch, ks);
  const int t = threadIdx.x;
  const int lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = (w % WM) * (TM / WM), wn = (w / WM) * (TN / WN);
  const int64_t kt0 = ks * d.kt_per_split;
  int64_t kt1 = kt0 + d.kt_per_split;
  if (kt1 > d.tiles_k) kt1 = d.tiles_k;
  const uint8_t* ga = (const uint8_t*)(d.a + ((batch * d.tiles_m + tm) * d.tiles_k) * (int64_t)pk_tile_elems(TM, NP));
  const uint8_t* gb = (const uint8_t*)(d.b + ((batch * d.tiles_n + tn) * d.tiles_k) * (int64_t)pk_tile_elems(TN, NP));
  const int c0 = w * CPW;
  auto issue = [&](int64_t kt) __attribute__((always_inline)) {
    const int sbase = (int)(kt & 1) * STAGE;
    const uint8_t* sa = ga + kt * A_BYTES + lane * 16;
    const uint8_t* sbb = gb + kt * B_BYTES + lane * 16;
#pragma unroll
    for (int i = 0; i < CPW; ++i) {
      const int c = c0 + i;
      const uint8_t* src = c < A_CH ? sa + c * 1024 : sbb + (c - A_CH) * 1024;
      const int loff = __builtin_amdgcn_readfirstlane(sbase + c * 1024);
      __builtin_amdgcn_global_load_lds((const void*)src,
                                       (__attribute__((address_space(3))) void*)(lds + loff),
                                       16, 0, 0);
    }
  };
  const int jl = lane >> 5;
  // fragments of sub-step h of the step in stage sb: fn[p][nb] from the B tile (rows n, the
  // MFMA's first operand), fm[p][mb] from the A tile (rows m), plane p
  auto rd = [&](int sb, int h, bf16x8_t (&fn)[NP][NB], bf16x8_t (&fm)[NP][MB])
      __attribute__((always_inline)) {
    const int base = sb * STAGE;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int r = wn + nb * 32 + (lane & 31);
      const int off = base + A_BYTES + r * 64 + 16 * ((2 * h + jl) ^ ((r >> 2) & 3));
#pragma unroll
      for (int p = 0; p < NP; ++p) fn[p][nb] = *(const bf16x8_t*)(lds + off + p * B_PLANE);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int r = wm + mb * 32 + (lane & 31);
      const int off = base + r * 64 + 16 * ((2 * h + jl) ^ ((r >> 2) & 3));
#pragma unroll
      for (int p = 0; p < NP; ++p) fm[p][mb] = *(const bf16x8_t*)(lds + off + p * A_PLANE);
    }
  };
  floatx16 acc[NB][MB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < MB; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  auto mm = [&](const bf16x8_t (&fn)[NP][NB], const bf16x8_t (&fm)[NP][MB])
      __attribute__((always_inline)) {
    constexpr int PN[6] = {2, 1, 0, 1, 0, 0}, PM[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
    for (int q = 0; q < 6; ++q)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
          acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fn[PN[q]][nb], fm[PM[q]][mb],
                                                                acc[nb][mb], 0, 0, 0);
  };
  bf16x8_t fn0[NP][NB], fm0[NP][MB], fn1[NP][NB], fm1[NP][MB];
  if (kt0 < kt1) {
    issue(kt0);
    if (kt0 + 1 < kt1) issue(kt0 + 1);
    if (kt0 + 1 < kt1) __builtin_amdgcn_s_waitcnt(waitcnt_vm(CPW));
    else __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    raw_barrier();
    rd((int)(kt0 & 1), 0, fn0, fm0);
  }
  constexpr int N_MFMA = 6 * NB * MB, N_RD = NP * (NB + MB);
  int64_t kt = kt0;
  for (; kt + 2 < kt1; ++kt) {
    const int sb = (int)(kt & 1);
    __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
    __builtin_amdgcn_sched_barrier(0);
    rd(sb, 1, fn1, fm1);
    __builtin_amdgcn_sched_barrier(0);
    mm(fn0, fm0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0) & WAIT_LGKM0);
    raw_barrier();
    issue(kt + 2);
    rd(sb ^ 1, 0, fn0, fm0);
    mm(fn1, fm1);
    sched_interleave<N_MFMA, CPW, N_RD>();
    __builtin_amdgcn_sched_barrier(0);
  }
  for (; kt < kt1; ++kt) {
    const int sb = (int)(kt & 1);
    __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
    __builtin_amdgcn_sched_barrier(0);
    rd(sb, 1, fn1, fm1);
    __builtin_amdgcn_sched_barrier(0);
    mm(fn0, fm0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0) & WAIT_LGKM0);
    raw_barrier();
    if (kt + 1 < kt1) rd(sb ^ 1, 0, fn0, fm0);
    __builtin_amdgcn_sched_barrier(0);
    mm(fn1, fm1);
    __builtin_amdgcn_sched_barrier(0);
  }
  const int64_t m0 = tm * TM + wm, n0 = tn * TN + wn;
  if (d.work) {
    float* W = d.work + ((ks * d.B + batch) * d.M) * d.N;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
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
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
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
  float* Y = (float*)d.y + 2 * batch * d.syb;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
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

// Workgroup tile (TM x TN) for a packed GEMM.  OME_PACKED_TILE=128|256x128|128x256|256|256x256
// forces one for measurements.
static void packed_tiles(int64_t M, int64_t N, int& TM, int& TN, int np) {
  // three planes: one tile shape (two stages of 256 x 128 fill the LDS)
  if (np == 3) {
    TM = 256;
    TN = 128;
    return;
  }
  static int forced = -1;
  if (forced < 0) {
    const char* e = getenv("OME_PACKED_TILE");
    forced = 0;
    if (e) {
      if (!strcmp(e, "128")) forced = 1;
      else if (!strcmp(e, "256x128")) forced = 2;
      else if (!strcmp(e, "128x256")) forced = 3;
      else if (!strcmp(e, "256")) forced = 4;
      else if (!strcmp(e, "256x256")) forced = 5;
    }
  }
  switch (forced) {
    case 1: TM = 128; TN = 128; return;
    case 2: TM = 256; TN = 128; return;
    case 3: TM = 128; TN = 256; return;
    case 4: TM = 256; TN = 128; return;
    case 5: TM = 256; TN = 256; return;
  }
  // 256 x 256 workgroup tiles (8 waves of 128 x 64) halve the LDS traffic per MFMA; use them
  // when they still give at least half a tile per CU, else 128 x 128
  if (((M + 255) / 256) * ((N + 255) / 256) >= 128) {
    TM = 256;
    TN = 256;
    return;
  }
  TM = 128;
  TN = 128;
}

int64_t packed_bytes_a(int64_t M, int64_t N, int64_t K, int64_t batch, int np) {
  int TM, TN;
  packed_tiles(M, N, TM, TN, np);
  return batch * ((M + TM - 1) / TM) * ((K + PK_K - 1) / PK_K) * (int64_t)pk_tile_elems(TM, np) * 2;
}
int64_t packed_bytes_b(int64_t M, int64_t N, int64_t K, int64_t batch, int np) {
  int TM, TN;
  packed_tiles(M, N, TM, TN, np);
  return batch * ((N + TN - 1) / TN) * ((K + PK_K - 1) / PK_K) * (int64_t)pk_tile_elems(TN, np) * 2;
}
int64_t packed_tiles_mn(int64_t M, int64_t N, int np) {
  int TM, TN;
  packed_tiles(M, N, TM, TN, np);
  return ((M + TM - 1) / TM) * ((N + TN - 1) / TN);
}

static PackDesc pack_desc(const void* src, uint16_t* dst, int64_t R, int64_t K, int64_t sr,
                          int64_t sk, int64_t sb, int mode, int64_t R0, int64_t K0, int TR) {
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
  p.tiles_r = (R + TR - 1) / TR;
  p.tiles_k = (K + PK_K - 1) / PK_K;
  p.rfast = (sr == 1 && sk != 1) ? 1 : 0;
  return p;
}

static void launch_pack(const PackDesc& p, int64_t B, int TR, hipStream_t st) {
  const int64_t blocks = p.tiles_r * p.tiles_k * B;
  TimedScope ts(KC_PACK, st);
  if (TR == 256) hipLaunchKernelGGL(k_pack_bf16x3<256>, dim3((uint32_t)blocks), dim3(256), 0, st, p);
  else hipLaunchKernelGGL(k_pack_bf16x3<128>, dim3((uint32_t)blocks), dim3(256), 0, st, p);
}

static void launch_pack2(const PackDesc& a, const PackDesc& b, int64_t B, int TR, hipStream_t st) {
  const int64_t na = a.tiles_r * a.tiles_k * B, nb = b.tiles_r * b.tiles_k * B;
  TimedScope ts(KC_PACK, st);
  if (TR == 256)
    hipLaunchKernelGGL(k_pack2_bf16x3<256>, dim3((uint32_t)(na + nb)), dim3(256), 0, st, a, b, na);
  else
    hipLaunchKernelGGL(k_pack2_bf16x3<128>, dim3((uint32_t)(na + nb)), dim3(256), 0, st, a, b, na);
}

// Real F32 or complex C32 GEMM on packed operands of np bf16 planes (2: BF16x3, 3: BF16x6).
// pa/pb: workspace of packed_bytes(...) each; work: split-K partials (ksplit*B*M'*N' floats)
// or null.
void launch_gemm_packed(int dt, const GemmDesc& g, void* pa, void* pb, float* work, int ksplit,
                        hipStream_t st, int np) {
  const bool cplx = (dt == C32);
  PackedGemmDesc d{};
  const int64_t M = g.M, N = cplx ? 2 * g.N : g.N, K = cplx ? 2 * g.K : g.K;
  int TM, TN;
  packed_tiles(M, N, TM, TN, np);
  OME_DRY_RETURN((ksplit > 1 ? 2 : 1) + (TM == TN || np == 3 ? 1 : 2));
  // A' rows m (stride sam), k (stride sak); B' rows n (stride sbn), k (stride sbk)
  const PackDesc pka = pack_desc(g.a, (uint16_t*)pa, M, K, g.sam, g.sak, g.sab, cplx ? 1 : 0, g.M, g.K, TM);
  const PackDesc pkb = pack_desc(g.b, (uint16_t*)pb, N, K, g.sbn, g.sbk, g.sbb, cplx ? 2 : 0, g.N, g.K, TN);
  if (np == 3) {
    const int64_t na = pka.tiles_r * pka.tiles_k * g.B, nb = pkb.tiles_r * pkb.tiles_k * g.B;
    TimedScope ts(KC_PACK, st);
    hipLaunchKernelGGL((k_pack2_bf16x6<256, 128>), dim3((uint32_t)(na + nb)), dim3(256), 0, st,
                       pka, pkb, na);
  } else if (TM == TN) {
    launch_pack2(pka, pkb, g.B, TM, st);
  } else {
    launch_pack(pka, g.B, TM, st);
    launch_pack(pkb, g.B, TN, st);
  }
  d.a = (const uint16_t*)pa;
  d.b = (const uint16_t*)pb;
  d.y = g.y;
  d.M = M;
  d.N = N;
  d.K = K;
  d.B = g.B;
  d.tiles_m = (M + TM - 1) / TM;
  d.tiles_n = (N + TN - 1) / TN;
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
  static int stages = -1;
  if (stages < 0) {
    const char* e = getenv("OME_PACKED_STAGES");
    // 1: single LDS buffer (default); 2, 3: ring of LDS buffers; 4: software-pipelined
    stages = (e && atoi(e) >= 1 && atoi(e) <= 4) ? atoi(e) : 1;
  }
  const dim3 grid((uint32_t)blocks), blk(256);
  {
  TimedScope ts(KC_GEMM_PACKED, st);
  if (np == 3) {
    hipLaunchKernelGGL((k_gemm_packed_bf16x6_pipe<256, 128, 4, 2>), grid, dim3(512), 0, st, d);
  } else if (TM == 256 && TN == 128) {
    if (stages != 2) hipLaunchKernelGGL((k_gemm_packed_bf16x3<256, 128, 1>), grid, blk, 0, st, d);
    else hipLaunchKernelGGL((k_gemm_packed_bf16x3<256, 128, 2>), grid, blk, 0, st, d);
  } else if (TN == 256 && TM == 128) {
    if (stages != 2) hipLaunchKernelGGL((k_gemm_packed_bf16x3<128, 256, 1>), grid, blk, 0, st, d);
    else hipLaunchKernelGGL((k_gemm_packed_bf16x3<128, 256, 2>), grid, blk, 0, st, d);
  } else if (TM == 256 && TN == 256) {
    hipLaunchKernelGGL((k_gemm_packed_bf16x3_pipe<256, 256, 2, 4>), grid, dim3(512), 0, st, d);
  } else if (stages == 4) {
    hipLaunchKernelGGL((k_gemm_packed_bf16x3_pipe<128, 128, 2, 2>), grid, blk, 0, st, d);
  } else {
    if (stages == 1) hipLaunchKernelGGL((k_gemm_packed_bf16x3<128, 128, 1>), grid, blk, 0, st, d);
    else if (stages == 3) hipLaunchKernelGGL((k_gemm_packed_bf16x3<128, 128, 3>), grid, blk, 0, st, d);
    else hipLaunchKernelGGL((k_gemm_packed_bf16x3<128, 128, 2>), grid, blk, 0, st, d);
  }
  }
  if (via_work) {
    TimedScope ts(KC_OTHER, st);
    const int64_t outs = cplx ? g.M * g.N * g.B : M * N * g.B;
    int64_t gr = (outs + 255) / 256;
    if (gr > 65536) gr = 65536;
    hipLaunchKernelGGL(k_packed_reduce, dim3((uint32_t)gr), dim3(256), 0, st, d);
  }
}

}  // namespace ome
