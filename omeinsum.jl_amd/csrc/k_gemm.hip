// Semiring GEMM on the vector ALUs: exact Int64 (wrapping) and tropical (max-plus)
// products stay off the matrix cores, and small or oddly-strided float products use it too.
//   y[m,n,b] = epilogue(sx, sy, sum_k a[m,k,b] * b[k,n,b], y[m,n,b])
// with arbitrary element strides on every operand.
#include <type_traits>

#include "kernels.h"

namespace ome {

static inline int grid_cap(int64_t n, int per = 256) {
  int64_t g = (n + per - 1) / per;
  return (int)(g < 1 ? 1 : (g > 1048576 ? 1048576 : g));
}

// 64x64 output tile, BK=16, 256 threads each owning a 4x4 sub-tile (rows tm+16i, cols tn+16j).
template <class O>
__global__ __launch_bounds__(256) void k_gemm_tile(GemmDesc d, int64_t tiles_m, int64_t tiles_n,
                                                   int64_t kchunk) {
  using T = typename O::T;
  constexpr int BM = 64, BN = 64, BK = 16;
  __shared__ T As[BK][BM + 1];
  __shared__ T Bs[BK][BN + 1];
  const T* __restrict__ A = (const T*)d.a;
  const T* __restrict__ B = (const T*)d.b;
  int64_t bid = blockIdx.x;
  const int64_t tm_i = bid % tiles_m; bid /= tiles_m;
  const int64_t tn_i = bid % tiles_n; bid /= tiles_n;
  const int64_t batch = bid % d.B; bid /= d.B;
  const int64_t ks = bid;  // split-K index
  const int64_t m0 = tm_i * BM, n0 = tn_i * BN;
  const int64_t kbeg = ks * kchunk;
  int64_t kend = kbeg + kchunk;
  if (kend > d.K) kend = d.K;
  const int t = threadIdx.x, tm = t & 15, tn = t >> 4;
  const bool a_mfast = (d.sam == 1) || (d.K == 1);
  const bool b_kfast = (d.sbk == 1) || (d.N == 1);
  const T* Ab = A + batch * d.sab;
  const T* Bb = B + batch * d.sbb;
  // Float32 / ComplexF32 products and sums in FP64 (AccOps), rounded once per output
  using W = AccOps<O>;
  using AO = typename W::A;
  typename AO::T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = AO::zero();
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    // all eight loads of the step are issued before any is staged: lanes outside the operand
    // load its first element and stage a zero
    T av[4], bv[4];
    uint32_t ok = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + 256 * q;
      const int mm = a_mfast ? (e & 63) : (e >> 4), kk = a_mfast ? (e >> 6) : (e & 15);
      const int64_t gm = m0 + mm, gk = k0 + kk;
      const bool oka = gm < d.M && gk < kend;
      av[q] = Ab[oka ? gm * d.sam + gk * d.sak : 0];
      const int nn = b_kfast ? (e >> 4) : (e & 63), kb = b_kfast ? (e & 15) : (e >> 6);
      const int64_t gn = n0 + nn, gkb = k0 + kb;
      const bool okb = gn < d.N && gkb < kend;
      bv[q] = Bb[okb ? gkb * d.sbk + gn * d.sbn : 0];
      ok |= (oka ? 1u : 0u) << q | (okb ? 16u : 0u) << q;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + 256 * q;
      const int mm = a_mfast ? (e & 63) : (e >> 4), kk = a_mfast ? (e >> 6) : (e & 15);
      As[kk][mm] = (ok >> q) & 1u ? av[q] : O::zero();
      const int nn = b_kfast ? (e >> 4) : (e & 63), kb = b_kfast ? (e & 15) : (e >> 6);
      Bs[kb][nn] = (ok >> (4 + q)) & 1u ? bv[q] : O::zero();
    }
    __syncthreads();
    const int kn = (int)((kend - k0) < BK ? (kend - k0) : BK);
    for (int kk = 0; kk < kn; ++kk) {
      typename AO::T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = W::widen(As[kk][tm + 16 * i]);
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = W::widen(Bs[kk][tn + 16 * j]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = AO::fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  if (d.ksplit > 1) {
    T* Wp = (T*)d.work + ((ks * d.B + batch) * d.M) * d.N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int64_t m = m0 + tm + 16 * i, n = n0 + tn + 16 * j;
        if (m < d.M && n < d.N) Wp[m + n * d.M] = W::narrow(acc[i][j]);
      }
    return;
  }
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  T* Y = (T*)d.y + batch * d.syb;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int64_t m = m0 + tm + 16 * i, n = n0 + tn + 16 * j;
      if (m < d.M && n < d.N) {
        T* yp = Y + m * d.sym + n * d.syn;
        *yp = epilogue<O>(d.mode, sx, sy, W::narrow(acc[i][j]), yp);
      }
    }
}

// Sum the ksplit partial products in split order and apply the epilogue.
template <class O>
__global__ __launch_bounds__(256) void k_gemm_splitk_reduce(GemmDesc d) {
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

template <class O>
__global__ __launch_bounds__(256) void k_gemm_strided(StridedGemmDesc d) {
  using T = typename O::T;
  __shared__ int32_t ska[SG_KCAP], skb[SG_KCAP];
  for (int k = threadIdx.x; k < d.nk; k += 256) {
    uint32_t r = (uint32_t)k;
    int32_t oa = 0, ob = 0;
    for (int q = 0; q < d.ni; ++q) {
      const uint32_t qq = fdiv_q(r, d.kdiv[q]);
      const int32_t v = (int32_t)(r - qq * d.kdiv[q].d);
      oa += v * d.ka[q];
      ob += v * d.kb[q];
      r = qq;
    }
    ska[k] = oa;
    skb[k] = ob;
  }
  __syncthreads();
  const T* __restrict__ A = (const T*)d.a;
  const T* __restrict__ B = (const T*)d.b;
  T* __restrict__ Y = (T*)d.y;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  for (int64_t o = blockIdx.x * 256ll + threadIdx.x; o < d.nout; o += (int64_t)gridDim.x * 256) {
    uint32_t r = (uint32_t)o;
    int32_t pa = 0, pb = 0, py = 0;
    for (int q = 0; q < d.no; ++q) {
      const uint32_t qq = fdiv_q(r, d.odiv[q]);
      const int32_t v = (int32_t)(r - qq * d.odiv[q].d);
      pa += v * d.oa[q];
      pb += v * d.ob[q];
      py += v * d.oy[q];
      r = qq;
    }
    using W = AccOps<O>;
    typename W::A::T acc = W::A::zero();
    for (int k = 0; k < d.nk; ++k)
      acc = W::A::fma(W::widen(A[pa + ska[k]]), W::widen(B[pb + skb[k]]), acc);
    T* yp = Y + py;
    *yp = epilogue<O>(d.mode, sx, sy, W::narrow(acc), yp);
  }
}

// One subtree of fused steps (steps [s0, s1) in order) by the whole workgroup; leaf(k) gives
// the pointer of leaf k.
template <class O, class Leaf>
__device__ __forceinline__ void fused_subtree(const FusedStep* __restrict__ steps, int s0, int s1,
                                              Leaf leaf, char* arena, int32_t* ska, int32_t* skb) {
  using T = typename O::T;
  for (int s = s0; s < s1; ++s) {
    const FusedStep& d = steps[s];
    const T* __restrict__ A = d.a_leaf >= 0 ? (const T*)leaf(d.a_leaf) : (const T*)(arena + d.a_off);
    const T* __restrict__ B = d.b_leaf >= 0 ? (const T*)leaf(d.b_leaf) : (const T*)(arena + d.b_off);
    T* __restrict__ Y = (T*)(arena + d.y_off);
    const int nk = d.nk;
    for (int k = threadIdx.x; k < nk; k += SF_THREADS) {
      uint32_t r = (uint32_t)k;
      int32_t oa = 0, ob = 0;
      for (int q = 0; q < d.ni; ++q) {
        const uint32_t qq = fdiv_q(r, d.kdiv[q]);
        const int32_t v = (int32_t)(r - qq * d.kdiv[q].d);
        oa += v * d.ka[q];
        ob += v * d.kb[q];
        r = qq;
      }
      ska[k] = oa;
      skb[k] = ob;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < d.nout; o += SF_THREADS) {
      uint32_t r = (uint32_t)o;
      int32_t pa = 0, pb = 0, py = 0;
      for (int q = 0; q < d.no; ++q) {
        const uint32_t qq = fdiv_q(r, d.odiv[q]);
        const int32_t v = (int32_t)(r - qq * d.odiv[q].d);
        pa += v * d.oa[q];
        pb += v * d.ob[q];
        py += v * d.oy[q];
        r = qq;
      }
      using W = AccOps<O>;
      typename W::A::T acc = W::A::zero();
      for (int k = 0; k < nk; ++k)
        acc = W::A::fma(W::widen(A[pa + ska[k]]), W::widen(B[pb + skb[k]]), acc);
      Y[py] = W::narrow(acc);
    }
    // this step's outputs are the next steps' operands; the offset table is rewritten next
    __syncthreads();
  }
}

template <class O>
__global__ __launch_bounds__(SF_THREADS) void k_fused_steps(const FusedStep* __restrict__ steps,
                                                            const int32_t* __restrict__ first,
                                                            FusedLeaves lv, char* arena) {
  __shared__ int32_t ska[SF_MAXNK], skb[SF_MAXNK];
  fused_subtree<O>(steps, first[blockIdx.x], first[blockIdx.x + 1],
                   [&](int k) { return lv.p[k]; }, arena, ska, skb);
}

// Workgroup b runs subtree b - sub_begin of the job whose range holds b.
template <class O>
__global__ __launch_bounds__(SF_THREADS) void k_fused_jobs(const FusedJob* __restrict__ jobs,
                                                           int njobs) {
  __shared__ int32_t ska[SF_MAXNK], skb[SF_MAXNK];
  const int b = blockIdx.x;
  int j = 0;
  while (j + 1 < njobs && jobs[j + 1].sub_begin <= b) ++j;
  const FusedJob& jb = jobs[j];
  const int q = b - jb.sub_begin;
  const void* const* __restrict__ leaves = jb.leaves;
  fused_subtree<O>(jb.steps, jb.first[q], jb.first[q + 1], [&](int k) { return leaves[k]; },
                   jb.arena, ska, skb);
}

// One thread per output element, K walked directly from global memory: for tiny M/N or K
// (tensor-network steps over dimension-2 labels) where a 64x64 tile would idle.
template <class O>
__global__ __launch_bounds__(256) void k_gemm_direct(GemmDesc d, int order_m_fast) {
  using T = typename O::T;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  const T* __restrict__ A = (const T*)d.a;
  const T* __restrict__ B = (const T*)d.b;
  const int64_t per = d.M * d.N, total = per * d.B;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t batch = e / per, r = e - batch * per;
    int64_t m, n;
    if (order_m_fast) { n = r / d.M; m = r - n * d.M; } else { m = r / d.N; n = r - m * d.N; }
    const T* ap = A + batch * d.sab + m * d.sam;
    const T* bp = B + batch * d.sbb + n * d.sbn;
    using W = AccOps<O>;
    typename W::A::T acc = W::A::zero();
    for (int64_t k = 0; k < d.K; ++k)
      acc = W::A::fma(W::widen(ap[k * d.sak]), W::widen(bp[k * d.sbk]), acc);
    T* yp = (T*)d.y + batch * d.syb + m * d.sym + n * d.syn;
    *yp = epilogue<O>(d.mode, sx, sy, W::narrow(acc), yp);
  }
}

// Few outputs and a long K (the last steps of a network contracted to a scalar: inner
// products of two large tensors).  Block (output e, part p) of 256 threads sums k over its
// chunk, strided by the block size, then reduces in LDS in a fixed tree order; with parts > 1
// the partials go to work[p][e] and k_dot_reduce finishes in part order.  Float32 and
// ComplexF32 products and sums run in FP64 (AccOps): the kernel is bound by HBM, and these
// sums (the final scalar of a network) are the deepest of a contraction.
template <class O>
__global__ __launch_bounds__(256) void k_gemm_dot(GemmDesc d, int64_t kchunk) {
  using T = typename O::T;
  using W = AccOps<O>;
  using A = typename W::A;
  using AT = typename A::T;
  __shared__ AT red[256];
  const int t = threadIdx.x;
  const int64_t e = blockIdx.x, part = blockIdx.y;
  const int64_t per = d.M * d.N;
  const int64_t batch = e / per, r = e - batch * per;
  const int64_t n = r / d.M, m = r - n * d.M;
  const T* __restrict__ ap = (const T*)d.a + batch * d.sab + m * d.sam;
  const T* __restrict__ bp = (const T*)d.b + batch * d.sbb + n * d.sbn;
  const int64_t k0 = part * kchunk;
  const int64_t k1 = (k0 + kchunk < d.K) ? k0 + kchunk : d.K;
  AT acc = A::zero();
  for (int64_t k = k0 + t; k < k1; k += 256)
    acc = A::fma(W::widen(ap[k * d.sak]), W::widen(bp[k * d.sbk]), acc);
  red[t] = acc;
  __syncthreads();
#pragma unroll
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] = A::add(red[t], red[t + s]);
    __syncthreads();
  }
  if (t != 0) return;
  if (d.ksplit > 1) {
    ((AT*)d.work)[part * per * d.B + e] = red[0];
    return;
  }
  T* yp = (T*)d.y + batch * d.syb + m * d.sym + n * d.syn;
  *yp = epilogue<O>(d.mode, O::from(d.sx), O::from(d.sy), W::narrow(red[0]), yp);
}

// Sum the dot kernel's partials (accumulation type) in part order, round once, epilogue.
template <class O>
__global__ __launch_bounds__(256) void k_dot_reduce(GemmDesc d) {
  using T = typename O::T;
  using W = AccOps<O>;
  using A = typename W::A;
  using AT = typename A::T;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  const int64_t per = d.M * d.N, total = per * d.B;
  const AT* Wk = (const AT*)d.work;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t batch = e / per, r = e - batch * per;
    const int64_t n = r / d.M, m = r - n * d.M;
    AT acc = A::zero();
    for (int s = 0; s < d.ksplit; ++s) acc = A::add(acc, Wk[(int64_t)s * total + e]);
    T* yp = (T*)d.y + batch * d.syb + m * d.sym + n * d.syn;
    *yp = epilogue<O>(d.mode, sx, sy, W::narrow(acc), yp);
  }
}

// The same for many parts per output (a few outputs over a very long K): one workgroup per
// output; thread t sums parts t, t + 256, ... in order, then a fixed LDS tree, so the 256
// partial loads of a round are in flight together instead of one dependent chain.
template <class O>
__global__ __launch_bounds__(256) void k_dot_reduce_wg(GemmDesc d) {
  using T = typename O::T;
  using W = AccOps<O>;
  using A = typename W::A;
  using AT = typename A::T;
  __shared__ AT red[256];
  const int t = threadIdx.x;
  const int64_t per = d.M * d.N, total = per * d.B;
  const int64_t e = blockIdx.x;
  const AT* Wk = (const AT*)d.work;
  AT acc = A::zero();
  for (int s = t; s < d.ksplit; s += 256) acc = A::add(acc, Wk[(int64_t)s * total + e]);
  red[t] = acc;
  __syncthreads();
#pragma unroll
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] = A::add(red[t], red[t + s]);
    __syncthreads();
  }
  if (t != 0) return;
  const int64_t batch = e / per, r = e - batch * per;
  const int64_t n = r / d.M, m = r - n * d.M;
  T* yp = (T*)d.y + batch * d.syb + m * d.sym + n * d.syn;
  *yp = epilogue<O>(d.mode, O::from(d.sx), O::from(d.sy), W::narrow(red[0]), yp);
}

// Shuffle of any trivially copyable value across the lanes of a wave, word by word.
template <class V>
__device__ __forceinline__ V shfl_xor_any(V v, int mask) {
  constexpr int NW = (int)((sizeof(V) + 3) / 4);
  uint32_t w[NW] = {};
  __builtin_memcpy(w, &v, sizeof(V));
#pragma unroll
  for (int i = 0; i < NW; ++i) w[i] = (uint32_t)__shfl_xor((int)w[i], mask, 64);
  V r;
  __builtin_memcpy(&r, w, sizeof(V));
  return r;
}

// Few outputs, long K, in one launch: workgroup b owns a TM x TN output tile over all of K.
// Its PN_THREADS threads stride K, each summing the tile's TM*TN products in the
// accumulation arithmetic; the loads of UNR strides (as many as about 96 registers hold) are
// issued before their products, so a thread waits on memory once per UNR * PN_THREADS values
// of k (once in all for Float32 4 x 4 tiles and K <= 4096).  The sums are then reduced in a
// fixed tree (lanes of a wave, then the waves), so the result does not depend on timing.
// Rows past M (columns past N) read the last row (column) and are not written.
constexpr int PN_THREADS = 512;
template <class O, int TM, int TN>
__global__ __launch_bounds__(PN_THREADS) void k_gemm_panel(GemmDesc d, int64_t tiles_m,
                                                           int64_t tiles_n) {
  using T = typename O::T;
  using W = AccOps<O>;
  using A = typename W::A;
  using AT = typename A::T;
  constexpr int NWAVE = PN_THREADS / 64;
  constexpr int STEP_REGS = (TM + TN) * (int)((sizeof(T) + 3) / 4);
  constexpr int UNR0 = 96 / STEP_REGS;
  constexpr int PN_UNROLL = UNR0 >= 8 ? 8 : UNR0 >= 4 ? 4 : UNR0 >= 2 ? 2 : 1;
  __shared__ AT part[NWAVE][TM * TN];
  int64_t bid = blockIdx.x;
  const int64_t tm = bid % tiles_m;
  bid /= tiles_m;
  const int64_t tn = bid % tiles_n;
  const int64_t batch = bid / tiles_n;
  const int64_t m0 = tm * TM, n0 = tn * TN;
  const T* __restrict__ ap = (const T*)d.a + batch * d.sab;
  const T* __restrict__ bp = (const T*)d.b + batch * d.sbb;
  int64_t ao[TM], bo[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) ao[i] = (m0 + i < d.M ? m0 + i : d.M - 1) * d.sam;
#pragma unroll
  for (int j = 0; j < TN; ++j) bo[j] = (n0 + j < d.N ? n0 + j : d.N - 1) * d.sbn;
  AT acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = A::zero();
  const int t = threadIdx.x;
  // A tile rows contiguous (sam == 1) and 16-byte aligned at every k: one vector load per k
  // instead of TM loads of the same cache line (likewise B's columns when sbn == 1)
  const bool va = TM > 1 && d.sam == 1 && m0 + TM <= d.M && (sizeof(T) * TM) % 16 == 0 &&
                  (((uintptr_t)(ap + m0)) % 16) == 0 && (d.sak * sizeof(T)) % 16 == 0;
  const bool vb = TN > 1 && d.sbn == 1 && n0 + TN <= d.N && (sizeof(T) * TN) % 16 == 0 &&
                  (((uintptr_t)(bp + n0)) % 16) == 0 && (d.sbk * sizeof(T)) % 16 == 0;
  auto run = [&](auto VA, auto VB) __attribute__((always_inline)) {
    for (int64_t kb = t; kb < d.K; kb += PN_UNROLL * PN_THREADS) {
      T av[PN_UNROLL][TM], bv[PN_UNROLL][TN];
#pragma unroll
      for (int u = 0; u < PN_UNROLL; ++u) {
        const int64_t k = kb + u * PN_THREADS;
        if (k < d.K) {
          const T* ak = ap + k * d.sak;
          const T* bk = bp + k * d.sbk;
          if constexpr (decltype(VA)::value)
            __builtin_memcpy(av[u], __builtin_assume_aligned(ak + m0, 16), sizeof(av[u]));
          else
#pragma unroll
            for (int i = 0; i < TM; ++i) av[u][i] = ak[ao[i]];
          if constexpr (decltype(VB)::value)
            __builtin_memcpy(bv[u], __builtin_assume_aligned(bk + n0, 16), sizeof(bv[u]));
          else
#pragma unroll
            for (int j = 0; j < TN; ++j) bv[u][j] = bk[bo[j]];
        }
      }
#pragma unroll
      for (int u = 0; u < PN_UNROLL; ++u) {
        if (kb + u * PN_THREADS < d.K) {
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[i][j] = A::fma(W::widen(av[u][i]), W::widen(bv[u][j]), acc[i][j]);
        }
      }
    }
  };
  if (va && vb) run(std::true_type{}, std::true_type{});
  else if (va) run(std::true_type{}, std::false_type{});
  else if (vb) run(std::false_type{}, std::true_type{});
  else run(std::false_type{}, std::false_type{});
#pragma unroll
  for (int s = 32; s > 0; s >>= 1)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = A::add(acc[i][j], shfl_xor_any(acc[i][j], s));
  const int lane = t & 63, w = t >> 6;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) part[w][i * TN + j] = acc[i][j];
  }
  __syncthreads();
  if (t >= TM * TN) return;
  const int i = t / TN, j = t - (t / TN) * TN;
  if (m0 + i >= d.M || n0 + j >= d.N) return;
  AT v = part[0][t];
#pragma unroll
  for (int q = 1; q < NWAVE; ++q) v = A::add(v, part[q][t]);
  T* yp = (T*)d.y + batch * d.syb + (m0 + i) * d.sym + (n0 + j) * d.syn;
  *yp = epilogue<O>(d.mode, O::from(d.sx), O::from(d.sy), W::narrow(v), yp);
}

bool panel_fits(int dt, const GemmDesc& d) {
  // complex products keep their matrix-core routes (the realified kernels)
  if (dt == C32 || dt == C64) return false;
  const int64_t outputs = d.M * d.N * d.B;
  return outputs > 1024 && outputs <= 16384 && d.K >= 512 && outputs * d.K <= (1ll << 26);
}

void launch_gemm_panel(int dt, GemmDesc d, hipStream_t st) {
  OME_DRY_RETURN(1);
  // tile shape: 4 x 4, or a line along the long side when the other has fewer than 4
  const int shape = d.M < 4 ? 1 : (d.N < 4 ? 2 : 0);
  d.ksplit = 1;
  OME_DISPATCH(dt, O, {
    d.mode = epi_mode<O>(d.sx, d.sy);
    TimedScope ts(KC_GEMM_VALU, st);
    if (shape == 0) {
      const int64_t tmn = (d.M + 3) / 4, tnn = (d.N + 3) / 4;
      hipLaunchKernelGGL((k_gemm_panel<O, 4, 4>), dim3((uint32_t)(tmn * tnn * d.B)),
                         dim3(PN_THREADS), 0, st, d, tmn, tnn);
    } else if (shape == 1) {
      const int64_t tnn = (d.N + 15) / 16;
      hipLaunchKernelGGL((k_gemm_panel<O, 1, 16>), dim3((uint32_t)(d.M * tnn * d.B)),
                         dim3(PN_THREADS), 0, st, d, d.M, tnn);
    } else {
      const int64_t tmn = (d.M + 15) / 16;
      hipLaunchKernelGGL((k_gemm_panel<O, 16, 1>), dim3((uint32_t)(tmn * d.N * d.B)),
                         dim3(PN_THREADS), 0, st, d, tmn, d.N);
    }
  });
}

int64_t dot_workspace(int dt, const GemmDesc& d, int parts) {
  if (parts <= 1) return 0;
  const int64_t es = (dt == F32 || dt == C32) ? 2 * (int64_t)elem_size(dt) : (int64_t)elem_size(dt);
  return (int64_t)parts * d.M * d.N * d.B * es;
}

int dot_plan_parts(const GemmDesc& d) {
  const int64_t outputs = d.M * d.N * d.B;
  int64_t parts = 1024 / (outputs > 0 ? outputs : 1);
  const int64_t kmax = (d.K + 1023) / 1024;
  if (parts > kmax) parts = kmax;
  if (parts < 1) parts = 1;
  if (parts > 4096) parts = 4096;
  return (int)parts;
}

void launch_gemm_dot(int dt, GemmDesc d, int parts, hipStream_t st) {
  const int64_t outputs = d.M * d.N * d.B;
  if (outputs <= 0) return;
  OME_DRY_RETURN(parts > 1 ? 2 : 1);
  if (parts < 1) parts = 1;
  const int64_t kchunk = (d.K + parts - 1) / parts;
  d.ksplit = parts;
  OME_DISPATCH(dt, O, {
    d.mode = epi_mode<O>(d.sx, d.sy);
    {
      TimedScope ts(KC_GEMM_VALU, st);
      hipLaunchKernelGGL(k_gemm_dot<O>, dim3((uint32_t)outputs, (uint32_t)parts), dim3(256), 0,
                         st, d, kchunk);
    }
    if (parts > 1) {
      TimedScope ts(KC_OTHER, st);
      if (parts >= 32)
        hipLaunchKernelGGL(k_dot_reduce_wg<O>, dim3((uint32_t)outputs), dim3(256), 0, st, d);
      else
        hipLaunchKernelGGL(k_dot_reduce<O>, dim3(grid_cap(outputs)), dim3(256), 0, st, d);
    }
  });
}

int64_t gemm_valu_workspace(int dt, const GemmDesc& d, int ksplit) {
  return ksplit > 1 ? (int64_t)ksplit * d.B * d.M * d.N * (int64_t)elem_size(dt) : 0;
}

void launch_gemm_valu(int dt, GemmDesc d, int kernel, int ksplit, hipStream_t st) {
  const int64_t outputs = d.M * d.N * d.B;
  if (outputs <= 0) return;
  OME_DRY_RETURN(ksplit > 1 ? 2 : 1);
  OME_DISPATCH(dt, O, {
    d.mode = epi_mode<O>(d.sx, d.sy);
    if (kernel == GEMM_VALU_DIRECT) {
      d.ksplit = 1;
      int order_m_fast = (d.sym == 1 || d.N == 1) ? 1 : 0;
      TimedScope ts(KC_GEMM_VALU, st);
      hipLaunchKernelGGL(k_gemm_direct<O>, dim3(grid_cap(outputs)), dim3(256), 0, st, d,
                         order_m_fast);
    } else {
      if (ksplit < 1) ksplit = 1;
      int64_t kchunk = (d.K + ksplit - 1) / ksplit;
      kchunk = ((kchunk + 15) / 16) * 16;
      if (kchunk < 16) kchunk = 16;
      ksplit = (int)((d.K + kchunk - 1) / kchunk);
      if (ksplit < 1) ksplit = 1;
      d.ksplit = ksplit;
      const int64_t tiles_m = (d.M + 63) / 64, tiles_n = (d.N + 63) / 64;
      const int64_t blocks = tiles_m * tiles_n * d.B * ksplit;
      {
        TimedScope ts(KC_GEMM_VALU, st);
        hipLaunchKernelGGL(k_gemm_tile<O>, dim3((uint32_t)blocks), dim3(256), 0, st, d, tiles_m,
                           tiles_n, kchunk);
      }
      TimedScope ts(KC_OTHER, st);
      if (ksplit > 1)
        hipLaunchKernelGGL(k_gemm_splitk_reduce<O>, dim3(grid_cap(outputs)), dim3(256), 0, st, d);
    }
  });
}

void launch_gemm_strided(int dt, const StridedGemmDesc& d, hipStream_t st) {
  if (d.nout <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_GEMM_VALU, st);
  int64_t g = (d.nout + 255) / 256;
  if (g > 4096) g = 4096;
  OME_DISPATCH(dt, O, {
    StridedGemmDesc dd = d;
    dd.mode = epi_mode<O>(d.sx, d.sy);
    hipLaunchKernelGGL(k_gemm_strided<O>, dim3((uint32_t)g), dim3(256), 0, st, dd);
  });
}

void launch_fused_jobs(int dt, const FusedJob* jobs, int njobs, int nsub_total, hipStream_t st) {
  if (njobs <= 0 || nsub_total <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_GEMM_VALU, st);
  OME_DISPATCH(dt, O, {
    hipLaunchKernelGGL(k_fused_jobs<O>, dim3((uint32_t)nsub_total), dim3(SF_THREADS), 0, st, jobs,
                       njobs);
  });
}

void launch_fused_steps(int dt, const FusedStep* steps, const int32_t* first, int nsub,
                        const FusedLeaves& leaves, char* arena, hipStream_t st) {
  if (nsub <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_GEMM_VALU, st);
  OME_DISPATCH(dt, O, {
    hipLaunchKernelGGL(k_fused_steps<O>, dim3((uint32_t)nsub), dim3(SF_THREADS), 0, st, steps, first,
                       leaves, arena);
  });
}

}  // namespace ome
