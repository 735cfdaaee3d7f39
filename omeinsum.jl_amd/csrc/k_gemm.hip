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

// Loads and stores through global-memory pointers (address space 1) of any element type, as
// raw words of its size.
template <class T>
struct GlobalRef {
  using V = std::conditional_t<
      sizeof(T) == 16, unsigned int __attribute__((ext_vector_type(4))),
      std::conditional_t<sizeof(T) == 8, unsigned int __attribute__((ext_vector_type(2))),
                         std::conditional_t<sizeof(T) == 4, unsigned int, unsigned char>>>;
  static_assert(sizeof(V) == sizeof(T), "element of 1, 4, 8 or 16 bytes");
  const __attribute__((address_space(1))) V* p;
  __device__ __forceinline__ T operator[](int i) const { return __builtin_bit_cast(T, p[i]); }
};
template <class T>
__device__ __forceinline__ void store_global(void* base, int i, T v) {
  using V = typename GlobalRef<T>::V;
  ((__attribute__((address_space(1))) V*)base)[i] = __builtin_bit_cast(V, v);
}

// What a wave needs of one step of a chunk: its operand and output locations (global base
// pointers, or element offsets in the LDS arena as flags say) and its item range.
struct FusedHead {
  const void* a;
  const void* b;
  void* y;
  int32_t la, lb, ly;
  int32_t ibase, nout, nk, flags, pad_;
};
// A chunk's step heads and contracted-offset tables, staged in LDS.
struct FusedStage {
  FusedHead h[SF_LMAX];
  int32_t total, lg;
  uint16_t ka[SF_LMAX][SF_MAXNK], kb[SF_LMAX][SF_MAXNK];
};

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

constexpr int SF_MAXCH = 64;  // chunks of a subtree whose chunk tables are held in LDS

// The LDS a fused subtree works in.
struct FusedShared {
  FusedStage stage[2];
  const void* leaves[SF_MAXLEAF];
  int32_t tab[SF_MAXCH + 1], itab[SF_MAXCH + 1];
  __align__(16) char arena[SF_LDS];
};

// One subtree of fused steps by the whole workgroup: chunks [c0, c1), chunk c being steps
// [tab[c], tab[c + 1]) with its item table at words [itab[c], itab[c + 1]) of words; leaf(k)
// gives the pointer of leaf k.  The prologue copies the leaf pointers and the subtree's chunk
// tables to LDS.  While a chunk runs, the next chunk's step heads (thread j < ns resolves step
// j's pointers) and contracted-offset tables (one 32-bit word per thread) are loaded, and written
// to the other stage buffer at the end; each thread works on two items at a time and loads the
// item-table words of its next two before working on the current ones.  A wave's items belong
// to one step, found from the staged item bases.
template <class O, class Leaf>
__device__ __forceinline__ void fused_subtree(const FusedStep* __restrict__ steps,
                                              const int32_t* __restrict__ gtab,
                                              const int32_t* __restrict__ gitab,
                                              const uint32_t* __restrict__ words, int c0, int c1,
                                              Leaf leaf, char* arena, FusedShared& sh) {
  using T = typename O::T;
  using W = AccOps<O>;
  constexpr int KW = SF_MAXNK / 2;  // 32-bit words of one offset table
  static_assert(SF_LMAX * 2 * KW <= SF_THREADS && SF_MAXLEAF <= SF_THREADS, "one word per thread");
  if (c0 >= c1) return;
  T* L = reinterpret_cast<T*>(sh.arena);
  const int t = threadIdx.x;
  const bool held = c1 - c0 <= SF_MAXCH;
  if (t < SF_MAXLEAF) sh.leaves[t] = leaf(t);
  if (held && t <= c1 - c0) {
    sh.tab[t] = gtab[c0 + t];
    sh.itab[t] = gitab[c0 + t];
  }
  __syncthreads();
  auto tab = [&](int c) { return held ? sh.tab[c - c0] : gtab[c]; };
  auto itab = [&](int c) { return held ? sh.itab[c - c0] : gitab[c]; };
  // thread t copies word t % KW of table (t / KW) % 2 of step t / (2 KW)
  const int cj = t / (2 * KW), ct = (t / KW) & 1, cw = t % KW;
  uint32_t w = 0;
  FusedHead hd{};
  int32_t ctotal = 0, clg = 0;
  auto fetch = [&](int c) {
    const int s0 = tab(c), ns = tab(c + 1) - s0;
    if (cj < ns) {
      const FusedStep& d = steps[s0 + cj];
      w = reinterpret_cast<const uint32_t*>(ct ? d.kb : d.ka)[cw];
    }
    if (t < ns) {
      const FusedStep& d = steps[s0 + t];
      hd.flags = d.flags;
      hd.ibase = d.ibase;
      hd.nout = d.nout;
      hd.nk = d.nk;
      hd.a = d.a_leaf >= 0 ? sh.leaves[d.a_leaf] : (const void*)(arena + d.a_off);
      hd.b = d.b_leaf >= 0 ? sh.leaves[d.b_leaf] : (const void*)(arena + d.b_off);
      hd.y = arena + d.y_off;
      hd.la = (int32_t)(d.a_off / (int64_t)sizeof(T));
      hd.lb = (int32_t)(d.b_off / (int64_t)sizeof(T));
      hd.ly = (int32_t)(d.y_off / (int64_t)sizeof(T));
      if (t == 0) {
        ctotal = d.total;
        clg = d.lg;
      }
    }
  };
  auto put = [&](int c, FusedStage& sg) {
    const int ns = tab(c + 1) - tab(c);
    if (cj < ns) reinterpret_cast<uint32_t*>(ct ? sg.kb[cj] : sg.ka[cj])[cw] = w;
    if (t < ns) sg.h[t] = hd;
    if (t == 0) {
      sg.total = ctotal;
      sg.lg = clg;
    }
  };
  auto item = [&](int c, int i) -> uint32_t {
    const int b = itab(c);
    return i < itab(c + 1) - b ? words[b + i] : 0u;
  };
  fetch(c0);
  put(c0, sh.stage[0]);
  uint32_t cur0 = item(c0, t), cur1 = item(c0, t + SF_THREADS);
  __syncthreads();
  for (int c = c0; c < c1; ++c) {
    const FusedStage& sg = sh.stage[(c - c0) & 1];
    const bool more = c + 1 < c1;
    if (more) fetch(c + 1);  // in flight while this chunk runs
    const int ns = tab(c + 1) - tab(c);
    const int lg = __builtin_amdgcn_readfirstlane(sg.lg), G = 1 << lg;
    const int items = __builtin_amdgcn_readfirstlane(sg.total);
    // item wi of the chunk, whose operand offsets are in word
    auto work = [&](int wi, uint32_t word) {
      const int wb = __builtin_amdgcn_readfirstlane(wi & ~63);
      if (wb >= items) return;  // the whole wave
      int j = 0;
#pragma unroll
      for (int u = 1; u < SF_LMAX; ++u) j += (u < ns && wb >= sg.h[u].ibase) ? 1 : 0;
      j = __builtin_amdgcn_readfirstlane(j);
      const FusedHead& h = sg.h[j];
      const int flags = __builtin_amdgcn_readfirstlane(h.flags);
      const int nk = __builtin_amdgcn_readfirstlane(h.nk);
      const int rel = wi - h.ibase;
      const int o = rel >> lg, g = rel & (G - 1);
      const bool live = o < h.nout;
      const int ia = (int)(word & 0xffffu), ib = (int)(word >> 16);
      typename W::A::T acc = W::A::zero();
      if (live) {
        // operands in LDS or in global memory (wave-uniform), one code path per pair
        const T* La = L + h.la + ia;
        const T* Lb = L + h.lb + ib;
        // h.a / h.b / h.y are global memory (leaves or the plan's arena)
        using GV = typename GlobalRef<T>::V;
        const GlobalRef<T> Ag{(const __attribute__((address_space(1))) GV*)h.a + ia};
        const GlobalRef<T> Bg{(const __attribute__((address_space(1))) GV*)h.b + ib};
        const uint16_t* ka = sg.ka[j];
        const uint16_t* kb = sg.kb[j];
        auto kloop = [&](auto A, auto B) {
          int k = g;
          // four independent operand pairs in flight
          for (; k + 3 * G < nk; k += 4 * G) {
            const T a0 = A[ka[k]], a1 = A[ka[k + G]], a2 = A[ka[k + 2 * G]], a3 = A[ka[k + 3 * G]];
            const T b0 = B[kb[k]], b1 = B[kb[k + G]], b2 = B[kb[k + 2 * G]], b3 = B[kb[k + 3 * G]];
            acc = W::A::fma(W::widen(a0), W::widen(b0), acc);
            acc = W::A::fma(W::widen(a1), W::widen(b1), acc);
            acc = W::A::fma(W::widen(a2), W::widen(b2), acc);
            acc = W::A::fma(W::widen(a3), W::widen(b3), acc);
          }
          for (; k < nk; k += G) acc = W::A::fma(W::widen(A[ka[k]]), W::widen(B[kb[k]]), acc);
        };
        if (flags & FS_COPY) {
          if (g == 0) acc = W::widen(Ag[0]);
        } else if (flags & FS_A_LDS) {
          if (flags & FS_B_LDS)
            kloop(La, Lb);
          else
            kloop(La, Bg);
        } else {
          if (flags & FS_B_LDS)
            kloop(Ag, Lb);
          else
            kloop(Ag, Bg);
        }
      }
      for (int m = G >> 1; m >= 1; m >>= 1) acc = W::A::add(acc, shfl_xor_any(acc, m));
      if (live && g == 0) {
        if (flags & FS_Y_LDS)
          L[h.ly + o] = W::narrow(acc);
        else
          store_global<T>(h.y, o, W::narrow(acc));
      }
    };
    for (int base = 0; base < items; base += 2 * SF_THREADS) {
      uint32_t n0 = 0, n1 = 0;
      if (base + 2 * SF_THREADS < items) {
        n0 = item(c, base + 2 * SF_THREADS + t);
        n1 = item(c, base + 3 * SF_THREADS + t);
      } else if (more) {
        n0 = item(c + 1, t);
        n1 = item(c + 1, t + SF_THREADS);
      }
      work(base + t, cur0);
      work(base + SF_THREADS + t, cur1);
      cur0 = n0;
      cur1 = n1;
    }
    if (!more) break;  // the kernel's end makes the last chunk's outputs visible
    put(c + 1, sh.stage[(c + 1 - c0) & 1]);  // the buffer chunk c - 1 used
    // this chunk's outputs are later chunks' operands
    __syncthreads();
  }
}

template <class O>
__global__ __launch_bounds__(SF_THREADS) void k_fused_steps(const FusedStep* __restrict__ steps,
                                                            const int32_t* __restrict__ first,
                                                            int nsub, FusedLeaves lv, char* arena) {
  __shared__ FusedShared sh;
  const int32_t* tab = first + nsub + 1;
  const int32_t* itab = tab + first[nsub] + 1;
  fused_subtree<O>(steps, tab, itab, (const uint32_t*)first, first[blockIdx.x],
                   first[blockIdx.x + 1], [&](int k) { return lv.p[k]; }, arena, sh);
}

// Workgroup b runs subtree b - sub_begin of the job whose range holds b.
template <class O>
__global__ __launch_bounds__(SF_THREADS) void k_fused_jobs(const FusedJob* __restrict__ jobs,
                                                           int njobs) {
  __shared__ FusedShared sh;
  const int b = blockIdx.x;
  int j = 0;
  while (j + 1 < njobs && jobs[j + 1].sub_begin <= b) ++j;
  const FusedJob& jb = jobs[j];
  const int q = b - jb.sub_begin;
  const void* const* __restrict__ leaves = jb.leaves;
  const int32_t* tab = jb.first + jb.nsub + 1;
  const int32_t* itab = tab + jb.first[jb.nsub] + 1;
  fused_subtree<O>(jb.steps, tab, itab, (const uint32_t*)jb.first, jb.first[q], jb.first[q + 1],
                   [&](int k) { return leaves[k]; }, jb.arena, sh);
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
                       nsub, leaves, arena);
  });
}

}  // namespace ome
