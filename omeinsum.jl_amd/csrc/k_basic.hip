// Bandwidth kernels: elementwise scaling/fill/conj/accumulate, the strided permute
// (tensorpermute!) and the generic index walk (loop_einsum!) that also carries the unary
// rules Diag / Sum / Repeat / Duplicate / Tr in one pass.
#include <cstdlib>

#include "kernels.h"

namespace ome {

LaunchMode& launch_mode() {
  static thread_local LaunchMode m;
  return m;
}

template <class O>
__device__ __forceinline__ typename O::T conj_if(typename O::T v, int c) {
  return v;
}
template <>
__device__ __forceinline__ c32 conj_if<OpsC32>(c32 v, int c) {
  if (c) v.y = -v.y;
  return v;
}
template <>
__device__ __forceinline__ c64 conj_if<OpsC64>(c64 v, int c) {
  if (c) v.y = -v.y;
  return v;
}

// ---- elementwise -----------------------------------------------------------------------

// y[i] <- epilogue(sx, sy, conj?(x[i]), y[i]);  x may be null (then v = zero).
template <class O>
__global__ __launch_bounds__(256) void k_axpby(const typename O::T* __restrict__ x,
                                               typename O::T* __restrict__ y, int64_t n,
                                               Scal sx_, Scal sy_, int mode, int conj) {
  using T = typename O::T;
  const T sx = O::from(sx_), sy = O::from(sy_);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T v = x ? conj_if<O>(x[i], conj) : O::zero();
    y[i] = epilogue<O>(mode, sx, sy, v, y + i);
  }
}

template <class O>
__global__ __launch_bounds__(256) void k_fill(typename O::T* __restrict__ y, int64_t n, Scal v_) {
  const typename O::T v = O::from(v_);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}

static inline int grid_for(int64_t n, int per_block = 256, int cap = 65536) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

void launch_axpby(int dt, const void* x, void* y, int64_t n, const Scal& sx, const Scal& sy,
                  int conj, hipStream_t st) {
  if (n <= 0) return;
  OME_DRY_RETURN(1);
  OME_DISPATCH(dt, O, {
    int mode = epi_mode<O>(sx, sy);
    hipLaunchKernelGGL(k_axpby<O>, dim3(grid_for(n)), dim3(256), 0, st,
                       (const typename O::T*)x, (typename O::T*)y, n, sx, sy, mode, conj);
  });
}

// One lane writes the word; the fence makes it visible to the host before the kernel ends.
__global__ void k_status_set(int* status, int code) {
  if (threadIdx.x == 0) {
    *status = code;
    __threadfence_system();
  }
}

void launch_status_set(int* status, int code, hipStream_t st) {
  hipLaunchKernelGGL(k_status_set, dim3(1), dim3(64), 0, st, status, code);
}

void launch_fill(int dt, void* y, int64_t n, const Scal& v, hipStream_t st) {
  if (n <= 0) return;
  OME_DRY_RETURN(1);
  OME_DISPATCH(dt, O, {
    hipLaunchKernelGGL(k_fill<O>, dim3(grid_for(n)), dim3(256), 0, st, (typename O::T*)y, n, v);
  });
}

// acc[i] += (double) x[i]   (complex: both parts) — FP64 cross-slice accumulation.
template <class O>
__global__ __launch_bounds__(256) void k_acc_f64(const typename O::T* __restrict__ x,
                                                 double* __restrict__ acc, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if constexpr (O::DT == C32 || O::DT == C64) {
      acc[2 * i] += (double)x[i].x;
      acc[2 * i + 1] += (double)x[i].y;
    } else {
      acc[i] += (double)x[i];
    }
  }
}
// y[i] <- epilogue(sx, sy, round(acc[i]), y[i])
template <class O>
__global__ __launch_bounds__(256) void k_from_f64(const double* __restrict__ acc,
                                                  typename O::T* __restrict__ y, int64_t n,
                                                  Scal sx_, Scal sy_, int mode) {
  using T = typename O::T;
  const T sx = O::from(sx_), sy = O::from(sy_);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    T v;
    if constexpr (O::DT == C32) {
      v = {(float)acc[2 * i], (float)acc[2 * i + 1]};
    } else if constexpr (O::DT == C64) {
      v = {acc[2 * i], acc[2 * i + 1]};
    } else {
      v = (T)acc[i];
    }
    y[i] = epilogue<O>(mode, sx, sy, v, y + i);
  }
}

void launch_acc_f64(int dt, const void* x, double* acc, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  OME_DRY_RETURN(1);
  switch (dt) {
    case F32: hipLaunchKernelGGL(k_acc_f64<OpsF32>, dim3(grid_for(n)), dim3(256), 0, st, (const float*)x, acc, n); break;
    case F64: hipLaunchKernelGGL(k_acc_f64<OpsF64>, dim3(grid_for(n)), dim3(256), 0, st, (const double*)x, acc, n); break;
    case C32: hipLaunchKernelGGL(k_acc_f64<OpsC32>, dim3(grid_for(n)), dim3(256), 0, st, (const c32*)x, acc, n); break;
    case C64: hipLaunchKernelGGL(k_acc_f64<OpsC64>, dim3(grid_for(n)), dim3(256), 0, st, (const c64*)x, acc, n); break;
    default: break;
  }
}

void launch_from_f64(int dt, const double* acc, void* y, int64_t n, const Scal& sx,
                     const Scal& sy, hipStream_t st) {
  if (n <= 0) return;
  OME_DRY_RETURN(1);
  switch (dt) {
    case F32: hipLaunchKernelGGL(k_from_f64<OpsF32>, dim3(grid_for(n)), dim3(256), 0, st, acc, (float*)y, n, sx, sy, epi_mode<OpsF32>(sx, sy)); break;
    case F64: hipLaunchKernelGGL(k_from_f64<OpsF64>, dim3(grid_for(n)), dim3(256), 0, st, acc, (double*)y, n, sx, sy, epi_mode<OpsF64>(sx, sy)); break;
    case C32: hipLaunchKernelGGL(k_from_f64<OpsC32>, dim3(grid_for(n)), dim3(256), 0, st, acc, (c32*)y, n, sx, sy, epi_mode<OpsC32>(sx, sy)); break;
    case C64: hipLaunchKernelGGL(k_from_f64<OpsC64>, dim3(grid_for(n)), dim3(256), 0, st, acc, (c64*)y, n, sx, sy, epi_mode<OpsC64>(sx, sy)); break;
    default: break;
  }
}

// ---- max-abs reduction for log-scale normalisation --------------------------------------
// Block partial maxima, then one block folds them: out[0] = max_i |x_i| (as double).
template <class O>
__device__ __forceinline__ double absval(typename O::T v) {
  if constexpr (O::DT == C32 || O::DT == C64) return hypot((double)v.x, (double)v.y);
  else return fabs((double)v);
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}

template <class O>
__global__ __launch_bounds__(256) void k_maxabs_partial(const typename O::T* __restrict__ x,
                                                        int64_t n, double* __restrict__ part) {
  __shared__ double sm[4];
  double m = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, absval<O>(x[i]));
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// Folds nparts maxima into scale[0] = m and adds log(m) to logacc[0] (when m > 0); the
// scaled tensor is x / m.
__global__ __launch_bounds__(256) void k_maxabs_final(const double* __restrict__ part, int nparts,
                                                      double* __restrict__ scale,
                                                      double* __restrict__ logacc) {
  __shared__ double sm[4];
  double m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) m = fmax(m, part[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    if (!(v > 0.0) || !isfinite(v)) v = 1.0;
    scale[0] = v;
    logacc[0] += log(v);
  }
}

template <class O>
__global__ __launch_bounds__(256) void k_scale_by_inv(typename O::T* __restrict__ x, int64_t n,
                                                      const double* __restrict__ scale) {
  const double inv = 1.0 / scale[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if constexpr (O::DT == C32) {
      x[i].x = (float)(x[i].x * inv); x[i].y = (float)(x[i].y * inv);
    } else if constexpr (O::DT == C64) {
      x[i].x *= inv; x[i].y *= inv;
    } else {
      x[i] = (typename O::T)(x[i] * inv);
    }
  }
}

void launch_normalize(int dt, void* x, int64_t n, double* work, double* logacc, hipStream_t st) {
  if (n <= 0) return;
  OME_DRY_RETURN(3);
  int g = grid_for(n, 256, 1024);
  double* part = work + 1;
  switch (dt) {
    case F32: hipLaunchKernelGGL(k_maxabs_partial<OpsF32>, dim3(g), dim3(256), 0, st, (const float*)x, n, part); break;
    case F64: hipLaunchKernelGGL(k_maxabs_partial<OpsF64>, dim3(g), dim3(256), 0, st, (const double*)x, n, part); break;
    case C32: hipLaunchKernelGGL(k_maxabs_partial<OpsC32>, dim3(g), dim3(256), 0, st, (const c32*)x, n, part); break;
    case C64: hipLaunchKernelGGL(k_maxabs_partial<OpsC64>, dim3(g), dim3(256), 0, st, (const c64*)x, n, part); break;
    default: return;
  }
  hipLaunchKernelGGL(k_maxabs_final, dim3(1), dim3(256), 0, st, part, g, work, logacc);
  switch (dt) {
    case F32: hipLaunchKernelGGL(k_scale_by_inv<OpsF32>, dim3(grid_for(n)), dim3(256), 0, st, (float*)x, n, work); break;
    case F64: hipLaunchKernelGGL(k_scale_by_inv<OpsF64>, dim3(grid_for(n)), dim3(256), 0, st, (double*)x, n, work); break;
    case C32: hipLaunchKernelGGL(k_scale_by_inv<OpsC32>, dim3(grid_for(n)), dim3(256), 0, st, (c32*)x, n, work); break;
    case C64: hipLaunchKernelGGL(k_scale_by_inv<OpsC64>, dim3(grid_for(n)), dim3(256), 0, st, (c64*)x, n, work); break;
    default: break;
  }
}

// ---- permute -------------------------------------------------------------------------
// Thread per output element (coalesced writes); the output index is decomposed with fast
// division by the output extents and mapped to an input offset through the per-dim input
// strides.  Each thread handles ELEMS consecutive outputs along the fastest output dim so the
// decomposition cost is amortised.

template <class O, int ELEMS>
__global__ __launch_bounds__(256) void k_permute(PermDesc d) {
  using T = typename O::T;
  const T* __restrict__ x = (const T*)d.x;
  T* __restrict__ y = (T*)d.y;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  const uint32_t d0 = d.odiv[0].d;
  const int64_t s0 = d.istride[0];
  // outputs are processed in runs of ELEMS along dim 0; runs do not cross dim-0 rows
  const uint32_t runs_per_row = (d0 + ELEMS - 1) / ELEMS;
  const int64_t rows = d.n / d0;
  const int64_t total_runs = rows * runs_per_row;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < total_runs;
       r += (int64_t)gridDim.x * blockDim.x) {
    uint32_t row = (uint32_t)(r / runs_per_row);
    uint32_t i0 = (uint32_t)(r - (int64_t)row * runs_per_row) * ELEMS;
    int64_t off = 0;
    uint32_t rem = row;
    for (int p = 1; p < d.rank; ++p) {
      uint32_t q = fdiv_q(rem, d.odiv[p]);
      off += (int64_t)(rem - q * d.odiv[p].d) * d.istride[p];
      rem = q;
    }
    const int64_t ybase = (int64_t)row * d0;
    // the run's loads are all issued before its stores (elements past the row load the run's
    // first element, which is in the row)
    T v[ELEMS];
#pragma unroll
    for (int e = 0; e < ELEMS; ++e) {
      const uint32_t i = i0 + e;
      v[e] = x[off + (int64_t)(i < d0 ? i : i0) * s0];
    }
#pragma unroll
    for (int e = 0; e < ELEMS; ++e) {
      const uint32_t i = i0 + e;
      if (i < d0) y[ybase + i] = epilogue<O>(d.mode, sx, sy, conj_if<O>(v[e], d.conj), y + ybase + i);
    }
  }
}

// LDS-tiled transpose for the case where the input's fastest dim is not the output's
// fastest dim: a 64x64 tile over (output dim 0, output dim t) where output dim t is the
// input's contiguous dim; both the reads and the writes are coalesced.
template <class O>
__global__ __launch_bounds__(256) void k_permute_tiled(PermDesc d, int t) {
  using T = typename O::T;
  __shared__ T tile[64][65];
  const T* __restrict__ x = (const T*)d.x;
  T* __restrict__ y = (T*)d.y;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  const uint32_t d0 = d.odiv[0].d, dt = d.odiv[t].d;
  const uint32_t tiles0 = (d0 + 63) / 64, tilest = (dt + 63) / 64;
  // blockIdx.x enumerates (tile0, tilet, rest) with rest over the other output dims
  uint32_t b = blockIdx.x;
  const uint32_t tb0 = b % tiles0; b /= tiles0;
  const uint32_t tbt = b % tilest; b /= tilest;
  uint32_t rest = b;
  // decompose rest over output dims other than 0 and t; compute input / output base
  int64_t ioff = 0, ooff = 0, ostride = d0;
  for (int p = 1; p < d.rank; ++p) {
    if (p == t) { ostride *= d.odiv[p].d; continue; }
    uint32_t q = fdiv_q(rest, d.odiv[p]);
    uint32_t v = rest - q * d.odiv[p].d;
    rest = q;
    ioff += (int64_t)v * d.istride[p];
    ooff += (int64_t)v * ostride;
    ostride *= d.odiv[p].d;
  }
  int64_t ostride_t = d0;
  for (int p = 1; p < t; ++p) ostride_t *= d.odiv[p].d;
  const int64_t s0 = d.istride[0];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // read: lanes along output dim t (input contiguous), rows along output dim 0
  for (int r = ty; r < 64; r += 4) {
    uint32_t i0 = tb0 * 64 + r, it = tbt * 64 + tx;
    if (i0 < d0 && it < dt) tile[r][tx] = x[ioff + (int64_t)i0 * s0 + (int64_t)it * d.istride[t]];
  }
  __syncthreads();
  // write: lanes along output dim 0 (output contiguous)
  for (int r = ty; r < 64; r += 4) {
    uint32_t i0 = tb0 * 64 + tx, it = tbt * 64 + r;
    if (i0 < d0 && it < dt) {
      int64_t o = ooff + (int64_t)it * ostride_t + i0;
      T v = conj_if<O>(tile[tx][r], d.conj);
      y[o] = epilogue<O>(d.mode, sx, sy, v, y + o);
    }
  }
}

// Permutation of a tensor whose grouped extents are all powers of two (bond-dimension-2
// networks: every label of extent 2).  Element offsets are then bit strings, and the
// permutation maps output offset bits to input offset bits.  A workgroup moves one tile: the
// elements that agree on every bit outside a set T of input bits, where T holds the input's
// lowest bits (a contiguous run to read) and the images of the output's lowest bits (a
// contiguous run to write), 2^8 .. 2^12 elements.  Read phase: local index i (T's bits in
// input order) = thread + 256 k, so 64 lanes read a run; write phase: local index j (T's bits
// in output order), so 64 lanes write a run, taking element i(j) from LDS.  Every offset is a
// sum of per-bit weights over disjoint bits, so a thread's part (its 8 index bits), an
// iteration's part and the tile's part are added.  LDS slots are xor-swizzled,
// slot = i ^ (((i >> 5) ^ (i >> 10)) & 31), so the write phase's gathered reads spread over
// the banks.
struct BitPermDesc {
  const void* x;
  void* y;
  int tbits;                 // |T| (8 .. 12)
  int rbits;                 // bits outside T: 2^rbits tiles
  int64_t t_in[8], t_out[8];  // thread bit q -> input offset (read phase), output offset
  int32_t t_lds[8];           // thread bit q -> input-order local index (write phase)
  int64_t k_in[16], k_out[16];  // iteration k -> input / output offset of its index bits
  int32_t k_lds[16];
  int64_t r_in[32], r_out[32];  // tile bit -> input / output offset
  Scal sx, sy;
  int mode;
  int conj;
};

__device__ __forceinline__ int bp_swz(int i) { return i ^ (((i >> 5) ^ (i >> 10)) & 31); }

// One tile (workgroup-local index b of the permute's 2^rbits tiles) of the permute d.
template <class O>
__device__ __forceinline__ void permute_bits_tile(const BitPermDesc& d, uint32_t b) {
  using T = typename O::T;
  constexpr int CAP = sizeof(T) > 8 ? 2048 : 4096;
  __shared__ T tile[CAP];
  const T* __restrict__ x = (const T*)d.x;
  T* __restrict__ y = (T*)d.y;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  const int t = threadIdx.x;
  int64_t ib = 0, ob = 0;
  for (int r = 0; r < d.rbits; ++r)
    if ((b >> r) & 1u) { ib += d.r_in[r]; ob += d.r_out[r]; }
  int64_t ti = 0, to = 0;
  int tl = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if ((t >> q) & 1) { ti += d.t_in[q]; to += d.t_out[q]; tl += d.t_lds[q]; }
  const int iters = 1 << (d.tbits - 8);
  T v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if (k < iters) v[k] = x[ib + ti + d.k_in[k]];
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if (k < iters) tile[bp_swz(t + 256 * k)] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if (k < iters) {
      const int64_t o = ob + to + d.k_out[k];
      const T e = conj_if<O>(tile[bp_swz(tl + d.k_lds[k])], d.conj);
      y[o] = epilogue<O>(d.mode, sx, sy, e, y + o);
    }
}

template <class O>
__global__ __launch_bounds__(256) void k_permute_bits(BitPermDesc d) {
  permute_bits_tile<O>(d, blockIdx.x);
}

// Two permutes in one launch: workgroups [0, 2^d0.rbits) move d0's tiles, the rest d1's.
template <class O>
__global__ __launch_bounds__(256) void k_permute_bits2(BitPermDesc d0, BitPermDesc d1) {
  const uint32_t n0 = 1u << d0.rbits;
  if (blockIdx.x < n0)
    permute_bits_tile<O>(d0, blockIdx.x);
  else
    permute_bits_tile<O>(d1, blockIdx.x - n0);
}

// The bit-permutation form of d, or false when some extent is not a power of two or the
// tensor has fewer than 2^8 elements.
static bool bit_perm_desc(int dt, const PermDesc& d, BitPermDesc& bp) {
  if (d.n < 256 || d.n >= (1ll << 32)) return false;
  int ib_of_ob[32];  // output bit -> input bit
  int nb = 0;
  for (int p = 0; p < d.rank; ++p) {
    const uint32_t e = d.odiv[p].d;
    const int64_t s = d.istride[p];
    if (e == 0 || (e & (e - 1)) || s <= 0 || (s & (s - 1))) return false;
    const int eb = __builtin_ctz(e), sb = __builtin_ctzll((uint64_t)s);
    for (int j = 0; j < eb; ++j) {
      if (nb >= 32) return false;
      ib_of_ob[nb++] = sb + j;
    }
  }
  if ((1ll << nb) != d.n) return false;
  int ob_of_ib[32];
  for (int j = 0; j < nb; ++j) ob_of_ib[j] = -1;
  for (int j = 0; j < nb; ++j) {
    if (ib_of_ob[j] >= nb || ob_of_ib[ib_of_ob[j]] >= 0) return false;
    ob_of_ib[ib_of_ob[j]] = j;
  }
  const int esz = (int)elem_size(dt);
  const int tmax = esz > 8 ? 11 : 12;
  const int target = nb < tmax ? nb : tmax;
  // runs of 128 bytes where the extents allow, then grow both runs alternately to the tile
  int run = 0;
  while ((esz << (run + 1)) <= 128) ++run;
  bool inT[32] = {false};
  int tcount = 0, a = 0, c = 0;
  auto add = [&](int ib) { if (!inT[ib] && tcount < target) { inT[ib] = true; ++tcount; } };
  while (a < run && a < nb) add(a++);
  while (c < run && c < nb) add(ib_of_ob[c++]);
  while (tcount < target) {
    if (a < nb) add(a++);
    if (c < nb) add(ib_of_ob[c++]);
  }
  int Tin[32], Tout[32], R[32], nt = 0, nr = 0, no = 0;
  int pos_in[32];
  for (int j = 0; j < nb; ++j) {
    if (inT[j]) { pos_in[j] = nt; Tin[nt++] = j; }
    else R[nr++] = j;
  }
  for (int ob = 0; ob < nb; ++ob)
    if (inT[ib_of_ob[ob]]) Tout[no++] = ib_of_ob[ob];
  bp = BitPermDesc{};
  bp.x = d.x;
  bp.y = d.y;
  bp.tbits = nt;
  bp.rbits = nr;
  for (int q = 0; q < 8; ++q) {
    bp.t_in[q] = 1ll << Tin[q];
    bp.t_out[q] = 1ll << ob_of_ib[Tout[q]];
    bp.t_lds[q] = 1 << pos_in[Tout[q]];
  }
  for (int k = 0; k < (1 << (nt - 8)); ++k) {
    int64_t ki = 0, ko = 0;
    int32_t kl = 0;
    for (int q = 8; q < nt; ++q)
      if ((k >> (q - 8)) & 1) {
        ki += 1ll << Tin[q];
        ko += 1ll << ob_of_ib[Tout[q]];
        kl += 1 << pos_in[Tout[q]];
      }
    bp.k_in[k] = ki;
    bp.k_out[k] = ko;
    bp.k_lds[k] = kl;
  }
  for (int r = 0; r < nr; ++r) {
    bp.r_in[r] = 1ll << R[r];
    bp.r_out[r] = 1ll << ob_of_ib[R[r]];
  }
  bp.sx = d.sx;
  bp.sy = d.sy;
  bp.mode = d.mode;
  bp.conj = d.conj;
  return true;
}

static bool permute_bits_on() {
  static int on = -1;
  if (on < 0) {
    const char* e = getenv("OME_PERMUTE_BITS");
    on = (e && e[0] == '0') ? 0 : 1;
  }
  return on == 1;
}

void launch_permute2(int dt, const PermDesc& d0, const PermDesc& d1, hipStream_t st) {
  BitPermDesc b0, b1;
  if (d0.n > 0 && d1.n > 0 && permute_bits_on() && bit_perm_desc(dt, d0, b0) &&
      bit_perm_desc(dt, d1, b1) && b0.rbits < 31 && b1.rbits < 31) {
    OME_DRY_RETURN(1);
    TimedScope ts(KC_PERMUTE, st);
    const uint32_t blocks = (1u << b0.rbits) + (1u << b1.rbits);
    OME_DISPATCH(dt, O, {
      hipLaunchKernelGGL(k_permute_bits2<O>, dim3(blocks), dim3(256), 0, st, b0, b1);
    });
    return;
  }
  launch_permute(dt, d0, st);
  launch_permute(dt, d1, st);
}

void launch_permute(int dt, const PermDesc& d, hipStream_t st) {
  if (d.n <= 0) return;
  OME_DRY_RETURN(1);
  TimedScope ts(KC_PERMUTE, st);
  // power-of-two extents (all of a bond-dimension-2 network's tensors): the bit-permutation
  // tile.  OME_PERMUTE_BITS=0 turns it off for measurements.
  BitPermDesc bp;
  if (permute_bits_on() && bit_perm_desc(dt, d, bp)) {
    const uint32_t blocks = 1u << bp.rbits;
    OME_DISPATCH(dt, O, {
      hipLaunchKernelGGL(k_permute_bits<O>, dim3(blocks), dim3(256), 0, st, bp);
    });
    return;
  }
  // tiled path when some output dim t > 0 is the input's contiguous dim and both extents
  // are large enough to fill a tile
  int t = -1;
  for (int p = 1; p < d.rank; ++p)
    if (d.istride[p] == 1) t = p;
  const uint32_t d0 = d.odiv[0].d;
  if (t > 0 && d.istride[0] != 1 && d0 >= 16 && d.odiv[t].d >= 16) {
    int64_t blocks = ((d0 + 63) / 64) * (int64_t)((d.odiv[t].d + 63) / 64);
    for (int p = 1; p < d.rank; ++p)
      if (p != t) blocks *= d.odiv[p].d;
    if (blocks < (1ll << 31)) {
      OME_DISPATCH(dt, O, {
        hipLaunchKernelGGL(k_permute_tiled<O>, dim3((uint32_t)blocks), dim3(256), 0, st, d, t);
      });
      return;
    }
  }
  const int64_t rows = d.n / d0;
  if (d0 >= 4) {
    int64_t runs = rows * ((d0 + 3) / 4);
    OME_DISPATCH(dt, O, {
      hipLaunchKernelGGL((k_permute<O, 4>), dim3(grid_for(runs)), dim3(256), 0, st, d);
    });
  } else {
    OME_DISPATCH(dt, O, {
      hipLaunchKernelGGL((k_permute<O, 1>), dim3(grid_for(rows * d0)), dim3(256), 0, st, d);
    });
  }
}

// ---- index walk ------------------------------------------------------------------------
// One worker per (y element, inner chunk).  The y linear index is decomposed into y-dim
// values, mapped to the unique output labels (a repeated label whose values disagree marks
// an off-diagonal element, which receives no contribution); the inner chunk is walked with
// an odometer that updates every input offset incrementally.

template <class O>
__device__ __forceinline__ typename AccOps<O>::A::T walk_point(const WalkDesc& d, int64_t yi, int64_t ibeg,
                                                    int64_t iend, bool& on_diag) {
  using T = typename O::T;
  int64_t off[WALK_MAXIN];
  for (int k = 0; k < d.nin; ++k) off[k] = 0;
  uint32_t val[WALK_MAXL];
  for (int u = 0; u < d.nouter; ++u) val[u] = 0xffffffffu;
  on_diag = true;
  uint32_t rem = (uint32_t)yi;
  for (int p = 0; p < d.ny_dims; ++p) {
    uint32_t q = fdiv_q(rem, d.ydiv[p]);
    uint32_t v = rem - q * d.ydiv[p].d;
    rem = q;
    int u = d.yslot[p];
    if (val[u] == 0xffffffffu) {
      val[u] = v;
      for (int k = 0; k < d.nin; ++k) off[k] += (int64_t)v * d.stride[u * d.nin + k];
    } else if (val[u] != v) {
      on_diag = false;
    }
  }
  using W = AccOps<O>;
  using AO = typename W::A;
  typename AO::T acc = AO::zero();
  if (!on_diag || ibeg >= iend) return acc;
  // inner odometer start
  uint32_t ival[WALK_MAXL];
  uint32_t r2 = (uint32_t)ibeg;
  const int sbase = d.nouter * d.nin;
  for (int j = 0; j < d.ninner; ++j) {
    uint32_t q = fdiv_q(r2, d.idiv[j]);
    ival[j] = r2 - q * d.idiv[j].d;
    r2 = q;
    for (int k = 0; k < d.nin; ++k) off[k] += (int64_t)ival[j] * d.stride[sbase + j * d.nin + k];
  }
  for (int64_t ip = ibeg; ip < iend; ++ip) {
    typename AO::T prod = W::widen(((const T*)d.x[0])[off[0]]);
    for (int k = 1; k < d.nin; ++k) prod = AO::mul(prod, W::widen(((const T*)d.x[k])[off[k]]));
    acc = AO::add(acc, prod);
    // advance the odometer
    for (int j = 0; j < d.ninner; ++j) {
      const int32_t* sj = d.stride + sbase + j * d.nin;
      if (++ival[j] < d.idiv[j].d) {
        for (int k = 0; k < d.nin; ++k) off[k] += sj[k];
        break;
      }
      ival[j] = 0;
      for (int k = 0; k < d.nin; ++k) off[k] -= (int64_t)sj[k] * (d.idiv[j].d - 1);
    }
  }
  return acc;
}

template <class O>
__global__ __launch_bounds__(256) void k_walk(WalkDesc d) {
  using T = typename O::T;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  T* y = (T*)d.y;
  const int64_t total = d.ny * d.nsplit;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < total;
       w += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = w / d.ny;
    const int64_t yi = w - s * d.ny;
    const int64_t ibeg = s * d.chunk;
    int64_t iend = ibeg + d.chunk;
    if (iend > d.ninner_pts) iend = d.ninner_pts;
    bool on_diag;
    const T acc = AccOps<O>::narrow(walk_point<O>(d, yi, ibeg, iend, on_diag));
    if (d.nsplit == 1) {
      y[yi] = epilogue<O>(d.mode, sx, sy, acc, y + yi);
    } else {
      ((T*)d.partial)[s * d.ny + yi] = acc;
    }
  }
}

// Fold the nsplit partial sums of every y element in split order, then apply the epilogue.
template <class O>
__global__ __launch_bounds__(256) void k_walk_reduce(WalkDesc d) {
  using T = typename O::T;
  const T sx = O::from(d.sx), sy = O::from(d.sy);
  T* y = (T*)d.y;
  const T* part = (const T*)d.partial;
  for (int64_t yi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; yi < d.ny;
       yi += (int64_t)gridDim.x * blockDim.x) {
    using W = AccOps<O>;
    typename W::A::T acc = W::A::zero();
    for (int s = 0; s < d.nsplit; ++s) acc = W::A::add(acc, W::widen(part[(int64_t)s * d.ny + yi]));
    y[yi] = epilogue<O>(d.mode, sx, sy, W::narrow(acc), y + yi);
  }
}

void launch_walk(int dt, const WalkDesc& d, hipStream_t st) {
  if (d.ny <= 0) return;
  OME_DRY_RETURN((d.nsplit > 1 ? 2 : 1));
  const int64_t total = d.ny * d.nsplit;
  OME_DISPATCH(dt, O, {
    {
      TimedScope ts(KC_WALK, st);
      hipLaunchKernelGGL(k_walk<O>, dim3(grid_for(total, 256, 1 << 20)), dim3(256), 0, st, d);
    }
    if (d.nsplit > 1) {
      TimedScope ts(KC_OTHER, st);
      hipLaunchKernelGGL(k_walk_reduce<O>, dim3(grid_for(d.ny)), dim3(256), 0, st, d);
    }
  });
}

}  // namespace ome

namespace ome {

// y = widen(x) / scale[0]: a Float32 / ComplexF32 tensor written as Float64 / ComplexF64,
// divided by its largest magnitude (the per-tensor normalisation of a network's inputs).
template <class O>
__global__ __launch_bounds__(256) void k_widen_scaled(const typename O::T* __restrict__ x,
                                                      typename AccOps<O>::A::T* __restrict__ y,
                                                      int64_t n, const double* __restrict__ scale) {
  const double inv = 1.0 / scale[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const auto v = AccOps<O>::widen(x[i]);
    if constexpr (O::DT == C32) {
      y[i] = {v.x * inv, v.y * inv};
    } else {
      y[i] = v * inv;
    }
  }
}

void launch_widen_normalized(int dt, const void* x, void* y, int64_t n, double* work,
                             double* logacc, hipStream_t st) {
  if (n <= 0) return;
  OME_DRY_RETURN(3);
  const int g = grid_for(n, 256, 1024);
  double* part = work + 1;
  if (dt == F32) {
    hipLaunchKernelGGL(k_maxabs_partial<OpsF32>, dim3(g), dim3(256), 0, st, (const float*)x, n, part);
    hipLaunchKernelGGL(k_maxabs_final, dim3(1), dim3(256), 0, st, part, g, work, logacc);
    hipLaunchKernelGGL(k_widen_scaled<OpsF32>, dim3(grid_for(n)), dim3(256), 0, st, (const float*)x,
                       (double*)y, n, work);
  } else if (dt == C32) {
    hipLaunchKernelGGL(k_maxabs_partial<OpsC32>, dim3(g), dim3(256), 0, st, (const c32*)x, n, part);
    hipLaunchKernelGGL(k_maxabs_final, dim3(1), dim3(256), 0, st, part, g, work, logacc);
    hipLaunchKernelGGL(k_widen_scaled<OpsC32>, dim3(grid_for(n)), dim3(256), 0, st, (const c32*)x,
                       (c64*)y, n, work);
  }
}

}  // namespace ome
