// Shared definitions of the MI355X contraction backend: element types, semiring
// arithmetic, fast integer division for index decomposition, and the kernel descriptors.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace ome {

enum DType : int {
  F32 = 0, F64 = 1, C32 = 2, C64 = 3, I64 = 4, TF32 = 5, TF64 = 6, DD = 7, CT64 = 8, TBOOL = 9
};
constexpr int DT_LAST = TBOOL;  // element type codes run 0..DT_LAST

inline size_t elem_size(int dt) {
  switch (dt) {
    case F32: case TF32: return 4;
    case F64: case I64: case TF64: case C32: return 8;
    case C64: case DD: case CT64: return 16;
    case TBOOL: return 1;
  }
  return 0;
}
inline bool is_complex(int dt) { return dt == C32 || dt == C64; }
inline bool is_tropical(int dt) { return dt == TF32 || dt == TF64; }
// element types whose values are real or complex numbers (FP64 slice sums, normalisation)
inline bool is_numeric_float(int dt) { return dt == F32 || dt == F64 || dt == C32 || dt == C64; }

struct Scal {  // layout of ome_scalar
  double re, im;
  int64_t i;
};

struct c32 { float x, y; };
struct c64 { double x, y; };
struct dd { double hi, lo; };  // double-word number hi + lo, |lo| <= ulp(hi) / 2
struct ct64 { double n, c; };  // CountingTropical{Float64,Float64}: tropical value n, count c

// ---- semiring arithmetic, one struct per element type ----------------------------------
// mul/add are the semiring's product and sum; zero/one its neutral elements.

struct OpsF32 {
  using T = float;
  static constexpr int DT = F32;
  __host__ __device__ static T zero() { return 0.f; }
  __host__ __device__ static T one() { return 1.f; }
  __host__ __device__ static T mul(T a, T b) { return a * b; }
  __host__ __device__ static T add(T a, T b) { return a + b; }
  __host__ __device__ static T fma(T a, T b, T c) { return a * b + c; }
  __host__ __device__ static T from(const Scal& s) { return (float)s.re; }
  static bool is_zero(const Scal& s) { return (float)s.re == 0.f; }
};
struct OpsF64 {
  using T = double;
  static constexpr int DT = F64;
  __host__ __device__ static T zero() { return 0.0; }
  __host__ __device__ static T one() { return 1.0; }
  __host__ __device__ static T mul(T a, T b) { return a * b; }
  __host__ __device__ static T add(T a, T b) { return a + b; }
  __host__ __device__ static T fma(T a, T b, T c) { return a * b + c; }
  __host__ __device__ static T from(const Scal& s) { return s.re; }
  static bool is_zero(const Scal& s) { return s.re == 0.0; }
};
struct OpsC32 {
  using T = c32;
  static constexpr int DT = C32;
  __host__ __device__ static T zero() { return {0.f, 0.f}; }
  __host__ __device__ static T one() { return {1.f, 0.f}; }
  __host__ __device__ static T mul(T a, T b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
  }
  __host__ __device__ static T add(T a, T b) { return {a.x + b.x, a.y + b.y}; }
  __host__ __device__ static T fma(T a, T b, T c) {
    return {c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x)};
  }
  __host__ __device__ static T from(const Scal& s) { return {(float)s.re, (float)s.im}; }
  static bool is_zero(const Scal& s) { return (float)s.re == 0.f && (float)s.im == 0.f; }
};
struct OpsC64 {
  using T = c64;
  static constexpr int DT = C64;
  __host__ __device__ static T zero() { return {0.0, 0.0}; }
  __host__ __device__ static T one() { return {1.0, 0.0}; }
  __host__ __device__ static T mul(T a, T b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
  }
  __host__ __device__ static T add(T a, T b) { return {a.x + b.x, a.y + b.y}; }
  __host__ __device__ static T fma(T a, T b, T c) {
    return {c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x)};
  }
  __host__ __device__ static T from(const Scal& s) { return {s.re, s.im}; }
  static bool is_zero(const Scal& s) { return s.re == 0.0 && s.im == 0.0; }
};
// Accumulation arithmetic of the long reductions that are bound by HBM (the few-outputs dot
// kernel and the split-K partial sums): Float32 and ComplexF32 sum in FP64, which costs
// nothing there, and round once; every other type sums in its own arithmetic.
template <class O>
struct AccOps {
  using A = O;
  __host__ __device__ static typename O::T widen(typename O::T x) { return x; }
  __host__ __device__ static typename O::T narrow(typename O::T x) { return x; }
};
template <>
struct AccOps<OpsF32> {
  using A = OpsF64;
  __host__ __device__ static double widen(float x) { return (double)x; }
  __host__ __device__ static float narrow(double x) { return (float)x; }
};
template <>
struct AccOps<OpsC32> {
  using A = OpsC64;
  __host__ __device__ static c64 widen(c32 x) { return {(double)x.x, (double)x.y}; }
  __host__ __device__ static c32 narrow(c64 x) { return {(float)x.x, (float)x.y}; }
};

// Int64: two's-complement wrapping arithmetic (Julia Int64), computed on uint64 so that
// overflow is defined.
struct OpsI64 {
  using T = int64_t;
  static constexpr int DT = I64;
  __host__ __device__ static T zero() { return 0; }
  __host__ __device__ static T one() { return 1; }
  __host__ __device__ static T mul(T a, T b) { return (T)((uint64_t)a * (uint64_t)b); }
  __host__ __device__ static T add(T a, T b) { return (T)((uint64_t)a + (uint64_t)b); }
  __host__ __device__ static T fma(T a, T b, T c) { return add(mul(a, b), c); }
  __host__ __device__ static T from(const Scal& s) { return s.i; }
  static bool is_zero(const Scal& s) { return s.i == 0; }
};
// Tropical (max-plus): a (+) b = max(a, b), a (*) b = a + b, zero = -Inf, one = 0.
struct OpsTF32 {
  using T = float;
  static constexpr int DT = TF32;
  __host__ __device__ static T zero() { return -INFINITY; }
  __host__ __device__ static T one() { return 0.f; }
  __host__ __device__ static T mul(T a, T b) { return a + b; }
  __host__ __device__ static T add(T a, T b) { return fmaxf(a, b); }
  __host__ __device__ static T fma(T a, T b, T c) { return fmaxf(a + b, c); }
  __host__ __device__ static T from(const Scal& s) { return (float)s.re; }
  static bool is_zero(const Scal& s) { return std::isinf(s.re) && s.re < 0; }
};
struct OpsTF64 {
  using T = double;
  static constexpr int DT = TF64;
  __host__ __device__ static T zero() { return -INFINITY; }
  __host__ __device__ static T one() { return 0.0; }
  __host__ __device__ static T mul(T a, T b) { return a + b; }
  __host__ __device__ static T add(T a, T b) { return fmax(a, b); }
  __host__ __device__ static T fma(T a, T b, T c) { return fmax(a + b, c); }
  __host__ __device__ static T from(const Scal& s) { return s.re; }
  static bool is_zero(const Scal& s) { return std::isinf(s.re) && s.re < 0; }
};
// Double-word numbers (DoubleFloats' Double64): value = hi + lo.  The operations are the
// published double-word algorithms of Joldes, Muller and Popescu, "Tight and rigorous error
// bounds for basic building blocks of double-word arithmetic", ACM TOMS 44(2), 2017:
//   add: Algorithm 6 (AccurateDWPlusDW), relative error <= 3u^2 + 13u^3;
//   mul: Algorithm 12 (DWTimesDW3), relative error <= 4u^2;   u = 2^-53.
// Every rounding is the one the algorithms specify: FP contraction is off in these bodies and
// the one fused operation used is the explicit fma of the error-free product.
struct OpsDD {
  using T = dd;
  static constexpr int DT = DD;
  __host__ __device__ static T zero() { return {0.0, 0.0}; }
  __host__ __device__ static T one() { return {1.0, 0.0}; }
  // s + e = a + b exactly (Knuth's TwoSum)
  __host__ __device__ static void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bv = s - a;
    e = (a - (s - bv)) + (b - bv);
  }
  // s + e = a + b exactly when the exponent of a is at least that of b (Dekker's Fast2Sum)
  __host__ __device__ static void fast_two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    s = a + b;
    e = b - (s - a);
  }
  __host__ __device__ static T add(T x, T y) {
#pragma clang fp contract(off)
    double sh, sl, th, tl, vh, vl;
    two_sum(x.hi, y.hi, sh, sl);
    two_sum(x.lo, y.lo, th, tl);
    const double c = sl + th;
    fast_two_sum(sh, c, vh, vl);
    const double w = tl + vl;
    T z;
    fast_two_sum(vh, w, z.hi, z.lo);
    return z;
  }
  __host__ __device__ static T mul(T x, T y) {
#pragma clang fp contract(off)
    const double ch = x.hi * y.hi;
    const double cl1 = __builtin_fma(x.hi, y.hi, -ch);  // ch + cl1 = x.hi * y.hi exactly
    const double tl0 = x.lo * y.lo;
    const double tl1 = __builtin_fma(x.hi, y.lo, tl0);
    const double cl2 = __builtin_fma(x.lo, y.hi, tl1);
    const double cl3 = cl1 + cl2;
    T z;
    fast_two_sum(ch, cl3, z.hi, z.lo);
    return z;
  }
  __host__ __device__ static T fma(T a, T b, T c) { return add(mul(a, b), c); }
  __host__ __device__ static T from(const Scal& s) { return {s.re, 0.0}; }
  static bool is_zero(const Scal& s) { return s.re == 0.0; }
};

// CountingTropical{Float64,Float64} (TropicalNumbers): a (*) b = (a.n + b.n, a.c * b.c);
// a (+) b keeps the larger n with its count and adds the counts of equal n; zero = (-Inf, 0),
// one = (0, 1).  A scalar's re is n and im is c.
struct OpsCT64 {
  using T = ct64;
  static constexpr int DT = CT64;
  __host__ __device__ static T zero() { return {-INFINITY, 0.0}; }
  __host__ __device__ static T one() { return {0.0, 1.0}; }
  __host__ __device__ static T mul(T a, T b) { return {a.n + b.n, a.c * b.c}; }
  __host__ __device__ static T add(T a, T b) {
    if (a.n > b.n) return a;
    if (b.n > a.n) return b;
    return {a.n, a.c + b.c};
  }
  __host__ __device__ static T fma(T a, T b, T c) { return add(mul(a, b), c); }
  __host__ __device__ static T from(const Scal& s) { return {s.re, s.im}; }
  static bool is_zero(const Scal& s) { return std::isinf(s.re) && s.re < 0; }
};
// TropicalAndOr (TropicalNumbers' boolean semiring): a (+) b = a | b, a (*) b = a & b,
// zero = false, one = true; one byte per element (0 or 1).  A scalar's i is the value.
struct OpsTBool {
  using T = uint8_t;
  static constexpr int DT = TBOOL;
  __host__ __device__ static T zero() { return 0; }
  __host__ __device__ static T one() { return 1; }
  __host__ __device__ static T mul(T a, T b) { return a & b; }
  __host__ __device__ static T add(T a, T b) { return a | b; }
  __host__ __device__ static T fma(T a, T b, T c) { return (a & b) | c; }
  __host__ __device__ static T from(const Scal& s) { return s.i != 0 ? 1 : 0; }
  static bool is_zero(const Scal& s) { return s.i == 0; }
};

// Dispatch a functor templated on the Ops struct by runtime dtype.
#define OME_DISPATCH(dt, OPS, ...)                          \
  switch (dt) {                                             \
    case F32: { using OPS = OpsF32; __VA_ARGS__; } break;   \
    case F64: { using OPS = OpsF64; __VA_ARGS__; } break;   \
    case C32: { using OPS = OpsC32; __VA_ARGS__; } break;   \
    case C64: { using OPS = OpsC64; __VA_ARGS__; } break;   \
    case I64: { using OPS = OpsI64; __VA_ARGS__; } break;   \
    case TF32: { using OPS = OpsTF32; __VA_ARGS__; } break; \
    case TF64: { using OPS = OpsTF64; __VA_ARGS__; } break; \
    case DD: { using OPS = OpsDD; __VA_ARGS__; } break;     \
    case CT64: { using OPS = OpsCT64; __VA_ARGS__; } break; \
    case TBOOL: { using OPS = OpsTBool; __VA_ARGS__; } break; \
    default: break;                                         \
  }

// BF16x3 split of two floats on gfx950's packed float -> bfloat16 conversion (round to
// nearest even): hi = the two bf16(x) (a in the low half), lo = the two bf16(x - hi).
typedef float split_f2 __attribute__((ext_vector_type(2)));
typedef __bf16 split_b2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split2_bf16(float a, float b, uint32_t& hi, uint32_t& lo) {
  const split_f2 x = {a, b};
  const split_b2 h = __builtin_convertvector(x, split_b2);
  uint32_t hb;
  __builtin_memcpy(&hb, &h, 4);
  const split_f2 hf = {__uint_as_float(hb << 16), __uint_as_float(hb & 0xffff0000u)};
  const split_b2 l = __builtin_convertvector(x - hf, split_b2);
  __builtin_memcpy(&lo, &l, 4);
  hi = hb;
}

// Three-plane split of two floats (six-product form): hi = bf16(x), mid = bf16(x - hi),
// lo = bf16(x - hi - mid), each rounded to nearest even.  Both differences are exact in FP32
// (each subtrahend is the nearest bf16 of its minuend), so |x - hi - mid - lo| <= 2^-24 |x|.
__device__ __forceinline__ void split3_bf16(float a, float b, uint32_t& hi, uint32_t& mid,
                                            uint32_t& lo) {
  const split_f2 x = {a, b};
  const split_b2 h = __builtin_convertvector(x, split_b2);
  uint32_t hb;
  __builtin_memcpy(&hb, &h, 4);
  const split_f2 hf = {__uint_as_float(hb << 16), __uint_as_float(hb & 0xffff0000u)};
  const split_f2 r = x - hf;
  const split_b2 m = __builtin_convertvector(r, split_b2);
  uint32_t mb;
  __builtin_memcpy(&mb, &m, 4);
  const split_f2 mf = {__uint_as_float(mb << 16), __uint_as_float(mb & 0xffff0000u)};
  const split_b2 l = __builtin_convertvector(r - mf, split_b2);
  __builtin_memcpy(&lo, &l, 4);
  hi = hb;
  mid = mb;
}

// Epilogue mode:  y <- sx*v (sy zero: y not read) | sy*y + sx*v | sy*y (sx zero).
enum EpiMode : int { EPI_OVERWRITE = 0, EPI_AXPBY = 1, EPI_SCALE_ONLY = 2 };

template <class O>
__device__ __forceinline__ typename O::T epilogue(int mode, typename O::T sx, typename O::T sy,
                                                  typename O::T v, const typename O::T* yp) {
  if (mode == EPI_OVERWRITE) return O::mul(sx, v);
  typename O::T y = *yp;
  if (mode == EPI_SCALE_ONLY) return O::mul(sy, y);
  return O::add(O::mul(sy, y), O::mul(sx, v));
}

template <class O>
inline int epi_mode(const Scal& sx, const Scal& sy) {
  if (O::is_zero(sx)) return EPI_SCALE_ONLY;
  if (O::is_zero(sy)) return EPI_OVERWRITE;
  return EPI_AXPBY;
}

// ---- fast unsigned division by an invariant divisor (Granlund & Montgomery 1994) -------
// q = (umulhi(n, m) + n) >> s, valid for all 32-bit n with the 33-bit intermediate.
struct FDiv {
  uint32_t d, m, s;
};
inline FDiv make_fdiv(uint32_t d) {
  FDiv f;
  f.d = d ? d : 1;
  uint32_t l = 0;
  while ((1ull << l) < f.d) ++l;
  f.s = l;
  f.m = (uint32_t)((((1ull << l) - f.d) << 32) / f.d + 1);
  if (f.d == 1) { f.m = 0; f.s = 0; }
  return f;
}
__device__ __forceinline__ uint32_t fdiv_q(uint32_t n, const FDiv& f) {
  return (uint32_t)(((uint64_t)__umulhi(n, f.m) + n) >> f.s);
}

constexpr int MAXR = 40;        // dims per permute after grouping
constexpr int WALK_MAXIN = 16;  // inputs of one index walk
constexpr int WALK_MAXL = 40;   // output dims / outer labels / inner labels
constexpr int WALK_MAXS = 192;  // stride table entries (labels x inputs)

struct PermDesc {
  const void* x;
  void* y;
  int64_t n;          // elements
  int rank;           // output dims
  FDiv odiv[MAXR];    // output extents, first dim fastest
  int64_t istride[MAXR];  // input stride of each output dim
  Scal sx, sy;
  int mode;
  int conj;
};

struct WalkDesc {
  const void* x[WALK_MAXIN];
  void* y;
  void* partial;      // [nsplit][ny] accumulators when the inner space is split
  int nin;
  int ny_dims;        // dims of y (repeats allowed)
  int nouter;         // unique output labels
  int ninner;
  int64_t ny;         // elements of y
  int64_t ninner_pts; // points of the inner space
  int64_t chunk;      // inner points per split
  int nsplit;
  int mode;
  Scal sx, sy;
  FDiv ydiv[WALK_MAXL];
  int8_t yslot[WALK_MAXL];    // y dim -> outer slot
  FDiv idiv[WALK_MAXL];       // inner label extents (first fastest)
  int32_t stride[WALK_MAXS];  // [label slot][input]; outer slots first, then inner
};

struct GemmDesc {
  const void* a;
  const void* b;
  void* y;
  int64_t M, N, K, B;
  int64_t sam, sak, sab;
  int64_t sbk, sbn, sbb;
  int64_t sym, syn, syb;
  Scal sx, sy;
  int mode;
  int ksplit;       // >1: partial products into work[ksplit][B][M][N], reduced after
  void* work;
};

}  // namespace ome
