/*
 * TEST INFRASTRUCTURE ONLY — the correctness oracle.  Nothing in the product path links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * CPU restatement of OMEinsum.loop_einsum! — the brute-force index walk that the
 * reference's own test-suite uses as its oracle.
 *
 *   loop_einsum!            src/loop_einsum.jl:21-32   (y <- 0 if iszero(sy), y <- sy*y
 *                                                       unless isone(sy); then reduce)
 *   reduce_einarray!        src/loop_einsum.jl:34-45   (for each outer point: yi = zero(T);
 *                                                       yi += map_prod over inner points;
 *                                                       y[iy] += sx*yi)
 *   indices_and_locs        src/Core.jl:256-268        (outer = unique(iy); inner = labels of
 *                                                       ixs not in outer, first-appearance)
 *   einarray / EinIndexer   src/Core.jl:122-145,201-224 (column-major offsets)
 *   map_prod                src/Core.jl:240-243        (left-fold product x1*x2*...*xn)
 *
 * Iteration order is the reference's: outer points column-major over the unique output
 * labels, inner points column-major over the inner labels (first label fastest), so for
 * Int64 and tropical elements the result is bit-identical and for floats the summation
 * order is the reference's sequential order.
 *
 * Element types: 0 f32, 1 f64, 2 c32, 3 c64, 4 i64 (wrapping), 5 tropical f32,
 * 6 tropical f64 (TropicalNumbers semantics: + is max, * is +, zero -Inf, one 0),
 * 7 double-word (DoubleFloats' Double64, the element type of test/roceinsum.jl:56-72: an
 * interleaved (hi, lo) pair of doubles).  DoubleFloats is a dependency of the reference's
 * tests that is not part of the reference tree; its + and * are restated here as the
 * published double-word algorithms of Joldes, Muller and Popescu, ACM TOMS 44(2), 2017:
 * Algorithm 6 (AccurateDWPlusDW) and Algorithm 12 (DWTimesDW3).  They are pinned against
 * exact rational arithmetic and those papers' error bounds in tests/test_oracle_double64.py;
 * 8 CountingTropical{Float64,Float64} (an interleaved (n, c) pair: * adds n and multiplies
 * c, + keeps the larger n and adds the counts of equal n, zero (-Inf, 0), one (0, 1));
 * 9 TropicalAndOr (one byte, + is or, * is and).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { D_F32 = 0, D_F64 = 1, D_C32 = 2, D_C64 = 3, D_I64 = 4, D_TF32 = 5, D_TF64 = 6, D_DD = 7,
       D_CT64 = 8, D_BOOL = 9 };

typedef struct { double re, im; int64_t i; } oscalar;

#define MAXL 128
#define MAXIN 256

/* a generic value: enough for every element type */
typedef struct { double re, im; uint64_t i; } val;

static val load(int dt, const void* p, int64_t off) {
  val v = {0, 0, 0};
  switch (dt) {
    case D_F32: case D_TF32: v.re = ((const float*)p)[off]; break;
    case D_F64: case D_TF64: v.re = ((const double*)p)[off]; break;
    case D_C32: v.re = ((const float*)p)[2 * off]; v.im = ((const float*)p)[2 * off + 1]; break;
    case D_C64: case D_DD: case D_CT64: v.re = ((const double*)p)[2 * off]; v.im = ((const double*)p)[2 * off + 1]; break;
    case D_I64: v.i = ((const uint64_t*)p)[off]; break;
    case D_BOOL: v.i = ((const uint8_t*)p)[off] != 0; break;
  }
  return v;
}

static void store(int dt, void* p, int64_t off, val v) {
  switch (dt) {
    case D_F32: case D_TF32: ((float*)p)[off] = (float)v.re; break;
    case D_F64: case D_TF64: ((double*)p)[off] = v.re; break;
    case D_C32: ((float*)p)[2 * off] = (float)v.re; ((float*)p)[2 * off + 1] = (float)v.im; break;
    case D_C64: case D_DD: case D_CT64: ((double*)p)[2 * off] = v.re; ((double*)p)[2 * off + 1] = v.im; break;
    case D_I64: ((uint64_t*)p)[off] = v.i; break;
    case D_BOOL: ((uint8_t*)p)[off] = v.i ? 1 : 0; break;
  }
}

/* Arithmetic is performed in the element type: Float32 results are rounded to float after
 * every operation, exactly as Julia evaluates Float32 expressions. */
static val round_t(int dt, val v) {
  if (dt == D_F32 || dt == D_TF32) v.re = (float)v.re;
  if (dt == D_C32) { v.re = (float)v.re; v.im = (float)v.im; }
  return v;
}

static val vzero(int dt) {
  val v = {0, 0, 0};
  if (dt == D_TF32 || dt == D_TF64 || dt == D_CT64) v.re = -INFINITY;
  return v;
}

/* ---- double-word arithmetic (value = re + im for D_DD; |im| <= ulp(re) / 2) */
/* s + e = a + b exactly (TwoSum) */
static void two_sum(double a, double b, double* s, double* e) {
  double bv;
  *s = a + b;
  bv = *s - a;
  *e = (a - (*s - bv)) + (b - bv);
}
/* s + e = a + b exactly when exponent(a) >= exponent(b) (Fast2Sum) */
static void fast_two_sum(double a, double b, double* s, double* e) {
  *s = a + b;
  *e = b - (*s - a);
}
/* Algorithm 6, AccurateDWPlusDW */
static val dw_add(val x, val y) {
  double sh, sl, th, tl, c, vh, vl, w;
  val z = {0, 0, 0};
  two_sum(x.re, y.re, &sh, &sl);
  two_sum(x.im, y.im, &th, &tl);
  c = sl + th;
  fast_two_sum(sh, c, &vh, &vl);
  w = tl + vl;
  fast_two_sum(vh, w, &z.re, &z.im);
  return z;
}
/* Algorithm 12, DWTimesDW3 (fma gives the exact error of the leading product) */
static val dw_mul(val x, val y) {
  double ch, cl1, tl0, tl1, cl2, cl3;
  val z = {0, 0, 0};
  ch = x.re * y.re;
  cl1 = fma(x.re, y.re, -ch);
  tl0 = x.im * y.im;
  tl1 = fma(x.re, y.im, tl0);
  cl2 = fma(x.im, y.re, tl1);
  cl3 = cl1 + cl2;
  fast_two_sum(ch, cl3, &z.re, &z.im);
  return z;
}

static val vmul(int dt, val a, val b) {
  val r = {0, 0, 0};
  switch (dt) {
    case D_F32: case D_F64: r.re = a.re * b.re; break;
    case D_C32: case D_C64:
      /* Julia: (a+bi)(c+di) = (ac - bd) + (ad + bc)i  (base/complex.jl) */
      if (dt == D_C32) {
        float ar = (float)a.re, ai = (float)a.im, br = (float)b.re, bi = (float)b.im;
        r.re = (float)(ar * br - ai * bi);
        r.im = (float)(ar * bi + ai * br);
        /* the float expression above is evaluated in float by C on this target only when
           FLT_EVAL_METHOD == 0 (x86-64 SSE), which gcc guarantees here */
      } else {
        r.re = a.re * b.re - a.im * b.im;
        r.im = a.re * b.im + a.im * b.re;
      }
      return r;
    case D_I64: r.i = a.i * b.i; break;
    case D_TF32: case D_TF64: r.re = a.re + b.re; break;
    case D_DD: return dw_mul(a, b);
    case D_CT64: r.re = a.re + b.re; r.im = a.im * b.im; break;
    case D_BOOL: r.i = a.i & b.i; break;
  }
  return round_t(dt, r);
}

static val vadd(int dt, val a, val b) {
  val r = {0, 0, 0};
  switch (dt) {
    case D_F32: case D_F64: r.re = a.re + b.re; break;
    case D_C32: case D_C64: r.re = a.re + b.re; r.im = a.im + b.im; break;
    case D_I64: r.i = a.i + b.i; break;
    case D_TF32: case D_TF64: r.re = a.re > b.re ? a.re : b.re; break;
    case D_DD: return dw_add(a, b);
    case D_CT64:
      if (a.re > b.re) return a;
      if (b.re > a.re) return b;
      r.re = a.re;
      r.im = a.im + b.im;
      break;
    case D_BOOL: r.i = a.i | b.i; break;
  }
  return round_t(dt, r);
}

static val from_scalar(int dt, const oscalar* s) {
  val v = {0, 0, 0};
  v.re = s->re; v.im = s->im; v.i = (uint64_t)s->i;
  if (dt == D_BOOL) v.i = s->i != 0;
  if (dt == D_F32 || dt == D_F64 || dt == D_TF32 || dt == D_TF64 || dt == D_DD) v.im = 0;
  return round_t(dt, v);
}

static int is_zero(int dt, val v) {
  switch (dt) {
    case D_F32: case D_F64: return v.re == 0.0;
    case D_C32: case D_C64: case D_DD: return v.re == 0.0 && v.im == 0.0;
    case D_I64: case D_BOOL: return v.i == 0;
    default: return v.re == -INFINITY;
  }
}
static int is_one(int dt, val v) {
  switch (dt) {
    case D_F32: case D_F64: return v.re == 1.0;
    case D_C32: case D_C64: case D_DD: return v.re == 1.0 && v.im == 0.0;
    case D_I64: case D_BOOL: return v.i == 1;
    case D_CT64: return v.re == 0.0 && v.im == 1.0;
    default: return v.re == 0.0;
  }
}

/* Returns 0 on success, 1 on invalid arguments. */
int oracle_loop_einsum(int dt, int nin, const void* const* xs, const int32_t* nix,
                       const int32_t* ixs, void* y, int niy, const int32_t* iy, int nlabels,
                       const int64_t* sizes, const oscalar* sx_, const oscalar* sy_) {
  if (nin < 1 || nin > MAXIN || nlabels > MAXL || niy > MAXL) return 1;
  /* ---- indices_and_locs (src/Core.jl:256-268) */
  int outer[MAXL], nouter = 0, inner[MAXL], ninner = 0;
  int is_outer[MAXL];
  memset(is_outer, 0, sizeof is_outer);
  for (int p = 0; p < niy; ++p) {
    int l = iy[p];
    if (l < 0 || l >= nlabels) return 1;
    if (!is_outer[l]) { is_outer[l] = 1; outer[nouter++] = l; }
  }
  int seen_inner[MAXL];
  memset(seen_inner, 0, sizeof seen_inner);
  int pos = 0;
  for (int k = 0; k < nin; ++k) {
    for (int p = 0; p < nix[k]; ++p) {
      int l = ixs[pos + p];
      if (l < 0 || l >= nlabels) return 1;
      if (!is_outer[l] && !seen_inner[l]) { seen_inner[l] = 1; inner[ninner++] = l; }
    }
    pos += nix[k];
  }
  /* ---- EinIndexer cumsizes (src/Core.jl:132-135): per input, per position stride */
  int64_t* xstride = (int64_t*)malloc(sizeof(int64_t) * (size_t)(pos > 0 ? pos : 1));
  pos = 0;
  for (int k = 0; k < nin; ++k) {
    int64_t s = 1;
    for (int p = 0; p < nix[k]; ++p) { xstride[pos + p] = s; s *= sizes[ixs[pos + p]]; }
    pos += nix[k];
  }
  int64_t ystride[MAXL];
  {
    int64_t s = 1;
    for (int p = 0; p < niy; ++p) { ystride[p] = s; s *= sizes[iy[p]]; }
  }
  int64_t ylen = 1;
  for (int p = 0; p < niy; ++p) ylen *= sizes[iy[p]];

  val sx = from_scalar(dt, sx_), sy = from_scalar(dt, sy_);
  /* ---- loop_einsum! pre-scaling (src/loop_einsum.jl:26-30) */
  if (is_zero(dt, sy)) {
    for (int64_t o = 0; o < ylen; ++o) store(dt, y, o, vzero(dt));
  } else if (!is_one(dt, sy)) {
    for (int64_t o = 0; o < ylen; ++o) store(dt, y, o, vmul(dt, sy, load(dt, y, o)));
  }
  /* ---- reduce_einarray! (src/loop_einsum.jl:34-45) */
  int64_t value[MAXL];
  memset(value, 0, sizeof value);
  int64_t nouter_pts = 1, ninner_pts = 1;
  for (int u = 0; u < nouter; ++u) nouter_pts *= sizes[outer[u]];
  for (int u = 0; u < ninner; ++u) ninner_pts *= sizes[inner[u]];
  if (nouter_pts == 0 || ninner_pts == 0) { free(xstride); return 0; }
  for (int u = 0; u < nouter; ++u) value[outer[u]] = 0;
  for (int64_t op = 0; op < nouter_pts; ++op) {
    for (int u = 0; u < ninner; ++u) value[inner[u]] = 0;
    val yi = vzero(dt);
    for (int64_t ip = 0; ip < ninner_pts; ++ip) {
      /* map_prod: left-fold product over inputs (src/Core.jl:240-243) */
      val prod = {0, 0, 0};
      pos = 0;
      for (int k = 0; k < nin; ++k) {
        int64_t off = 0;
        for (int p = 0; p < nix[k]; ++p) off += value[ixs[pos + p]] * xstride[pos + p];
        pos += nix[k];
        val xv = round_t(dt, load(dt, xs[k], off));
        prod = (k == 0) ? xv : vmul(dt, prod, xv);
      }
      yi = vadd(dt, yi, prod);
      /* advance the inner odometer, first inner label fastest */
      for (int u = 0; u < ninner; ++u) {
        int l = inner[u];
        if (++value[l] < sizes[l]) break;
        value[l] = 0;
      }
    }
    int64_t yoff = 0;
    for (int p = 0; p < niy; ++p) yoff += value[iy[p]] * ystride[p];
    store(dt, y, yoff, vadd(dt, load(dt, y, yoff), vmul(dt, sx, yi)));
    for (int u = 0; u < nouter; ++u) {
      int l = outer[u];
      if (++value[l] < sizes[l]) break;
      value[l] = 0;
    }
  }
  free(xstride);
  return 0;
}

/* The element arithmetic alone, for the oracle's tests: op 0 = +, 1 = *, on n pairs. */
int oracle_dw_op(int op, int64_t n, const double* x, const double* y, double* z) {
  for (int64_t i = 0; i < n; ++i) {
    val a = {x[2 * i], x[2 * i + 1], 0}, b = {y[2 * i], y[2 * i + 1], 0};
    val r = op == 0 ? dw_add(a, b) : dw_mul(a, b);
    z[2 * i] = r.re;
    z[2 * i + 1] = r.im;
  }
  return 0;
}
