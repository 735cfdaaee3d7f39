/*
 * TEST INFRASTRUCTURE ONLY — drives the C oracle (loop_einsum.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle sanitize`).  Every case checks the oracle's
 * result against an answer computed here by a direct loop or fixed by hand, so the run pins
 * behaviour as well as memory safety: element types 0-9, empty dimensions, scalar outputs,
 * repeated input labels (diagonals), repeated output labels (embedding), three operands,
 * rank 12, the sx / sy scaling of loop_einsum! and the argument checks.  Exit status 0 and
 * "ok" on success.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; int64_t i; } oscalar;
int oracle_loop_einsum(int dt, int nin, const void* const* xs, const int32_t* nix,
                       const int32_t* ixs, void* y, int niy, const int32_t* iy, int nlabels,
                       const int64_t* sizes, const oscalar* sx, const oscalar* sy);
int oracle_dw_op(int op, int64_t n, const double* x, const double* y, double* z);

enum { F32 = 0, F64 = 1, C32 = 2, C64 = 3, I64 = 4, TF32 = 5, TF64 = 6, DD = 7, CT64 = 8,
       TBOOL = 9 };

static int failures = 0, cases = 0;
#define CHECK(cond, what)                                  \
  do {                                                     \
    ++cases;                                               \
    if (!(cond)) {                                         \
      ++failures;                                          \
      fprintf(stderr, "FAIL %s (line %d)\n", what, __LINE__); \
    }                                                      \
  } while (0)

static const oscalar ONE = {1, 0, 1}, ZERO = {0, 0, 0};

/* y[i,k] = sum_j a[i,j] b[j,k] (column-major), labels i=0 j=1 k=2 */
static int matmul(int dt, const void* a, const void* b, void* y, int64_t m, int64_t n,
                  int64_t p, const oscalar* sx, const oscalar* sy) {
  const void* xs[2] = {a, b};
  const int32_t nix[2] = {2, 2}, ixs[4] = {0, 1, 1, 2}, iy[2] = {0, 2};
  const int64_t sizes[3] = {m, n, p};
  return oracle_loop_einsum(dt, 2, xs, nix, ixs, y, 2, iy, 3, sizes, sx, sy);
}

static void case_f64_matmul_and_scaling(void) {
  enum { M = 3, N = 4, P = 5 };
  double a[M * N], b[N * P], y[M * P], ref[M * P], y0[M * P];
  for (int t = 0; t < M * N; ++t) a[t] = (t % 7) - 3;
  for (int t = 0; t < N * P; ++t) b[t] = (t % 5) - 2;
  for (int t = 0; t < M * P; ++t) y0[t] = t + 1;
  for (int i = 0; i < M; ++i)
    for (int k = 0; k < P; ++k) {
      double s = 0;
      for (int j = 0; j < N; ++j) s += a[i + M * j] * b[j + N * k];
      ref[i + M * k] = s;
    }
  CHECK(matmul(F64, a, b, y, M, N, P, &ONE, &ZERO) == 0, "f64 matmul returns 0");
  CHECK(memcmp(y, ref, sizeof y) == 0, "f64 matmul equals the direct loop");
  /* y <- 2 y + 3 a b */
  const oscalar sx = {3, 0, 3}, sy = {2, 0, 2};
  memcpy(y, y0, sizeof y);
  matmul(F64, a, b, y, M, N, P, &sx, &sy);
  int ok = 1;
  for (int t = 0; t < M * P; ++t) ok &= y[t] == 2 * y0[t] + 3 * ref[t];
  CHECK(ok, "f64 sx/sy scaling");
}

static void case_c64_and_c32_matmul(void) {
  enum { M = 2, N = 3, P = 2 };
  double a[2 * M * N], b[2 * N * P], y[2 * M * P];
  for (int t = 0; t < M * N; ++t) { a[2 * t] = t + 1; a[2 * t + 1] = -t; }
  for (int t = 0; t < N * P; ++t) { b[2 * t] = 2 - t; b[2 * t + 1] = t % 2; }
  matmul(C64, a, b, y, M, N, P, &ONE, &ZERO);
  int ok = 1;
  for (int i = 0; i < M; ++i)
    for (int k = 0; k < P; ++k) {
      double re = 0, im = 0;
      for (int j = 0; j < N; ++j) {
        const double ar = a[2 * (i + M * j)], ai = a[2 * (i + M * j) + 1];
        const double br = b[2 * (j + N * k)], bi = b[2 * (j + N * k) + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
      }
      ok &= y[2 * (i + M * k)] == re && y[2 * (i + M * k) + 1] == im;
    }
  CHECK(ok, "c64 matmul equals the direct loop");
  float af[2 * M * N], bf[2 * N * P], yf[2 * M * P];
  for (int t = 0; t < 2 * M * N; ++t) af[t] = (float)a[t];
  for (int t = 0; t < 2 * N * P; ++t) bf[t] = (float)b[t];
  matmul(C32, af, bf, yf, M, N, P, &ONE, &ZERO);
  ok = 1;
  for (int t = 0; t < 2 * M * P; ++t) ok &= yf[t] == (float)y[t];  /* small integers: exact */
  CHECK(ok, "c32 matmul equals c64 on small integers");
}

static void case_f32_trace_and_sum(void) {
  float x[16];
  for (int t = 0; t < 16; ++t) x[t] = (float)t;
  const void* xs[1] = {x};
  const int32_t nix[1] = {2}, ixs[2] = {0, 0};
  const int64_t sizes[1] = {4};
  float y = -1;
  oracle_loop_einsum(F32, 1, xs, nix, ixs, &y, 0, NULL, 1, sizes, &ONE, &ZERO);
  CHECK(y == 0 + 5 + 10 + 15, "f32 trace ii->");
  const int32_t nix1[1] = {1}, ixs1[1] = {0};
  const int64_t sizes1[1] = {16};
  oracle_loop_einsum(F32, 1, xs, nix1, ixs1, &y, 0, NULL, 1, sizes1, &ONE, &ZERO);
  CHECK(y == 120, "f32 sum i->");
}

static void case_i64_three_operands(void) {
  /* ij,jk,kl->il with sizes 2,3,2,2 */
  int64_t a[6], b[6], c[4], y[4], ref[4] = {0};
  for (int t = 0; t < 6; ++t) { a[t] = t - 2; b[t] = 3 - t; }
  for (int t = 0; t < 4; ++t) c[t] = t + 1;
  for (int i = 0; i < 2; ++i)
    for (int l = 0; l < 2; ++l)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 2; ++k) ref[i + 2 * l] += a[i + 2 * j] * b[j + 3 * k] * c[k + 2 * l];
  const void* xs[3] = {a, b, c};
  const int32_t nix[3] = {2, 2, 2}, ixs[6] = {0, 1, 1, 2, 2, 3}, iy[2] = {0, 3};
  const int64_t sizes[4] = {2, 3, 2, 2};
  oracle_loop_einsum(I64, 3, xs, nix, ixs, y, 2, iy, 4, sizes, &ONE, &ZERO);
  CHECK(memcmp(y, ref, sizeof y) == 0, "i64 three operands");
}

static void case_tropical_and_counting(void) {
  /* max-plus: y[i,k] = max_j a[i,j] + b[j,k] */
  double a[4] = {0, 1, 2, -1}, b[4] = {3, 0, -2, 5}, y[4];
  matmul(TF64, a, b, y, 2, 2, 2, &(oscalar){0, 0, 0}, &(oscalar){-INFINITY, 0, 0});
  int ok = 1;
  for (int i = 0; i < 2; ++i)
    for (int k = 0; k < 2; ++k) {
      double m = -INFINITY;
      for (int j = 0; j < 2; ++j) m = fmax(m, a[i + 2 * j] + b[j + 2 * k]);
      ok &= y[i + 2 * k] == m;
    }
  CHECK(ok, "tropical f64 matmul");
  float af[4] = {0, 1, 2, -1}, bf[4] = {3, 0, -2, 5}, yf[4];
  matmul(TF32, af, bf, yf, 2, 2, 2, &(oscalar){0, 0, 0}, &(oscalar){-INFINITY, 0, 0});
  ok = 1;
  for (int t = 0; t < 4; ++t) ok &= yf[t] == (float)y[t];
  CHECK(ok, "tropical f32 equals f64");
  /* counting tropical: i,i-> over equal values counts the ties: (n, c) pairs */
  double x[6] = {1, 1, 2, 1, 2, 1}, ones[6] = {0, 1, 0, 1, 0, 1}, r[2];
  const void* xs[2] = {x, ones};
  const int32_t nix[2] = {1, 1}, ixs[2] = {0, 0};
  const int64_t sizes[1] = {3};
  oracle_loop_einsum(CT64, 2, xs, nix, ixs, r, 0, NULL, 1, sizes, &(oscalar){0, 1, 0},
                     &(oscalar){-INFINITY, 0, 0});
  CHECK(r[0] == 2 && r[1] == 2, "counting tropical keeps the max and counts its ties");
}

static void case_bool_and_or(void) {
  uint8_t a[4] = {1, 0, 0, 1}, b[4] = {0, 1, 1, 0}, y[4];
  matmul(TBOOL, a, b, y, 2, 2, 2, &ONE, &ZERO);
  /* a is the identity pattern, so y = b */
  CHECK(memcmp(y, b, 4) == 0, "and-or matmul with the identity");
}

static void case_double_word(void) {
  /* hi = 1, lo = 2^-60 summed with itself: exactly 2 + 2^-59 as a double-word */
  double x[4] = {1, ldexp(1, -60), 1, ldexp(1, -60)}, y[2];
  const void* xs[1] = {x};
  const int32_t nix[1] = {1}, ixs[1] = {0};
  const int64_t sizes[1] = {2};
  oracle_loop_einsum(DD, 1, xs, nix, ixs, y, 0, NULL, 1, sizes, &ONE, &ZERO);
  CHECK(y[0] == 2 && y[1] == ldexp(1, -59), "double-word sum keeps the low word");
  double p[2], q[2] = {3, 0}, z[2];
  p[0] = 1;
  p[1] = ldexp(1, -60);
  oracle_dw_op(1, 1, p, q, z);
  CHECK(z[0] == 3 && z[1] == 3 * ldexp(1, -60), "double-word product");
}

static void case_empty_and_embedding(void) {
  /* ij,jk->ik with j = 0: y = sx * 0 + sy * y; sy = 0 clears y */
  double a[1] = {7}, b[1] = {7}, y[6];
  for (int t = 0; t < 6; ++t) y[t] = 5;
  CHECK(matmul(F64, a, b, y, 2, 0, 3, &ONE, &ZERO) == 0, "empty contraction returns 0");
  int ok = 1;
  for (int t = 0; t < 6; ++t) ok &= y[t] == 0;
  CHECK(ok, "empty contraction with sy = 0 gives zeros");
  /* i->ii: the diagonal embedding */
  double x[3] = {1, 2, 3}, d[9];
  const void* xs[1] = {x};
  const int32_t nix[1] = {1}, ixs[1] = {0}, iy[2] = {0, 0};
  const int64_t sizes[1] = {3};
  oracle_loop_einsum(F64, 1, xs, nix, ixs, d, 2, iy, 1, sizes, &ONE, &ZERO);
  ok = 1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ok &= d[i + 3 * j] == (i == j ? x[i] : 0);
  CHECK(ok, "i->ii embeds the diagonal");
}

static void case_rank12_permutation(void) {
  enum { R = 12 };
  int64_t sizes[R];
  int32_t ix[R], iy[R];
  int64_t n = 1;
  for (int q = 0; q < R; ++q) { sizes[q] = 2; ix[q] = q; iy[q] = R - 1 - q; n *= 2; }
  double* x = malloc(sizeof(double) * n);
  double* y = malloc(sizeof(double) * n);
  for (int64_t t = 0; t < n; ++t) x[t] = (double)t;
  const void* xs[1] = {x};
  const int32_t nix[1] = {R};
  oracle_loop_einsum(F64, 1, xs, nix, ix, y, R, iy, R, sizes, &ONE, &ZERO);
  /* reversing the labels reverses the bits of the linear index */
  int ok = 1;
  for (int64_t t = 0; t < n; ++t) {
    int64_t rev = 0;
    for (int q = 0; q < R; ++q) rev |= ((t >> q) & 1) << (R - 1 - q);
    ok &= y[rev] == x[t];
  }
  CHECK(ok, "rank-12 label reversal");
  free(x);
  free(y);
}

static void case_argument_checks(void) {
  double a[1] = {1}, y[1];
  const void* xs[1] = {a};
  const int32_t nix[1] = {1}, bad[1] = {5}, iy[1] = {0};
  const int64_t sizes[1] = {1};
  CHECK(oracle_loop_einsum(F64, 0, xs, nix, bad, y, 1, iy, 1, sizes, &ONE, &ZERO) == 1,
        "no operands is rejected");
  CHECK(oracle_loop_einsum(F64, 1, xs, nix, bad, y, 1, iy, 1, sizes, &ONE, &ZERO) == 1,
        "an input label out of range is rejected");
  const int32_t badout[1] = {-1}, ok_ix[1] = {0};
  CHECK(oracle_loop_einsum(F64, 1, xs, nix, ok_ix, y, 1, badout, 1, sizes, &ONE, &ZERO) == 1,
        "an output label out of range is rejected");
}

int main(void) {
  case_f64_matmul_and_scaling();
  case_c64_and_c32_matmul();
  case_f32_trace_and_sum();
  case_i64_three_operands();
  case_tropical_and_counting();
  case_bool_and_or();
  case_double_word();
  case_empty_and_embedding();
  case_rank12_permutation();
  case_argument_checks();
  if (failures) {
    fprintf(stderr, "%d of %d checks failed\n", failures, cases);
    return 1;
  }
  printf("ok %d checks\n", cases);
  return 0;
}
