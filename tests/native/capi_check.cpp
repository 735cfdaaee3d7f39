// Host-sanitizer check of the library's host code (planner, tree executor, grouped launches,
// request loop, memory pool): this program and the library sources are built with
// AddressSanitizer and UndefinedBehaviorSanitizer on the host side only (tests/native/Makefile)
// and drive the C ABI on the GPU.  Every result is checked against a direct loop on the host;
// a sanitizer finding aborts the run.  Prints "ok <checks>" and exits 0 on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/omeinsum_mi355x.h"

static int g_checks = 0, g_fail = 0;
#define CHECK(cond, what)                                                  \
  do {                                                                     \
    ++g_checks;                                                            \
    if (!(cond)) {                                                         \
      ++g_fail;                                                            \
      std::fprintf(stderr, "FAIL %s (line %d)\n", what, __LINE__);         \
    }                                                                      \
  } while (0)

static ome_ctx* g_ctx = nullptr;
static const ome_scalar ONE = {1, 0, 1}, ZERO = {0, 0, 0};

template <class T>
static void* up(const std::vector<T>& h) {
  void* d = nullptr;
  if (ome_alloc(g_ctx, h.size() * sizeof(T) + 16, &d) != OME_OK) return nullptr;
  ome_memcpy_h2d(g_ctx, d, h.data(), h.size() * sizeof(T));
  return d;
}
template <class T>
static std::vector<T> down(const void* d, size_t n) {
  std::vector<T> h(n);
  ome_synchronize(g_ctx);
  ome_memcpy_d2h(g_ctx, h.data(), d, n * sizeof(T));
  return h;
}
static double val(int64_t s, int64_t t) { return (double)((s * 7919 + t * 104729) % 17) - 8.0; }

// C[i,k] = sum_j A[i,j] B[j,k], column-major, in double
static std::vector<double> host_mm(const std::vector<double>& a, const std::vector<double>& b,
                                   int64_t m, int64_t n, int64_t p) {
  std::vector<double> c(m * p, 0.0);
  for (int64_t k = 0; k < p; ++k)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < m; ++i) c[i + m * k] += a[i + m * j] * b[j + n * k];
  return c;
}

static void binary_and_unary() {
  const int64_t M = 37, N = 29, P = 23;
  std::vector<double> a(M * N), b(N * P);
  for (int64_t t = 0; t < M * N; ++t) a[t] = val(1, t);
  for (int64_t t = 0; t < N * P; ++t) b[t] = val(2, t);
  const std::vector<double> ref = host_mm(a, b, M, N, P);
  void *da = up(a), *db = up(b), *dy = up(std::vector<double>(M * P));
  const void* xs[2] = {da, db};
  const int32_t nix[2] = {2, 2}, ixs[4] = {0, 1, 1, 2}, iy[2] = {0, 2};
  const int64_t sizes[3] = {M, N, P};
  CHECK(ome_einsum(g_ctx, OME_F64, 2, xs, nix, ixs, dy, 2, iy, 3, sizes, &ONE, &ZERO) == OME_OK,
        "f64 ij,jk->ik");
  CHECK(down<double>(dy, M * P) == ref, "f64 matmul equals the host loop (integer values)");
  // Float32 at BF16x3 precision: norm-wise error against the host loop
  std::vector<float> af(a.begin(), a.end()), bf(b.begin(), b.end());
  void *fa = up(af), *fb = up(bf), *fy = up(std::vector<float>(M * P)), *fz = up(std::vector<float>(M * P));
  ome_set_option(g_ctx, OME_OPT_PRECISION, OME_PREC_BF16X3);
  const void* fxs[2] = {fa, fb};
  CHECK(ome_einsum(g_ctx, OME_F32, 2, fxs, nix, ixs, fy, 2, iy, 3, sizes, &ONE, &ZERO) == OME_OK,
        "f32 bf16x3 ij,jk->ik");
  CHECK(ome_einsum_noalloc(g_ctx, OME_F32, 2, fxs, nix, ixs, fz, 2, iy, 3, sizes, &ONE, &ZERO) == OME_OK,
        "f32 bf16x3 without allocation");
  ome_set_option(g_ctx, OME_OPT_PRECISION, OME_PREC_FP32);
  const std::vector<float> gy = down<float>(fy, M * P), gz = down<float>(fz, M * P);
  double num = 0, den = 0;
  for (int64_t t = 0; t < M * P; ++t) {
    num += (gy[t] - ref[t]) * (gy[t] - ref[t]);
    den += ref[t] * ref[t];
  }
  CHECK(std::sqrt(num / den) <= 1e-5, "bf16x3 within 1e-5 norm-wise");
  CHECK(gy == gz, "the allocation-free entry gives the same bits");
  // unary: ii->i (diagonal) and ij->j (sum), on a square operand
  std::vector<double> sq(N * N);
  for (int64_t t = 0; t < N * N; ++t) sq[t] = val(3, t);
  void *dsq = up(sq), *dd = up(std::vector<double>(N));
  const void* ux[1] = {dsq};
  const int32_t unix_[1] = {2}, uix[2] = {0, 0}, uiy[1] = {0}, six[2] = {0, 1}, siy[1] = {1};
  const int64_t usz[2] = {N, N};
  ome_einsum(g_ctx, OME_F64, 1, ux, unix_, uix, dd, 1, uiy, 1, usz, &ONE, &ZERO);
  std::vector<double> gd = down<double>(dd, N);
  bool ok = true;
  for (int64_t i = 0; i < N; ++i) ok = ok && gd[i] == sq[i + N * i];
  CHECK(ok, "ii->i");
  ome_einsum(g_ctx, OME_F64, 1, ux, unix_, six, dd, 1, siy, 2, usz, &ONE, &ZERO);
  gd = down<double>(dd, N);
  ok = true;
  for (int64_t j = 0; j < N; ++j) {
    double s = 0;
    for (int64_t i = 0; i < N; ++i) s += sq[i + N * j];
    ok = ok && gd[j] == s;
  }
  CHECK(ok, "ij->j");
  // a size mismatch between an operand and the size table is reported, not run
  const int64_t bad[3] = {M, N, -1};
  CHECK(ome_einsum(g_ctx, OME_F64, 2, xs, nix, ixs, dy, 2, iy, 3, bad, &ONE, &ZERO) != OME_OK,
        "negative size rejected");
  CHECK(ome_last_error(g_ctx) && std::strlen(ome_last_error(g_ctx)) > 0, "error message set");
  for (void* p : {da, db, dy, fa, fb, fy, fz, dsq, dd}) ome_free(g_ctx, p);
}

static void nary_and_permute() {
  // ij,jk,kl->il in Int64 (exact)
  const int64_t I = 5, J = 6, K = 4, L = 3;
  std::vector<int64_t> a(I * J), b(J * K), c(K * L), ref(I * L, 0);
  for (int64_t t = 0; t < I * J; ++t) a[t] = t % 5 - 2;
  for (int64_t t = 0; t < J * K; ++t) b[t] = 3 - t % 7;
  for (int64_t t = 0; t < K * L; ++t) c[t] = t % 3 + 1;
  for (int64_t i = 0; i < I; ++i)
    for (int64_t l = 0; l < L; ++l)
      for (int64_t j = 0; j < J; ++j)
        for (int64_t k = 0; k < K; ++k) ref[i + I * l] += a[i + I * j] * b[j + J * k] * c[k + K * l];
  void *da = up(a), *db = up(b), *dc = up(c), *dy = up(std::vector<int64_t>(I * L));
  const void* xs[3] = {da, db, dc};
  const int32_t nix[3] = {2, 2, 2}, ixs[6] = {0, 1, 1, 2, 2, 3}, iy[2] = {0, 3};
  const int64_t sizes[4] = {I, J, K, L};
  CHECK(ome_einsum(g_ctx, OME_I64, 3, xs, nix, ixs, dy, 2, iy, 4, sizes, &ONE, &ZERO) == OME_OK,
        "i64 three operands");
  CHECK(down<int64_t>(dy, I * L) == ref, "i64 three operands exact");
  // a 4-D permute, c dim p = a dim perm[p]
  const int64_t dims[4] = {3, 4, 5, 6};
  const int32_t perm[4] = {2, 0, 3, 1};
  std::vector<int64_t> x(360);
  for (int64_t t = 0; t < 360; ++t) x[t] = t * 3 - 7;
  void *dx = up(x), *dp = up(std::vector<int64_t>(360));
  CHECK(ome_permute(g_ctx, OME_I64, 4, dims, perm, dx, dp, &ONE, &ZERO) == OME_OK, "permute");
  const std::vector<int64_t> gp = down<int64_t>(dp, 360);
  bool ok = true;
  for (int64_t i0 = 0; i0 < 3; ++i0)
    for (int64_t i1 = 0; i1 < 4; ++i1)
      for (int64_t i2 = 0; i2 < 5; ++i2)
        for (int64_t i3 = 0; i3 < 6; ++i3) {
          const int64_t src = i0 + 3 * (i1 + 4 * (i2 + 5 * i3));
          const int64_t dst = i2 + 5 * (i0 + 3 * (i3 + 6 * i1));  // c dims (5, 3, 6, 4)
          ok = ok && gp[dst] == x[src];
        }
  CHECK(ok, "permute exact");
  for (void* p : {da, db, dc, dy, dx, dp}) ome_free(g_ctx, p);
}

// The chain ((A B) C) over labels i j k l as a NestedEinsum; slice: optional sliced label.
static ome_tree* chain_tree(const int64_t* sizes, int nslice, const int32_t* slice) {
  const std::vector<int32_t> desc = {3, 2,
                                     2, 0, 2, 0, 1, 1, 2, 1, 2, 2, 0, 2,
                                     2, -1, 2, 0, 2, 2, 2, 2, 3, 2, 0, 3};
  ome_tree* t = nullptr;
  if (ome_tree_create(g_ctx, OME_F64, desc.data(), (int64_t)desc.size(), 4, sizes, nslice, slice,
                      &t) != OME_OK)
    return nullptr;
  return t;
}

static void trees_and_requests() {
  const int64_t I = 24, J = 40, K = 33, L = 17;
  const int64_t sizes[4] = {I, J, K, L};
  std::vector<double> a(I * J), b(J * K), c(K * L);
  for (int64_t t = 0; t < I * J; ++t) a[t] = val(4, t);
  for (int64_t t = 0; t < J * K; ++t) b[t] = val(5, t);
  for (int64_t t = 0; t < K * L; ++t) c[t] = val(6, t);
  const std::vector<double> ref = host_mm(host_mm(a, b, I, J, K), c, I, K, L);
  void *da = up(a), *db = up(b), *dc = up(c), *dy = up(std::vector<double>(I * L));
  const void* xs[3] = {da, db, dc};
  ome_tree* t = chain_tree(sizes, 0, nullptr);
  CHECK(t != nullptr, "tree created");
  for (int rep = 0; rep < 2; ++rep) {  // the second evaluation replays the recorded graph
    CHECK(ome_tree_execute(t, xs, dy, nullptr, nullptr) == OME_OK, "tree executes");
    CHECK(down<double>(dy, I * L) == ref, "tree equals the host loop");
  }
  const int32_t sl[1] = {1};
  ome_tree* ts = chain_tree(sizes, 1, sl);
  CHECK(ts != nullptr, "sliced tree created");
  CHECK(ome_tree_execute(ts, xs, dy, nullptr, nullptr) == OME_OK, "sliced tree executes");
  CHECK(down<double>(dy, I * L) == ref, "sliced tree equals the host loop");
  ome_tree_destroy(ts);
  ome_tree_destroy(t);
  // the request loop: four plans, the third request writes status 9 into its slot's word
  ome_set_option(g_ctx, OME_OPT_TREE_DEFER_MEMORY, 1);
  std::vector<ome_tree*> plans;
  for (int q = 0; q < 4; ++q) plans.push_back(chain_tree(sizes, 0, nullptr));
  ome_set_option(g_ctx, OME_OPT_TREE_DEFER_MEMORY, 0);
  ome_server* srv = nullptr;
  CHECK(ome_server_create(g_ctx, (size_t)1 << 30, 2, &srv) == OME_OK, "server created");
  std::vector<void*> ys;
  std::vector<int64_t> tickets;
  for (int q = 0; q < 4; ++q) {
    ys.push_back(up(std::vector<double>(I * L)));
    if (q == 2) ome_server_inject_status(srv, 9);
    int64_t tk = 0;
    CHECK(ome_server_submit(srv, plans[q], xs, ys.back(), &tk) == OME_OK, "submitted");
    tickets.push_back(tk);
  }
  for (int q = 0; q < 4; ++q) {
    const int rc = ome_server_wait(srv, tickets[q], nullptr);
    if (q == 2) {
      CHECK(rc == OME_ERR_DEVICE, "the injected request fails with a device error");
      CHECK(std::string(ome_last_error(g_ctx)).find("status 9") != std::string::npos,
            "its message names the status");
    } else {
      CHECK(rc == OME_OK, "the other requests complete");
      CHECK(down<double>(ys[q], I * L) == ref, "and equal the host loop");
    }
  }
  ome_server_stats st{};
  ome_server_get_stats(srv, &st);
  CHECK(st.submitted == 4 && st.completed == 3, "loop statistics");
  ome_server_destroy(srv);
  for (auto* p : plans) ome_tree_destroy(p);
  for (void* p : ys) ome_free(g_ctx, p);
  for (void* p : {da, db, dc, dy}) ome_free(g_ctx, p);
}

static void grouped() {
  // 24 Float32 problems of uneven shapes (FP32 precision) and 12 permutes
  const int G = 24;
  std::vector<int64_t> dims;
  std::vector<std::vector<float>> A, B;
  std::vector<void*> da, db, dc;
  for (int g = 0; g < G; ++g) {
    const int64_t m = 1 + (g * 37) % 90, n = 1 + (g * 53) % 70, k = 1 + (g * 29) % 130,
                  bt = 1 + g % 3;
    dims.insert(dims.end(), {m, n, k, bt});
    A.emplace_back(m * k * bt);
    B.emplace_back(k * n * bt);
    for (size_t t = 0; t < A.back().size(); ++t) A.back()[t] = (float)val(7 + g, (int64_t)t);
    for (size_t t = 0; t < B.back().size(); ++t) B.back()[t] = (float)val(8 + g, (int64_t)t);
    da.push_back(up(A.back()));
    db.push_back(up(B.back()));
    dc.push_back(up(std::vector<float>(m * n * bt)));
  }
  CHECK(ome_grouped_gemm(g_ctx, OME_F32, G, dims.data(), da.data(), db.data(), dc.data(), &ONE,
                         &ZERO) == OME_OK,
        "grouped GEMM");
  bool ok = true;
  for (int g = 0; g < G; ++g) {
    const int64_t m = dims[4 * g], n = dims[4 * g + 1], k = dims[4 * g + 2], bt = dims[4 * g + 3];
    const std::vector<float> got = down<float>(dc[g], m * n * bt);
    for (int64_t l = 0; l < bt; ++l)
      for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
          double s = 0;
          for (int64_t q = 0; q < k; ++q)
            s += (double)A[g][i + m * q + m * k * l] * B[g][q + k * j + k * n * l];
          ok = ok && std::fabs(got[i + m * j + m * n * l] - s) <= 1e-5 * (std::fabs(s) + 8 * k);
        }
  }
  CHECK(ok, "grouped GEMM equals the host loop");
  const int P = 12;
  std::vector<int32_t> nd, perms;
  std::vector<int64_t> pd;
  std::vector<std::vector<float>> X;
  std::vector<void*> dx, dy;
  for (int p = 0; p < P; ++p) {
    const int64_t d0 = 2 + p % 5, d1 = 3 + p % 4, d2 = 1 + p % 6;
    nd.push_back(3);
    pd.insert(pd.end(), {d0, d1, d2});
    perms.insert(perms.end(), {2, 0, 1});
    X.emplace_back(d0 * d1 * d2);
    for (size_t t = 0; t < X.back().size(); ++t) X.back()[t] = (float)t + 0.5f * p;
    dx.push_back(up(X.back()));
    dy.push_back(up(std::vector<float>(X.back().size())));
  }
  CHECK(ome_grouped_permute(g_ctx, OME_F32, P, nd.data(), pd.data(), perms.data(), dx.data(),
                            dy.data(), &ONE, &ZERO) == OME_OK,
        "grouped permute");
  ok = true;
  for (int p = 0; p < P; ++p) {
    const int64_t d0 = pd[3 * p], d1 = pd[3 * p + 1], d2 = pd[3 * p + 2];
    const std::vector<float> got = down<float>(dy[p], d0 * d1 * d2);
    for (int64_t i0 = 0; i0 < d0; ++i0)
      for (int64_t i1 = 0; i1 < d1; ++i1)
        for (int64_t i2 = 0; i2 < d2; ++i2)  // c dims (d2, d0, d1)
          ok = ok && got[i2 + d2 * (i0 + d0 * i1)] == X[p][i0 + d0 * (i1 + d1 * i2)];
  }
  CHECK(ok, "grouped permute exact");
  for (auto* v : {&da, &db, &dc, &dx, &dy})
    for (void* p : *v) ome_free(g_ctx, p);
}

int main() {
  if (ome_init(0, 0, nullptr, &g_ctx) != OME_OK) {
    std::fprintf(stderr, "ome_init failed: %s\n", ome_last_error(nullptr));
    return 2;
  }
  binary_and_unary();
  nary_and_permute();
  trees_and_requests();
  grouped();
  ome_shutdown(g_ctx);
  if (g_fail) {
    std::fprintf(stderr, "%d of %d checks failed\n", g_fail, g_checks);
    return 1;
  }
  std::printf("ok %d checks\n", g_checks);
  return 0;
}
