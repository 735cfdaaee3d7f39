// The einsum planner: turns one flat EinCode call into kernel launches.
//
//   arity 1  -> identity scale-copy | strided permute (tensorpermute!) | one index-walk pass
//               that fuses Diag, Sum, Permutedims, Repeat and Duplicate
//               (OMEinsum runs these as a pipeline of separate passes, src/einsum.jl:52-101)
//   arity 2  -> label groups of analyze_binary (src/einsum.jl:139-211); each operand is read
//               in place by a strided GEMM when its groups are contiguous runs, otherwise it is
//               permuted (or diag/summed) first, as simplifyto does (src/einsum.jl:127-134);
//               the result is written in place or through a post-pass (src/einsum.jl:116-119)
//   arity>=3 -> the index walk (loop_einsum!, src/einsum.jl:47-50)
#include <algorithm>
#include <cstdio>
#include <numeric>

#include "runtime.h"

namespace ome {

Scal scal_one(int dt) {
  Scal s{1.0, 0.0, 1};
  if (is_tropical(dt)) s.re = 0.0;
  if (dt == CT64) s = Scal{0.0, 1.0, 1};
  return s;
}
Scal scal_zero(int dt) {
  Scal s{0.0, 0.0, 0};
  if (is_tropical(dt)) s.re = -INFINITY;
  if (dt == CT64) s.re = -INFINITY;
  return s;
}

static int64_t prod_sizes(const std::vector<int>& ls, const std::vector<int64_t>& sz) {
  int64_t p = 1;
  for (int l : ls) p *= sz[l];
  return p;
}

static std::vector<int> uniq(const std::vector<int>& v) {
  std::vector<int> out;
  for (int l : v)
    if (std::find(out.begin(), out.end(), l) == out.end()) out.push_back(l);
  return out;
}
static bool contains(const std::vector<int>& v, int l) {
  return std::find(v.begin(), v.end(), l) != v.end();
}

static std::string fmt_labels(const std::vector<int>& v) {
  std::string s = "(";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + ")";
}

// ---- permute -----------------------------------------------------------------------------

// y[labels oy] <- sy*y + sx*conj?(x[labels ix]) for a permutation oy of the unique labels ix.
// With `out`, a permute that needs the permute kernel is not launched: its descriptor is
// written to *out (and *deferred set) for the caller to launch.
static Status do_permute(int dt, const void* x, const std::vector<int>& ix, void* y,
                         const std::vector<int>& oy, const std::vector<int64_t>& sizes,
                         const Scal& sx, const Scal& sy, int conj, hipStream_t st,
                         PermDesc* out = nullptr, bool* deferred = nullptr) {
  const int r = (int)ix.size();
  int64_t n = prod_sizes(ix, sizes);
  if (n == 0) return Status::ok();
  // input strides
  std::vector<int64_t> xstride(r);
  int64_t s = 1;
  for (int p = 0; p < r; ++p) { xstride[p] = s; s *= sizes[ix[p]]; }
  // output dims: (input position, extent), dropping extent-1 dims; then merge runs of
  // consecutive input positions (the grouping of tensorpermute!, src/utils.jl:132-144)
  std::vector<int> pos;
  for (int l : oy) {
    int p = (int)(std::find(ix.begin(), ix.end(), l) - ix.begin());
    if (sizes[l] != 1) pos.push_back(p);
  }
  std::vector<int64_t> odim, ostr;
  for (size_t i = 0; i < pos.size(); ++i) {
    int p = pos[i];
    if (!odim.empty() && i > 0) {
      // previous output dim ends at input position pos[i-1]; extend if contiguous in input
      int prev = pos[i - 1];
      bool consecutive = true;
      for (int q = prev + 1; q < p; ++q)
        if (sizes[ix[q]] != 1) consecutive = false;
      if (p > prev && consecutive) {
        odim.back() *= sizes[ix[p]];
        continue;
      }
    }
    odim.push_back(sizes[ix[p]]);
    ostr.push_back(xstride[p]);
  }
  if (odim.empty() || (odim.size() == 1 && ostr[0] == 1)) {
    launch_axpby(dt, x, y, n, sx, sy, conj, st);
    return Status::ok();
  }
  if ((int)odim.size() > MAXR)
    return Status::err(OME_ERR_UNSUPPORTED, "permute of rank " + std::to_string(odim.size()) +
                                                " after grouping exceeds " + std::to_string(MAXR));
  if (n >= (1ll << 32)) return Status::err(OME_ERR_UNSUPPORTED, "permute of >= 2^32 elements");
  PermDesc d{};
  d.x = x;
  d.y = y;
  d.n = n;
  d.rank = (int)odim.size();
  for (int p = 0; p < d.rank; ++p) {
    d.odiv[p] = make_fdiv((uint32_t)odim[p]);
    d.istride[p] = ostr[p];
  }
  d.sx = sx;
  d.sy = sy;
  OME_DISPATCH(dt, O, { d.mode = epi_mode<O>(sx, sy); });
  d.conj = conj;
  if (out) {
    *out = d;
    *deferred = true;
    return Status::ok();
  }
  launch_permute(dt, d, st);
  return Status::ok();
}

// Whether y[iy] <- x[ix] is a plain permutation (unique labels, the same set).
static bool is_permutation(const std::vector<int>& ix, const std::vector<int>& iy);

// ---- index walk ----------------------------------------------------------------------------

static Status do_walk(ome_ctx* ctx, int dt, const std::vector<const void*>& xs,
                      const std::vector<std::vector<int>>& ixs, void* y, const std::vector<int>& iy,
                      const std::vector<int64_t>& sizes, const Scal& sx, const Scal& sy,
                      hipStream_t st, std::vector<Temp>& keep) {
  const int nin = (int)xs.size();
  if (nin > WALK_MAXIN)
    return Status::err(OME_ERR_UNSUPPORTED, "index walk over more than 16 inputs");
  std::vector<int> outer = uniq(iy);
  std::vector<int> inner;
  for (auto& ix : ixs)
    for (int l : ix)
      if (!contains(outer, l) && !contains(inner, l)) inner.push_back(l);
  // extent-1 inner labels contribute a single point: drop them from the walk
  std::vector<int> inner_nz;
  for (int l : inner)
    if (sizes[l] != 1) inner_nz.push_back(l);
  inner = inner_nz;
  if ((int)iy.size() > WALK_MAXL || (int)outer.size() > WALK_MAXL || (int)inner.size() > WALK_MAXL ||
      (int)(outer.size() + inner.size()) * nin > WALK_MAXS)
    return Status::err(OME_ERR_UNSUPPORTED, "index walk: too many labels x inputs");
  const int64_t ny = prod_sizes(iy, sizes);
  const int64_t ninner = prod_sizes(inner, sizes);
  if (ny >= (1ll << 32) || ninner >= (1ll << 32))
    return Status::err(OME_ERR_UNSUPPORTED, "index walk over >= 2^32 points");
  WalkDesc d{};
  d.nin = nin;
  for (int k = 0; k < nin; ++k) {
    d.x[k] = xs[k];
    if (prod_sizes(ixs[k], sizes) >= (1ll << 31))
      return Status::err(OME_ERR_UNSUPPORTED, "index walk input with >= 2^31 elements");
  }
  d.y = y;
  d.ny_dims = (int)iy.size();
  d.nouter = (int)outer.size();
  d.ninner = (int)inner.size();
  d.ny = ny;
  d.ninner_pts = ninner;
  for (int p = 0; p < d.ny_dims; ++p) {
    d.ydiv[p] = make_fdiv((uint32_t)sizes[iy[p]]);
    d.yslot[p] = (int8_t)(std::find(outer.begin(), outer.end(), iy[p]) - outer.begin());
  }
  for (int j = 0; j < d.ninner; ++j) d.idiv[j] = make_fdiv((uint32_t)sizes[inner[j]]);
  auto label_stride = [&](int k, int l) {
    int64_t s = 1, acc = 0;
    for (int l2 : ixs[k]) {
      if (l2 == l) acc += s;
      s *= sizes[l2];
    }
    return acc;
  };
  for (int u = 0; u < d.nouter; ++u)
    for (int k = 0; k < nin; ++k) d.stride[u * nin + k] = (int32_t)label_stride(k, outer[u]);
  for (int j = 0; j < d.ninner; ++j)
    for (int k = 0; k < nin; ++k)
      d.stride[(d.nouter + j) * nin + k] = (int32_t)label_stride(k, inner[j]);
  d.sx = sx;
  d.sy = sy;
  OME_DISPATCH(dt, O, { d.mode = epi_mode<O>(sx, sy); });
  // split the inner space when y alone cannot fill the GPU
  int nsplit = 1;
  const int64_t want = 1 << 18;
  if (ny < want && ninner > 256) {
    int64_t s = (want + ny - 1) / ny;
    int64_t smax = ninner / 64;
    if (s > smax) s = smax;
    if (s > 4096) s = 4096;
    if (s > 1) nsplit = (int)s;
  }
  d.chunk = (ninner + nsplit - 1) / nsplit;
  nsplit = (int)((ninner + d.chunk - 1) / d.chunk);
  if (nsplit < 1) nsplit = 1;
  d.nsplit = nsplit;
  if (nsplit > 1) {
    keep.emplace_back(ctx, (size_t)nsplit * ny * elem_size(dt));
    if (!keep.back().p) return Status::err(OME_ERR_OOM, "device pool exhausted (index walk partials)");
    d.partial = keep.back().p;
  }
  launch_walk(dt, d, st);
  return Status::ok();
}

// ---- unary -----------------------------------------------------------------------------------

static bool is_permutation(const std::vector<int>& ix, const std::vector<int>& iy) {
  std::vector<int> ux = uniq(ix), uy = uniq(iy);
  bool perm = ux.size() == ix.size() && uy.size() == iy.size() && ux.size() == uy.size();
  if (perm)
    for (int l : iy)
      if (!contains(ix, l)) perm = false;
  return perm;
}

static Status do_unary(ome_ctx* ctx, int dt, const void* x, const std::vector<int>& ix, void* y,
                       const std::vector<int>& iy, const std::vector<int64_t>& sizes,
                       const Scal& sx, const Scal& sy, int conj, hipStream_t st,
                       std::vector<Temp>& keep) {
  if (ix == iy) {  // empty pipeline: @flatten_addmul! sy*y + sx*x (src/einsum.jl:97-99)
    launch_axpby(dt, x, y, prod_sizes(iy, sizes), sx, sy, conj, st);
    return Status::ok();
  }
  if (is_permutation(ix, iy)) return do_permute(dt, x, ix, y, iy, sizes, sx, sy, conj, st);
  const void* src = x;
  if (conj) {
    keep.emplace_back(ctx, (size_t)prod_sizes(ix, sizes) * elem_size(dt));
    if (!keep.back().p) return Status::err(OME_ERR_OOM, "device pool exhausted");
    launch_axpby(dt, x, keep.back().p, prod_sizes(ix, sizes), scal_one(dt), scal_zero(dt), 1, st);
    src = keep.back().p;
  }
  return do_walk(ctx, dt, {src}, {ix}, y, iy, sizes, sx, sy, st, keep);
}

// ---- GEMM dispatch -------------------------------------------------------------------------

// BF16x3 on packed operands: Float32 and ComplexF32 products large enough that the one-pass
// pack (split + tile-major relayout) is amortised by the GEMM.
static Status run_gemm_packed(ome_ctx* ctx, int dt, const GemmDesc& d0, hipStream_t st, int np) {
  const bool cplx = dt == C32;
  // The kernel's lanes run along m in its stores (32 consecutive m per half-wave).  When y is
  // contiguous along n instead, compute y^T = B^T A^T (swap the operands' roles) so that every
  // store instruction writes whole 128-byte lines rather than 64 scattered words.
  GemmDesc d = d0;
  if (!cplx && d.syn == 1 && d.sym != 1) {
    std::swap(d.a, d.b);
    std::swap(d.M, d.N);
    std::swap(d.sam, d.sbn);
    std::swap(d.sak, d.sbk);
    std::swap(d.sab, d.sbb);
    std::swap(d.sym, d.syn);
  }
  const int64_t M = d.M, N = cplx ? 2 * d.N : d.N, K = cplx ? 2 * d.K : d.K;
  const int64_t tiles = packed_tiles_mn(M, N, np) * d.B;
  const int64_t ktiles = (K + 31) / 32;
  int ks = 1;
  while (tiles * ks < 256 && ktiles / (ks * 2) >= 8) ks *= 2;
  Temp pa(ctx, (size_t)packed_bytes_a(M, N, K, d.B, np));
  Temp pb(ctx, (size_t)packed_bytes_b(M, N, K, d.B, np));
  const bool via_work = ks > 1 || (cplx && !packed_plain_overwrite(d));
  Temp work(ctx, via_work ? (size_t)ks * d.B * M * N * 4 : 0);
  if (!pa.p || !pb.p || (via_work && !work.p))
    return Status::err(OME_ERR_OOM, "device pool exhausted (packed GEMM operands)");
  launch_gemm_packed(dt, d, pa.p, pb.p, (float*)work.p, ks, st, np);
  return Status::ok();
}

// Complex products as real GEMMs (4M / 3M) and Float32 widened to Float64: prepare dense
// real operands, run the real GEMM (which picks its own kernel), combine into y.
static Status run_gemm_via_real(ome_ctx* ctx, int dt, const GemmDesc& d, hipStream_t st,
                                int precision, int algo) {
  const bool rdouble = (dt == C64) || (precision == PREC_FP64);
  const RealShape rs = real_shape(algo, d);
  const size_t esz = rdouble ? 8 : 4;
  Temp ta(ctx, (size_t)rs.a_elems * esz), tb(ctx, (size_t)rs.b_elems * esz);
  Temp tc(ctx, (size_t)rs.c_elems * esz);
  if (!ta.p || !tb.p || !tc.p)
    return Status::err(OME_ERR_OOM, "device pool exhausted (real GEMM operands)");
  launch_prep_a(dt, rdouble, algo, d, ta.p, st);
  launch_prep_b(dt, rdouble, algo, d, tb.p, st);
  GemmDesc g{};
  g.a = ta.p;
  g.b = tb.p;
  g.y = tc.p;
  g.M = rs.M;
  g.N = rs.N;
  g.K = rs.K;
  g.B = rs.B;
  g.sam = 1; g.sak = rs.M; g.sab = rs.M * rs.K;
  g.sbk = 1; g.sbn = rs.K; g.sbb = rs.K * rs.N;
  g.sym = 1; g.syn = rs.M; g.syb = rs.M * rs.N;
  const int rdt = rdouble ? F64 : F32;
  g.sx = scal_one(rdt);
  g.sy = scal_zero(rdt);
  Status s = run_gemm(ctx, rdt, g, st, rdouble ? PREC_FP64 : precision);
  if (!s.good()) return s;
  launch_combine(dt, rdouble, algo, d, tc.p, st);
  return Status::ok();
}

Status run_gemm(ome_ctx* ctx, int dt, GemmDesc d, hipStream_t st, int precision) {
  if (d.M <= 0 || d.N <= 0 || d.B <= 0) return Status::ok();
  int kernel = ctx->opt.gemm_kernel;
  // few outputs over a long K (up to 2^26 products): one launch, latency-bound, summed in the
  // accumulation arithmetic (FP64 for Float32), so at least as precise as every path below
  if (kernel == GEMM_AUTO && panel_fits(dt, d)) {
    launch_gemm_panel(dt, d, st);
    return Status::ok();
  }
  if (precision == PREC_BF16X3 && (dt == F32 || dt == C32) &&
      (kernel == GEMM_AUTO || kernel == GEMM_MFMA)) {
    const double flops = 2.0 * d.M * d.N * d.K * d.B * (dt == C32 ? 4 : 1);
    if (d.M >= 64 && d.N >= 64 && d.K >= 64 && flops >= 3.0e7)
      return run_gemm_packed(ctx, dt, d, st, 2);
  }
  // six products on three planes: FP32-level error at a lower MFMA cost than FP32 operands.
  // The pack reads and writes every operand element once more (about 10 bytes per element),
  // which the cheaper products repay when M'N' / (M' + N') (output elements per operand
  // element of one k) is a few hundred; smaller products run FP32 MFMA.  Split-K makes up for
  // few 256 x 128 tiles over a long K.
  if (precision == PREC_BF16X6 && (dt == F32 || dt == C32) &&
      (kernel == GEMM_AUTO || kernel == GEMM_MFMA)) {
    const double rm = (double)d.M, rn = (double)(dt == C32 ? 2 * d.N : d.N);
    const double rk = (double)(dt == C32 ? 2 * d.K : d.K);
    const double flops = 2.0 * rm * rn * rk * d.B;
    if (std::max(rm, rn) >= 256 && std::min(rm, rn) >= 128 && rk >= 256 &&
        rm * rn / (rm + rn) >= 200 && flops >= 2.0e8)
      return run_gemm_packed(ctx, dt, d, st, 3);
  }
  if (kernel == GEMM_AUTO || kernel == GEMM_MFMA) {
    const bool fits = d.M * d.K * d.B < (1ll << 31) && d.K * d.N * d.B < (1ll << 31) &&
                      d.M * d.N * d.B < (1ll << 31) && d.M < (1ll << 31) && d.N < (1ll << 31) &&
                      d.K < (1ll << 31);
    // Float32 at FP64 precision: widened operands on the FP64 matrix cores
    if (dt == F32 && precision == PREC_FP64 && fits)
      return run_gemm_via_real(ctx, dt, d, st, precision, RA_WIDEN);
    // complex: real GEMMs on the matrix cores once the product is large enough to pay for
    // the operand preparation (otherwise the complex VALU kernels below)
    const bool cx = (dt == C32 || dt == C64);
    const int algo = ctx->opt.complex_algo ? RA_3M : RA_4M;
    const int64_t rm = d.M, rn = algo == RA_4M ? 2 * d.N : d.N, rk = algo == RA_4M ? 2 * d.K : d.K;
    const bool big = fits && rm >= 32 && rn >= 32 && rk >= 16 && d.M * d.N * d.K * d.B >= (1ll << 16);
    // ComplexF32 in the four-real-products form with unit-stride operands: one kernel builds
    // the real operands while staging them and writes the complex result itself
    if (dt == C32 && precision != PREC_FP64 && algo == RA_4M && big &&
        (d.sam == 1 || d.sak == 1 || d.M == 1 || d.K == 1) &&
        (d.sbk == 1 || d.sbn == 1 || d.N == 1 || d.K == 1)) {
      if (d.M == 1 && d.sak != 1) d.sam = 1;
      if (d.K == 1) { if (d.sam != 1) d.sak = 1; if (d.sbn != 1) d.sbk = 1; }
      if (d.N == 1 && d.sbk != 1) d.sbn = 1;
      const int ks = mfma_c32_plan_ksplit(d);
      Temp ws;
      if (ks > 1) {
        ws = Temp(ctx, (size_t)mfma_c32_workspace(d, ks));
        if (!ws.p) return Status::err(OME_ERR_OOM, "device pool exhausted (split-K workspace)");
        d.work = ws.p;
      }
      launch_gemm_mfma_c32(d, ks, st);
      return Status::ok();
    }
    if (cx && big)
      return run_gemm_via_real(ctx, dt, d, st, precision, algo);
    if (dt == C32 && precision == PREC_FP64 && fits)
      return run_gemm_via_real(ctx, dt, d, st, precision, algo);
  }
  const bool mfma_type = (dt == F32 || dt == F64);
  const bool a_unit = d.sam == 1 || d.sak == 1 || d.M == 1 || d.K == 1;
  const bool b_unit = d.sbk == 1 || d.sbn == 1 || d.N == 1 || d.K == 1;
  if (kernel == GEMM_AUTO) {
    if (mfma_type && a_unit && b_unit && d.M >= 32 && d.N >= 32 && d.K >= 16)
      kernel = GEMM_MFMA;
    else if (d.M >= 16 && d.N >= 16 && d.K >= 8)
      kernel = GEMM_VALU_TILE;
    else
      kernel = GEMM_VALU_DIRECT;
  }
  if (kernel == GEMM_MFMA && (!mfma_type || !a_unit || !b_unit)) kernel = GEMM_VALU_TILE;
  if (kernel == GEMM_MFMA) {
    // normalise unit strides so the kernel sees which dim is contiguous
    if (d.M == 1 && d.sak != 1) d.sam = 1;
    if (d.K == 1) { if (d.sam != 1) d.sak = 1; if (d.sbn != 1) d.sbk = 1; }
    if (d.N == 1 && d.sbk != 1) d.sbn = 1;
    // BF16x3 pays on the packed path (large products, above); smaller products run FP32 MFMA,
    // as fast for them and 8 bits more precise.  A forced MFMA kernel keeps the split for tests.
    const bool split = precision == PREC_BF16X3 && ctx->opt.gemm_kernel == GEMM_MFMA;
    const int prec = (dt == F64) ? PREC_FP64 : (split ? PREC_BF16X3 : PREC_FP32);
    int ks = mfma_plan_ksplit(d, 0);
    Temp ws;
    if (ks > 1) {
      ws = Temp(ctx, (size_t)mfma_workspace(dt, d, ks));
      if (!ws.p) return Status::err(OME_ERR_OOM, "device pool exhausted (split-K workspace)");
      d.work = ws.p;
    }
    launch_gemm_mfma(prec, d, ks, st);
    return Status::ok();
  }
  // few outputs and a long K: parallel over K instead of one thread (or one tile) per output
  if (d.M * d.N * d.B <= 1024 && d.K >= 2048 && ctx->opt.gemm_kernel == GEMM_AUTO) {
    const int parts = dot_plan_parts(d);
    Temp ws;
    if (parts > 1) {
      ws = Temp(ctx, (size_t)dot_workspace(dt, d, parts));
      if (!ws.p) return Status::err(OME_ERR_OOM, "device pool exhausted (split-K workspace)");
      d.work = ws.p;
    }
    launch_gemm_dot(dt, d, parts, st);
    return Status::ok();
  }
  int ks = 1;
  if (kernel == GEMM_VALU_TILE) {
    const int64_t tiles = ((d.M + 63) / 64) * ((d.N + 63) / 64) * d.B;
    while (tiles * ks < 512 && d.K / (ks * 2) >= 32) ks *= 2;
  }
  Temp ws;
  if (ks > 1) {
    ws = Temp(ctx, (size_t)gemm_valu_workspace(dt, d, ks));
    if (!ws.p) return Status::err(OME_ERR_OOM, "device pool exhausted (split-K workspace)");
    d.work = ws.p;
  }
  launch_gemm_valu(dt, d, kernel, ks, st);
  return Status::ok();
}

// ---- binary ----------------------------------------------------------------------------------

// Column-major stride of label l in a unique-label list, and whether `grp` is a contiguous
// run of `ls` in order (empty groups are always fine).
static bool group_run(const std::vector<int>& ls, const std::vector<int>& grp,
                      const std::vector<int64_t>& sizes, int64_t& stride, int& first_pos) {
  if (grp.empty()) { stride = 0; first_pos = -1; return true; }
  auto it = std::find(ls.begin(), ls.end(), grp[0]);
  if (it == ls.end()) return false;
  int p = (int)(it - ls.begin());
  if (p + grp.size() > ls.size()) return false;
  for (size_t i = 0; i < grp.size(); ++i)
    if (ls[p + i] != grp[i]) return false;
  int64_t s = 1;
  for (int q = 0; q < p; ++q) s *= sizes[ls[q]];
  stride = s;
  first_pos = p;
  return true;
}

struct BinLayout {
  std::vector<int> o1, in, o2, ba;  // group label orders
};

// Orders of a group's labels as they appear in ls (if all present), else empty candidate.
static std::vector<int> order_in(const std::vector<int>& grp, const std::vector<int>& ls) {
  std::vector<int> out;
  for (int l : ls)
    if (contains(grp, l) && !contains(out, l)) out.push_back(l);
  if (out.size() != grp.size()) return {};
  return out;
}

// a binary step goes through the strided kernel when its contracted space has at most this
// many points and output points x contracted points stays below the work bound
constexpr int64_t STRIDED_MAX_K = 64;
constexpr int64_t STRIDED_MAX_OUT = 1 << 16;  // larger outputs: tiled permutes + GEMM stream better
constexpr int64_t STRIDED_MAX_WORK = 1 << 20;

static Status do_binary(ome_ctx* ctx, const EinsumCall& c, const std::vector<int>& ix1,
                        const std::vector<int>& ix2, const std::vector<int>& iy, const void* x1,
                        const void* x2, hipStream_t st, std::vector<Temp>& keep) {
  const int dt = c.dt;
  const auto& sz = c.sizes;
  const Scal one = scal_one(dt), zero = scal_zero(dt);
  // label sets after the pre-passes (diag + dangling sums) and before the post-pass
  std::vector<int> u1 = uniq(ix1), u2 = uniq(ix2), uy = uniq(iy);
  std::vector<int> a_lab, b_lab;
  for (int l : u1)
    if (contains(u2, l) || contains(uy, l)) a_lab.push_back(l);
  for (int l : u2)
    if (contains(u1, l) || contains(uy, l)) b_lab.push_back(l);
  const bool a_pre = a_lab.size() != ix1.size();  // repeated or dangling labels
  const bool b_pre = b_lab.size() != ix2.size();
  std::vector<int> o1, in, o2, ba;
  for (int l : a_lab) {
    if (contains(b_lab, l)) (contains(uy, l) ? ba : in).push_back(l);
    else o1.push_back(l);
  }
  for (int l : b_lab)
    if (!contains(a_lab, l)) o2.push_back(l);
  std::vector<int> y_lab;  // labels of the GEMM output
  for (int l : uy)
    if (contains(a_lab, l) || contains(b_lab, l)) y_lab.push_back(l);
  const bool y_post = y_lab.size() != iy.size();  // repeated or broadcast output labels

  // candidate group orders (from A, B, Y order of appearance); search for the cheapest layout
  auto cands = [&](const std::vector<int>& g) {
    std::vector<std::vector<int>> out;
    for (auto* ls : {&ix1, &ix2, &iy}) {
      auto o = order_in(g, *ls);
      if (!o.empty() && std::find(out.begin(), out.end(), o) == out.end()) out.push_back(o);
    }
    if (out.empty()) out.push_back(g);
    return out;
  };
  auto co1 = cands(o1), cin = cands(in), co2 = cands(o2), cba = cands(ba);
  const int64_t na = prod_sizes(a_lab, sz), nb = prod_sizes(b_lab, sz), nyy = prod_sizes(y_lab, sz);
  const bool direct = ctx->opt.direct_gemm != 0;
  struct Choice {
    int64_t cost = INT64_MAX;
    BinLayout L;
    bool a_inplace = false, b_inplace = false, y_inplace = false;
  } best;
  for (auto& vo1 : co1)
    for (auto& vin : cin)
      for (auto& vo2 : co2)
        for (auto& vba : cba) {
          int64_t sm, sk, sn, sbt;
          int pm, pk, pn, pb;
          bool ai = false, bi = false, yi = false;
          if (direct && !a_pre && group_run(ix1, vo1, sz, sm, pm) && group_run(ix1, vin, sz, sk, pk) &&
              group_run(ix1, vba, sz, sbt, pb)) {
            // A needs m or k to be the fastest group
            ai = (pm == 0) || (pk == 0) || o1.empty() || in.empty();
          }
          if (direct && !b_pre && group_run(ix2, vin, sz, sk, pk) && group_run(ix2, vo2, sz, sn, pn) &&
              group_run(ix2, vba, sz, sbt, pb)) {
            bi = (pk == 0) || (pn == 0) || o2.empty() || in.empty();
          }
          if (direct && !y_post && group_run(iy, vo1, sz, sm, pm) && group_run(iy, vo2, sz, sn, pn) &&
              group_run(iy, vba, sz, sbt, pb))
            yi = true;
          int64_t cost = (ai ? 0 : 2 * na) + (bi ? 0 : 2 * nb) + (yi ? 0 : 3 * nyy);
          if (cost < best.cost) {
            best.cost = cost;
            best.L = {vo1, vin, vo2, vba};
            best.a_inplace = ai;
            best.b_inplace = bi;
            best.y_inplace = yi;
          }
        }
  // Small steps of tensor-network trees (a short contracted space, labels scattered over the
  // operands): one strided kernel instead of permute + GEMM (+ permute).
  {
    const int64_t nk = prod_sizes(in, sz);
    const bool all_inplace = best.a_inplace && best.b_inplace && best.y_inplace;
    // (Float32 / ComplexF32 at FP64 precision keep the widened FP64 GEMM: they round once)
    const bool fp64_widened = ctx->opt.precision == PREC_FP64 && (dt == F32 || dt == C32);
    if (ctx->opt.gemm_kernel == GEMM_AUTO && !fp64_widened && !all_inplace && !a_pre && !b_pre &&
        !y_post && nk <= STRIDED_MAX_K && nyy <= STRIDED_MAX_OUT && nyy * nk <= STRIDED_MAX_WORK &&
        (int)in.size() <= SG_MAXK &&
        (int)iy.size() <= SG_MAXO && na < (1ll << 31) && nb < (1ll << 31) && nyy < (1ll << 31)) {
      auto stride_in = [&](const std::vector<int>& ls, int l) -> int32_t {
        int64_t s = 1;
        for (int l2 : ls) {
          if (l2 == l) return (int32_t)s;
          s *= sz[l2];
        }
        return 0;
      };
      StridedGemmDesc d{};
      d.a = x1;
      d.b = x2;
      d.y = c.y;
      d.nout = nyy;
      d.nk = (int32_t)nk;
      d.no = (int32_t)iy.size();
      d.ni = (int32_t)in.size();
      for (int q = 0; q < d.no; ++q) {
        const int l = iy[q];
        d.odiv[q] = make_fdiv((uint32_t)sz[l]);
        d.oa[q] = stride_in(ix1, l);
        d.ob[q] = stride_in(ix2, l);
        d.oy[q] = stride_in(iy, l);
      }
      for (int q = 0; q < d.ni; ++q) {
        const int l = in[q];
        d.kdiv[q] = make_fdiv((uint32_t)sz[l]);
        d.ka[q] = stride_in(ix1, l);
        d.kb[q] = stride_in(ix2, l);
      }
      d.sx = c.sx;
      d.sy = c.sy;
      launch_gemm_strided(dt, d, st);
      return Status::ok();
    }
  }
  const BinLayout& L = best.L;
  auto cat = [](std::initializer_list<const std::vector<int>*> parts) {
    std::vector<int> v;
    for (auto* p : parts) v.insert(v.end(), p->begin(), p->end());
    return v;
  };
  // materialise A as [o1, in, ba] / B as [in, o2, ba] when not read in place; when both are
  // plain permutes needing the permute kernel, they go out as one launch
  const void* pa = x1;
  std::vector<int> la = ix1;
  const void* pb = x2;
  std::vector<int> lb = ix2;
  if (!best.a_inplace) la = cat({&L.o1, &L.in, &L.ba});
  if (!best.b_inplace) lb = cat({&L.in, &L.o2, &L.ba});
  const bool pair = !best.a_inplace && !best.b_inplace && is_permutation(ix1, la) &&
                    is_permutation(ix2, lb);
  PermDesc pda{}, pdb{};
  bool dfa = false, dfb = false;
  if (!best.a_inplace) {
    keep.emplace_back(ctx, (size_t)na * elem_size(dt));
    if (!keep.back().p && na) return Status::err(OME_ERR_OOM, "device pool exhausted");
    Status s = pair ? do_permute(dt, x1, ix1, keep.back().p, la, sz, one, zero, 0, st, &pda, &dfa)
                    : do_unary(ctx, dt, x1, ix1, keep.back().p, la, sz, one, zero, 0, st, keep);
    if (!s.good()) return s;
    pa = keep.back().p;
  }
  if (!best.b_inplace) {
    keep.emplace_back(ctx, (size_t)nb * elem_size(dt));
    if (!keep.back().p && nb) return Status::err(OME_ERR_OOM, "device pool exhausted");
    Status s = pair ? do_permute(dt, x2, ix2, keep.back().p, lb, sz, one, zero, 0, st, &pdb, &dfb)
                    : do_unary(ctx, dt, x2, ix2, keep.back().p, lb, sz, one, zero, 0, st, keep);
    if (!s.good()) return s;
    pb = keep.back().p;
  }
  if (dfa && dfb)
    launch_permute2(dt, pda, pdb, st);
  else if (dfa)
    launch_permute(dt, pda, st);
  else if (dfb)
    launch_permute(dt, pdb, st);
  void* py = c.y;
  std::vector<int> ly = iy;
  Scal gsx = c.sx, gsy = c.sy;
  if (!best.y_inplace) {
    ly = cat({&L.o1, &L.o2, &L.ba});
    keep.emplace_back(ctx, (size_t)nyy * elem_size(dt));
    if (!keep.back().p && nyy) return Status::err(OME_ERR_OOM, "device pool exhausted");
    py = keep.back().p;
    gsx = one;
    gsy = zero;
  }
  GemmDesc d{};
  d.a = pa;
  d.b = pb;
  d.y = py;
  d.M = prod_sizes(L.o1, sz);
  d.K = prod_sizes(L.in, sz);
  d.N = prod_sizes(L.o2, sz);
  d.B = prod_sizes(L.ba, sz);
  int64_t s;
  int p;
  group_run(la, L.o1, sz, d.sam, p);
  group_run(la, L.in, sz, d.sak, p);
  group_run(la, L.ba, sz, d.sab, p);
  group_run(lb, L.in, sz, d.sbk, p);
  group_run(lb, L.o2, sz, d.sbn, p);
  group_run(lb, L.ba, sz, d.sbb, p);
  group_run(ly, L.o1, sz, d.sym, p);
  group_run(ly, L.o2, sz, d.syn, p);
  group_run(ly, L.ba, sz, d.syb, p);
  (void)s;
  d.sx = gsx;
  d.sy = gsy;
  Status st2 = run_gemm(ctx, dt, d, st, ctx->opt.precision);
  if (!st2.good()) return st2;
  if (!best.y_inplace)
    return do_unary(ctx, dt, py, ly, c.y, iy, sz, c.sx, c.sy, 0, st, keep);
  return Status::ok();
}

// ---- entry -------------------------------------------------------------------------------------

Status run_einsum(ome_ctx* ctx, const EinsumCall& c0, hipStream_t st) {
  EinsumCall c = c0;
  const int dt = c.dt;
  if (dt < F32 || dt > DT_LAST) return Status::err(OME_ERR_INVALID_ARG, "unknown element type");
  const int nin = (int)c.xs.size();
  if (nin < 1) return Status::err(OME_ERR_INVALID_ARG, "number of input tensors must be greater than 0");
  const int nl = (int)c.sizes.size();
  auto check = [&](const std::vector<int>& ls) {
    if ((int)ls.size() > OME_MAX_RANK) return false;
    for (int l : ls)
      if (l < 0 || l >= nl) return false;
    return true;
  };
  for (auto& ix : c.ixs)
    if (!check(ix)) return Status::err(OME_ERR_INVALID_ARG, "label id out of range or rank too large");
  if (!check(c.iy)) return Status::err(OME_ERR_INVALID_ARG, "output label id out of range");
  for (int64_t s : c.sizes)
    if (s < 0) return Status::err(OME_ERR_INVALID_ARG, "negative label size");
  // every output label must have a size (labels only in iy come from size_info)
  const int64_t ny = prod_sizes(c.iy, c.sizes);
  if (ny == 0) return Status::ok();
  // zero-element tensors may come with null pointers; all others must not
  if (!c.y) return Status::err(OME_ERR_INVALID_ARG, "null output pointer");
  for (int k = 0; k < nin; ++k)
    if (!c.xs[k] && prod_sizes(c.ixs[k], c.sizes) != 0)
      return Status::err(OME_ERR_INVALID_ARG, "null input pointer");
  if (c.force_loop || nin >= 3) {
    if (!ctx->opt.allow_loops) {
      std::string m = "using `loop_einsum` is forbidden: code: ";
      for (auto& ix : c.ixs) m += fmt_labels(ix);
      m += " -> " + fmt_labels(c.iy);
      return Status::err(OME_ERR_LOOPS_FORBIDDEN, m);
    }
  }
  std::vector<Temp> keep;
  // conjugated inputs are materialised once
  for (int k = 0; k < nin; ++k)
    if (k < (int)c.conj.size() && c.conj[k] && is_complex(dt)) {
      int64_t n = prod_sizes(c.ixs[k], c.sizes);
      keep.emplace_back(ctx, (size_t)n * elem_size(dt));
      if (!keep.back().p && n) return Status::err(OME_ERR_OOM, "device pool exhausted");
      launch_axpby(dt, c.xs[k], keep.back().p, n, scal_one(dt), scal_zero(dt), 1, st);
      c.xs[k] = keep.back().p;
    }
  // an empty input makes the contraction an empty sum: y <- sy*y + sx*zero
  for (int k = 0; k < nin; ++k)
    if (prod_sizes(c.ixs[k], c.sizes) == 0) {
      launch_axpby(dt, nullptr, c.y, ny, c.sx, c.sy, 0, st);
      return Status::ok();
    }
  // extent-1 labels change no layout: drop them everywhere
  auto drop1 = [&](const std::vector<int>& ls) {
    std::vector<int> out;
    for (int l : ls)
      if (c.sizes[l] != 1) out.push_back(l);
    return out;
  };
  std::vector<std::vector<int>> ixs;
  for (auto& ix : c.ixs) ixs.push_back(drop1(ix));
  std::vector<int> iy = drop1(c.iy);
  Status s;
  if (c.force_loop || nin >= 3) {
    s = do_walk(ctx, dt, c.xs, ixs, c.y, iy, c.sizes, c.sx, c.sy, st, keep);
  } else if (nin == 1) {
    s = do_unary(ctx, dt, c.xs[0], ixs[0], c.y, iy, c.sizes, c.sx, c.sy, 0, st, keep);
  } else {
    s = do_binary(ctx, c, ixs[0], ixs[1], iy, c.xs[0], c.xs[1], st, keep);
  }
  if (hipPeekAtLastError() != hipSuccess) {
    hipError_t e = hipGetLastError();
    return Status::err(OME_ERR_DEVICE, std::string("HIP launch failed: ") + hipGetErrorString(e));
  }
  return s;
}

}  // namespace ome
