// Host side of the grouped launches (ome_grouped_gemm, ome_grouped_permute) and the
// autotuner that picks the grouped GEMM's tile variant and grid per shape class.
//
// Problem descriptors are built on the host, staged in pinned memory and copied to the device
// once per distinct problem list (cache of 32 lists per context, keyed by the full list of
// shapes and pointers), so a repeated launch over the same buffers costs one kernel launch.
// A list used while the stream is being captured may be read by every replay of that graph
// (its copy is a graph node, its device block a kernel argument): such an entry is pinned and
// kept until the context is shut down.  Under capture nothing is uploaded (an upload would
// allocate) and nothing is evicted (an eviction waits for the stream, which a capture
// forbids): a list is captured after it has run once on the context.
#include <algorithm>
#include <cstring>
#include <list>
#include <map>

#include "runtime.h"

namespace ome {

namespace {

struct GroupEntry {
  std::vector<int64_t> key;
  void* dev = nullptr;
  void* host = nullptr;
  size_t bytes = 0;
  size_t first_off = 0;
  size_t map_off = 0;
  int nprob = 0;
  int64_t total = 0;
  int64_t split = 0;  // grouped permute: blocks of the shared-dim problems, which come first
  bool pinned = false;  // used under graph capture: kept for the context's life
  int64_t xfirst[GG_XCDS + 1] = {};  // grouped GEMM: first tile of each XCD's range
  bool offsets32 = false;  // grouped GEMM: every problem fits 32-bit buffer offsets
};

inline int ilog2(double v) {
  int l = 0;
  while (v >= 2.0 && l < 62) { v *= 0.5; ++l; }
  return l;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    hipGetLastError();
    return true;
  }
  return cs != hipStreamCaptureStatusNone;
}

bool is_zero_scalar(int dt, const Scal& s) {
  bool z = false;
  OME_DISPATCH(dt, O, { z = O::is_zero(s); });
  return z;
}

}  // namespace

struct GroupState {
  std::list<GroupEntry> cache;               // most recent first
  std::map<std::vector<int64_t>, int> tuned;  // shape class -> candidate index
};

static GroupState* group_state(ome_ctx* ctx) {
  if (!ctx->gstate) ctx->gstate = new GroupState();
  return ctx->gstate;
}

void destroy_group_state(ome_ctx* ctx) {
  if (!ctx->gstate) return;
  for (auto& e : ctx->gstate->cache) {
    ctx->pool->free(e.dev);
    hipHostFree(e.host);
  }
  delete ctx->gstate;
  ctx->gstate = nullptr;
}

// Find or build the device copy of a descriptor list.  On a miss `fill` writes at most
// `bytes` bytes of descriptors into pinned host memory and sets the entry's table offset,
// problem count and total; the whole buffer is then copied to the device.
template <class Fill>
static GroupEntry* group_entry(ome_ctx* ctx, hipStream_t st, const std::vector<int64_t>& key,
                               size_t bytes, Fill fill) {
  GroupState* gs = group_state(ctx);
  const bool cap = capturing(st);
  for (auto it = gs->cache.begin(); it != gs->cache.end(); ++it)
    if (it->key == key) {
      if (cap) it->pinned = true;
      if (it != gs->cache.begin()) gs->cache.splice(gs->cache.begin(), gs->cache, it);
      return &gs->cache.front();
    }
  // a new list is not uploaded under capture (the upload would allocate): the caller reports
  // that the list must run once on the context before the capture
  if (cap) return nullptr;
  GroupEntry e;
  e.key = key;
  e.bytes = bytes;
  if (hipHostMalloc(&e.host, bytes, hipHostMallocDefault) != hipSuccess) {
    hipGetLastError();
    return nullptr;
  }
  e.dev = ctx->pool->alloc(bytes);
  if (!e.dev) {
    hipHostFree(e.host);
    return nullptr;
  }
  fill((char*)e.host, e);
  if (hipMemcpyAsync(e.dev, e.host, bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
    hipGetLastError();
    ctx->pool->free(e.dev);
    hipHostFree(e.host);
    return nullptr;
  }
  gs->cache.push_front(std::move(e));
  if (gs->cache.size() > 32 && !cap) {
    // evict the least recently used entry that no graph may replay
    for (auto it = std::prev(gs->cache.end()); it != gs->cache.begin(); --it) {
      if (it->pinned) continue;
      if (hipStreamSynchronize(st) != hipSuccess) {  // the evicted copy may still be in flight
        hipGetLastError();
        break;  // keep the entry rather than free memory a kernel may still read
      }
      ctx->pool->free(it->dev);
      hipHostFree(it->host);
      gs->cache.erase(it);
      break;
    }
  }
  return &gs->cache.front();
}

// ---- grouped GEMM --------------------------------------------------------------------------

struct GGProblem {
  int64_t m, n, k, batch;
  const void *a, *b;
  void* c;
};

static GroupEntry* gg_entry(ome_ctx* ctx, hipStream_t st, int dt, int variant,
                            const std::vector<GGProblem>& ps) {
  int tm, tn;
  gg_tile_shape(dt, variant, tm, tn);
  std::vector<int64_t> key;
  key.reserve(4 + ps.size() * 7);
  key.push_back(1);  // grouped GEMM
  key.push_back(dt);
  key.push_back(tm);
  key.push_back(tn);
  int64_t ntiles = 0;
  for (const auto& p : ps) {
    key.push_back(p.m); key.push_back(p.n); key.push_back(p.k); key.push_back(p.batch);
    key.push_back((int64_t)(uintptr_t)p.a); key.push_back((int64_t)(uintptr_t)p.b);
    key.push_back((int64_t)(uintptr_t)p.c);
    if (p.m > 0 && p.n > 0 && p.batch > 0)
      ntiles += ((p.m + tm - 1) / tm) * ((p.n + tn - 1) / tn) * p.batch;
  }
  // the descriptors are built only when the list is not cached:
  // [problems | first tile of each problem | problem of each tile]
  const size_t pbytes = (ps.size() * sizeof(GGProb) + 255) & ~(size_t)255;
  const size_t fbytes = (std::max<size_t>(ps.size(), 1) * sizeof(int64_t) + 255) & ~(size_t)255;
  const size_t bytes = pbytes + fbytes + std::max<int64_t>(ntiles, 1) * sizeof(int32_t);
  return group_entry(ctx, st, key, bytes, [&](char* h, GroupEntry& e) {
    // longest K first, then the larger output: tiles of similar cost sit together
    std::vector<int> sorted(ps.size());
    for (size_t i = 0; i < ps.size(); ++i) sorted[i] = (int)i;
    std::stable_sort(sorted.begin(), sorted.end(), [&](int x, int y) {
      if (ps[x].k != ps[y].k) return ps[x].k > ps[y].k;
      return ps[x].m * ps[x].n > ps[y].m * ps[y].n;
    });
    // deal the problems in that order to the least loaded XCD (cost: tiles x (K + a fixed
    // per-tile share)); each XCD's problems keep the order, and its tiles form one range
    std::vector<int> per_xcd[GG_XCDS];
    double load[GG_XCDS] = {};
    for (int i : sorted) {
      const GGProblem& p = ps[i];
      if (p.m <= 0 || p.n <= 0 || p.batch <= 0) continue;
      const double nt = (double)((p.m + tm - 1) / tm) * ((p.n + tn - 1) / tn) * p.batch;
      int best = 0;
      for (int c = 1; c < GG_XCDS; ++c)
        if (load[c] < load[best]) best = c;
      load[best] += nt * (double)(std::max<int64_t>(p.k, 1) + 32);
      per_xcd[best].push_back(i);
    }
    std::vector<int> order;
    order.reserve(ps.size());
    std::vector<size_t> xcd_np(GG_XCDS + 1, 0);
    for (int c = 0; c < GG_XCDS; ++c) {
      xcd_np[c] = order.size();
      order.insert(order.end(), per_xcd[c].begin(), per_xcd[c].end());
    }
    xcd_np[GG_XCDS] = order.size();
    GGProb* probs = (GGProb*)h;
    int64_t* first = (int64_t*)(h + pbytes);
    int32_t* map = (int32_t*)(h + pbytes + fbytes);
    int np = 0;
    int64_t total = 0;
    int xc = 0;
    for (size_t oi = 0; oi < order.size(); ++oi) {
      while (xc < GG_XCDS && oi == xcd_np[xc]) e.xfirst[xc++] = total;
      const GGProblem& p = ps[order[oi]];
      GGProb g{};
      g.a = p.a;
      g.b = p.b;
      g.c = p.c;
      g.m = p.m;
      g.n = p.n;
      g.k = p.k;
      g.tiles_m = (int32_t)((p.m + tm - 1) / tm);
      g.tiles_n = (int32_t)((p.n + tn - 1) / tn);
      g.vec_a = (p.m % 4 == 0 && aligned16(p.a)) ? 1 : 0;
      g.vec_b = (p.k % 4 == 0 && aligned16(p.b)) ? 1 : 0;
      g.per_div = make_fdiv((uint32_t)((int64_t)g.tiles_m * g.tiles_n));
      g.tm_div = make_fdiv((uint32_t)g.tiles_m);
      g.first_tile = total;
      probs[np] = g;
      first[np] = total;
      const int64_t nt = (int64_t)g.tiles_m * g.tiles_n * p.batch;
      for (int64_t q = 0; q < nt; ++q) map[total + q] = np;
      ++np;
      total += nt;
    }
    while (xc <= GG_XCDS) e.xfirst[xc++] = total;
    e.offsets32 = true;
    for (const auto& p : ps) {
      if (p.m <= 0 || p.n <= 0 || p.batch <= 0) continue;
      const int64_t lim = (int64_t)1 << 22, slim = (int64_t)1 << 31;
      const int64_t ptiles = ((p.m + tm - 1) / tm) * ((p.n + tn - 1) / tn) * p.batch;
      if (p.m >= lim || p.n >= lim || p.k >= lim || p.m * p.k * 4 >= slim ||
          p.k * p.n * 4 >= slim || p.m * p.n * 4 >= slim || ptiles >= slim)
        e.offsets32 = false;
    }
    e.first_off = pbytes;
    e.map_off = pbytes + fbytes;
    e.nprob = np;
    e.total = total;
  });
}

// Candidates: (tile variant, grid).  Grid 0: one workgroup per tile (the hardware dispatcher
// balances uneven tiles); grid 1: persistent, a few workgroups per CU walking the list.
struct GGCand {
  int variant;
  int persistent;
};

static std::vector<GGCand> gg_candidates(int dt, int precision, double mean_mn) {
  std::vector<GGCand> c;
  if (dt == F32 && precision == PREC_BF16X3) {
    c.push_back({GG_BF16X3_W32, 1});
    c.push_back({GG_BF16X3_W32, 0});
    c.push_back({GG_BF16X3_64, 0});
    c.push_back({GG_BF16X3_64, 1});
    c.push_back({GG_BF16X3_64_BK64, 0});
    c.push_back({GG_BF16X3_RUN, 0});
    if (mean_mn >= 96) c.push_back({GG_BF16X3_128, 0});
  } else if (dt == F32 && precision == PREC_BF16X6) {
    // FP32-accurate: the six-product split against the FP32 kernels, the faster one kept
    c.push_back({GG_BF16X6_64, 0});
    c.push_back({GG_BF16X6_64, 1});
    c.push_back({GG_F32_64_BK32, 0});
    c.push_back({GG_F32_RUN, 0});
  } else if (dt == F32) {
    c.push_back({GG_F32_W32, 1});
    c.push_back({GG_F32_W32, 0});
    c.push_back({GG_F32_64, 0});
    c.push_back({GG_F32_64, 1});
    c.push_back({GG_F32_64_BK32, 0});
    c.push_back({GG_F32_RUN, 0});
    if (mean_mn >= 96) c.push_back({GG_F32_128, 0});
  } else {
    c.push_back({0, 0});
    c.push_back({0, 1});
  }
  return c;
}

// A multiple of GG_XCDS: every XCD gets the same number of workgroups for its range.
static int gg_grid(ome_ctx* ctx, int dt, const GGCand& c, const GroupEntry& e) {
  // wave-per-tile variants: four tiles per workgroup
  const bool wave = dt == F32 && (c.variant == GG_F32_W32 || c.variant == GG_BF16X3_W32);
  int64_t widest = 0;
  for (int x = 0; x < GG_XCDS; ++x) widest = std::max(widest, e.xfirst[x + 1] - e.xfirst[x]);
  int64_t per = wave ? (widest + 3) / 4 : widest;  // workgroups one XCD's range can use
  if (per < 1) per = 1;
  if (c.persistent) {
    const int per_cu = (dt == F32 && (c.variant == GG_F32_128 || c.variant == GG_BF16X3_128)) ? 2
                       : wave ? 8 : 4;
    per = std::min<int64_t>(per, std::max<int64_t>(1, (int64_t)ctx->info.compute_units * per_cu /
                                                          GG_XCDS));
  }
  return (int)std::min<int64_t>(per * GG_XCDS, (int64_t)1 << 30);
}

static Status gg_launch(ome_ctx* ctx, hipStream_t st, int dt, const GGCand& c,
                        const std::vector<GGProblem>& ps, const Scal& alpha, const Scal& beta) {
  GroupEntry* e = gg_entry(ctx, st, dt, c.variant, ps);
  if (!e && capturing(st))
    return Status::err(OME_ERR_UNSUPPORTED, "grouped GEMM under graph capture: run the same problem "
                                            "list once on this context before capturing it");
  if (!e) return Status::err(OME_ERR_OOM, "no memory for the grouped GEMM descriptors");
  GGDesc g{};
  g.probs = (const GGProb*)e->dev;
  g.first = (const int64_t*)((char*)e->dev + e->first_off);
  g.map = (const int32_t*)((char*)e->dev + e->map_off);
  g.nprob = e->nprob;
  g.total_tiles = e->total;
  g.alpha = alpha;
  g.beta = beta;
  g.mode = dt == F32 ? epi_mode<OpsF32>(alpha, beta) : 0;
  for (int x = 0; x <= GG_XCDS; ++x) g.xcd_first[x] = e->xfirst[x];
  g.offsets32 = e->offsets32 && ctx->opt.grouped_buffers ? 1 : 0;
  GGCand cc = c;
  if (cc.variant == GG_F32_RUN || cc.variant == GG_BF16X3_RUN) {
    if (dt == F32 && g.offsets32) {
      // every workgroup resident (as many per CU as fit), no more than the widest range needs
      int64_t widest = 0;
      for (int x = 0; x < GG_XCDS; ++x) widest = std::max(widest, e->xfirst[x + 1] - e->xfirst[x]);
      const int per_cu = gg_run_per_cu(cc.variant == GG_BF16X3_RUN);
      const int64_t per = std::max<int64_t>(
          1, std::min<int64_t>(widest, (int64_t)ctx->info.compute_units * per_cu / GG_XCDS));
      launch_grouped_gemm(dt, cc.variant, g, (int)(per * GG_XCDS), st);
      return Status::ok();
    }
    // the run kernels read through buffer resources only: the same tiles, one per workgroup
    cc = {cc.variant == GG_F32_RUN ? GG_F32_64_BK32 : GG_BF16X3_64, 0};
  }
  launch_grouped_gemm(dt, cc.variant, g, gg_grid(ctx, dt, cc, *e), st);
  return Status::ok();
}

Status run_grouped_gemm(ome_ctx* ctx, int dt, const std::vector<GGProblem>& ps, const Scal& alpha,
                        const Scal& beta, hipStream_t st) {
  RangeScope rg("grouped GEMM");
  double flops = 0, smn = 0, sm = 0, sn = 0, sk = 0, w = 0;
  for (const auto& p : ps) {
    const double f = (double)p.m * p.n * std::max<int64_t>(p.k, 1) * p.batch;
    flops += f;
    sm += f * p.m; sn += f * p.n; sk += f * p.k; w += f;
  }
  if (w <= 0) w = 1;
  smn = (sm + sn) / (2 * w);
  // Float32 follows the context's precision (FP32, BF16x3 or BF16x6 products); FP64
  // precision has no grouped form and runs as FP32
  const int prec = ctx->opt.precision;
  const auto cands = gg_candidates(dt, prec, smn);
  int pick = -1;
  if (ctx->opt.grouped_variant >= 0) {
    // forced tile variant (one workgroup per tile)
    GGCand c{dt == F32 ? ctx->opt.grouped_variant : 0, 0};
    ctx->last_grouped_choice = c.variant * 2;
    return gg_launch(ctx, st, dt, c, ps, alpha, beta);
  }
  std::vector<int64_t> cls = {dt, prec, ilog2((double)ps.size()), ilog2(flops), ilog2(sm / w),
                              ilog2(sn / w), ilog2(sk / w)};
  GroupState* gs = group_state(ctx);
  auto it = gs->tuned.find(cls);
  if (it != gs->tuned.end()) pick = it->second;
  const bool can_tune = ctx->opt.autotune && !launch_mode().dry && !capturing(st) &&
                        is_zero_scalar(dt, beta) && cands.size() > 1;
  if (pick < 0 && can_tune) {
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float best = 1e30f;
    for (size_t i = 0; i < cands.size(); ++i) {
      Status s = gg_launch(ctx, st, dt, cands[i], ps, alpha, beta);  // warm-up (and upload)
      if (!s.good()) { hipEventDestroy(e0); hipEventDestroy(e1); return s; }
      float ms = 0, sum = 0;
      for (int r = 0; r < 2; ++r) {
        hipEventRecord(e0, st);
        gg_launch(ctx, st, dt, cands[i], ps, alpha, beta);
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        hipEventElapsedTime(&ms, e0, e1);
        sum += ms;
      }
      if (sum < best) { best = sum; pick = (int)i; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    gs->tuned[cls] = pick;
  }
  if (pick < 0) {
    // untuned default: 64x64 tiles, one workgroup per tile; 128x128 for large outputs
    pick = 0;
    if (dt == F32 && smn >= 192)
      for (size_t i = 0; i < cands.size(); ++i)
        if (cands[i].variant == GG_F32_128 || cands[i].variant == GG_BF16X3_128) pick = (int)i;
  }
  ctx->last_grouped_choice = cands[pick].variant * 2 + cands[pick].persistent;
  return gg_launch(ctx, st, dt, cands[pick], ps, alpha, beta);
}

// ---- grouped permute ------------------------------------------------------------------------

// Tile plan of one permute whose output dims od (fastest first, size-1 dims dropped, runs of
// consecutive input dims merged) have input strides os: fills pr (x, y, n, the plain walk and,
// when one applies, the two-group or the shared-dim tile) and returns its block count.
static int64_t gp_plan(int dt, const std::vector<int64_t>& od, const std::vector<int64_t>& os,
                       const void* x, void* y, int64_t n, GPProb& pr) {
  pr = GPProb{};
  pr.x = x;
  pr.y = y;
  pr.n = n;
  pr.rank = (int32_t)od.size();
  for (int q = 0; q < pr.rank; ++q) {
    pr.odiv[q] = make_fdiv((uint32_t)od[q]);
    pr.istride[q] = os[q];
  }
  // tiled path: I = the input's leading dims (by input stride) until their product
  // reaches 64; O = the output's leading dims, up to the first one in I, until 64
  const int R = pr.rank;
  std::vector<int64_t> ost(R);
  {
    int64_t acc = 1;
    for (int q = 0; q < R; ++q) { ost[q] = acc; acc *= od[q]; }
  }
  std::vector<int> byin(R);
  for (int q = 0; q < R; ++q) byin[q] = q;
  std::sort(byin.begin(), byin.end(), [&](int u, int v) { return os[u] < os[v]; });
  std::vector<int> gi, go;
  int64_t pb = 1, pa = 1;
  // choose how many leading input dims form I (then O = the leading output dims before the
  // first dim of I): maximise the shorter of the contiguous read run (along b, at most 64)
  // and the contiguous write run (along a, at most what fills a 4096-element tile)
  int best_score = -1;
  for (int ni = 1; ni <= std::min(R, GP_MAXG); ++ni) {
    std::vector<int> ci(byin.begin(), byin.begin() + ni);
    int64_t cb = 1;
    for (int q : ci) cb *= od[q];
    std::vector<int> co;
    int64_t ca = 1;
    for (int q = 0; q < R; ++q) {
      if (ca >= 256 || (int)co.size() == GP_MAXG) break;
      if (std::find(ci.begin(), ci.end(), q) != ci.end()) break;
      co.push_back(q);
      ca *= od[q];
    }
    const int64_t rb = std::min<int64_t>(cb, 64);
    const int64_t ra = std::min<int64_t>(std::min<int64_t>(ca, 4096 / rb), 256);
    const int score = co.empty() ? -1 : (int)std::min(rb, ra);
    if (score > best_score) {
      best_score = score;
      gi = ci;
      go = co;
      pb = cb;
      pa = ca;
    }
    if (cb >= 64) break;
  }
  // shared-dim tile: a dim S that follows a whole group I of leading input dims in the input
  // and a whole group O of leading output dims in the output; a block takes all of I and O
  // and a chunk of S, so its reads run over pb * tsc and its writes over pa * tsc elements
  const int64_t scap = GP_SCAP_BYTES / (int64_t)elem_size(dt);
  int sq = -1;
  std::vector<int> s_gi;
  int64_t s_pa = 0, s_pb = 0, s_tsc = 0, s_score = -1;
  for (int q = 1; q < R; ++q) {
    std::vector<int> ci;
    int64_t cb = 1;
    for (int u = 0; u < R; ++u)
      if (os[u] < os[q]) { ci.push_back(u); cb *= od[u]; }
    if (ci.empty() || cb != os[q]) continue;  // S must follow I contiguously in the input
    bool overlap = false;
    for (int u : ci)
      if (u < q) overlap = true;               // I and O = output dims 0..q-1 disjoint
    const int64_t ca = ost[q];
    if (overlap || q > GP_MAXG || (int)ci.size() > GP_MAXG || ca > 256 || cb > 256 ||
        ca * cb * 2 > scap)
      continue;
    std::sort(ci.begin(), ci.end(), [&](int u, int v) { return os[u] < os[v]; });
    const int64_t tsc = std::min<int64_t>(od[q], scap / (ca * cb));
    const int64_t score = std::min(ca, cb) * tsc;
    if (score > s_score) {
      s_score = score;
      sq = q;
      s_gi = ci;
      s_pa = ca;
      s_pb = cb;
      s_tsc = tsc;
    }
  }
  int64_t blocks;
  if (sq >= 0 && s_score > best_score) {
    pr.tiled = 2;
    pr.no = sq;
    pr.ni = (int32_t)s_gi.size();
    pr.pa = s_pa;
    pr.pb = s_pb;
    pr.tla = (int32_t)s_pa;
    pr.tlb = (int32_t)s_pb;
    pr.tla_div = make_fdiv((uint32_t)s_pa);
    pr.tlb_div = make_fdiv((uint32_t)s_pb);
    pr.tsc = (int32_t)s_tsc;
    pr.tsc_div = make_fdiv((uint32_t)s_tsc);
    pr.st_div = make_fdiv((uint32_t)(s_tsc * s_pb));
    pr.s_ext = od[sq];
    pr.s_is = os[sq];
    pr.s_os = ost[sq];
    pr.ts_tiles = (od[sq] + s_tsc - 1) / s_tsc;
    for (int u = 0; u < sq; ++u) {
      pr.odv[u] = make_fdiv((uint32_t)od[u]);
      pr.o_is[u] = os[u];
    }
    for (int u = 0; u < pr.ni; ++u) {
      pr.idv[u] = make_fdiv((uint32_t)od[s_gi[u]]);
      pr.i_os[u] = ost[s_gi[u]];
    }
    int nr = 0;
    int64_t restn = 1;
    for (int q = 0; q < R; ++q) {
      if (q <= sq || std::find(s_gi.begin(), s_gi.end(), q) != s_gi.end()) continue;
      pr.rdv[nr] = make_fdiv((uint32_t)od[q]);
      pr.r_is[nr] = os[q];
      pr.r_os[nr] = ost[q];
      restn *= od[q];
      ++nr;
    }
    pr.nr = nr;
    blocks = pr.ts_tiles * restn;
  } else if (!gi.empty() && !go.empty() && os[gi[0]] == 1 && R > 1) {
    pr.tiled = 1;
    pr.no = (int32_t)go.size();
    pr.ni = (int32_t)gi.size();
    pr.pa = pa;
    pr.pb = pb;
    // tile: up to 64 along b, and along a as many as fill 4096 elements (at most 256)
    const int64_t tlb = std::min<int64_t>(pb, 64);
    const int64_t tla = std::min<int64_t>(std::min<int64_t>(pa, 4096 / tlb), 256);
    pr.tla = (int32_t)tla;
    pr.tlb = (int32_t)tlb;
    pr.tla_div = make_fdiv((uint32_t)tla);
    pr.tlb_div = make_fdiv((uint32_t)tlb);
    pr.ta = (pa + tla - 1) / tla;
    pr.tb = (pb + tlb - 1) / tlb;
    for (int u = 0; u < pr.no; ++u) {
      pr.odv[u] = make_fdiv((uint32_t)od[go[u]]);
      pr.o_is[u] = os[go[u]];
    }
    for (int u = 0; u < pr.ni; ++u) {
      pr.idv[u] = make_fdiv((uint32_t)od[gi[u]]);
      pr.i_os[u] = ost[gi[u]];
    }
    int nr = 0;
    int64_t restn = 1;
    for (int q = 0; q < R; ++q) {
      if (std::find(gi.begin(), gi.end(), q) != gi.end() ||
          std::find(go.begin(), go.end(), q) != go.end())
        continue;
      pr.rdv[nr] = make_fdiv((uint32_t)od[q]);
      pr.r_is[nr] = os[q];
      pr.r_os[nr] = ost[q];
      restn *= od[q];
      ++nr;
    }
    pr.nr = nr;
    blocks = pr.ta * pr.tb * restn;
  } else {
    blocks = (n + 1023) / 1024;
  }
  return blocks;
}

Status run_grouped_permute(ome_ctx* ctx, int dt, int np, const int32_t* ndims, const int64_t* dims,
                           const int32_t* perms, const void* const* a, void* const* c,
                           const Scal& sx, const Scal& sy, hipStream_t st) {
  RangeScope rg("grouped permute");
  std::vector<int64_t> key = {2, dt};
  std::vector<GPProb> probs;
  std::vector<int64_t> first;
  int64_t total = 0;
  int64_t off = 0;
  for (int g = 0; g < np; ++g) {
    const int r = ndims[g];
    if (r < 0 || r > OME_MAX_RANK) return Status::err(OME_ERR_INVALID_ARG, "invalid rank");
    const int64_t* d = dims + off;
    const int32_t* pm = perms + off;
    off += r;
    std::vector<int> seen(r, 0);
    int64_t n = 1;
    for (int q = 0; q < r; ++q) {
      if (pm[q] < 0 || pm[q] >= r || seen[pm[q]]) return Status::err(OME_ERR_INVALID_ARG, "invalid permutation");
      seen[pm[q]] = 1;
      if (d[q] < 0) return Status::err(OME_ERR_INVALID_ARG, "negative extent");
      n *= d[q];
    }
    key.push_back(r);
    for (int q = 0; q < r; ++q) { key.push_back(d[q]); key.push_back(pm[q]); }
    key.push_back((int64_t)(uintptr_t)a[g]);
    key.push_back((int64_t)(uintptr_t)c[g]);
    if (n == 0) continue;
    if (!a[g] || !c[g]) return Status::err(OME_ERR_INVALID_ARG, "null tensor pointer");
    if (n >= (1ll << 32)) return Status::err(OME_ERR_UNSUPPORTED, "permute of >= 2^32 elements");
    // output dims (fastest first) with input strides; size-1 dims dropped and runs of
    // consecutive input dims merged (the grouping of tensorpermute!, src/utils.jl:132-144)
    std::vector<int64_t> xs(r);
    int64_t s = 1;
    for (int q = 0; q < r; ++q) { xs[q] = s; s *= d[q]; }
    std::vector<int64_t> od, os;
    int prev = -2;
    for (int q = 0; q < r; ++q) {
      const int p = pm[q];
      if (d[p] == 1) continue;
      bool merge = false;
      if (!od.empty() && p > prev) {
        merge = true;
        for (int u = prev + 1; u < p; ++u)
          if (d[u] != 1) merge = false;
      }
      if (merge) od.back() *= d[p];
      else { od.push_back(d[p]); os.push_back(xs[p]); }
      prev = p;
    }
    if (od.empty()) { od.push_back(1); os.push_back(1); }
    if ((int)od.size() > GP_MAXR)
      return Status::err(OME_ERR_UNSUPPORTED, "grouped permute of rank > 24 after grouping");
    GPProb pr;
    const int64_t blocks = gp_plan(dt, od, os, a[g], c[g], n, pr);
    probs.push_back(pr);
    first.push_back(blocks);  // block count for now; turned into first blocks below
  }
  if (probs.empty()) return Status::ok();
  // shared-dim problems first (a kernel instantiation of their own), then the others, each
  // in call order
  std::stable_partition(probs.begin(), probs.end(),
                        [](const GPProb& q) { return q.tiled == 2; });
  // the partition moved the problems: recompute each one's block count in the new order
  for (size_t q = 0; q < probs.size(); ++q) {
    const GPProb& pr = probs[q];
    int64_t restn = 1;
    for (int u = 0; u < pr.nr; ++u) restn *= pr.rdv[u].d;
    first[q] = pr.tiled == 2 ? pr.ts_tiles * restn
             : pr.tiled == 1 ? pr.ta * pr.tb * restn
                             : (pr.n + 1023) / 1024;
  }
  int64_t split = 0;
  for (size_t q = 0; q < probs.size(); ++q) {
    const int64_t nb = first[q];
    first[q] = total;
    if (probs[q].tiled == 2) split += nb;
    total += nb;
  }
  const size_t pbytes = (probs.size() * sizeof(GPProb) + 255) & ~(size_t)255;
  const size_t fbytes = (first.size() * sizeof(int64_t) + 255) & ~(size_t)255;
  const size_t bytes = pbytes + fbytes + total * sizeof(int32_t);
  GroupEntry* e = group_entry(ctx, st, key, bytes, [&](char* h, GroupEntry& en) {
    std::memcpy(h, probs.data(), probs.size() * sizeof(GPProb));
    std::memcpy(h + pbytes, first.data(), first.size() * sizeof(int64_t));
    int32_t* map = (int32_t*)(h + pbytes + fbytes);
    for (size_t q = 0; q < first.size(); ++q) {
      const int64_t end = q + 1 < first.size() ? first[q + 1] : total;
      for (int64_t b = first[q]; b < end; ++b) map[b] = (int32_t)q;
    }
    en.first_off = pbytes;
    en.map_off = pbytes + fbytes;
    en.nprob = (int)probs.size();
    en.total = total;
    en.split = split;
  });
  if (!e && capturing(st))
    return Status::err(OME_ERR_UNSUPPORTED, "grouped permute under graph capture: run the same "
                                            "problem list once on this context before capturing it");
  if (!e) return Status::err(OME_ERR_OOM, "no memory for the grouped permute descriptors");
  GPDesc g{};
  g.probs = (const GPProb*)e->dev;
  g.first = (const int64_t*)((char*)e->dev + e->first_off);
  g.map = (const int32_t*)((char*)e->dev + e->map_off);
  g.nprob = e->nprob;
  g.total_blocks = e->total;
  g.sx = sx;
  g.sy = sy;
  launch_grouped_permute(dt, g, e->split, st);
  return Status::ok();
}

}  // namespace ome

using namespace ome;

static int gfail(ome_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

extern "C" {

int ome_grouped_gemm(ome_ctx* ctx, int dtype, int ngroups, const int64_t* dims,
                     const void* const* a, const void* const* b, void* const* c,
                     const ome_scalar* alpha, const ome_scalar* beta) {
  if (!ctx) return OME_ERR_INVALID_ARG;
  if (ngroups < 0 || (ngroups && (!dims || !a || !b || !c)))
    return gfail(ctx, OME_ERR_INVALID_ARG, "invalid grouped GEMM arguments");
  if (dtype < F32 || dtype > DT_LAST) return gfail(ctx, OME_ERR_INVALID_ARG, "unknown element type");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  std::vector<GGProblem> ps(ngroups);
  for (int g = 0; g < ngroups; ++g) {
    GGProblem& p = ps[g];
    p.m = dims[4 * g];
    p.n = dims[4 * g + 1];
    p.k = dims[4 * g + 2];
    p.batch = dims[4 * g + 3];
    p.a = a[g];
    p.b = b[g];
    p.c = c[g];
    if (p.m < 0 || p.n < 0 || p.k < 0 || p.batch < 0)
      return gfail(ctx, OME_ERR_INVALID_ARG, "negative extent in grouped GEMM");
    if (p.m * p.n * p.batch > 0 && (!p.c || (p.k > 0 && (!p.a || !p.b))))
      return gfail(ctx, OME_ERR_INVALID_ARG, "null pointer in grouped GEMM");
  }
  const Scal al = alpha ? Scal{alpha->re, alpha->im, alpha->i} : scal_one(dtype);
  const Scal be = beta ? Scal{beta->re, beta->im, beta->i} : scal_zero(dtype);
  Status s = run_grouped_gemm(ctx, dtype, ps, al, be, ctx->stream);
  if (!s.good()) return gfail(ctx, s.code, s.msg);
  return OME_OK;
}

int ome_grouped_permute(ome_ctx* ctx, int dtype, int nperm, const int32_t* ndims,
                        const int64_t* dims, const int32_t* perms, const void* const* a,
                        void* const* c, const ome_scalar* sx, const ome_scalar* sy) {
  if (!ctx) return OME_ERR_INVALID_ARG;
  if (nperm < 0 || (nperm && (!ndims || !a || !c)))
    return gfail(ctx, OME_ERR_INVALID_ARG, "invalid grouped permute arguments");
  if (dtype < F32 || dtype > DT_LAST) return gfail(ctx, OME_ERR_INVALID_ARG, "unknown element type");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  const Scal vx = sx ? Scal{sx->re, sx->im, sx->i} : scal_one(dtype);
  const Scal vy = sy ? Scal{sy->re, sy->im, sy->i} : scal_zero(dtype);
  Status s = run_grouped_permute(ctx, dtype, nperm, ndims, dims, perms, a, c, vx, vy, ctx->stream);
  if (!s.good()) return gfail(ctx, s.code, s.msg);
  return OME_OK;
}

}  // extern "C"
