// Contraction-tree plans: NestedEinsum / SlicedEinsum evaluation on the device.
//
//   * steps in post-order (src/einsequence.jl:279-287); each step is one flat einsum planned
//     by run_einsum (permutes / GEMM / index walk);
//   * node outputs live in ONE arena whose offsets come from a static liveness plan
//     (first fit; a step's output may not overlap its own inputs; a child's block is
//     released once its parent has run) — no allocation during evaluation;
//   * planner scratch (permute copies, split-K partials) comes from a bump region sized by a
//     dry run, so a recorded hipGraph keeps valid addresses;
//   * SlicedEinsum (src/slicing.jl:112-137): per slice, leaves carrying sliced labels are
//     copied (take_slice), the reduced tree runs, and the slice result is added into an FP64
//     accumulator (complex FP64 for complex types; Int64 and tropical accumulate exactly in
//     their own type); the output gets sy*y + sx*acc once at the end;
//   * evaluation is recorded once into a hipGraph per (inputs, output, sx, sy) and replayed;
//   * cost_and_gradient (src/bp.jl:122-160): forward with every node output kept, then the
//     reverse sweep with einsum_grad (src/autodiff.jl:24-30) using the same planner.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <functional>
#include <list>
#include <map>

#include "runtime.h"

using namespace ome;

namespace ome {

// ---- strided accumulate (fill_slice!, src/slicing.jl:86-94) ------------------------------

struct AccDesc {
  const void* x;  // dense slice result
  void* acc;      // accumulator (double / double2 in f64 mode, else the element type)
  int64_t n;
  int rank;
  FDiv div[MAXR];
  int64_t dstride[MAXR];
  int64_t base;
};

template <class O, bool F64>
__global__ __launch_bounds__(256) void k_acc_strided(AccDesc d) {
  using T = typename O::T;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < d.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t rem = (uint32_t)i;
    int64_t off = d.base;
    for (int p = 0; p < d.rank; ++p) {
      uint32_t q = fdiv_q(rem, d.div[p]);
      off += (int64_t)(rem - q * d.div[p].d) * d.dstride[p];
      rem = q;
    }
    T v = ((const T*)d.x)[i];
    if constexpr (F64) {
      if constexpr (O::DT == C32 || O::DT == C64) {
        ((double*)d.acc)[2 * off] += (double)v.x;
        ((double*)d.acc)[2 * off + 1] += (double)v.y;
      } else {
        ((double*)d.acc)[off] += (double)v;
      }
    } else {
      T* a = (T*)d.acc + off;
      *a = O::add(*a, v);
    }
  }
}

static void launch_acc_strided(int dt, bool f64, const AccDesc& d, hipStream_t st) {
  if (d.n <= 0) return;
  OME_DRY_RETURN(1);
  int64_t g = (d.n + 255) / 256;
  if (g > 65536) g = 65536;
  OME_DISPATCH(dt, O, {
    if (f64) {
      if constexpr (O::DT == F32 || O::DT == F64 || O::DT == C32 || O::DT == C64)
        hipLaunchKernelGGL((k_acc_strided<O, true>), dim3((uint32_t)g), dim3(256), 0, st, d);
    } else {
      hipLaunchKernelGGL((k_acc_strided<O, false>), dim3((uint32_t)g), dim3(256), 0, st, d);
    }
  });
}

}  // namespace ome

// ---- plan ------------------------------------------------------------------------------------

struct TreeArg {
  int ref;  // >= 0 leaf tensor index, < 0 internal node -(ref+1)
  std::vector<int> ix;
};
struct TreeStep {
  std::vector<TreeArg> args;
  std::vector<int> iy;
  int parent = -1;
  int64_t out_off = 0, out_elems = 0;
  int64_t grad_off = 0;
};

struct GraphEntry {
  char* mem;  // the plan's memory when the graph was recorded (its own or lent)
  bool fused_elsewhere;  // recorded without the fused subtrees (run by the request loop)
  std::vector<const void*> xs;
  void* y;
  Scal sx, sy;
  hipGraphExec_t exec = nullptr;
};

struct ome_tree {
  ome_ctx* ctx = nullptr;
  int dt = 0;
  int n_leaves = 0;
  std::vector<int64_t> sizes;     // original label sizes
  std::vector<int64_t> sizes_sl;  // sliced labels set to 1
  std::vector<int> slicing;
  int64_t nslices = 1;
  std::vector<std::vector<int>> leaf_ix;  // labels of every leaf (from the tree)
  std::vector<TreeStep> steps;
  std::vector<int> root_iy;  // output labels (full, with sliced labels)
  // memory: [arena | scratch | slice staging | accumulator | normalisation work]
  int64_t arena_bytes = 0, scratch_bytes = 0, stage_bytes = 0, acc_bytes = 0, norm_bytes = 0;
  std::vector<int64_t> stage_off;  // per leaf (-1 when not staged)
  int64_t slice_out_off = 0;       // per-slice root output (inside the arena)
  char* mem = nullptr;
  bool acc_f64 = false;
  ome_tree_info info{};
  std::list<GraphEntry> graphs;  // most recent first
  hipStream_t cap_stream = nullptr;
  ome::Options opt;  // planning options frozen at creation
  int64_t mem_bytes = 0;  // size of mem
  bool owns_mem = true;   // mem came from the pool for this tree (else a request loop lent it)
  // gradient plan (lazily built)
  char* gmem = nullptr;
  int64_t g_arena = 0, g_scratch = 0;
  // small steps fused into one launch (one workgroup per subtree of small steps)
  std::vector<char> fused;           // per step
  FusedStep* fdev = nullptr;         // device descriptors, subtree by subtree
  int32_t* ffirst = nullptr;         // device: subtree, chunk and item tables
  int fsub = 0;                      // subtrees
  // normalised evaluation of Float32 / ComplexF32 trees: the FP64 twin of this plan (same
  // description, Float64 / ComplexF64 elements) and its widened, normalised inputs
  std::vector<int32_t> desc;
  ome_tree* wide = nullptr;
  char* wleaves = nullptr;
  std::vector<int64_t> wleaf_off;
  // the request loop's admission: a sliced twin whose memory fits fit_cap bytes
  ome_tree* fit = nullptr;
  int64_t fit_cap = 0;
  // set while the request loop evaluates the plan after running its fused subtrees in a
  // launch shared with other requests: the evaluation leaves them out
  bool fused_elsewhere = false;
};

namespace {

// The planning options of a tree are frozen when it is created: its scratch is sized, its
// graphs recorded and every evaluation planned with the same kernel choices, whatever the
// context's options are at the time of a later call.
struct OptScope {
  ome_ctx* ctx;
  ome::Options saved;
  OptScope(ome_ctx* c, const ome::Options& o) : ctx(c), saved(c->opt) {
    ctx->opt.precision = o.precision;
    ctx->opt.complex_algo = o.complex_algo;
    ctx->opt.allow_loops = o.allow_loops;
    ctx->opt.direct_gemm = o.direct_gemm;
    ctx->opt.gemm_kernel = o.gemm_kernel;
  }
  ~OptScope() { ctx->opt = saved; }
};

int64_t prod(const std::vector<int>& ls, const std::vector<int64_t>& sz) {
  int64_t p = 1;
  for (int l : ls) p *= sz[l];
  return p;
}
bool has(const std::vector<int>& v, int l) { return std::find(v.begin(), v.end(), l) != v.end(); }
int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// First-fit placement over live blocks [off, off+size).
struct Arena {
  std::map<int64_t, int64_t> live;  // off -> size
  int64_t high = 0;
  int64_t place(int64_t size) {
    size = align256(size ? size : 1);
    int64_t cur = 0;
    for (auto& kv : live) {
      if (kv.first - cur >= size) break;
      cur = std::max(cur, kv.first + kv.second);
    }
    live[cur] = size;
    high = std::max(high, cur + size);
    return cur;
  }
  void release(int64_t off) { live.erase(off); }
};

std::vector<int> drop_sliced(const std::vector<int>& ls, const std::vector<int>& sl) {
  std::vector<int> out;
  for (int l : ls)
    if (!has(sl, l)) out.push_back(l);
  return out;
}

}  // namespace

// Plans the arena: every step output placed first-fit; children released after the parent.
// Fused steps (all evaluated by one launch before the other steps) are placed first: the
// outputs inside a fused subtree keep their blocks for the whole evaluation and a subtree's
// root output is released after its (unfused) parent, so the plan also holds when every step
// runs one by one in order (the normalised evaluation).
static void plan_arena(ome_tree* t, bool keep_all, int64_t* out_bytes) {
  Arena ar;
  const size_t es = elem_size(t->dt);
  const int n = (int)t->steps.size();
  auto is_fused = [&](int s) { return !t->fused.empty() && t->fused[s]; };
  for (int s = 0; s < n; ++s) {
    TreeStep& st = t->steps[s];
    st.out_elems = prod(drop_sliced(st.iy, t->slicing), t->sizes_sl);
    if (is_fused(s)) st.out_off = ar.place(st.out_elems * es);
  }
  for (int s = 0; s < n; ++s) {
    if (is_fused(s)) continue;
    TreeStep& st = t->steps[s];
    // the root of an unsliced, non-scalar-mode tree writes straight into y; it still gets
    // a block so scalar/normalised and sliced evaluation have somewhere to write
    st.out_off = ar.place(st.out_elems * es);
    if (!keep_all)
      for (auto& a : st.args)
        if (a.ref < 0) ar.release(t->steps[-(a.ref + 1)].out_off);
  }
  *out_bytes = align256(ar.high);
}

// Marks the steps to fuse: binary, not the root, unique labels on both operands and the
// output with none dangling or broadcast (after dropping sliced labels), a short contracted
// space and a small output; fused are the maximal subtrees made only of such steps.
static void plan_fusion(ome_tree* t) {
  const int n = (int)t->steps.size();
  t->fused.assign(n, 0);
  if (t->opt.gemm_kernel != GEMM_AUTO || t->n_leaves > SF_MAXLEAF ||
      (t->opt.precision == PREC_FP64 && (t->dt == F32 || t->dt == C32)))
    return;
  auto unique = [](const std::vector<int>& v) {
    for (size_t i = 0; i < v.size(); ++i)
      for (size_t j = i + 1; j < v.size(); ++j)
        if (v[i] == v[j]) return false;
    return true;
  };
  std::vector<char> ok(n, 0);
  for (int s = 0; s + 1 < n; ++s) {
    const TreeStep& st = t->steps[s];
    if (st.args.size() != 2) continue;
    const std::vector<int> x1 = drop_sliced(st.args[0].ix, t->slicing);
    const std::vector<int> x2 = drop_sliced(st.args[1].ix, t->slicing);
    const std::vector<int> y = drop_sliced(st.iy, t->slicing);
    if (!unique(x1) || !unique(x2) || !unique(y)) continue;
    bool cover = true;
    std::vector<int> inner;
    for (int l : x1) {
      if (!has(x2, l) && !has(y, l)) cover = false;
      if (has(x2, l) && !has(y, l)) inner.push_back(l);
    }
    for (int l : x2)
      if (!has(x1, l) && !has(y, l)) cover = false;
    for (int l : y)
      if (!has(x1, l) && !has(x2, l)) cover = false;
    if (!cover) continue;
    const int64_t nout = prod(y, t->sizes_sl), nk = prod(inner, t->sizes_sl);
    // (operands then hold at most 2^15 elements: 16-bit offsets in the item tables)
    ok[s] = nout <= SF_MAXOUT && nk <= SF_MAXNK && nout * nk <= (1 << 15);
  }
  int count = 0;
  for (int s = 0; s < n; ++s) {  // post-order: children first
    bool f = ok[s];
    for (auto& a : t->steps[s].args)
      if (a.ref < 0 && !t->fused[-(a.ref + 1)]) f = false;
    t->fused[s] = f;
    count += f;
  }
  if (count < 2) t->fused.assign(n, 0);  // nothing to gain from one fused step
}

// Device descriptors of the fused subtrees (after plan_arena has placed their outputs).
static Status upload_fusion(ome_tree* t) {
  const int n = (int)t->steps.size();
  const int64_t es = (int64_t)elem_size(t->dt);
  std::vector<FusedStep> desc;
  // per descriptor: the operand offsets of each output (A's | B's << 16)
  std::vector<std::vector<uint32_t>> offs;
  std::vector<int32_t> first, table;
  std::vector<int> level(n, 0), members;
  std::function<int(int)> collect = [&](int s) {  // children first
    int lv = 0;
    for (auto& a : t->steps[s].args)
      if (a.ref < 0) lv = std::max(lv, collect(-(a.ref + 1)));
    members.push_back(s);
    return level[s] = lv + 1;
  };
  for (int r = 0; r < n; ++r) {
    const int parent = t->steps[r].parent;
    if (!t->fused[r] || (parent >= 0 && t->fused[parent])) continue;  // subtree roots only
    members.clear();
    collect(r);
    // the LDS arena: outputs inside the subtree, then the leaves read above level 1, while
    // they fit (the rest stay in the plan's arena / in place)
    std::map<int, int64_t> out_lds, leaf_lds;
    std::map<int, int64_t> leaf_elems;
    int64_t used = 0;
    auto fit = [&](int64_t bytes, int64_t* off) {
      bytes = (bytes + 15) & ~(int64_t)15;
      if (used + bytes > SF_LDS) return false;
      *off = used;
      used += bytes;
      return true;
    };
    for (int s : members) {
      int64_t off;
      if (s != r && fit(t->steps[s].out_elems * es, &off)) out_lds[s] = off;
    }
    for (int s : members)
      for (auto& a : t->steps[s].args) {
        if (a.ref < 0 || level[s] < 2) continue;
        const int64_t e = prod(drop_sliced(a.ix, t->slicing), t->sizes_sl);
        int64_t off;
        if (fit(e * es, &off)) {
          leaf_lds[a.ref] = off;
          leaf_elems[a.ref] = e;
        }
      }
    // (level, step, leaf): a step, or the copy of a leaf into the LDS arena at level 1
    std::vector<std::array<int, 3>> items;
    for (int s : members) items.push_back({level[s], s, -1});
    for (auto& kv : leaf_lds) items.push_back({1, -1, kv.first});
    std::stable_sort(items.begin(), items.end(),
                     [](const std::array<int, 3>& x, const std::array<int, 3>& y) {
                       return x[0] < y[0];
                     });
    first.push_back((int32_t)table.size());
    for (size_t i = 0; i < items.size(); ++i) {
      if (i == 0 || items[i][0] != items[i - 1][0] || (int)desc.size() - table.back() == SF_LMAX)
        table.push_back((int32_t)desc.size());
      FusedStep d{};
      d.a_leaf = d.b_leaf = -1;
      std::vector<uint32_t> off;
      if (items[i][2] >= 0) {  // copy step: output o is element o of the leaf
        const int leaf = items[i][2];
        d.a_leaf = leaf;
        d.flags = FS_COPY | FS_Y_LDS;
        d.y_off = leaf_lds[leaf];
        d.nout = (int32_t)leaf_elems[leaf];
        d.nk = 1;
        for (int o = 0; o < d.nout; ++o) off.push_back((uint32_t)o);
      } else {
        const int s = items[i][1];
        const TreeStep& st = t->steps[s];
        const std::vector<int> x1 = drop_sliced(st.args[0].ix, t->slicing);
        const std::vector<int> x2 = drop_sliced(st.args[1].ix, t->slicing);
        const std::vector<int> y = drop_sliced(st.iy, t->slicing);
        auto stride_in = [&](const std::vector<int>& ls, int l) -> int64_t {
          int64_t v = 1;
          for (int l2 : ls) {
            if (l2 == l) return v;
            v *= t->sizes_sl[l2];
          }
          return 0;
        };
        auto operand = [&](const TreeArg& a, int32_t* leaf, int64_t* aoff, int32_t lds_flag) {
          if (a.ref >= 0) {
            auto it = leaf_lds.find(a.ref);
            if (it != leaf_lds.end()) {
              d.flags |= lds_flag;
              *aoff = it->second;
            } else {
              *leaf = a.ref;
            }
          } else {
            const int c = -(a.ref + 1);
            auto it = out_lds.find(c);
            if (it != out_lds.end()) {
              d.flags |= lds_flag;
              *aoff = it->second;
            } else {
              *aoff = t->steps[c].out_off;
            }
          }
        };
        operand(st.args[0], &d.a_leaf, &d.a_off, FS_A_LDS);
        operand(st.args[1], &d.b_leaf, &d.b_off, FS_B_LDS);
        auto yo = out_lds.find(s);
        if (yo != out_lds.end()) {
          d.flags |= FS_Y_LDS;
          d.y_off = yo->second;
        } else {
          d.y_off = st.out_off;
        }
        std::vector<int> inner;
        for (int l : x1)
          if (has(x2, l) && !has(y, l)) inner.push_back(l);
        d.nout = (int32_t)prod(y, t->sizes_sl);
        d.nk = (int32_t)prod(inner, t->sizes_sl);
        // operands hold at most nout * nk <= 32768 elements: offsets fit 16 bits
        auto digits = [&](int64_t idx, const std::vector<int>& ls, int64_t* oa, int64_t* ob) {
          *oa = *ob = 0;
          for (int l : ls) {
            const int64_t v = idx % t->sizes_sl[l];
            idx /= t->sizes_sl[l];
            *oa += v * stride_in(x1, l);
            *ob += v * stride_in(x2, l);
          }
        };
        for (int k = 0; k < d.nk; ++k) {  // contracted point k, first inner label fastest
          int64_t oa, ob;
          digits(k, inner, &oa, &ob);
          if (oa > 65535 || ob > 65535)
            return Status::err(OME_ERR_UNSUPPORTED, "fused step operand beyond 16-bit offsets");
          d.ka[k] = (uint16_t)oa;
          d.kb[k] = (uint16_t)ob;
        }
        for (int o = 0; o < d.nout; ++o) {  // output o, first label of y fastest
          int64_t oa, ob;
          digits(o, y, &oa, &ob);
          if (oa > 65535 || ob > 65535)
            return Status::err(OME_ERR_UNSUPPORTED, "fused step operand beyond 16-bit offsets");
          off.push_back((uint32_t)oa | ((uint32_t)ob << 16));
        }
      }
      desc.push_back(d);
      offs.push_back(std::move(off));
    }
  }
  first.push_back((int32_t)table.size());
  table.push_back((int32_t)desc.size());
  t->fsub = (int)first.size() - 1;
  if (t->fsub <= 0) return Status::ok();
  // per chunk: the work items per output (2^lg lanes splitting its contracted points: a power
  // of two, at most 16 and at most the chunk's largest contracted extent, while the workgroup
  // holds all the items), each step's items from a wave boundary so that a wave works on one
  // step; then the chunk's item table
  const int nch = (int)table.size() - 1;
  const size_t head = first.size() + table.size() + (size_t)nch + 1;
  std::vector<int32_t> ibases;
  std::vector<uint32_t> words;
  auto items_of = [](int64_t nout, int lg) { return ((nout << lg) + 63) & ~(int64_t)63; };
  for (int c = 0; c < nch; ++c) {
    int32_t maxnk = 1;
    for (int i = table[c]; i < table[c + 1]; ++i) maxnk = std::max(maxnk, desc[i].nk);
    auto count = [&](int lg) {
      int64_t m = 0;
      for (int i = table[c]; i < table[c + 1]; ++i) m += items_of(desc[i].nout, lg);
      return m;
    };
    int lg = 0;
    while (lg < 4 && (2 << lg) <= maxnk && count(lg + 1) <= SF_THREADS) ++lg;
    ibases.push_back((int32_t)(head + words.size()));
    int32_t ibase = 0;
    for (int i = table[c]; i < table[c + 1]; ++i) {
      desc[i].ibase = ibase;
      desc[i].lg = lg;
      const int64_t m = items_of(desc[i].nout, lg);
      for (int64_t it = 0; it < m; ++it) {
        const int64_t o = it >> lg;
        words.push_back(o < desc[i].nout ? offs[i][o] : 0u);
      }
      ibase += (int32_t)m;
    }
    for (int i = table[c]; i < table[c + 1]; ++i) desc[i].total = ibase;
  }
  ibases.push_back((int32_t)(head + words.size()));
  if (head + words.size() > (size_t)INT32_MAX)
    return Status::err(OME_ERR_UNSUPPORTED, "fused step tables beyond 2^31 words");
  first.insert(first.end(), table.begin(), table.end());
  first.insert(first.end(), ibases.begin(), ibases.end());
  for (uint32_t w : words) first.push_back((int32_t)w);
  ome_ctx* ctx = t->ctx;
  t->fdev = (FusedStep*)ctx->pool->alloc(desc.size() * sizeof(FusedStep));
  t->ffirst = (int32_t*)ctx->pool->alloc(first.size() * sizeof(int32_t));
  if (!t->fdev || !t->ffirst) return Status::err(OME_ERR_OOM, "device pool exhausted (fused steps)");
  if (hipMemcpyAsync(t->fdev, desc.data(), desc.size() * sizeof(FusedStep), hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess ||
      hipMemcpyAsync(t->ffirst, first.data(), first.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return Status::err(OME_ERR_DEVICE, "upload of the fused step descriptors failed");
  return Status::ok();
}

// Enqueue one evaluation of the (reduced) tree for one slice.  xs are the leaf pointers to
// use (staged copies for sliced leaves); the root writes into `root_out` with rsx/rsy.
static Status run_steps(ome_tree* t, const std::vector<const void*>& xs, void* root_out,
                        const Scal& rsx, const Scal& rsy, hipStream_t st, Bump* bump,
                        bool normalize, double* norm_work, double* logacc) {
  ome_ctx* ctx = t->ctx;
  const Scal one = scal_one(t->dt), zero = scal_zero(t->dt);
  const int n = (int)t->steps.size();
  // the fused subtrees first, in one launch (the normalised evaluation runs every step alone)
  const bool fuse = t->fsub > 0 && !normalize;
  if (fuse && !t->fused_elsewhere) {
    RangeScope rg("tree: fused small steps");
    FusedLeaves lv{};
    for (int k = 0; k < t->n_leaves; ++k) lv.p[k] = xs[k];
    launch_fused_steps(t->dt, t->fdev, t->ffirst, t->fsub, lv, t->mem, st);
  }
  for (int s = 0; s < n; ++s) {
    if (fuse && t->fused[s]) continue;
    TreeStep& step = t->steps[s];
    EinsumCall c;
    c.dt = t->dt;
    c.sizes = t->sizes_sl;
    for (auto& a : step.args) {
      c.ixs.push_back(drop_sliced(a.ix, t->slicing));
      if (a.ref >= 0) c.xs.push_back(xs[a.ref]);
      else c.xs.push_back(t->mem + t->steps[-(a.ref + 1)].out_off);
    }
    c.iy = drop_sliced(step.iy, t->slicing);
    std::string rname;
    if (ranges_enabled()) {  // "tree step s: ixs -> iy" with the labels as numbers
      rname = "tree step " + std::to_string(s) + ":";
      for (size_t q = 0; q < c.ixs.size(); ++q) {
        rname += q ? ", (" : " (";
        for (size_t l = 0; l < c.ixs[q].size(); ++l) rname += (l ? " " : "") + std::to_string(c.ixs[q][l]);
        rname += ")";
      }
      rname += " -> (";
      for (size_t l = 0; l < c.iy.size(); ++l) rname += (l ? " " : "") + std::to_string(c.iy[l]);
      rname += ")";
    }
    RangeScope rg(rname.c_str());
    const bool root = (s == n - 1);
    c.y = root ? root_out : (void*)(t->mem + step.out_off);
    c.sx = root ? rsx : one;
    c.sy = root ? rsy : zero;
    bump->reset();
    ctx->bump = bump;
    Status r = run_einsum(ctx, c, st);
    ctx->bump = nullptr;
    if (!r.good()) return r;
    if (normalize && !root)
      launch_normalize(t->dt, t->mem + step.out_off, step.out_elems, norm_work, logacc, st);
  }
  return Status::ok();
}

// Stage the slice `sidx` of every leaf that carries sliced labels (take_slice).
static void stage_leaves(ome_tree* t, const std::vector<const void*>& xs, int64_t sidx,
                         std::vector<const void*>& out, hipStream_t st) {
  out = xs;
  if (t->slicing.empty()) return;
  // slice values, first sliced label fastest (CartesianIndices order)
  std::vector<int64_t> val(t->sizes.size(), 0);
  int64_t r = sidx;
  for (int l : t->slicing) {
    val[l] = r % t->sizes[l];
    r /= t->sizes[l];
  }
  for (int k = 0; k < t->n_leaves; ++k) {
    if (t->stage_off[k] < 0) continue;
    const auto& ix = t->leaf_ix[k];
    int64_t stride = 1, base = 0;
    PermDesc d{};
    std::vector<int64_t> odim, ostr;
    for (int l : ix) {
      if (has(t->slicing, l)) base += val[l] * stride;
      else if (t->sizes[l] != 1) { odim.push_back(t->sizes[l]); ostr.push_back(stride); }
      stride *= t->sizes[l];
    }
    const size_t es = elem_size(t->dt);
    void* dst = t->mem + t->stage_off[k];
    const char* src = (const char*)xs[k] + base * es;
    int64_t nel = 1;
    for (auto v : odim) nel *= v;
    if (odim.empty()) { odim.push_back(1); ostr.push_back(1); }
    d.x = src;
    d.y = dst;
    d.n = nel;
    d.rank = (int)odim.size();
    for (int p = 0; p < d.rank; ++p) {
      d.odiv[p] = make_fdiv((uint32_t)odim[p]);
      d.istride[p] = ostr[p];
    }
    d.sx = scal_one(t->dt);
    d.sy = scal_zero(t->dt);
    d.mode = EPI_OVERWRITE;
    launch_permute(t->dt, d, st);
    out[k] = dst;
  }
}

// Add the slice result into the accumulator at the slice's position (fill_slice!).
static void accumulate_slice(ome_tree* t, int64_t sidx, void* acc, hipStream_t st) {
  std::vector<int64_t> val(t->sizes.size(), 0);
  int64_t r = sidx;
  for (int l : t->slicing) {
    val[l] = r % t->sizes[l];
    r /= t->sizes[l];
  }
  AccDesc d{};
  d.x = t->mem + t->slice_out_off;
  d.acc = acc;
  int64_t stride = 1;
  std::vector<int64_t> dim, dstr;
  for (int l : t->root_iy) {
    if (has(t->slicing, l)) d.base += val[l] * stride;
    else if (t->sizes[l] != 1) { dim.push_back(t->sizes[l]); dstr.push_back(stride); }
    stride *= t->sizes[l];
  }
  d.n = 1;
  for (auto v : dim) d.n *= v;
  d.rank = (int)dim.size();
  for (int p = 0; p < d.rank; ++p) {
    d.div[p] = make_fdiv((uint32_t)dim[p]);
    d.dstride[p] = dstr[p];
  }
  launch_acc_strided(t->dt, t->acc_f64, d, st);
}

// Give the plan its own device memory if it has none (plans created with deferred memory,
// or whose memory a request loop took back).  Recorded graphs are keyed by the memory they
// were recorded with and replay only while the plan holds that memory again.
static Status ensure_mem(ome_tree* t) {
  if (t->mem) return Status::ok();
  t->mem = (char*)t->ctx->pool->alloc((size_t)t->mem_bytes);
  if (!t->mem)
    return Status::err(OME_ERR_OOM, "device pool exhausted (tree arena of " +
                                        std::to_string(t->mem_bytes) + " bytes)");
  t->owns_mem = true;
  return Status::ok();
}

// One full evaluation (all slices) enqueued on st.
static Status enqueue_eval(ome_tree* t, const std::vector<const void*>& xs, void* y,
                           const Scal& sx, const Scal& sy, hipStream_t st) {
  Bump bump;
  bump.base = t->mem + t->arena_bytes;
  bump.cap = (size_t)t->scratch_bytes;
  const Scal one = scal_one(t->dt), zero = scal_zero(t->dt);
  if (t->slicing.empty())
    return run_steps(t, xs, y, sx, sy, st, &bump, false, nullptr, nullptr);
  void* acc = t->mem + t->arena_bytes + t->scratch_bytes + t->stage_bytes;
  const int64_t ny = prod(t->root_iy, t->sizes);
  if (t->acc_f64) {
    launch_fill(F64, acc, ny * (is_complex(t->dt) ? 2 : 1), Scal{0, 0, 0}, st);
  } else {
    launch_fill(t->dt, acc, ny, zero, st);
  }
  std::vector<const void*> sx_leaves;
  for (int64_t s = 0; s < t->nslices; ++s) {
    stage_leaves(t, xs, s, sx_leaves, st);
    Status r = run_steps(t, sx_leaves, t->mem + t->slice_out_off, one, zero, st, &bump, false,
                         nullptr, nullptr);
    if (!r.good()) return r;
    accumulate_slice(t, s, acc, st);
  }
  if (t->acc_f64) launch_from_f64(t->dt, (const double*)acc, y, ny, sx, sy, st);
  else launch_axpby(t->dt, acc, y, ny, sx, sy, 0, st);
  return Status::ok();
}

// One full evaluation as a graph launched on `st`: the graph recorded for the plan's current
// memory, inputs, output and scalars is replayed (LRU of 8 per plan), or recorded first on
// the plan's capture stream.  Falls back to direct launches when capture is unavailable.
static Status launch_graphed(ome_tree* t, const std::vector<const void*>& xv, void* y,
                             const Scal& sx, const Scal& sy, hipStream_t st) {
  auto same = [&](const GraphEntry& e) {
    return e.mem == t->mem && e.fused_elsewhere == t->fused_elsewhere && e.xs == xv && e.y == y &&
           std::memcmp(&e.sx, &sx, sizeof sx) == 0 && std::memcmp(&e.sy, &sy, sizeof sy) == 0;
  };
  auto it = std::find_if(t->graphs.begin(), t->graphs.end(), same);
  if (it == t->graphs.end()) {
    RangeScope rg("tree: graph record");
    hipGraph_t graph;
    if (hipStreamBeginCapture(t->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      hipGetLastError();
      return enqueue_eval(t, xv, y, sx, sy, st);
    }
    Status r = enqueue_eval(t, xv, y, sx, sy, t->cap_stream);
    hipError_t e = hipStreamEndCapture(t->cap_stream, &graph);
    if (!r.good()) {
      if (e == hipSuccess) hipGraphDestroy(graph);
      return r;
    }
    if (e != hipSuccess) {
      hipGetLastError();
      return enqueue_eval(t, xv, y, sx, sy, st);
    }
    GraphEntry ge;
    ge.mem = t->mem;
    ge.fused_elsewhere = t->fused_elsewhere;
    ge.xs = xv;
    ge.y = y;
    ge.sx = sx;
    ge.sy = sy;
    e = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      hipGetLastError();
      return enqueue_eval(t, xv, y, sx, sy, st);
    }
    t->graphs.push_front(ge);
    if (t->graphs.size() > 8) {
      hipGraphExecDestroy(t->graphs.back().exec);
      t->graphs.pop_back();
    }
    it = t->graphs.begin();
  } else if (it != t->graphs.begin()) {
    t->graphs.splice(t->graphs.begin(), t->graphs, it);
    it = t->graphs.begin();
  }
  RangeScope rg("tree: graph replay");
  const hipError_t e = hipGraphLaunch(it->exec, st);
  if (e != hipSuccess)
    return Status::err(OME_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
  return Status::ok();
}

static int tree_fail(ome_ctx* ctx, const Status& s) {
  if (ctx) ctx->last_error = s.msg;
  return s.code;
}

extern "C" {

int ome_tree_create(ome_ctx* ctx, int dtype, const int32_t* desc, int64_t desc_len, int nlabels,
                    const int64_t* sizes, int nslice, const int32_t* slice_labels, ome_tree** out) {
  if (!ctx || !desc || !out || desc_len < 2 || nlabels < 0 || (nlabels && !sizes))
    return tree_fail(ctx, Status::err(OME_ERR_INVALID_ARG, "invalid tree arguments"));
  if (dtype < F32 || dtype > DT_LAST) return tree_fail(ctx, Status::err(OME_ERR_INVALID_ARG, "unknown element type"));
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  ome_tree* t = new ome_tree();
  t->ctx = ctx;
  t->dt = dtype;
  t->opt = ctx->opt;
  t->sizes.assign(sizes, sizes + nlabels);
  for (int64_t s : t->sizes)
    if (s < 0) { delete t; return tree_fail(ctx, Status::err(OME_ERR_INVALID_ARG, "negative size")); }
  // ---- parse the post-order description
  int64_t p = 0;
  auto rd = [&](int32_t& v) {
    if (p >= desc_len) return false;
    v = desc[p++];
    return true;
  };
  int32_t nleaves, nnodes;
  bool ok = rd(nleaves) && rd(nnodes) && nleaves >= 1 && nnodes >= 1;
  t->n_leaves = nleaves;
  t->leaf_ix.assign(nleaves > 0 ? nleaves : 0, {});
  std::vector<int> leaf_seen(nleaves > 0 ? nleaves : 0, 0);
  for (int s = 0; ok && s < nnodes; ++s) {
    TreeStep st;
    int32_t nargs;
    ok = rd(nargs) && nargs >= 1 && nargs <= OME_MAX_INPUTS;
    for (int a = 0; ok && a < nargs; ++a) {
      TreeArg ta;
      int32_t ref, nix;
      ok = rd(ref) && rd(nix) && nix >= 0 && nix <= OME_MAX_RANK;
      for (int q = 0; ok && q < nix; ++q) {
        int32_t l;
        ok = rd(l) && l >= 0 && l < nlabels;
        ta.ix.push_back(l);
      }
      if (!ok) break;
      ta.ref = ref;
      if (ref >= 0) {
        ok = ref < nleaves && !leaf_seen[ref];
        if (ok) { leaf_seen[ref] = 1; t->leaf_ix[ref] = ta.ix; }
      } else {
        int child = -(ref + 1);
        ok = child < s && t->steps[child].parent < 0 && t->steps[child].iy.size() == ta.ix.size();
        if (ok) {
          for (size_t q = 0; q < ta.ix.size(); ++q) ok = ok && t->steps[child].iy[q] == ta.ix[q];
          t->steps[child].parent = s;
        }
      }
      st.args.push_back(ta);
    }
    int32_t niy;
    ok = ok && rd(niy) && niy >= 0 && niy <= OME_MAX_RANK;
    for (int q = 0; ok && q < niy; ++q) {
      int32_t l;
      ok = rd(l) && l >= 0 && l < nlabels;
      st.iy.push_back(l);
    }
    t->steps.push_back(st);
  }
  for (int k = 0; ok && k < nleaves; ++k) ok = leaf_seen[k];
  for (int s = 0; ok && s + 1 < nnodes; ++s) ok = t->steps[s].parent >= 0;
  if (!ok || p != desc_len) {
    delete t;
    return tree_fail(ctx, Status::err(OME_ERR_INVALID_ARG, "malformed tree description"));
  }
  t->root_iy = t->steps.back().iy;
  t->desc.assign(desc, desc + desc_len);
  // ---- slicing
  for (int q = 0; q < nslice; ++q) {
    int l = slice_labels[q];
    if (l < 0 || l >= nlabels || has(t->slicing, l)) {
      delete t;
      return tree_fail(ctx, Status::err(OME_ERR_INVALID_ARG, "invalid slice label"));
    }
    t->slicing.push_back(l);
    t->nslices *= t->sizes[l];
  }
  t->sizes_sl = t->sizes;
  for (int l : t->slicing) t->sizes_sl[l] = 1;
  const size_t es = elem_size(dtype);
  // ---- memory plan
  plan_fusion(t);
  plan_arena(t, false, &t->arena_bytes);
  {
    Status f = upload_fusion(t);
    if (!f.good()) {
      if (t->fdev) ctx->pool->free(t->fdev);
      if (t->ffirst) ctx->pool->free(t->ffirst);
      delete t;
      return tree_fail(ctx, f);
    }
  }
  t->slice_out_off = t->steps.back().out_off;
  t->stage_off.assign(nleaves, -1);
  int64_t stage = 0;
  if (!t->slicing.empty()) {
    for (int k = 0; k < nleaves; ++k) {
      bool sl = false;
      for (int l : t->leaf_ix[k]) sl = sl || has(t->slicing, l);
      if (!sl) continue;
      t->stage_off[k] = stage;
      stage += align256((int64_t)prod(drop_sliced(t->leaf_ix[k], t->slicing), t->sizes_sl) * es);
    }
    t->acc_f64 = is_numeric_float(dtype);
    t->acc_bytes = align256(prod(t->root_iy, t->sizes) * (t->acc_f64 ? (is_complex(dtype) ? 16 : 8) : (int64_t)es));
  }
  t->stage_bytes = stage;
  t->norm_bytes = align256((1 + 1024 + 2) * 8);
  // ---- size the scratch with a dry run (nothing is launched)
  {
    LaunchMode& lm = launch_mode();
    lm.dry = true;
    lm.count = 0;
    Bump bump;
    bump.base = (char*)(uintptr_t)4096;  // never dereferenced in a dry run
    bump.cap = (size_t)1 << 60;
    std::vector<const void*> fake(nleaves, (const void*)(uintptr_t)4096);
    t->mem = (char*)(uintptr_t)4096;
    Status r = run_steps(t, fake, (void*)(uintptr_t)4096, scal_one(dtype), scal_zero(dtype),
                         ctx->stream, &bump, false, nullptr, nullptr);
    lm.dry = false;
    t->mem = nullptr;
    if (!r.good()) {
      if (t->fdev) ctx->pool->free(t->fdev);
      if (t->ffirst) ctx->pool->free(t->ffirst);
      delete t;
      return tree_fail(ctx, r);
    }
    t->scratch_bytes = align256((int64_t)bump.high + 256);
    t->info.n_kernels = lm.count * t->nslices +
                        (t->slicing.empty() ? 0 : t->nslices * 2 + 2);
  }
  // stage offsets are relative to the staging region: rebase them
  const int64_t stage_base = t->arena_bytes + t->scratch_bytes;
  for (auto& o : t->stage_off)
    if (o >= 0) o += stage_base;
  const int64_t total = t->arena_bytes + t->scratch_bytes + t->stage_bytes + t->acc_bytes + t->norm_bytes;
  t->mem_bytes = total;
  if (!ctx->opt.defer_tree_mem) {
    t->mem = (char*)ctx->pool->alloc((size_t)total);
    if (!t->mem) {
      if (t->fdev) ctx->pool->free(t->fdev);
      if (t->ffirst) ctx->pool->free(t->ffirst);
      delete t;
      return tree_fail(ctx, Status::err(OME_ERR_OOM, "device pool exhausted (tree arena of " +
                                                         std::to_string(total) + " bytes)"));
    }
  }
  // ---- info
  t->info.n_steps = nnodes;
  t->info.n_leaves = nleaves;
  t->info.n_slices = t->nslices;
  t->info.arena_bytes = total;
  double flops = 0, bytes = 0;
  const double per = is_complex(dtype) ? 8.0 : 2.0;
  for (auto& st : t->steps) {
    std::vector<int> all;
    for (auto& a : st.args)
      for (int l : drop_sliced(a.ix, t->slicing))
        if (!has(all, l)) all.push_back(l);
    for (int l : drop_sliced(st.iy, t->slicing))
      if (!has(all, l)) all.push_back(l);
    flops += per * (double)prod(all, t->sizes_sl);
    for (auto& a : st.args) bytes += (double)prod(drop_sliced(a.ix, t->slicing), t->sizes_sl) * es;
    bytes += (double)prod(drop_sliced(st.iy, t->slicing), t->sizes_sl) * es;
  }
  t->info.flops = flops * t->nslices;
  t->info.bytes = bytes * t->nslices;
  // peak_memory measure: earlier siblings' results live while the next is evaluated
  {
    std::vector<int64_t> peak(nnodes, 0);
    for (int s = 0; s < nnodes; ++s) {
      int64_t largest = 0, temp = 0;
      for (auto& a : t->steps[s].args) {
        int64_t m = prod(drop_sliced(a.ix, t->slicing), t->sizes_sl);
        int64_t sub = a.ref >= 0 ? m : peak[-(a.ref + 1)];
        largest = std::max(largest, sub + temp);
        temp += m;
      }
      peak[s] = std::max(largest, temp + prod(drop_sliced(t->steps[s].iy, t->slicing), t->sizes_sl));
    }
    t->info.peak_live_elems = peak[nnodes - 1];
  }
  if (hipStreamCreateWithFlags(&t->cap_stream, hipStreamNonBlocking) != hipSuccess) {
    hipGetLastError();
    t->cap_stream = nullptr;
  }
  *out = t;
  return OME_OK;
}

int ome_get_tree_info(ome_tree* t, ome_tree_info* info) {
  if (!t || !info) return OME_ERR_INVALID_ARG;
  *info = t->info;
  return OME_OK;
}

int ome_tree_execute(ome_tree* t, const void* const* xs, void* y, const ome_scalar* sx_,
                     const ome_scalar* sy_) {
  if (!t || !xs) return tree_fail(t ? t->ctx : nullptr, Status::err(OME_ERR_INVALID_ARG, "null argument"));
  if (prod(t->root_iy, t->sizes) == 0) return OME_OK;
  if (!y) return tree_fail(t->ctx, Status::err(OME_ERR_INVALID_ARG, "null output pointer"));
  ome_ctx* ctx = t->ctx;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  OptScope os(ctx, t->opt);
  hipSetDevice(ctx->device);
  const Scal sx = sx_ ? Scal{sx_->re, sx_->im, sx_->i} : scal_one(t->dt);
  const Scal sy = sy_ ? Scal{sy_->re, sy_->im, sy_->i} : scal_zero(t->dt);
  std::vector<const void*> xv(xs, xs + t->n_leaves);
  for (int k = 0; k < t->n_leaves; ++k)
    if (!xv[k] && prod(t->leaf_ix[k], t->sizes) != 0)
      return tree_fail(ctx, Status::err(OME_ERR_INVALID_ARG, "null input pointer"));
  {
    Status m = ensure_mem(t);
    if (!m.good()) return tree_fail(ctx, m);
  }
  RangeScope rg("tree evaluation");
  if (!ctx->opt.use_graphs || !t->cap_stream)
    return tree_fail(ctx, enqueue_eval(t, xv, y, sx, sy, ctx->stream)) ;
  return tree_fail(ctx, launch_graphed(t, xv, y, sx, sy, ctx->stream));
}

}  // extern "C"

static double* norm_work(ome_tree* t) {
  return (double*)(t->mem + t->arena_bytes + t->scratch_bytes + t->stage_bytes + t->acc_bytes);
}

// Every slice of t evaluated one step at a time, each intermediate divided by its largest
// magnitude with the logs summed in FP64; the slice values (of t's element type) and logs
// are combined on the host in FP64.  base_log is added to every slice's log.
static Status normalized_slices(ome_tree* t, const std::vector<const void*>& xv, double base_log,
                                double* value, double* log_scale) {
  ome_ctx* ctx = t->ctx;
  hipStream_t st = ctx->stream;
  {
    Status m = ensure_mem(t);
    if (!m.good()) return m;
  }
  double* work = norm_work(t);
  double* logacc = work + 1 + 1024;
  const size_t es = elem_size(t->dt);
  Bump bump;
  bump.base = t->mem + t->arena_bytes;
  bump.cap = (size_t)t->scratch_bytes;
  // per slice: value and log scale, combined on the host in FP64
  double sum_re = 0, sum_im = 0, lmax = -INFINITY;
  std::vector<double> vr(t->nslices), vi(t->nslices), vl(t->nslices);
  std::vector<const void*> leaves;
  for (int64_t s = 0; s < t->nslices; ++s) {
    launch_fill(F64, logacc, 1, Scal{0, 0, 0}, st);
    stage_leaves(t, xv, s, leaves, st);
    Status r = run_steps(t, leaves, t->mem + t->slice_out_off, scal_one(t->dt), scal_zero(t->dt),
                         st, &bump, true, work, logacc);
    if (!r.good()) return r;
    char host[16];
    double lg;
    if (hipMemcpyAsync(host, t->mem + t->slice_out_off, es, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&lg, logacc, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return Status::err(OME_ERR_DEVICE, "copy of the scalar result failed");
    double re = 0, im = 0;
    switch (t->dt) {
      case F32: re = *(float*)host; break;
      case F64: re = *(double*)host; break;
      case C32: re = ((float*)host)[0]; im = ((float*)host)[1]; break;
      case C64: re = ((double*)host)[0]; im = ((double*)host)[1]; break;
    }
    vr[s] = re; vi[s] = im; vl[s] = lg + base_log;
    if (re != 0 || im != 0) lmax = std::max(lmax, vl[s]);
  }
  if (!std::isfinite(lmax)) lmax = 0;
  for (int64_t s = 0; s < t->nslices; ++s) {
    if (vr[s] == 0 && vi[s] == 0) continue;  // a zero slice adds nothing (and its log may be any)
    double f = std::exp(vl[s] - lmax);
    sum_re += vr[s] * f;
    sum_im += vi[s] * f;
  }
  value[0] = sum_re;
  value[1] = sum_im;
  *log_scale = lmax;
  return Status::ok();
}

// Float32 / ComplexF32 trees: every input tensor is divided by its largest magnitude while it
// is widened to Float64 / ComplexF64 (the logs summed in FP64), and the plan's FP64 twin
// evaluates the network with per-step normalisation, the final scalar included, in FP64.
static Status normalized_wide(ome_tree* t, const void* const* xs, double* value, double* log_scale) {
  ome_ctx* ctx = t->ctx;
  hipStream_t st = ctx->stream;
  const int wdt = t->dt == F32 ? F64 : C64;
  if (!t->wide) {
    const ome::Options saved = ctx->opt;
    ctx->opt.defer_tree_mem = 0;
    const int rc = ome_tree_create(ctx, wdt, t->desc.data(), (int64_t)t->desc.size(),
                                   (int)t->sizes.size(), t->sizes.data(), (int)t->slicing.size(),
                                   t->slicing.data(), &t->wide);
    ctx->opt = saved;
    if (rc != OME_OK) {
      t->wide = nullptr;
      return Status::err(rc, ctx->last_error);
    }
  }
  ome_tree* w = t->wide;
  {
    Status m = ensure_mem(w);
    if (!m.good()) return m;
  }
  const size_t wes = elem_size(wdt);
  if (!t->wleaves) {
    int64_t off = 0;
    t->wleaf_off.assign(t->n_leaves, 0);
    for (int k = 0; k < t->n_leaves; ++k) {
      t->wleaf_off[k] = off;
      off += align256(prod(t->leaf_ix[k], t->sizes) * (int64_t)wes);
    }
    t->wleaves = (char*)ctx->pool->alloc((size_t)std::max<int64_t>(off, 256));
    if (!t->wleaves) return Status::err(OME_ERR_OOM, "device pool exhausted (widened inputs)");
  }
  double* work = norm_work(w);
  double* leaflog = work + 1 + 1024 + 1;
  launch_fill(F64, leaflog, 1, Scal{0, 0, 0}, st);
  std::vector<const void*> wx(t->n_leaves);
  for (int k = 0; k < t->n_leaves; ++k) {
    const int64_t n = prod(t->leaf_ix[k], t->sizes);
    wx[k] = t->wleaves + t->wleaf_off[k];
    if (n > 0 && !xs[k]) return Status::err(OME_ERR_INVALID_ARG, "null input pointer");
    launch_widen_normalized(t->dt, xs[k], t->wleaves + t->wleaf_off[k], n, work, leaflog, st);
  }
  double base = 0;
  if (hipMemcpyAsync(&base, leaflog, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return Status::err(OME_ERR_DEVICE, "copy of the input scales failed");
  OptScope os(ctx, w->opt);
  return normalized_slices(w, wx, base, value, log_scale);
}

extern "C" {

int ome_tree_execute_scalar_normalized(ome_tree* t, const void* const* xs, double* value,
                                       double* log_scale) {
  if (!t || !xs || !value || !log_scale)
    return tree_fail(t ? t->ctx : nullptr, Status::err(OME_ERR_INVALID_ARG, "null argument"));
  ome_ctx* ctx = t->ctx;
  if (prod(t->root_iy, t->sizes) != 1 || !is_numeric_float(t->dt))
    return tree_fail(ctx, Status::err(OME_ERR_INVALID_ARG,
                                      "normalised evaluation needs a float/complex scalar output"));
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  OptScope os(ctx, t->opt);
  hipSetDevice(ctx->device);
  Status r = (t->dt == F32 || t->dt == C32)
                 ? normalized_wide(t, xs, value, log_scale)
                 : normalized_slices(t, std::vector<const void*>(xs, xs + t->n_leaves), 0.0, value,
                                     log_scale);
  return r.good() ? OME_OK : tree_fail(ctx, r);
}

int ome_tree_gradient(ome_tree* t, const void* const* xs, const void* ybar, void* y,
                      void* const* xbars) {
  if (!t || !xs || !ybar || !xbars)
    return tree_fail(t ? t->ctx : nullptr, Status::err(OME_ERR_INVALID_ARG, "null argument"));
  ome_ctx* ctx = t->ctx;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  OptScope os(ctx, t->opt);
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const size_t es = elem_size(t->dt);
  const int n = (int)t->steps.size();
  const Scal one = scal_one(t->dt), zero = scal_zero(t->dt);
  // gradients run unsliced (src/bp.jl:26-29,73-76): the full sizes of every label
  const std::vector<int64_t>& sz = t->sizes;
  std::vector<int64_t> out_off(n), grad_off(n);
  int64_t off = 0;
  for (int s = 0; s < n; ++s) {
    out_off[s] = off;
    off += align256(prod(t->steps[s].iy, sz) * es);
  }
  for (int s = 0; s < n; ++s) {
    grad_off[s] = off;
    off += align256(prod(t->steps[s].iy, sz) * es);
  }
  Bump bump;
  char* mem = nullptr;
  auto run = [&](EinsumCall& c) {
    bump.reset();
    ctx->bump = &bump;
    Status r = run_einsum(ctx, c, st);
    ctx->bump = nullptr;
    return r;
  };
  auto ptr_of = [&](const TreeArg& a) -> const void* {
    return a.ref >= 0 ? xs[a.ref] : (const void*)(mem + out_off[-(a.ref + 1)]);
  };
  // one pass = cached forward (cached_einsum, src/bp.jl:33-48) then back_propagate
  // (src/bp.jl:79-96) with einsum_grad (src/autodiff.jl:24-30)
  auto pass = [&]() -> Status {
    Status r;
    for (int s = 0; r.good() && s < n; ++s) {
      EinsumCall c;
      c.dt = t->dt;
      c.sizes = sz;
      for (auto& a : t->steps[s].args) {
        c.ixs.push_back(a.ix);
        c.xs.push_back(ptr_of(a));
      }
      c.iy = t->steps[s].iy;
      c.y = mem + out_off[s];
      r = run(c);
    }
    if (r.good() && y)
      launch_axpby(t->dt, mem + out_off[n - 1], y, prod(t->root_iy, sz), one, zero, 0, st);
    if (r.good())
      launch_axpby(t->dt, ybar, mem + grad_off[n - 1], prod(t->root_iy, sz), one, zero, 0, st);
    for (int s = n - 1; r.good() && s >= 0; --s) {
      const TreeStep& step = t->steps[s];
      const void* dy = mem + grad_off[s];
      for (size_t i = 0; r.good() && i < step.args.size(); ++i) {
        EinsumCall c;
        c.dt = t->dt;
        c.sizes = sz;
        for (size_t j = 0; j < step.args.size(); ++j) {
          if (j == i) {
            c.ixs.push_back(step.iy);
            c.xs.push_back(dy);
            c.conj.push_back(1);
          } else {
            c.ixs.push_back(step.args[j].ix);
            c.xs.push_back(ptr_of(step.args[j]));
            c.conj.push_back(0);
          }
        }
        const TreeArg& a = step.args[i];
        c.iy = a.ix;
        void* dst = a.ref >= 0 ? xbars[a.ref] : (void*)(mem + grad_off[-(a.ref + 1)]);
        if (!dst) continue;
        c.y = dst;
        r = run(c);
        if (r.good() && is_complex(t->dt))
          launch_axpby(t->dt, dst, dst, prod(a.ix, sz), one, zero, 1, st);
      }
    }
    return r;
  };
  // size the scratch with a dry run, then run for real
  Status r;
  {
    LaunchMode& lm = launch_mode();
    lm.dry = true;
    mem = (char*)(uintptr_t)4096;
    bump.base = mem;
    bump.cap = (size_t)1 << 60;
    bump.high = 0;
    r = pass();
    lm.dry = false;
  }
  if (!r.good()) return tree_fail(ctx, r);
  const int64_t scratch = align256((int64_t)bump.high + 256);
  mem = (char*)ctx->pool->alloc((size_t)(off + scratch));
  if (!mem) return tree_fail(ctx, Status::err(OME_ERR_OOM, "device pool exhausted (gradient)"));
  bump.base = mem + off;
  bump.cap = (size_t)scratch;
  bump.high = 0;
  r = pass();
  // release after the stream drains: the pool reuses blocks in stream order, but graph-free
  // callers may switch streams, so wait here (gradients are not on the hot replay path)
  hipStreamSynchronize(st);
  ctx->pool->free(mem);
  return tree_fail(ctx, r);
}

int ome_tree_destroy(ome_tree* t) {
  if (!t) return OME_ERR_INVALID_ARG;
  ome_ctx* ctx = t->ctx;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (auto& ge : t->graphs) hipGraphExecDestroy(ge.exec);
  if (t->cap_stream) hipStreamDestroy(t->cap_stream);
  if (t->mem && t->owns_mem) ctx->pool->free(t->mem);
  if (t->fdev) ctx->pool->free(t->fdev);
  if (t->ffirst) ctx->pool->free(t->ffirst);
  if (t->wleaves) ctx->pool->free(t->wleaves);
  if (t->wide) ome_tree_destroy(t->wide);
  if (t->fit) ome_tree_destroy(t->fit);
  delete t;
  return OME_OK;
}

}  // extern "C"

// ---- request-loop access (csrc/server.hip) ------------------------------------------------

namespace ome {

int64_t tree_mem_bytes(const ome_tree* t) { return t->mem_bytes; }

// Evaluate t once on st in memory lent by the caller (mem_bytes bytes): the plan's own
// memory, if any, goes back to the pool and its graphs are dropped.  y <- result.
Status tree_run_lent(ome_tree* t, const void* const* xs, void* y, char* mem, hipStream_t st,
                     bool fused_done) {
  ome_ctx* ctx = t->ctx;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  OptScope os(ctx, t->opt);
  hipSetDevice(ctx->device);
  if (t->mem != mem) {
    if (t->mem && t->owns_mem) ctx->pool->free(t->mem);
    t->mem = mem;
    t->owns_mem = false;
  }
  if (prod(t->root_iy, t->sizes) == 0) return Status::ok();
  std::vector<const void*> xv(xs, xs + t->n_leaves);
  for (int k = 0; k < t->n_leaves; ++k)
    if (!xv[k] && prod(t->leaf_ix[k], t->sizes) != 0)
      return Status::err(OME_ERR_INVALID_ARG, "null input pointer");
  // a plan lent the same memory again (the loop hands blocks back to the plan that used them)
  // replays its recorded graph: no host planning of the steps
  t->fused_elsewhere = fused_done && t->fsub > 0 && t->slicing.empty();
  Status r = (t->opt.use_graphs && t->cap_stream)
                 ? launch_graphed(t, xv, y, scal_one(t->dt), scal_zero(t->dt), st)
                 : enqueue_eval(t, xv, y, scal_one(t->dt), scal_zero(t->dt), st);
  t->fused_elsewhere = false;
  return r;
}

// The plan's fused subtrees as a job of a shared launch (unsliced plans that have them).
bool tree_fused_job(const ome_tree* t, FusedJob* job) {
  if (t->fsub <= 0 || !t->slicing.empty() || t->n_leaves > SF_MAXLEAF) return false;
  job->steps = t->fdev;
  job->first = t->ffirst;
  job->nsub = t->fsub;
  return true;
}

int tree_dtype(const ome_tree* t) { return t->dt; }

// The plan the request loop admits for t under a per-request limit of cap bytes: t itself
// when its memory fits; else a twin of the same tree with more labels sliced (kept with t and
// reused while the limit is unchanged).  Labels are added one at a time, each time the label
// (not in the output, of size > 1) carried by the most elements of the step outputs, until
// the twin's memory fits; fails when slicing every such label does not make it fit.
Status tree_fit(ome_tree* t, int64_t cap, ome_tree** out) {
  *out = t;
  if (t->mem_bytes <= cap) return Status::ok();
  ome_ctx* ctx = t->ctx;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  if (t->fit && t->fit_cap == cap) {
    *out = t->fit;
    return Status::ok();
  }
  if (t->fit) {
    ome_tree_destroy(t->fit);
    t->fit = nullptr;
  }
  std::vector<int> slicing = t->slicing;
  std::vector<int64_t> sz = t->sizes;
  for (int l : slicing) sz[l] = 1;
  OptScope os(ctx, t->opt);
  const ome::Options saved = ctx->opt;
  ctx->opt.defer_tree_mem = 1;
  Status result = Status::err(OME_ERR_OOM, "no slicing of this tree fits " + std::to_string(cap) +
                                               " bytes of device memory");
  for (;;) {
    std::vector<double> score(sz.size(), 0.0);
    for (const auto& st : t->steps) {
      const double e = (double)prod(st.iy, sz);
      for (int l : st.iy)
        if (sz[l] > 1 && !has(t->root_iy, l)) score[l] += e;
    }
    int best = -1;
    for (size_t l = 0; l < score.size(); ++l)
      if (score[l] > 0 && (best < 0 || score[l] > score[best])) best = (int)l;
    if (best < 0) break;
    slicing.push_back(best);
    sz[best] = 1;
    ome_tree* c = nullptr;
    const int rc = ome_tree_create(ctx, t->dt, t->desc.data(), (int64_t)t->desc.size(),
                                   (int)t->sizes.size(), t->sizes.data(), (int)slicing.size(),
                                   slicing.data(), &c);
    if (rc != OME_OK) {
      result = Status::err(rc, ctx->last_error);
      break;
    }
    if (c->mem_bytes <= cap) {
      t->fit = c;
      t->fit_cap = cap;
      *out = c;
      result = Status::ok();
      break;
    }
    ome_tree_destroy(c);
  }
  ctx->opt = saved;
  return result;
}

// Forget lent memory (the lender frees it).
void tree_return_lent(ome_tree* t) {
  std::lock_guard<std::recursive_mutex> g(t->ctx->mu);
  if (!t->owns_mem) t->mem = nullptr;
}

}  // namespace ome
