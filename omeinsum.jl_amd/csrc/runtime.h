// Internal runtime of the backend: the context, the device memory pool and the planner
// entry points shared by the C ABI, the tree executor and the request loop.
#pragma once
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/omeinsum_mi355x.h"
#include "kernels.h"

namespace ome {

// Caching device allocator.  Requests are rounded to size classes (eighths of a power of
// two above 4 KiB), freed blocks are kept per class and reused; when hipMalloc fails the
// cached blocks are returned to the driver and the request is retried.  Blocks are reused
// in stream order on the owning context's stream.
class Pool {
 public:
  explicit Pool(int device) : device_(device) {}
  ~Pool();
  void* alloc(size_t bytes);  // nullptr when out of memory
  void free(void* p);
  void release_cached();
  ome_pool_stats stats();

 private:
  static size_t round_class(size_t bytes);
  int device_;
  std::mutex mu_;
  std::map<size_t, std::vector<void*>> free_;
  std::unordered_map<void*, size_t> live_;
  int64_t reserved_ = 0, in_use_ = 0, peak_ = 0, n_allocs_ = 0, n_mallocs_ = 0;
};

struct Options {
  int precision = PREC_FP32;
  int complex_algo = 0;  // 0: 4M, 1: 3M
  int allow_loops = 1;
  int direct_gemm = 1;
  int use_graphs = 1;
  int autotune = 1;
  int gemm_kernel = GEMM_AUTO;
  int grouped_variant = -1;  // -1: autotuned; else the grouped GEMM tile variant
  int defer_tree_mem = 0;    // 1: tree plans get device memory on first use, not at creation
  int grouped_buffers = 1;   // 1: grouped GEMMs may use the buffer-resource kernels
};

struct Status {
  int code = OME_OK;
  std::string msg;
  static Status ok() { return Status(); }
  static Status err(int c, std::string m) {
    Status s;
    s.code = c;
    s.msg = std::move(m);
    return s;
  }
  bool good() const { return code == OME_OK; }
};

}  // namespace ome

namespace ome {
// Bump allocator over a caller-owned device buffer.  While installed on a context, scratch
// requests of the planner come from here (used by the tree executor so that a recorded
// graph keeps valid scratch addresses); reset() between steps.
struct Bump {
  char* base = nullptr;
  size_t cap = 0, off = 0, high = 0;
  void* get(size_t bytes) {
    size_t o = (off + 255) & ~(size_t)255;
    if (o + bytes > cap) {
      high = o + bytes > high ? o + bytes : high;
      return nullptr;
    }
    off = o + bytes;
    if (off > high) high = off;
    return base + o;
  }
  void reset() { off = 0; }
};
}  // namespace ome

namespace ome {
struct GroupState;  // grouped-launch descriptor cache and autotuner (grouped.hip)
void destroy_group_state(ome_ctx* ctx);
}  // namespace ome

struct ome_ctx {
  ome::Bump* bump = nullptr;
  ome::GroupState* gstate = nullptr;
  int last_grouped_choice = -1;  // 2 * variant + persistent of the last grouped GEMM
  int device = 0;
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;
  std::string last_error;
  ome::Pool* pool = nullptr;
  ome::Options opt;
  ome_device_info info{};
};

namespace ome {

// Scoped pool allocation released at scope exit (stream-ordered reuse).
struct Temp {
  ome_ctx* ctx = nullptr;
  void* p = nullptr;
  Temp() = default;
  bool owned = false;
  Temp(ome_ctx* c, size_t bytes) : ctx(c) {
    if (!bytes) return;
    if (c->bump) {
      p = c->bump->get(bytes);
    } else {
      p = c->pool->alloc(bytes);
      owned = true;
    }
  }
  Temp(const Temp&) = delete;
  Temp& operator=(const Temp&) = delete;
  Temp(Temp&& o) noexcept : ctx(o.ctx), p(o.p), owned(o.owned) { o.p = nullptr; }
  Temp& operator=(Temp&& o) noexcept {
    if (this != &o) {
      if (p && owned) ctx->pool->free(p);
      ctx = o.ctx; p = o.p; owned = o.owned;
      o.p = nullptr;
    }
    return *this;
  }
  ~Temp() {
    if (p && owned) ctx->pool->free(p);
  }
};

// A flat einsum with dense column-major operands, labels as ids into sizes.
struct EinsumCall {
  int dt = F32;
  std::vector<const void*> xs;
  std::vector<std::vector<int>> ixs;
  void* y = nullptr;
  std::vector<int> iy;
  std::vector<int64_t> sizes;
  Scal sx{1, 0, 1}, sy{0, 0, 0};
  std::vector<int> conj;  // per input: read the complex conjugate
  bool force_loop = false;
};

// Plans and enqueues one einsum on `st` (temporaries from the context pool).
Status run_einsum(ome_ctx* ctx, const EinsumCall& c, hipStream_t st);
// Enqueues one GEMM choosing the kernel family (VALU / MFMA) for the element type.
Status run_gemm(ome_ctx* ctx, int dt, GemmDesc d, hipStream_t st, int precision);

Scal scal_one(int dt);
Scal scal_zero(int dt);

// Request-loop access to tree plans (tree.hip).
int64_t tree_mem_bytes(const ome_tree* t);
// fused_done: the plan's fused subtrees already ran in a shared launch (tree_fused_job)
Status tree_run_lent(ome_tree* t, const void* const* xs, void* y, char* mem, hipStream_t st,
                     bool fused_done = false);
bool tree_fused_job(const ome_tree* t, FusedJob* job);
int tree_dtype(const ome_tree* t);
void tree_return_lent(ome_tree* t);
// The plan to admit under a per-request memory limit: t, or a sliced twin that fits.
Status tree_fit(ome_tree* t, int64_t cap, ome_tree** out);

}  // namespace ome
