// The C ABI (include/omeinsum_mi355x.h): context, memory pool, flat einsum, permute and
// batched GEMM entry points.  Trees and the request loop live in tree.hip / server.hip.
#include <dlfcn.h>

#include <cstdio>
#include <cstring>

#include "runtime.h"

using namespace ome;

// ---- pool ----------------------------------------------------------------------------------

namespace ome {

Pool::~Pool() {
  release_cached();
  for (auto& kv : live_) hipFree(kv.first);
}

size_t Pool::round_class(size_t b) {
  if (b <= 4096) return (b + 255) & ~(size_t)255;
  int lg = 63 - __builtin_clzll(b);
  size_t step = (size_t)1 << (lg - 3);
  return (b + step - 1) & ~(step - 1);
}

void* Pool::alloc(size_t bytes) {
  if (bytes == 0) bytes = 1;
  const size_t cls = round_class(bytes);
  std::lock_guard<std::mutex> g(mu_);
  ++n_allocs_;
  void* p = nullptr;
  auto it = free_.find(cls);
  if (it != free_.end() && !it->second.empty()) {
    p = it->second.back();
    it->second.pop_back();
  } else {
    hipSetDevice(device_);
    if (hipMalloc(&p, cls) != hipSuccess) {
      hipGetLastError();
      // give cached blocks back to the driver and retry once
      for (auto& kv : free_)
        for (void* q : kv.second) {
          hipFree(q);
          reserved_ -= kv.first;
        }
      free_.clear();
      if (hipMalloc(&p, cls) != hipSuccess) {
        hipGetLastError();
        return nullptr;
      }
    }
    reserved_ += cls;
    ++n_mallocs_;
  }
  live_[p] = cls;
  in_use_ += cls;
  if (in_use_ > peak_) peak_ = in_use_;
  return p;
}

void Pool::free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_[it->second].push_back(p);
  in_use_ -= it->second;
  live_.erase(it);
}

void Pool::release_cached() {
  std::lock_guard<std::mutex> g(mu_);
  hipSetDevice(device_);
  for (auto& kv : free_)
    for (void* q : kv.second) {
      hipFree(q);
      reserved_ -= kv.first;
    }
  free_.clear();
}

ome_pool_stats Pool::stats() {
  std::lock_guard<std::mutex> g(mu_);
  ome_pool_stats s;
  s.reserved_bytes = reserved_;
  s.in_use_bytes = in_use_;
  s.peak_in_use_bytes = peak_;
  s.n_allocs = n_allocs_;
  s.n_device_mallocs = n_mallocs_;
  return s;
}

// ---- kernel timing ---------------------------------------------------------------------------

namespace {
struct TimingState {
  std::mutex mu;
  bool on = false;
  std::vector<hipEvent_t> spare;
  struct Rec {
    int kc;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  hipEvent_t get() {
    if (!spare.empty()) {
      hipEvent_t e = spare.back();
      spare.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};
TimingState& timing() {
  static TimingState t;
  return t;
}
}  // namespace

void timing_enable(bool on) {
  std::lock_guard<std::mutex> g(timing().mu);
  timing().on = on;
}
bool timing_enabled() { return timing().on; }

TimedScope::TimedScope(int k, hipStream_t s) : kc(k), st(s) {
  TimingState& t = timing();
  if (!t.on) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  std::lock_guard<std::mutex> g(t.mu);
  start = t.get();
  if (start && hipEventRecord(start, st) != hipSuccess) {
    t.spare.push_back(start);
    start = nullptr;
  }
}

TimedScope::~TimedScope() {
  if (!start) return;
  TimingState& t = timing();
  std::lock_guard<std::mutex> g(t.mu);
  hipEvent_t stop = t.get();
  if (stop && hipEventRecord(stop, st) == hipSuccess) {
    t.recs.push_back({kc, start, stop});
  } else {
    t.spare.push_back(start);
    if (stop) t.spare.push_back(stop);
  }
}

void timing_collect(int kc, double* total_ms, int64_t* launches) {
  TimingState& t = timing();
  std::lock_guard<std::mutex> g(t.mu);
  double ms = 0;
  int64_t n = 0;
  std::vector<TimingState::Rec> keep;
  for (auto& r : t.recs) {
    if (r.kc != kc) {
      keep.push_back(r);
      continue;
    }
    float e = 0.f;
    hipEventSynchronize(r.b);
    if (hipEventElapsedTime(&e, r.a, r.b) == hipSuccess) {
      ms += e;
      ++n;
    }
    t.spare.push_back(r.a);
    t.spare.push_back(r.b);
  }
  t.recs.swap(keep);
  *total_ms = ms;
  *launches = n;
}

// ---- roctx ranges ----------------------------------------------------------------------------
// The profiler's roctx library is opened on first use (no link-time dependency); its entry
// points are the roctx C API of rocprofiler-sdk.
namespace {
struct RangeApi {
  std::mutex mu;
  bool loaded = false, on = false;
  int (*push)(const char*) = nullptr;
  int (*pop)() = nullptr;
  uint64_t (*start)(const char*) = nullptr;
  void (*stop)(uint64_t) = nullptr;
};
RangeApi& range_api() {
  static RangeApi r;
  return r;
}
}  // namespace

bool ranges_enable(bool on) {
  RangeApi& r = range_api();
  std::lock_guard<std::mutex> g(r.mu);
  if (on && !r.loaded) {
    void* h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_GLOBAL);
    if (!h) h = dlopen("librocprofiler-sdk-roctx.so", RTLD_NOW | RTLD_GLOBAL);
    if (!h) return false;
    r.push = (int (*)(const char*))dlsym(h, "roctxRangePushA");
    r.pop = (int (*)())dlsym(h, "roctxRangePop");
    r.start = (uint64_t(*)(const char*))dlsym(h, "roctxRangeStartA");
    r.stop = (void (*)(uint64_t))dlsym(h, "roctxRangeStop");
    if (!r.push || !r.pop || !r.start || !r.stop) return false;
    r.loaded = true;
  }
  r.on = on;
  return true;
}
bool ranges_enabled() { return range_api().on; }
uint64_t range_start(const char* name) {
  RangeApi& r = range_api();
  return r.on ? r.start(name) : 0;
}
void range_stop(uint64_t id) {
  RangeApi& r = range_api();
  if (r.on && id) r.stop(id);
}
RangeScope::RangeScope(const char* name) : on(range_api().on) {
  if (on) range_api().push(name);
}
RangeScope::~RangeScope() {
  if (on) range_api().pop();
}

}  // namespace ome

// ---- error plumbing -------------------------------------------------------------------------

static thread_local std::string g_last_error;

static int fail(ome_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  g_last_error = msg;
  return code;
}
static int finish(ome_ctx* ctx, const Status& s) {
  if (!s.good()) return fail(ctx, s.code, s.msg);
  return OME_OK;
}

#define OME_CHECK_CTX(ctx) \
  if (!(ctx)) return fail(nullptr, OME_ERR_INVALID_ARG, "null context")

// ---- fixed scratch of ome_einsum_noalloc -----------------------------------------------------
// One device array per GPU the library runs on (the code object's own memory, held until the
// process ends).  In host code `scratch` names a host-side stand-in, never the device memory:
// its device address on each GPU comes from hipGetSymbolAddress at ome_init, with that GPU
// current, and only that address is passed to kernels.
static __device__ __align__(256) char scratch[OME_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};        // device address of `scratch`, per device
hipError_t g_scratch_err[kMaxDevices] = {};  // the lookup's failure, per device

// Resolves `scratch` on `device` (the current device) once; later calls reuse the address.
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    hipGetLastError();
    g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return;
  }
  g_scratch_ptr[device] = p;
}

void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace

extern "C" {

int ome_abi_version(void) { return OME_ABI_VERSION; }

const char* ome_last_error(const ome_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : g_last_error.c_str();
}

int ome_init(int device, size_t pool_bytes, void* stream, ome_ctx** out) {
  if (!out) return fail(nullptr, OME_ERR_INVALID_ARG, "null output pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    hipGetLastError();
    return fail(nullptr, OME_ERR_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= n) return fail(nullptr, OME_ERR_INVALID_ARG, "device index out of range");
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, OME_ERR_DEVICE, "hipSetDevice failed");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return fail(nullptr, OME_ERR_DEVICE, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, OME_ERR_DEVICE,
                std::string("this library is built for gfx950 (MI355X); device is ") + prop.gcnArchName);
  ome_ctx* ctx = new ome_ctx();
  ctx->device = device;
  ctx->stream = (hipStream_t)stream;
  ctx->pool = new Pool(device);
  std::snprintf(ctx->info.name, sizeof ctx->info.name, "%s", prop.name);
  std::snprintf(ctx->info.arch, sizeof ctx->info.arch, "%s", prop.gcnArchName);
  ctx->info.compute_units = prop.multiProcessorCount;
  ctx->info.xcds = 8;
  ctx->info.total_mem = (int64_t)prop.totalGlobalMem;
  ctx->info.lds_per_cu = (int64_t)prop.maxSharedMemoryPerMultiProcessor;
  ctx->info.clock_mhz = prop.clockRate / 1000;
  ctx->info.device = device;
  scratch_setup(device);  // device address of ome_einsum_noalloc's scratch on this GPU
  if (pool_bytes) {  // warm the pool with one reusable block
    void* p = ctx->pool->alloc(pool_bytes);
    ctx->pool->free(p);
  }
  *out = ctx;
  return OME_OK;
}

int ome_shutdown(ome_ctx* ctx) {
  OME_CHECK_CTX(ctx);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  destroy_group_state(ctx);
  delete ctx->pool;
  delete ctx;
  return OME_OK;
}

int ome_set_stream(ome_ctx* ctx, void* stream) {
  OME_CHECK_CTX(ctx);
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  // pooled blocks are reused in stream order: drain the old stream before switching
  hipStreamSynchronize(ctx->stream);
  ctx->stream = (hipStream_t)stream;
  return OME_OK;
}

int ome_set_option(ome_ctx* ctx, int key, int64_t v) {
  OME_CHECK_CTX(ctx);
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  switch (key) {
    case OME_OPT_PRECISION:
      if (v < 0 || v > 3) return fail(ctx, OME_ERR_INVALID_ARG, "precision must be 0, 1, 2 or 3");
      ctx->opt.precision = (int)v; break;
    case OME_OPT_COMPLEX_ALGO: ctx->opt.complex_algo = (int)(v != 0); break;
    case OME_OPT_ALLOW_LOOPS: ctx->opt.allow_loops = (int)(v != 0); break;
    case OME_OPT_DIRECT_GEMM: ctx->opt.direct_gemm = (int)(v != 0); break;
    case OME_OPT_USE_GRAPHS: ctx->opt.use_graphs = (int)(v != 0); break;
    case OME_OPT_AUTOTUNE: ctx->opt.autotune = (int)(v != 0); break;
    case OME_OPT_GEMM_KERNEL:
      if (v < -1 || v > 2) return fail(ctx, OME_ERR_INVALID_ARG, "gemm kernel must be -1..2");
      ctx->opt.gemm_kernel = (int)v; break;
    case OME_OPT_KERNEL_TIMING: timing_enable(v != 0); break;
    case OME_OPT_TREE_DEFER_MEMORY: ctx->opt.defer_tree_mem = (int)(v != 0); break;
    case OME_OPT_GROUPED_BUFFERS: ctx->opt.grouped_buffers = (int)(v != 0); break;
    case OME_OPT_TRACE_RANGES:
      if (!ranges_enable(v != 0))
        return fail(ctx, OME_ERR_UNSUPPORTED, "roctx library (librocprofiler-sdk-roctx) not found");
      break;
    case OME_OPT_GROUPED_VARIANT:
      if (v < -1 || v >= GG_NVARIANTS) return fail(ctx, OME_ERR_INVALID_ARG, "grouped variant must be -1..10");
      ctx->opt.grouped_variant = (int)v; break;
    default: return fail(ctx, OME_ERR_INVALID_ARG, "unknown option");
  }
  return OME_OK;
}

int ome_get_option(ome_ctx* ctx, int key, int64_t* v) {
  OME_CHECK_CTX(ctx);
  if (!v) return fail(ctx, OME_ERR_INVALID_ARG, "null output pointer");
  switch (key) {
    case OME_OPT_PRECISION: *v = ctx->opt.precision; break;
    case OME_OPT_COMPLEX_ALGO: *v = ctx->opt.complex_algo; break;
    case OME_OPT_ALLOW_LOOPS: *v = ctx->opt.allow_loops; break;
    case OME_OPT_DIRECT_GEMM: *v = ctx->opt.direct_gemm; break;
    case OME_OPT_USE_GRAPHS: *v = ctx->opt.use_graphs; break;
    case OME_OPT_AUTOTUNE: *v = ctx->opt.autotune; break;
    case OME_OPT_GEMM_KERNEL: *v = ctx->opt.gemm_kernel; break;
    case OME_OPT_KERNEL_TIMING: *v = timing_enabled() ? 1 : 0; break;
    case OME_OPT_GROUPED_VARIANT: *v = ctx->opt.grouped_variant; break;
    case OME_OPT_LAST_GROUPED_CHOICE: *v = ctx->last_grouped_choice; break;
    case OME_OPT_TREE_DEFER_MEMORY: *v = ctx->opt.defer_tree_mem; break;
    case OME_OPT_GROUPED_BUFFERS: *v = ctx->opt.grouped_buffers; break;
    case OME_OPT_TRACE_RANGES: *v = ranges_enabled() ? 1 : 0; break;
    default: return fail(ctx, OME_ERR_INVALID_ARG, "unknown option");
  }
  return OME_OK;
}

int ome_get_kernel_timing(ome_ctx* ctx, int kclass, double* total_ms, int64_t* launches) {
  OME_CHECK_CTX(ctx);
  if (!total_ms || !launches) return fail(ctx, OME_ERR_INVALID_ARG, "null output pointer");
  if (kclass < 0 || kclass >= KC_COUNT) return fail(ctx, OME_ERR_INVALID_ARG, "unknown kernel class");
  hipSetDevice(ctx->device);
  timing_collect(kclass, total_ms, launches);
  return OME_OK;
}

int ome_get_device_info(ome_ctx* ctx, ome_device_info* info) {
  OME_CHECK_CTX(ctx);
  if (!info) return fail(ctx, OME_ERR_INVALID_ARG, "null output pointer");
  *info = ctx->info;
  return OME_OK;
}

int ome_synchronize(ome_ctx* ctx) {
  OME_CHECK_CTX(ctx);
  hipSetDevice(ctx->device);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, OME_ERR_DEVICE, std::string("stream: ") + hipGetErrorString(e));
  return OME_OK;
}

int ome_alloc(ome_ctx* ctx, size_t bytes, void** ptr) {
  OME_CHECK_CTX(ctx);
  if (!ptr) return fail(ctx, OME_ERR_INVALID_ARG, "null output pointer");
  *ptr = ctx->pool->alloc(bytes);
  if (!*ptr) return fail(ctx, OME_ERR_OOM, "device pool exhausted");
  return OME_OK;
}

int ome_free(ome_ctx* ctx, void* ptr) {
  OME_CHECK_CTX(ctx);
  ctx->pool->free(ptr);
  return OME_OK;
}

int ome_get_pool_stats(ome_ctx* ctx, ome_pool_stats* s) {
  OME_CHECK_CTX(ctx);
  if (!s) return fail(ctx, OME_ERR_INVALID_ARG, "null output pointer");
  *s = ctx->pool->stats();
  return OME_OK;
}

static int memcpy_kind(ome_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind k) {
  OME_CHECK_CTX(ctx);
  if (bytes == 0) return OME_OK;
  if (!dst || !src) return fail(ctx, OME_ERR_INVALID_ARG, "null pointer");
  hipSetDevice(ctx->device);
  hipError_t e = hipMemcpyAsync(dst, src, bytes, k, ctx->stream);
  if (e == hipSuccess && k != hipMemcpyDeviceToDevice) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, OME_ERR_DEVICE, std::string("memcpy: ") + hipGetErrorString(e));
  return OME_OK;
}
int ome_memcpy_h2d(ome_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return memcpy_kind(ctx, dst, src, bytes, hipMemcpyHostToDevice);
}
int ome_memcpy_d2h(ome_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return memcpy_kind(ctx, dst, src, bytes, hipMemcpyDeviceToHost);
}
int ome_memcpy_d2d(ome_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return memcpy_kind(ctx, dst, src, bytes, hipMemcpyDeviceToDevice);
}

static int einsum_entry(ome_ctx* ctx, int dtype, int nin, const void* const* xs, const int32_t* nix,
                        const int32_t* ixs, void* y, int niy, const int32_t* iy, int nlabels,
                        const int64_t* sizes, const ome_scalar* sx, const ome_scalar* sy,
                        bool force_loop, bool noalloc = false) {
  OME_CHECK_CTX(ctx);
  if (nin < 1 || nin > OME_MAX_INPUTS)
    return fail(ctx, OME_ERR_INVALID_ARG, "number of input tensors must be in 1..64");
  if (!xs || !nix || (niy > 0 && !iy) || niy < 0 || nlabels < 0 || (nlabels > 0 && !sizes))
    return fail(ctx, OME_ERR_INVALID_ARG, "null or negative argument");
  EinsumCall c;
  c.dt = dtype;
  c.sizes.assign(sizes, sizes + nlabels);
  int pos = 0;
  for (int k = 0; k < nin; ++k) {
    if (nix[k] < 0) return fail(ctx, OME_ERR_INVALID_ARG, "negative rank");
    c.xs.push_back(xs[k]);
    c.ixs.emplace_back(ixs + pos, ixs + pos + nix[k]);
    pos += nix[k];
  }
  c.y = y;
  c.iy.assign(iy, iy + niy);
  c.sx = sx ? Scal{sx->re, sx->im, sx->i} : scal_one(dtype);
  c.sy = sy ? Scal{sy->re, sy->im, sy->i} : scal_zero(dtype);
  c.force_loop = force_loop;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  if (!noalloc) return finish(ctx, run_einsum(ctx, c, ctx->stream));
  // ---- allocation-free: the planner's temporaries come from `scratch` on this device
  char* base = (char*)scratch_ptr(ctx->device);
  if (!base) {
    const hipError_t e = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_err[ctx->device]
                                                                        : hipErrorInvalidDevice;
    return fail(ctx, OME_ERR_DEVICE, std::string("ome_einsum_noalloc: the scratch array has no "
                                                 "device address on this GPU (") +
                                         hipGetErrorString(e) + ")");
  }
  // size the temporaries with a dry run of the same plan (nothing is launched)
  Bump dry;
  dry.base = (char*)(uintptr_t)4096;  // never dereferenced: a dry run launches nothing
  dry.cap = (size_t)1 << 60;
  LaunchMode& lm = launch_mode();
  lm.dry = true;
  ctx->bump = &dry;
  Status r = run_einsum(ctx, c, ctx->stream);
  ctx->bump = nullptr;
  lm.dry = false;
  if (!r.good()) return finish(ctx, r);
  if (dry.high > (size_t)OME_NOALLOC_SCRATCH_BYTES)
    return fail(ctx, OME_ERR_OOM,
                "ome_einsum_noalloc: this contraction's temporaries need " + std::to_string(dry.high) +
                    " bytes, above the " + std::to_string((long long)OME_NOALLOC_SCRATCH_BYTES) +
                    "-byte scratch array (use ome_einsum)");
  Bump b;
  b.base = base;
  b.cap = (size_t)OME_NOALLOC_SCRATCH_BYTES;
  ctx->bump = &b;
  r = run_einsum(ctx, c, ctx->stream);
  ctx->bump = nullptr;
  return finish(ctx, r);
}

int ome_einsum(ome_ctx* ctx, int dtype, int nin, const void* const* xs, const int32_t* nix,
               const int32_t* ixs, void* y, int niy, const int32_t* iy, int nlabels,
               const int64_t* sizes, const ome_scalar* sx, const ome_scalar* sy) {
  return einsum_entry(ctx, dtype, nin, xs, nix, ixs, y, niy, iy, nlabels, sizes, sx, sy, false);
}

int ome_einsum_noalloc(ome_ctx* ctx, int dtype, int nin, const void* const* xs, const int32_t* nix,
                       const int32_t* ixs, void* y, int niy, const int32_t* iy, int nlabels,
                       const int64_t* sizes, const ome_scalar* sx, const ome_scalar* sy) {
  return einsum_entry(ctx, dtype, nin, xs, nix, ixs, y, niy, iy, nlabels, sizes, sx, sy, false,
                      true);
}

int ome_loop_einsum(ome_ctx* ctx, int dtype, int nin, const void* const* xs, const int32_t* nix,
                    const int32_t* ixs, void* y, int niy, const int32_t* iy, int nlabels,
                    const int64_t* sizes, const ome_scalar* sx, const ome_scalar* sy) {
  return einsum_entry(ctx, dtype, nin, xs, nix, ixs, y, niy, iy, nlabels, sizes, sx, sy, true);
}

int ome_permute(ome_ctx* ctx, int dtype, int ndim, const int64_t* dims, const int32_t* perm,
                const void* a, void* c, const ome_scalar* sx, const ome_scalar* sy) {
  OME_CHECK_CTX(ctx);
  if (ndim < 0 || ndim > OME_MAX_RANK || (ndim > 0 && (!dims || !perm)) || !a || !c)
    return fail(ctx, OME_ERR_INVALID_ARG, "invalid permute arguments");
  std::vector<int> seen(ndim, 0);
  for (int p = 0; p < ndim; ++p) {
    if (perm[p] < 0 || perm[p] >= ndim || seen[perm[p]])
      return fail(ctx, OME_ERR_INVALID_ARG, "perm is not a permutation");
    seen[perm[p]] = 1;
  }
  EinsumCall e;
  e.dt = dtype;
  e.xs = {a};
  e.ixs.resize(1);
  for (int p = 0; p < ndim; ++p) {
    e.ixs[0].push_back(p);
    e.sizes.push_back(dims[p]);
    e.iy.push_back(perm[p]);
  }
  e.y = c;
  e.sx = sx ? Scal{sx->re, sx->im, sx->i} : scal_one(dtype);
  e.sy = sy ? Scal{sy->re, sy->im, sy->i} : scal_zero(dtype);
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  return finish(ctx, run_einsum(ctx, e, ctx->stream));
}

int ome_addmul(ome_ctx* ctx, int dtype, int64_t n, const void* x, void* y, const ome_scalar* sx,
               const ome_scalar* sy, int conj_x) {
  OME_CHECK_CTX(ctx);
  if (n < 0 || !x || !y || dtype < F32 || dtype > DT_LAST)
    return fail(ctx, OME_ERR_INVALID_ARG, "invalid addmul arguments");
  const Scal a = sx ? Scal{sx->re, sx->im, sx->i} : scal_one(dtype);
  const Scal b = sy ? Scal{sy->re, sy->im, sy->i} : scal_zero(dtype);
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  launch_axpby(dtype, x, y, n, a, b, conj_x ? 1 : 0, ctx->stream);
  return OME_OK;
}

int ome_batched_gemm(ome_ctx* ctx, int dtype, char ta, char tb, int64_t m, int64_t n, int64_t k,
                     int64_t batch, const ome_scalar* alpha, const void* a, const void* b,
                     const ome_scalar* beta, void* c) {
  OME_CHECK_CTX(ctx);
  if ((ta != 'N' && ta != 'T') || (tb != 'N' && tb != 'T'))
    return fail(ctx, OME_ERR_INVALID_ARG, "transpose flags must be 'N' or 'T'");
  if (m < 0 || n < 0 || k < 0 || batch < 0 || !a || !b || !c)
    return fail(ctx, OME_ERR_INVALID_ARG, "invalid GEMM arguments");
  GemmDesc d{};
  d.a = a; d.b = b; d.y = c;
  d.M = m; d.N = n; d.K = k; d.B = batch;
  if (ta == 'N') { d.sam = 1; d.sak = m; } else { d.sak = 1; d.sam = k; }
  d.sab = m * k;
  if (tb == 'N') { d.sbk = 1; d.sbn = k; } else { d.sbn = 1; d.sbk = n; }
  d.sbb = k * n;
  d.sym = 1; d.syn = m; d.syb = m * n;
  d.sx = alpha ? Scal{alpha->re, alpha->im, alpha->i} : scal_one(dtype);
  d.sy = beta ? Scal{beta->re, beta->im, beta->i} : scal_zero(dtype);
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  if (m * n * batch == 0) return OME_OK;
  if (k == 0) {
    launch_axpby(dtype, nullptr, c, m * n * batch, d.sx, d.sy, 0, ctx->stream);
    return OME_OK;
  }
  return finish(ctx, run_gemm(ctx, dtype, d, ctx->stream, ctx->opt.precision));
}

}  // extern "C"
