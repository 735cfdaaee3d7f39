// Continuous batching of tensor-network contraction requests (ome_server_*).
//
// A request is a planned tree (ome_tree) with its input and output buffers.  One host
// scheduler thread per server runs the loop
//   1. retire: a slot whose completion event has fired returns its request's device memory
//      to the pool, marks the request done (completion time = when the event was seen) and
//      frees the slot;
//   2. admit: queued requests are scanned in arrival order and every one whose plan memory
//      (arena + scratch + staging + accumulators, known exactly at plan time) fits the
//      remaining HBM budget is started on a free slot: a memory block is lent to the plan
//      (the block this plan used last when it is free, else the smallest free block that
//      fits, else a new one from the pool), the slot's stream waits for the context's stream
//      as it stands at admission (the inputs are ready, and so is every earlier use of the
//      lent memory: the pool hands out blocks freed on the host as soon as work on them was
//      enqueued on the context's stream), the evaluation is launched (the plan's recorded
//      graph when it was lent this block before, so the steps are not planned again), and a
//      completion event is recorded;
//   3. sleep briefly while requests run, or block until the next submission when idle.
// Each slot has its own stream, so the steps of admitted requests overlap on the device, and
// requests that arrive while others run are admitted as soon as a slot and memory are free.
// Retired blocks stay with the loop (within the budget) for the next admissions.  A request
// whose plan needs more than the per-request limit (the whole budget unless set) is sliced at
// submission: the loop admits a twin of its tree with more labels sliced, whose memory fits
// (SlicedEinsum semantics, FP64 cross-slice sum).
// Cross-request batching: the requests admitted in one round run their fused small-step
// subtrees (the maximal subtrees of small steps a plan evaluates in one launch) together, in
// one launch shared by all of them, before each runs its other steps on its slot stream.
// Sliced requests (SlicedEinsum plans) keep their FP64 cross-slice accumulator in the lent
// memory.  Every slot owns a status word in coherent pinned memory, cleared at admission and
// read at retirement: a kernel that writes it non-zero fails its own request, and only that
// one (ome_server_inject_status makes the next request write it, for tests).  With roctx ranges on,
// each request is one range from admission to retirement.  A tree must not be used elsewhere while its request is queued or running, and the
// context's stream must not be under graph capture while the loop has queued requests.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "runtime.h"

using namespace ome;

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  int64_t ticket = 0;
  ome_tree* tree = nullptr;
  std::vector<const void*> xs;
  void* y = nullptr;
  hipEvent_t ready = nullptr;  // recorded on the submitter's stream
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = 0;  // 0 queued, 1 running, 2 done, < 0 failed (negated ome_status)
  std::string msg;
  size_t bytes = 0;     // plan memory
  size_t held = 0;      // bytes of the lent block (>= bytes)
  char* mem = nullptr;
  int inject = 0;       // non-zero: written into the slot's status word after the steps
  uint64_t range = 0;   // roctx range from admission to retirement
};

struct Block {
  char* p = nullptr;
  size_t bytes = 0;
  ome_tree* last = nullptr;  // the plan that used it last
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  Req* req = nullptr;
};

}  // namespace

struct ome_server {
  ome_ctx* ctx = nullptr;
  size_t budget = 0, in_use = 0;
  size_t req_limit = 0;            // per-request plan memory; larger requests are sliced
  std::vector<Block> free_blocks;  // retired blocks kept for the next admissions
  size_t free_bytes = 0;
  ome_server_stats stats{};
  // the shared launch of fused subtrees across requests
  bool batching = true;
  bool hold = false;                         // admission held (ome_server_hold)
  hipStream_t bst = nullptr;                 // batch stream
  hipEvent_t round_admit = nullptr, batch_done = nullptr, copy_ev[2] = {nullptr, nullptr};
  char* job_host[2] = {nullptr, nullptr};    // pinned staging of the job table, two buffers
  char* job_dev[2] = {nullptr, nullptr};
  int jb = 0;
  std::vector<Slot> slots;
  // one status word per slot in coherent pinned memory: kernels write it through status_dev,
  // the scheduler reads it when the slot's request retires
  int* status_host = nullptr;
  int* status_dev = nullptr;
  int inject_next = 0;  // status the next submitted request writes (ome_server_inject_status)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queue;
  std::unordered_map<int64_t, Req*> reqs;
  int64_t next = 1;
  int active = 0;
  bool stop = false;
  std::thread th;
};

namespace {

void retire(ome_server* s, Slot& sl, int status, const std::string& msg) {
  Req* r = sl.req;
  range_stop(r->range);
  r->range = 0;
  tree_return_lent(r->tree);
  if (r->mem) {
    if (status == 2) {  // kept for the next admissions, tagged with its last plan
      s->free_blocks.push_back(Block{r->mem, r->held, r->tree});
      s->free_bytes += r->held;
    } else {
      s->ctx->pool->free(r->mem);
    }
  }
  r->mem = nullptr;
  s->in_use -= r->held;
  if (status == 2) ++s->stats.completed;
  r->t_done = now_ms();
  r->status = status;
  r->msg = msg;
  sl.req = nullptr;
  --s->active;
}

// Launch the requests of one admission round (their slots claimed, memory lent).  Every slot
// stream waits for the submission event of its request and for the context's stream as it
// stands now.  With batching on and two or more such requests in the round, the fused
// small-step subtrees of the round's unsliced requests (one element type, at most SF_MAXJOBS
// requests) run first in one launch on the batch stream, from a job table copied to the
// device; those requests then run only their other steps, after that launch.  Returns one
// status per request.
std::vector<Status> launch_round(ome_server* s, const std::vector<Slot*>& round) {
  std::vector<Status> res(round.size());
  hipError_t ea;
  {
    std::lock_guard<std::recursive_mutex> g(s->ctx->mu);
    ea = hipEventRecord(s->round_admit, s->ctx->stream);
  }
  if (ea != hipSuccess) {
    hipGetLastError();
    for (auto& r : res)
      r = Status::err(OME_ERR_DEVICE, "could not order the request after the context's stream");
    return res;
  }
  std::vector<char> batched(round.size(), 0);
  if (s->batching) {
    std::vector<FusedJob> jobs;
    std::vector<size_t> who;
    int dt = -1, nsub = 0;
    for (size_t i = 0; i < round.size() && (int)jobs.size() < SF_MAXJOBS; ++i) {
      Req* r = round[i]->req;
      FusedJob jb{};
      if (!tree_fused_job(r->tree, &jb)) continue;
      if (dt < 0) dt = tree_dtype(r->tree);
      if (tree_dtype(r->tree) != dt) continue;
      jb.arena = r->mem;
      jb.sub_begin = nsub;
      nsub += jb.nsub;
      jobs.push_back(jb);
      who.push_back(i);
    }
    // one request alone gains nothing from the shared launch (its own graph holds its fused
    // launch) and would wait for the job table's copy: batch rounds of two or more
    if (jobs.size() >= 2) {
      const int b = s->jb;
      s->jb ^= 1;
      bool ok = hipEventSynchronize(s->copy_ev[b]) == hipSuccess;  // staging b was copied
      char* host = s->job_host[b];
      char* dev = s->job_dev[b];
      const size_t jbytes = SF_MAXJOBS * sizeof(FusedJob);
      for (size_t q = 0; q < jobs.size(); ++q) {
        Req* r = round[who[q]]->req;
        const void** lt = (const void**)(host + jbytes) + q * SF_MAXLEAF;
        for (size_t k = 0; k < r->xs.size() && k < (size_t)SF_MAXLEAF; ++k) lt[k] = r->xs[k];
        jobs[q].leaves = (const void* const*)(dev + jbytes) + q * SF_MAXLEAF;
      }
      std::memcpy(host, jobs.data(), jobs.size() * sizeof(FusedJob));
      const size_t bytes = jbytes + jobs.size() * SF_MAXLEAF * sizeof(void*);
      ok = ok && hipStreamWaitEvent(s->bst, s->round_admit, 0) == hipSuccess;
      for (size_t q = 0; ok && q < jobs.size(); ++q)
        ok = hipStreamWaitEvent(s->bst, round[who[q]]->req->ready, 0) == hipSuccess;
      ok = ok && hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s->bst) == hipSuccess;
      ok = ok && hipEventRecord(s->copy_ev[b], s->bst) == hipSuccess;
      if (ok) {
        RangeScope rg("request loop: shared launch of fused small steps");
        launch_fused_jobs(dt, (const FusedJob*)dev, (int)jobs.size(), nsub, s->bst);
        ok = hipGetLastError() == hipSuccess && hipEventRecord(s->batch_done, s->bst) == hipSuccess;
      }
      if (ok) {
        for (size_t q : who) batched[q] = 1;
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->stats.batched_launches;
        s->stats.batched_requests += (int64_t)jobs.size();
      } else {
        hipGetLastError();
        // the requests run their own fused launches; the batch stream is drained first so
        // that no partial batch writes into their memory afterwards
        hipStreamSynchronize(s->bst);
        hipGetLastError();
      }
    }
  }
  for (size_t i = 0; i < round.size(); ++i) {
    Slot& sl = *round[i];
    Req* r = sl.req;
    const size_t idx = (size_t)(&sl - s->slots.data());
    // the slot's previous request has retired, so no kernel writes its word any more
    __atomic_store_n(&s->status_host[idx], 0, __ATOMIC_RELEASE);
    if (ranges_enabled()) r->range = range_start(("request " + std::to_string(r->ticket)).c_str());
    hipStreamWaitEvent(sl.st, r->ready, 0);
    hipStreamWaitEvent(sl.st, s->round_admit, 0);
    if (batched[i]) hipStreamWaitEvent(sl.st, s->batch_done, 0);
    Status st = tree_run_lent(r->tree, r->xs.data(), r->y, r->mem, sl.st, batched[i] != 0);
    if (st.good() && r->inject) launch_status_set(s->status_dev + idx, r->inject, sl.st);
    const hipError_t e = hipEventRecord(sl.done, sl.st);
    if (!st.good() || e != hipSuccess) {
      if (e != hipSuccess) hipGetLastError();
      hipStreamSynchronize(sl.st);
      res[i] = st.good() ? Status::err(OME_ERR_DEVICE, "event record failed") : st;
    }
  }
  return res;
}

void scheduler(ome_server* s) {
  hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  std::vector<Slot*> round;
  for (;;) {
    bool progress = false;
    // 1. retire finished requests
    for (auto& sl : s->slots) {
      if (!sl.req || sl.req->status != 1) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      if (q == hipSuccess) {
        const int ks = __atomic_load_n(&s->status_host[&sl - s->slots.data()], __ATOMIC_ACQUIRE);
        if (ks != 0)
          retire(s, sl, -OME_ERR_DEVICE,
                 "the request's kernels reported status " + std::to_string(ks));
        else
          retire(s, sl, 2, "");
      } else {
        hipGetLastError();
        retire(s, sl, -OME_ERR_DEVICE, std::string("request failed on the device: ") + hipGetErrorString(q));
      }
      progress = true;
    }
    if (progress) s->cv_done.notify_all();
    // 2. admit queued requests that fit, in arrival order: claim a slot and memory (none
    // while admission is held, unless the server is stopping)
    const bool admit = !s->hold || s->stop;
    for (auto it = s->queue.begin(); admit && it != s->queue.end();) {
      Slot* free_slot = nullptr;
      for (auto& sl : s->slots)
        if (!sl.req) { free_slot = &sl; break; }
      if (!free_slot) break;
      Req* r = *it;
      if (r->bytes > s->budget - s->in_use) { ++it; continue; }
      it = s->queue.erase(it);
      char* mem = nullptr;
      size_t held = r->bytes;
      if (r->bytes) {
        // the block this plan used last; else a new block while the budget has room beside
        // the kept ones (so that other plans find theirs); else the smallest kept block
        // that fits
        int pick = -1;
        for (size_t b = 0; b < s->free_blocks.size(); ++b) {
          const Block& bl = s->free_blocks[b];
          if (bl.last == r->tree && bl.bytes >= r->bytes && bl.bytes <= s->budget - s->in_use) {
            pick = (int)b;
            break;
          }
        }
        if (pick < 0 && s->in_use + s->free_bytes + r->bytes > s->budget) {
          for (size_t b = 0; b < s->free_blocks.size(); ++b) {
            const Block& bl = s->free_blocks[b];
            if (bl.bytes < r->bytes || bl.bytes > s->budget - s->in_use || bl.bytes > 2 * r->bytes)
              continue;
            if (pick < 0 || bl.bytes < s->free_blocks[pick].bytes) pick = (int)b;
          }
        }
        if (pick >= 0) {
          mem = s->free_blocks[pick].p;
          held = s->free_blocks[pick].bytes;
          if (s->free_blocks[pick].last == r->tree) ++s->stats.blocks_reused;
          s->free_bytes -= held;
          s->free_blocks.erase(s->free_blocks.begin() + pick);
        } else {
          // a new block: kept blocks go back to the pool (oldest first) while the kept and
          // the running memory together would exceed the budget
          while (!s->free_blocks.empty() && s->in_use + s->free_bytes + r->bytes > s->budget) {
            s->ctx->pool->free(s->free_blocks.front().p);
            s->free_bytes -= s->free_blocks.front().bytes;
            s->free_blocks.erase(s->free_blocks.begin());
          }
          mem = (char*)s->ctx->pool->alloc(r->bytes);
        }
      }
      if (r->bytes && !mem) {
        // the pool cannot serve the budget it was promised: fail this request
        r->status = -OME_ERR_OOM;
        r->msg = "device pool exhausted for an admitted request";
        r->t_done = now_ms();
        s->cv_done.notify_all();
        continue;
      }
      r->mem = mem;
      r->held = held;
      r->status = 1;
      r->t_start = now_ms();
      s->in_use += held;
      ++s->active;
      free_slot->req = r;
      round.push_back(free_slot);
    }
    // 3. launch the round outside the server lock (submitters keep going; the slots are
    // claimed): the shared launch of the round's fused subtrees, then every request's steps
    if (!round.empty()) {
      lk.unlock();
      std::vector<Status> res = launch_round(s, round);
      lk.lock();
      for (size_t i = 0; i < round.size(); ++i) {
        if (res[i].good()) continue;
        retire(s, *round[i], -res[i].code, res[i].msg);
        s->cv_done.notify_all();
      }
      round.clear();
      progress = true;
    }
    if (s->stop && s->queue.empty() && s->active == 0) break;
    if (!progress) {
      if (s->active == 0) {
        s->cv_work.wait(lk, [s] { return s->stop || (!s->queue.empty() && !s->hold); });
      } else {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      }
    }
  }
}

// Streams, events and buffers of a server whose scheduler is not running (destroy, or a
// failed create): every stream is drained first.
void free_server_resources(ome_server* s) {
  for (auto& sl : s->slots) {
    if (sl.st) {
      hipStreamSynchronize(sl.st);
      hipStreamDestroy(sl.st);
    }
    if (sl.done) hipEventDestroy(sl.done);
  }
  if (s->bst) {
    hipStreamSynchronize(s->bst);
    hipStreamDestroy(s->bst);
  }
  if (s->round_admit) hipEventDestroy(s->round_admit);
  if (s->batch_done) hipEventDestroy(s->batch_done);
  for (int b = 0; b < 2; ++b) {
    if (s->copy_ev[b]) hipEventDestroy(s->copy_ev[b]);
    if (s->job_host[b]) hipHostFree(s->job_host[b]);
    if (s->job_dev[b]) s->ctx->pool->free(s->job_dev[b]);
  }
  for (auto& b : s->free_blocks) s->ctx->pool->free(b.p);
  s->free_blocks.clear();
  if (s->status_host) hipHostFree(s->status_host);
  s->status_host = s->status_dev = nullptr;
}

int sfail(ome_server* s, int code, const std::string& msg) {
  if (s && s->ctx) s->ctx->last_error = msg;
  return code;
}

}  // namespace

extern "C" {

int ome_server_create(ome_ctx* ctx, size_t hbm_budget, int max_active, ome_server** out) {
  if (!ctx || !out || max_active < 1 || hbm_budget == 0) {
    if (ctx) ctx->last_error = "invalid request-loop arguments";
    return OME_ERR_INVALID_ARG;
  }
  hipSetDevice(ctx->device);
  ome_server* s = new ome_server();
  s->ctx = ctx;
  s->budget = hbm_budget;
  s->req_limit = hbm_budget;
  s->slots.resize(max_active);
  bool ok = true;
  for (auto& sl : s->slots)
    ok = ok && hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess;
  const size_t jbytes = SF_MAXJOBS * (sizeof(FusedJob) + SF_MAXLEAF * sizeof(void*));
  ok = ok && hipStreamCreateWithFlags(&s->bst, hipStreamNonBlocking) == hipSuccess &&
       hipEventCreateWithFlags(&s->round_admit, hipEventDisableTiming) == hipSuccess &&
       hipEventCreateWithFlags(&s->batch_done, hipEventDisableTiming) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&s->status_host, (size_t)max_active * sizeof(int),
                           hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess &&
       hipHostGetDevicePointer((void**)&s->status_dev, s->status_host, 0) == hipSuccess;
  if (ok) std::memset(s->status_host, 0, (size_t)max_active * sizeof(int));
  for (int b = 0; ok && b < 2; ++b) {
    ok = hipEventCreateWithFlags(&s->copy_ev[b], hipEventDisableTiming) == hipSuccess &&
         hipHostMalloc((void**)&s->job_host[b], jbytes, hipHostMallocDefault) == hipSuccess;
    if (ok) {
      s->job_dev[b] = (char*)ctx->pool->alloc(jbytes);
      ok = s->job_dev[b] != nullptr;
    }
  }
  if (!ok) {
    hipGetLastError();
    free_server_resources(s);
    delete s;
    ctx->last_error = "could not create the request-loop streams and buffers";
    return OME_ERR_DEVICE;
  }
  s->th = std::thread(scheduler, s);
  *out = s;
  return OME_OK;
}

int ome_server_submit(ome_server* s, ome_tree* tree, const void* const* xs, void* y, int64_t* ticket) {
  if (!s || !tree || !ticket) return sfail(s, OME_ERR_INVALID_ARG, "null argument");
  ome_tree_info info{};
  ome_get_tree_info(tree, &info);
  size_t limit;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    limit = std::min(s->req_limit, s->budget);
  }
  // a plan larger than the per-request limit is sliced here (once per plan and limit)
  ome_tree* fit = nullptr;
  Status fs = tree_fit(tree, (int64_t)limit, &fit);
  if (!fs.good())
    return sfail(s, fs.code == OME_ERR_OOM ? OME_ERR_OOM : fs.code,
                 "request needs " + std::to_string(tree_mem_bytes(tree)) +
                     " bytes of device memory, above the per-request limit of " +
                     std::to_string(limit) + ", and slicing does not fit it: " + fs.msg);
  Req* r = new Req();
  r->tree = fit;
  r->xs.assign(xs, xs + info.n_leaves);
  r->y = y;
  r->bytes = (size_t)tree_mem_bytes(fit);
  hipSetDevice(s->ctx->device);
  hipStream_t submit_stream;
  {
    std::lock_guard<std::recursive_mutex> g(s->ctx->mu);
    submit_stream = s->ctx->stream;
  }
  if (hipEventCreateWithFlags(&r->ready, hipEventDisableTiming) != hipSuccess ||
      hipEventRecord(r->ready, submit_stream) != hipSuccess) {
    hipGetLastError();
    if (r->ready) hipEventDestroy(r->ready);
    delete r;
    return sfail(s, OME_ERR_DEVICE, "could not record the submission event");
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next++;
    r->t_submit = now_ms();
    r->inject = s->inject_next;
    s->inject_next = 0;
    ++s->stats.submitted;
    if (fit != tree) ++s->stats.sliced;
    s->reqs[r->ticket] = r;
    s->queue.push_back(r);
  }
  s->cv_work.notify_one();
  *ticket = r->ticket;
  return OME_OK;
}

int ome_server_poll(ome_server* s, int64_t ticket, int* status) {
  if (!s || !status) return sfail(s, OME_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return sfail(s, OME_ERR_INVALID_ARG, "unknown ticket");
  *status = it->second->status;
  return OME_OK;
}

int ome_server_wait(ome_server* s, int64_t ticket, double* latency_ms) {
  if (!s) return OME_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return sfail(s, OME_ERR_INVALID_ARG, "unknown ticket");
  Req* r = it->second;
  s->cv_done.wait(lk, [r] { return r->status >= 2 || r->status < 0; });
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  const int status = r->status;
  const std::string msg = r->msg;
  s->reqs.erase(it);
  lk.unlock();
  hipEventDestroy(r->ready);
  delete r;
  if (status < 0) return sfail(s, -status, msg);
  return OME_OK;
}

int ome_server_destroy(ome_server* s) {
  if (!s) return OME_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->th.join();
  hipSetDevice(s->ctx->device);
  for (auto& kv : s->reqs) {
    hipEventDestroy(kv.second->ready);
    delete kv.second;
  }
  free_server_resources(s);
  delete s;
  return OME_OK;
}

int ome_server_inject_status(ome_server* s, int code) {
  if (!s) return OME_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->inject_next = code;
  return OME_OK;
}

int ome_server_set_batching(ome_server* s, int on) {
  if (!s) return OME_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->batching = on != 0;
  return OME_OK;
}

int ome_server_hold(ome_server* s, int on) {
  if (!s) return OME_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->hold = on != 0;
  }
  s->cv_work.notify_all();
  return OME_OK;
}

int ome_server_set_request_limit(ome_server* s, size_t bytes) {
  if (!s || bytes == 0) return sfail(s, OME_ERR_INVALID_ARG, "invalid request limit");
  std::lock_guard<std::mutex> lk(s->mu);
  s->req_limit = bytes;
  return OME_OK;
}

int ome_server_get_stats(ome_server* s, ome_server_stats* stats) {
  if (!s || !stats) return sfail(s, OME_ERR_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *stats = s->stats;
  stats->kept_bytes = (int64_t)s->free_bytes;
  return OME_OK;
}

}  // extern "C"
