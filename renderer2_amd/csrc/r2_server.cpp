// Device-resident render server with continuous batching.
//
// The reference has no server; its nearest relative is the interactive preview
// loop (raytracer/pipeline.cu:41-78) that renders one frame at a time.  Here a
// host scheduler thread owns a HIP stream and
//   * admits queued requests earliest deadline first (requests without a
//     deadline after those with one, in arrival order) while the sum of their
//     device memory (scene copy, chunk sums, frame) fits the HBM budget and
//     fewer than max_inflight are resident;
//   * drops a request, queued or resident, that is still unfinished at its
//     deadline (checked at least once per slice; R2_ERR_DEADLINE);
//   * runs slices: one batched persistent launch (k_paths_batch) whose work is
//     an equal share of paths from every resident request, so requests that
//     arrive while others run join at the next slice and their work overlaps
//     on the device; two slices are kept queued on the stream so the GPU does
//     not idle while the host retires one and plans the next;
//   * reduces a request's chunk sums when all its items are done (on a second
//     stream, after the slice that finished it), copies the frame to the
//     caller's buffer and releases its memory, which admits the next.
// Request memory is carved out of one device buffer of the HBM budget,
// allocated when the server starts: admission and release are host-side
// bookkeeping, so the scheduler loop never calls hipMalloc or hipFree (hipFree
// would wait for the whole device and drain the slices queued on the stream).
// A request is split into the same work items (sample chunk x pixel) and the
// same chunk-ordered reduction as a standalone r2_render, so its result is the
// same alone and in any mix.
#include "r2_internal.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { RQ_QUEUED, RQ_ACTIVE, RQ_DONE, RQ_CANCELLED, RQ_FAILED, RQ_EXPIRED };

struct SrvReq {
  int64_t id = 0;
  r2_render_params params{};
  std::vector<r2_instance> inst;
  std::vector<r2_surface> surfs;
  float* out = nullptr;
  r2host::HostScene hs;             // the request's scene, instance table and hierarchy
  const r2::DScene& scene = hs.s;
  std::vector<unsigned char> scene_img;  // host image of the device copy (kept until the request ends)
  int chunk = 1, nchunks = 1;
  bool fp64 = false, single = true, surf = false, packed = false;
  uint64_t total_items = 0, next_item = 0;  // items in all, items handed to slices
  uint64_t done_items = 0;                   // items of completed slices
  int inflight = 0;                          // slices on the GPU that hold items of this request
  size_t bytes = 0;
  void* d_mem = nullptr;
  r2::DScene* d_scene = nullptr;
  void* d_partial = nullptr;
  float* d_frame = nullptr;
  ReqState state = RQ_QUEUED;
  bool cancel = false;
  bool expired = false;                      // dropped at its deadline (with cancel set)
  bool has_deadline = false;
  int rc = R2_OK;
  Clock::time_point t_submit, t_done, deadline;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// First-fit allocator over one device buffer: 256-byte aligned ranges, the free
// list kept by offset and merged with its neighbours on release.
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void reset(char* b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    if (n) free_[0] = n;
  }
  char* alloc(size_t bytes) {
    bytes = align256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      return base + off;
    }
    return nullptr;
  }
  void release(char* p, size_t bytes) {
    bytes = align256(bytes);
    size_t off = (size_t)(p - base);
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
};

}  // namespace

struct r2_server {
  r2_ctx* ctx = nullptr;
  r2_ctx shadow;  // error sink and stream of the scheduler thread
  r2_server_config cfg{};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> queue;
  std::map<int64_t, std::unique_ptr<SrvReq>> reqs;
  std::vector<int64_t> active;
  bool stopping = false;
  int64_t next_id = 0;
  int64_t inflight_bytes = 0, peak_bytes = 0;
  int64_t n_submitted = 0, n_completed = 0, n_rejected = 0, n_cancelled = 0, n_expired = 0;
  std::vector<double> latencies;
  int64_t paths_done = 0;
  bool have_first = false;
  Clock::time_point t_first, t_last;
  void* d_ents = nullptr;       // two halves: one per slice in flight
  uint64_t* d_counters = nullptr;
  hipStream_t copy_stream = nullptr;  // chunk reduction and frame copy-out
  void* d_pool = nullptr;             // the HBM budget, sub-allocated per request
  Arena arena;
};

namespace {

void finish_locked(r2_server* s, SrvReq& r, ReqState st, int rc) {
  if (r.d_mem) {
    // no slice in flight and no copy-out uses the range any more
    s->arena.release((char*)r.d_mem, r.bytes);
    r.d_mem = nullptr;
    s->inflight_bytes -= (int64_t)r.bytes;
  }
  r.state = st;
  r.rc = rc;
  r.t_done = Clock::now();
  s->active.erase(std::remove(s->active.begin(), s->active.end(), r.id), s->active.end());
  if (st == RQ_DONE) {
    ++s->n_completed;
    s->latencies.push_back(std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count());
    s->paths_done += (int64_t)r.scene.W * r.scene.H * r.scene.spp;
    s->t_last = r.t_done;
  } else if (st == RQ_CANCELLED) {
    ++s->n_cancelled;
  } else if (st == RQ_EXPIRED) {
    ++s->n_expired;
  }
  s->cv_done.notify_all();
}

// a request that no slice in flight uses any more and that was cancelled or has
// expired: release it with the matching outcome (lock held)
void drop_locked(r2_server* s, SrvReq& r) {
  if (r.expired) finish_locked(s, r, RQ_EXPIRED, R2_ERR_DEADLINE);
  else finish_locked(s, r, RQ_CANCELLED, R2_ERR_CANCELLED);
}

// mark requests whose deadline has passed; queued requests that expired or were
// cancelled leave the queue at once, resident ones are dropped once no slice in
// flight still uses them (lock held)
void expire_locked(r2_server* s) {
  const Clock::time_point now = Clock::now();
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    SrvReq& r = *s->reqs[*it];
    if (r.has_deadline && !r.cancel && now >= r.deadline) r.cancel = r.expired = true;
    if (r.cancel) {
      it = s->queue.erase(it);
      drop_locked(s, r);
    } else {
      ++it;
    }
  }
  for (int64_t id : s->active) {
    SrvReq& r = *s->reqs[id];
    if (r.has_deadline && !r.cancel && now >= r.deadline) r.cancel = r.expired = true;
  }
}

// the queued request to admit next: earliest deadline, then arrival order
std::deque<int64_t>::iterator next_queued(r2_server* s) {
  auto best = s->queue.begin();
  for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
    const SrvReq& r = *s->reqs[*it];
    const SrvReq& b = *s->reqs[*best];
    if (r.has_deadline && (!b.has_deadline || r.deadline < b.deadline)) best = it;
  }
  return best;
}

// admit queued requests, earliest deadline first, while memory and slots allow
// (lock held); the next request waits for memory rather than being overtaken
void admit_locked(r2_server* s) {
  expire_locked(s);
  while (!s->queue.empty() && (int)s->active.size() < s->cfg.max_inflight) {
    const auto qi = next_queued(s);
    SrvReq& r = *s->reqs[*qi];
    if (r.cancel) {
      s->queue.erase(qi);
      drop_locked(s, r);
      continue;
    }
    if (s->inflight_bytes + (int64_t)r.bytes > s->cfg.hbm_budget_bytes) break;
    void* mem = s->arena.alloc(r.bytes);
    if (!mem) {
      if (s->active.empty()) {  // nothing to wait for: fail the request
        s->queue.erase(qi);
        finish_locked(s, r, RQ_FAILED, R2_ERR_OOM);
        continue;
      }
      break;  // fragmented: wait for a release
    }
    s->queue.erase(qi);
    r.d_mem = mem;
    r.d_scene = (r2::DScene*)mem;
    r.d_partial = (char*)mem + align256(r.hs.device_bytes());
    const size_t part = (size_t)r.nchunks * r.scene.W * r.scene.H * 3 * (r.fp64 ? 8 : 4);
    r.d_frame = (float*)((char*)r.d_partial + align256(part));
    r.hs.pack(r.d_scene, r.scene_img);
    if (hipMemcpyAsync(r.d_scene, r.scene_img.data(), r.scene_img.size(), hipMemcpyHostToDevice, s->shadow.stream) !=
        hipSuccess) {
      finish_locked(s, r, RQ_FAILED, R2_ERR_HIP);
      continue;
    }
    s->inflight_bytes += (int64_t)r.bytes;
    s->peak_bytes = std::max(s->peak_bytes, s->inflight_bytes);
    r.state = RQ_ACTIVE;
    s->active.push_back(r.id);
    if (!s->have_first) {
      s->have_first = true;
      s->t_first = Clock::now();
    }
  }
}

struct Slice {
  std::vector<std::pair<int64_t, uint64_t>> took;  // (request, items)
  hipEvent_t done = nullptr;
  int half = 0;
  int rc = R2_OK;
};

// reduce a finished request on the copy stream (after its last slice) and copy it out
void complete_locked(r2_server* s, SrvReq& r, hipEvent_t after, std::unique_lock<std::mutex>& lk) {
  const uint32_t n3 = (uint32_t)((uint64_t)r.scene.W * r.scene.H * 3);
  int rr = hipStreamWaitEvent(s->copy_stream, after, 0) == hipSuccess ? R2_OK : R2_ERR_HIP;
  if (rr == R2_OK) rr = r2host::launch_reduce(s->copy_stream, r.d_partial, r.fp64, r.nchunks, n3, r.d_frame);
  if (rr == R2_OK &&
      hipMemcpyAsync(r.out, r.d_frame, (size_t)n3 * 4, hipMemcpyDeviceToHost, s->copy_stream) != hipSuccess)
    rr = R2_ERR_HIP;
  lk.unlock();
  if (rr == R2_OK && hipStreamSynchronize(s->copy_stream) != hipSuccess) rr = R2_ERR_HIP;
  lk.lock();
  finish_locked(s, r, rr == R2_OK ? RQ_DONE : RQ_FAILED, rr);
}

void scheduler(r2_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::deque<Slice> flight;  // at most two slices queued on the stream
  hipEvent_t ev[2];
  for (auto& e : ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
  int next_half = 0;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    admit_locked(s);
    // drop cancelled or expired requests that no slice in flight still uses
    for (int64_t id : std::vector<int64_t>(s->active)) {
      SrvReq& r = *s->reqs[id];
      if (r.cancel && r.inflight == 0) drop_locked(s, r);
    }
    // plan and launch a slice while fewer than two are in flight.  Half of the
    // slice's paths are shared equally by the resident requests with items left,
    // the other half goes to the requests with the least work left first
    // (shortest remaining work first), so that small requests finish within a
    // few slices while large ones keep progressing; whole 64-item tiles
    bool launched = false;
    if (flight.size() < 2) {
      std::vector<r2host::BatchSlice> slice;
      Slice sl;
      std::vector<int64_t> cand;
      for (int64_t id : s->active) {
        SrvReq& r = *s->reqs[id];
        if (!r.cancel && r.next_item < r.total_items) cand.push_back(id);
      }
      auto left_paths = [&](int64_t id) {
        const SrvReq& r = *s->reqs[id];
        return (r.total_items - r.next_item) * (uint64_t)r.chunk;
      };
      std::stable_sort(cand.begin(), cand.end(), [&](int64_t a, int64_t b) { return left_paths(a) < left_paths(b); });
      const int n_act = std::min<int>((int)cand.size(), r2host::kMaxBatchEntries / 2);
      if (n_act > 0) {
        const uint64_t budget = (uint64_t)s->cfg.slice_paths;
        const uint64_t share = std::max<uint64_t>(1, budget / (2 * (uint64_t)n_act));
        uint64_t extra = budget / 2;  // handed out in order of least work left
        for (int i = 0; i < n_act; ++i) {
          SrvReq& r = *s->reqs[cand[i]];
          const uint64_t left = r.total_items - r.next_item;
          uint64_t items = std::max<uint64_t>(64 * 64, share / (uint64_t)r.chunk);
          const uint64_t more = std::min<uint64_t>(extra / (uint64_t)r.chunk, left > items ? left - items : 0);
          items += more;
          extra -= more * (uint64_t)r.chunk;
          items = (items + 63) & ~(uint64_t)63;
          items = std::min<uint64_t>(items, left);
          slice.push_back(r2host::BatchSlice{r.d_scene, r.d_partial, r.next_item, items, r.chunk, r.single, r.surf,
                                             r.fp64, r.packed});
          sl.took.push_back({r.id, items});
          r.next_item += items;
          ++r.inflight;
        }
        sl.half = next_half;
        next_half ^= 1;
        sl.done = ev[sl.half];
        void* ents = (char*)s->d_ents + (size_t)sl.half * r2host::batch_entry_bytes() * (r2host::kMaxBatchEntries / 2);
        uint64_t* ctr = s->d_counters + (size_t)sl.half * (r2host::kMaxBatchEntries / 2);
        lk.unlock();
        sl.rc = r2host::launch_batch(&s->shadow, s->shadow.stream, slice, ents, ctr);
        if (sl.rc == R2_OK && hipEventRecord(sl.done, s->shadow.stream) != hipSuccess) sl.rc = R2_ERR_HIP;
        lk.lock();
        flight.push_back(std::move(sl));
        launched = true;
      }
    }
    if (flight.empty()) {
      if (s->stopping && s->active.empty() && s->queue.empty()) break;
      if (s->active.empty() || !launched) {
        bool work = false;
        for (int64_t id : s->active) work = work || s->reqs[id]->cancel;
        if (!work && !s->stopping) s->cv_work.wait(lk);
        else if (s->stopping && s->active.empty()) break;
      }
      continue;
    }
    if (launched && flight.size() < 2) continue;  // keep the stream two slices deep
    // retire the oldest slice
    Slice sl = std::move(flight.front());
    flight.pop_front();
    lk.unlock();
    int rc = sl.rc;
    if (rc == R2_OK && hipEventSynchronize(sl.done) != hipSuccess) rc = R2_ERR_HIP;
    lk.lock();
    for (auto& t : sl.took) {
      SrvReq& r = *s->reqs[t.first];
      --r.inflight;
      if (rc != R2_OK) {
        if (r.inflight == 0) finish_locked(s, r, RQ_FAILED, rc);
        else r.cancel = true;
        continue;
      }
      r.done_items += t.second;
      if (r.cancel || r.done_items < r.total_items) continue;
      complete_locked(s, r, sl.done, lk);
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
}

}  // namespace

extern "C" int r2_server_start(r2_ctx* ctx, const r2_server_config* cfg, r2_server** out) {
  if (!ctx || !out) return R2_ERR_INVALID;
  *out = nullptr;
  auto s = std::make_unique<r2_server>();
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (s->cfg.hbm_budget_bytes <= 0) s->cfg.hbm_budget_bytes = 8ll << 30;
  if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 16;
  s->cfg.max_inflight = std::min(s->cfg.max_inflight, r2host::kMaxBatchEntries / 2);
  if (s->cfg.slice_paths <= 0) s->cfg.slice_paths = 1 << 24;
  s->shadow.device = ctx->device;
  if (hipSetDevice(ctx->device) != hipSuccess) return r2host::fail(ctx, R2_ERR_HIP, "server: hipSetDevice");
  if (hipStreamCreate(&s->shadow.stream) != hipSuccess) return r2host::fail(ctx, R2_ERR_HIP, "server: stream");
  if (hipStreamCreate(&s->copy_stream) != hipSuccess) {
    (void)hipStreamDestroy(s->shadow.stream);
    return r2host::fail(ctx, R2_ERR_HIP, "server: stream");
  }
  if (hipMalloc(&s->d_ents, r2host::batch_entry_bytes() * r2host::kMaxBatchEntries) != hipSuccess ||
      hipMalloc((void**)&s->d_counters, sizeof(uint64_t) * r2host::kMaxBatchEntries) != hipSuccess ||
      hipMalloc(&s->d_pool, (size_t)s->cfg.hbm_budget_bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (s->d_ents) (void)hipFree(s->d_ents);
    if (s->d_counters) (void)hipFree(s->d_counters);
    (void)hipStreamDestroy(s->copy_stream);
    (void)hipStreamDestroy(s->shadow.stream);
    return r2host::fail(ctx, R2_ERR_OOM, "server: device tables or the HBM budget do not fit the device");
  }
  s->arena.reset((char*)s->d_pool, (size_t)s->cfg.hbm_budget_bytes);
  r2_server* raw = s.release();
  raw->th = std::thread(scheduler, raw);
  *out = raw;
  return R2_OK;
}

extern "C" int64_t r2_server_submit(r2_server* s, const r2_request* req) {
  if (!s || !req || !req->out_rgb) return R2_ERR_INVALID;
  if (!(req->deadline_ms >= 0.0) || !std::isfinite(req->deadline_ms)) return R2_ERR_INVALID;
  auto r = std::make_unique<SrvReq>();
  r->params = req->params;
  if (r->params.n_instances < 0 || r->params.n_instances > r2::kMaxInstances ||
      (r->params.n_instances > 0 && !r->params.instances))
    return R2_ERR_INVALID;
  r->inst.assign(r->params.instances, r->params.instances + r->params.n_instances);
  r->params.instances = r->inst.data();
  if (r->params.n_surfaces < 0 || r->params.n_surfaces > r2::kMaxSurfaces ||
      (r->params.n_surfaces > 0 && !r->params.surfaces))
    return R2_ERR_INVALID;
  r->surfs.assign(r->params.surfaces, r->params.surfaces + r->params.n_surfaces);
  r->params.surfaces = r->surfs.data();
  r->out = req->out_rgb;
  {
    // scene assembly on the caller's thread (validates the request)
    r2_ctx probe;
    probe.device = s->ctx->device;
    const int rc = r2host::build_scene(&probe, &r->params, &r->hs);
    if (rc != R2_OK) return rc;
  }
  r2host::render_chunking(r->scene, &r->chunk, &r->nchunks);
  r->fp64 = (r->params.flags & R2_FLAG_FP64_ACCUM) != 0;
  // scenes with surfaces run the general-instance form of the batched kernel
  r->surf = r->scene.n_surf > 0;
  r->single = r->scene.n_inst == 1 && !r->surf;
  if (r->hs.precision < 0) return R2_ERR_UNSUPPORTED;  // instances mix FP32 and packed grids
  r->packed = r->hs.precision == 1;
  const uint64_t npix = (uint64_t)r->scene.W * r->scene.H;
  r->total_items = (uint64_t)r->nchunks * npix;
  r->bytes = align256(r->hs.device_bytes()) + align256((size_t)r->nchunks * npix * 3 * (r->fp64 ? 8 : 4)) +
             align256((size_t)npix * 3 * 4);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) return R2_ERR_CANCELLED;
  ++s->n_submitted;
  if ((int64_t)r->bytes > s->cfg.hbm_budget_bytes) {
    ++s->n_rejected;
    return R2_ERR_OOM;
  }
  const int64_t id = s->next_id++;
  r->id = id;
  r->t_submit = Clock::now();
  if (req->deadline_ms > 0.0) {
    r->has_deadline = true;
    r->deadline = r->t_submit + std::chrono::duration_cast<Clock::duration>(
                                    std::chrono::duration<double, std::milli>(req->deadline_ms));
  }
  s->queue.push_back(id);
  s->reqs[id] = std::move(r);
  s->cv_work.notify_one();
  return id;
}

// state of a request; a finished one is reported once and then forgotten, so the
// table holds only requests whose outcome has not been collected yet
static int poll_locked(r2_server* s, int64_t id, double* lat) {
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return R2_ERR_INVALID;
  SrvReq& r = *it->second;
  if (r.state == RQ_QUEUED || r.state == RQ_ACTIVE) return 0;
  if (lat) *lat = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  const int rc = r.state == RQ_DONE ? 1 : (r.rc < 0 ? r.rc : R2_ERR_HIP);
  s->reqs.erase(it);
  return rc;
}

extern "C" int r2_server_poll(r2_server* s, int64_t id, double* latency_ms) {
  if (!s) return R2_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  return poll_locked(s, id, latency_ms);
}

extern "C" int r2_server_wait(r2_server* s, int64_t id, double* latency_ms) {
  if (!s) return R2_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    const int rc = poll_locked(s, id, latency_ms);
    if (rc != 0) return rc;
    s->cv_done.wait(lk);
  }
}

extern "C" int r2_server_cancel(r2_server* s, int64_t id) {
  if (!s) return R2_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return R2_ERR_INVALID;
  it->second->cancel = true;
  s->cv_work.notify_one();
  return R2_OK;
}

extern "C" int r2_server_get_stats(r2_server* s, r2_server_stats* out) {
  if (!s || !out) return R2_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  memset(out, 0, sizeof(*out));
  out->submitted = s->n_submitted;
  out->completed = s->n_completed;
  out->rejected = s->n_rejected;
  out->cancelled = s->n_cancelled;
  out->expired = s->n_expired;
  out->peak_bytes = s->peak_bytes;
  if (s->have_first && s->n_completed > 0) {
    const double span = std::chrono::duration<double>(s->t_last - s->t_first).count();
    out->paths_per_sec = span > 0 ? (double)s->paths_done / span : 0.0;
  }
  if (!s->latencies.empty()) {
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    // nearest-rank percentiles
    auto pct = [&](double q) {
      size_t k = (size_t)std::ceil(q * (double)l.size());
      k = std::min(std::max<size_t>(k, 1), l.size());
      return l[k - 1];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
  }
  return R2_OK;
}

extern "C" int r2_server_stop(r2_server* s) {
  if (!s) return R2_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    for (auto& kv : s->reqs)
      if (kv.second->state == RQ_QUEUED || kv.second->state == RQ_ACTIVE) kv.second->cancel = true;
    s->cv_work.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& kv : s->reqs) {
      SrvReq& r = *kv.second;
      if (r.state == RQ_QUEUED) finish_locked(s, r, RQ_CANCELLED, R2_ERR_CANCELLED);
    }
  }
  (void)hipFree(s->d_ents);
  (void)hipFree(s->d_counters);
  (void)hipFree(s->d_pool);
  (void)hipStreamDestroy(s->copy_stream);
  (void)hipStreamDestroy(s->shadow.stream);
  delete s;
  return R2_OK;
}
