// Volume integrator kernels for gfx950.
//
// Work item = (sample chunk, pixel), items numbered so that 64 consecutive items
// are one 8x8 pixel tile of one chunk (coherent primary rays).  A persistent
// grid of wavefronts runs the path state machine of r2_pathsm.hpp; each lane
// runs the K samples of its item in order and writes the chunk sum; a second
// kernel adds the chunk sums in chunk order.  Results are independent of
// scheduling: every path's random stream is keyed by (seed, stream, pixel,
// sample) and the reduction order is fixed.
#define R2_PREC_MODE 0
#define R2K_NS r2k_fp32
#include "r2_render_kernels.hpp"
#include <mutex>

using namespace r2;

namespace r2host {

static int g_num_cus = 0;

static int num_cus(r2_ctx* ctx) {
  if (!g_num_cus) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) v = 0;
    g_num_cus = v > 0 ? v : 256;
  }
  return g_num_cus;
}

static void launch_sm(r2_ctx* ctx, int prec, int blocks, uint64_t total, int chunk, void* partial, bool fp64, bool single,
                      bool surf, bool count, const int32_t* pp, const int32_t* ps, float* po, uint32_t* pc) {
  if (prec == 1)
    r2k_packed::launch_sm(ctx->stream, blocks, ctx->d_scene, (uint64_t*)ctx->d_counter, total, chunk, partial, fp64,
                          single, surf, count, pp, ps, po, pc);
  else
    r2k_fp32::launch_sm(ctx->stream, blocks, ctx->d_scene, (uint64_t*)ctx->d_counter, total, chunk, partial, fp64,
                        single, surf, count, pp, ps, po, pc);
}

size_t batch_entry_bytes() { return sizeof(r2::BatchEntry); }

// samples per work item: about 2^24 items per frame, so that a frame (and each
// share of one that the render server runs as a slice) has many short items for
// the 200 K lanes of the chip; the chunk sums take 12 B per item
void render_chunking(const r2::DScene& scene, int* chunk_out, int* nchunks_out) {
  const uint64_t npix = (uint64_t)scene.W * scene.H;
  const uint64_t target = 1ull << 24;
  const int chunk =
      (int)std::min<uint64_t>((uint64_t)scene.spp, std::max<uint64_t>(1, (npix * scene.spp + target - 1) / target));
  *chunk_out = chunk;
  *nchunks_out = (scene.spp + chunk - 1) / chunk;
}

int launch_reduce(hipStream_t st, const void* partial, bool fp64, int nchunks, uint32_t n3, float* out) {
  r2k_fp32::launch_reduce_k(st, partial, fp64, nchunks, n3, out);
  return hipGetLastError() == hipSuccess ? R2_OK : R2_ERR_HIP;
}

int launch_batch(r2_ctx* ctx, hipStream_t st, const std::vector<BatchSlice>& items, void* d_ents, uint64_t* d_counters) {
  using namespace r2;
  if (items.empty()) return R2_OK;
  // one launch per (single instance, surfaces, accumulator type, record precision) group
  std::vector<BatchEntry> host(items.size());
  int off = 0;
  for (int g = 0; g < 16; ++g) {
    const bool single = (g & 1) != 0, fp64 = (g & 2) != 0, packed = (g & 4) != 0, surf = (g & 8) != 0;
    if (single && surf) continue;  // surface scenes run the general-instance form, whatever their instance count
    const int first = off;
    uint64_t total = 0;
    for (const BatchSlice& b : items) {
      if (b.surf != surf || (!surf && b.single != single) || b.fp64 != fp64 || b.packed != packed || b.count == 0)
        continue;
      BatchEntry& e = host[off++];
      e.scene = b.d_scene;
      e.partial = b.d_partial;
      e.first = b.first;
      e.count = b.count;
      e.chunk = b.chunk;
      e.pad = 0;
      total += b.count;
    }
    const int n = off - first;
    if (n == 0) continue;
    BatchEntry* de = (BatchEntry*)d_ents + first;
    R2_HIP(ctx, hipMemcpyAsync(de, &host[first], sizeof(BatchEntry) * n, hipMemcpyHostToDevice, st));
    R2_HIP(ctx, hipMemsetAsync(d_counters + first, 0, sizeof(uint64_t) * n, st));
    const int blocks = (int)std::min<uint64_t>((uint64_t)num_cus(ctx) * 8, (total + 255) / 256);
    if (packed) r2k_packed::launch_batch_group(st, blocks, de, n, d_counters + first, single, surf, fp64);
    else r2k_fp32::launch_batch_group(st, blocks, de, n, d_counters + first, single, surf, fp64);
    R2_HIP(ctx, hipGetLastError());
  }
  return R2_OK;
}

// The frame's work items are numbered chunk-major (item = chunk * npix + pixel
// tile order), so the items of sample chunk c are the range [c npix, (c + 1) npix).
// A render runs them as one launch, or (with a progress callback) one launch per
// sample chunk, like the reference's one launch per sample: a launch takes the
// items [first, total) by starting the shared item counter at `first`.  Each item
// writes its own chunk sum, so the frame is the same bit for bit either way.
static int launch_items(r2_ctx* ctx, int prec, bool single, bool surf, bool fp64, uint64_t first, uint64_t total,
                        int chunk, void* partial) {
  hipStream_t st = ctx->stream;
  // the counter's low word = first (< 2^32: items of a launch fit 32 bits), high word 0
  R2_HIP(ctx, hipMemsetD32Async((hipDeviceptr_t)ctx->d_counter, (int)(uint32_t)first, 1, st));
  R2_HIP(ctx, hipMemsetD32Async((hipDeviceptr_t)(ctx->d_counter + 1), 0, 1, st));
  const int blocks = (int)std::min<uint64_t>((uint64_t)num_cus(ctx) * 8, (total - first + 255) / 256);
  launch_sm(ctx, prec, blocks, total, chunk, partial, fp64, single, surf, false, nullptr, nullptr, nullptr, nullptr);
  R2_HIP(ctx, hipGetLastError());
  return R2_OK;
}

int launch_render_ex(r2_ctx* ctx, const HostScene& hs, float* d_out, bool fp64, const RenderOpts& o) {
  using namespace r2;
  const DScene& scene = hs.s;
  const uint64_t npix = (uint64_t)scene.W * scene.H;
  int chunk = 1, nchunks = 1;
  render_chunking(scene, &chunk, &nchunks);
  const size_t acc_bytes = fp64 ? 8 : 4;
  void* partial = nullptr;
  if (o.partial) {
    // caller-provided chunk-sum buffer: fewer, longer items when the frame's
    // default chunking does not fit it
    const uint64_t cap_items = o.partial_bytes / (3 * acc_bytes);
    if (cap_items < npix) return fail(ctx, R2_ERR_UNSUPPORTED, "render: frame larger than the chunk-sum buffer");
    if ((uint64_t)nchunks * npix > cap_items) {
      const uint64_t max_chunks = cap_items / npix;
      chunk = (int)((scene.spp + max_chunks - 1) / max_chunks);
      nchunks = (scene.spp + chunk - 1) / chunk;
    }
    partial = o.partial;
  } else {
    const size_t part_bytes = (size_t)nchunks * npix * 3 * acc_bytes;
    int rc = ensure_scratch(ctx, part_bytes);
    if (rc) return rc;
    partial = ctx->scratch;
  }
  const uint64_t total = (uint64_t)nchunks * npix;
  hipStream_t st = ctx->stream;
  const bool single = scene.n_inst == 1, surf = scene.n_surf > 0;
  const int prec = hs.precision;
  if (prec < 0) return fail(ctx, R2_ERR_UNSUPPORTED, "render: instances mix FP32 and packed grids");
  {
    const int rc = upload_scene(ctx, hs, o.scene_alloc);
    if (rc) return rc;
  }
  R2_HIP(ctx, hipEventRecord(ctx->ev0, st));
  int done_chunks = nchunks;
  int launches = 1;
  if (!o.progress) {
    int rc = launch_items(ctx, prec, single, surf, fp64, 0, total, chunk, partial);
    if (rc) return rc;
  } else {
    // one launch per sample chunk, with up to two in flight; the callback runs
    // after each one completes and may stop the frame
    launches = nchunks;
    std::vector<hipEvent_t> evs(2, nullptr);
    for (auto& e : evs) R2_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    struct Ev {
      std::vector<hipEvent_t>& v;
      ~Ev() { for (auto e : v) if (e) (void)hipEventDestroy(e); }
    } guard{evs};
    int queued = 0, completed = 0, rc = R2_OK;
    bool stop = false;
    while (completed < nchunks) {
      while (!stop && queued < nchunks && queued - completed < 2) {
        rc = launch_items(ctx, prec, single, surf, fp64, (uint64_t)queued * npix, (uint64_t)(queued + 1) * npix,
                          chunk, partial);
        if (rc) break;
        R2_HIP(ctx, hipEventRecord(evs[queued & 1], st));
        ++queued;
      }
      if (rc) break;
      if (completed == queued) break;
      R2_HIP(ctx, hipEventSynchronize(evs[completed & 1]));
      ++completed;
      if (!stop && o.progress(o.user, std::min(scene.spp, completed * chunk), scene.spp) != 0) stop = true;
    }
    if (rc) {
      (void)hipStreamSynchronize(st);
      return rc;
    }
    done_chunks = completed;
    launches = completed;
  }
  R2_HIP(ctx, hipEventRecord(ctx->ev1, st));
  const uint32_t n3 = (uint32_t)(npix * 3);
  // the frame of the completed sample chunks: sum over paths of L / spp, as the
  // reference's framebuffer holds it when renderFrame's callback stops the loop
  r2k_fp32::launch_reduce_k(st, partial, fp64, done_chunks, n3, d_out);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipStreamSynchronize(st));
  float ms = 0.f;
  R2_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->stats.last_kernel_ms = ms;
#ifdef R2_SM_STATS
  {
    unsigned long long st_host[64];
    R2_HIP(ctx, hipMemcpyFromSymbol(st_host, HIP_SYMBOL(r2k_fp32::g_sm_stats), sizeof(st_host)));
    fprintf(stderr,
            "[sm-stats] paths %llu steps %llu | shade %llu passes %llu lanes %llu cyc | seg %llu passes %llu lanes "
            "%llu cyc | trav %llu steps %llu lanes %llu cyc\n",
            (unsigned long long)(npix * scene.spp), st_host[0], st_host[1], st_host[2], st_host[5], st_host[3],
            st_host[4], st_host[6], st_host[7], st_host[8], st_host[9]);
    const unsigned long long z[64] = {0};
    R2_HIP(ctx, hipMemcpyToSymbol(HIP_SYMBOL(r2k_fp32::g_sm_stats), z, sizeof(z)));
  }
#endif
  ctx->stats.last_paths = (int64_t)std::min<uint64_t>(npix * scene.spp, npix * (uint64_t)done_chunks * chunk);
  ctx->stats.last_launches = launches + 1;
  if (done_chunks < nchunks) return fail(ctx, R2_ERR_CANCELLED, "render: stopped by the progress callback");
  return R2_OK;
}

int launch_render(r2_ctx* ctx, const HostScene& scene, float* d_out, bool fp64) {
  return launch_render_ex(ctx, scene, d_out, fp64, RenderOpts{});
}

// Chunk sums of r2_render_noalloc, for callers that cannot allocate device
// memory.  Each GPU the library runs on holds its own copy of this array until
// the process ends.  In host code the array's name is a host-side stand-in,
// not a device address: host code uses only the per-device address that
// hipGetSymbolAddress returns (kept in g_scratch_dev).  Every item of a frame
// writes its own chunk sum before the reduction reads it, so the array needs no
// clearing between calls; calls that overlap on one GPU share it and race.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kScratchDevices] = {};

int noalloc_scratch_init(r2_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kScratchDevices)
    return fail(ctx, R2_ERR_UNSUPPORTED, "render_noalloc: device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[ctx->device]) return R2_OK;
  void* ptr = nullptr;
  // the calling thread has ctx->device current (r2_create)
  const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
  g_scratch_dev[ctx->device] = ptr;
  return R2_OK;
}

void* noalloc_scratch(const r2_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_dev[ctx->device];
}

int launch_paths(r2_ctx* ctx, const HostScene& hs, int n, const int32_t* d_pix, const int32_t* d_smp, float* d_out,
                 uint32_t* d_cnt) {
  hipStream_t st = ctx->stream;
  const r2::DScene& scene = hs.s;
  const int prec = hs.precision;
  if (prec < 0) return fail(ctx, R2_ERR_UNSUPPORTED, "render: instances mix FP32 and packed grids");
  {
    const int rc = upload_scene(ctx, hs, true);
    if (rc) return rc;
  }
  R2_HIP(ctx, hipMemsetAsync(ctx->d_counter, 0, 8, st));
  const int blocks = (int)std::min<uint64_t>((uint64_t)num_cus(ctx) * 8, ((uint64_t)n + 255) / 256);
  launch_sm(ctx, prec, blocks, (uint64_t)n, 1, nullptr, false, scene.n_inst == 1, scene.n_surf > 0, true, d_pix, d_smp,
            d_out, d_cnt);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipStreamSynchronize(st));
  return R2_OK;
}

}  // namespace r2host
