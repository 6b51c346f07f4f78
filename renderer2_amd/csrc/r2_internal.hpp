// Internal host-side declarations shared by the translation units of the library.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/r2.h"
#include "r2_types.hpp"

struct r2_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // second stream for work that overlaps the main one (the filter's record fill),
  // joined back through side_done
  hipStream_t side = nullptr;
  hipEvent_t side_done = nullptr;
  // marks the main stream's queued work, for the side stream to wait on
  hipEvent_t main_mark = nullptr;
  std::string err;
  r2_stats stats{};
  // reusable device scratch for renders
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // the scene image (scene, instance table, hierarchy) of the last render, and
  // its host copy, kept until the next upload
  r2::DScene* d_scene = nullptr;
  size_t scene_cap = 0;
  std::vector<unsigned char> scene_img;
  uint32_t* d_counter = nullptr;
  r2_filter_stats fstats{};
  // grow-only device buffers reused across calls (filter binning scratch)
  static constexpr int kPoolSlots = 24;
  void* pool[kPoolSlots] = {};
  size_t pool_bytes[kPoolSlots] = {};
};

struct r2_grid {
  r2::DGrid dev{};            // device view (pointers into `allocs`)
  r2_grid_info info{};
  std::vector<void*> allocs;  // owned device allocations
  uint32_t* range0 = nullptr;  // level-0 brick ranges (also in allocs)
  uint32_t* mip[3] = {nullptr, nullptr, nullptr};
  float voxel_size_filtering = 0.f;
  float sigma_a[3] = {0, 0, 0};
  float sigma_s[3] = {0, 0, 0};
};

struct r2_env {
  r2::DEnv dev{};
  std::vector<void*> allocs;
};

namespace r2 {
// device entry of a batched launch: a request's scene, chunk sums and item range
struct BatchEntry {
  const DScene* scene;  // device copy of the request's scene
  void* partial;        // chunk sums of the request (Acc[nchunks * npix * 3])
  uint64_t first;       // first item of this slice in the request's item space
  uint64_t count;       // items of this slice
  int chunk;            // samples per item
  int pad;
};
}  // namespace r2

// integrator launchers of the two record-precision instantiations
#define R2K_DECLARE(NS)                                                                                              \
  namespace NS {                                                                                                     \
  void launch_sm(hipStream_t st, int blocks, const r2::DScene* d_scene, uint64_t* counter, uint64_t total, int chunk, \
                 void* partial, bool fp64, bool single, bool surf, bool count, const int32_t* pp, const int32_t* ps, \
                 float* po, uint32_t* pc);                                                                           \
  void launch_batch_group(hipStream_t st, int blocks, const void* ents, int n, uint64_t* counters, bool single,      \
                          bool surf, bool fp64);                                                                     \
  void launch_reduce_k(hipStream_t st, const void* partial, bool fp64, int nchunks, uint32_t n3, float* out);        \
  }
R2K_DECLARE(r2k_fp32)
R2K_DECLARE(r2k_packed)
#undef R2K_DECLARE

namespace r2host {
int fail(r2_ctx* ctx, int code, const std::string& msg);
int hip_fail(r2_ctx* ctx, hipError_t e, const char* what);
// grid build (r2_grid.hip)
int grid_build(r2_ctx* ctx, const r2_grid_desc* desc, r2_grid* g);
int grid_lookup(r2_ctx* ctx, const r2_grid* g, int n, const int32_t* coords, uint8_t* out);
int grid_sample(r2_ctx* ctx, const r2_grid* g, int n, const float* pos, float* out);
int dda_cells(r2_ctx* ctx, int n, const float* rays, float sigma_t, int collide_every, int max_cells, int32_t* cells,
              int32_t* counts);
// A render's scene as the host assembles it: the device scene and the tables that
// follow it in device memory (the instances, and their hierarchy when there are
// more than kLinearInstances), and the record precision of its grids (0 FP32,
// 1 packed, -1 mixed).
struct HostScene {
  r2::DScene s{};
  std::vector<r2::DInstance> inst;
  std::vector<r2::DBvhNode> bvh;
  std::vector<int32_t> bvh_idx;
  int precision = 0;
  size_t device_bytes() const;
  // the device image for base address `dev`: the scene with its table pointers
  // set into the image, then the tables
  void pack(void* dev, std::vector<unsigned char>& out) const;
};
// instances a context's preallocated scene image holds (r2_render_noalloc's limit)
constexpr int kCtxSceneInstances = 4096;
size_t scene_bytes_for(int n_inst);
// copy the scene image to ctx->d_scene on ctx->stream, growing the buffer when
// may_alloc (else R2_ERR_UNSUPPORTED beyond kCtxSceneInstances)
int upload_scene(r2_ctx* ctx, const HostScene& hs, bool may_alloc);
// render kernels (r2_render.hip)
int launch_render(r2_ctx* ctx, const HostScene& scene, float* d_out, bool fp64_accum);
struct RenderOpts {
  // called after each sample chunk's launch (one launch per chunk); nonzero stops the frame
  r2_progress_fn progress = nullptr;
  void* user = nullptr;
  // caller-provided chunk-sum buffer (no allocation); nullptr: the context's scratch
  void* partial = nullptr;
  size_t partial_bytes = 0;
  // false: the scene image must fit the context's preallocated buffer
  bool scene_alloc = true;
};
int launch_render_ex(r2_ctx* ctx, const HostScene& scene, float* d_out, bool fp64_accum, const RenderOpts& o);
int launch_paths(r2_ctx* ctx, const HostScene& scene, int n, const int32_t* d_pix, const int32_t* d_smp, float* d_out,
                 uint32_t* d_cnt);
// chunk-sum scratch of r2_render_noalloc: a file-scope device array of
// kNoAllocScratchBytes in r2_render.hip, one copy per GPU for the process
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
// looks up the array's address on the current device once (called by r2_create)
int noalloc_scratch_init(r2_ctx* ctx);
// that address for ctx's device, or nullptr when the look-up failed
void* noalloc_scratch(const r2_ctx* ctx);
// one request's share of a batched launch (render server)
struct BatchSlice {
  const r2::DScene* d_scene;
  void* d_partial;
  uint64_t first, count;
  int chunk;
  bool single, surf, fp64, packed;
};
constexpr int kMaxBatchEntries = 256;
void render_chunking(const r2::DScene& scene, int* chunk, int* nchunks);
int launch_reduce(hipStream_t st, const void* partial, bool fp64, int nchunks, uint32_t n3, float* out);
// d_ents: kMaxBatchEntries entries of device scratch, d_counters: as many u64
int launch_batch(r2_ctx* ctx, hipStream_t st, const std::vector<BatchSlice>& items, void* d_ents, uint64_t* d_counters);
size_t batch_entry_bytes();
// scene assembly (r2_api.cpp)
int build_scene(r2_ctx* ctx, const r2_render_params* p, HostScene* out);
// scene-intersection probe (r2_grid.hip): n x (hit, t, instance, surface)
int scene_intersect_probe(r2_ctx* ctx, const HostScene& hs, int n, const float* rays7, float* out4);
int ensure_scratch(r2_ctx* ctx, size_t bytes);
// device buffer of at least `bytes` in pool slot `slot` (contents undefined)
hipError_t pool_get(r2_ctx* ctx, int slot, size_t bytes, void** out);
// getRotation (functions.hpp:128-142), host only
r2::m44 get_rotation(r2::v3 n);
// NanoVDB grid files of dense 32-byte records (r2_nvdb.cpp)
void nvdb_write(const std::string& path, int nx, int ny, int nz, const uint8_t* records, const float* map_mat,
                const float* map_vec);
void nvdb_read(const std::string& path, r2_nvdb_info* info, uint8_t* records, int64_t records_bytes);
// the file's active values as sparse 8^3 leaves (records NULL: count only); returns the leaf count
int64_t nvdb_read_leaves(const std::string& path, r2_nvdb_info* info, int32_t shift[3], int64_t max_leaves,
                         int32_t* origins, uint8_t* records, uint8_t* masks);
// image files of 8-bit RGB frames, rows bottom to top, written flipped (r2_image.cpp)
void write_png(const std::string& path, int w, int h, const uint8_t* rgb);
void write_jpg(const std::string& path, int w, int h, const uint8_t* rgb, int quality);
}  // namespace r2host

#define R2_HIP(ctx, call)                                   \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return r2host::hip_fail((ctx), _e, #call); \
  } while (0)
