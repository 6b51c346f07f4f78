// C ABI glue: contexts, grid / env upload, scene assembly and the render entry
// points of include/r2.h.  Host-side transforms follow the reference's host code
// (camera.cu:20-37, sceneobject.cu:108-114,375-377, aabb.hpp:9-13,
// environmentmap.cu:16-29, distribution_1d.cu:6-39, distribution_2d.cu:7-14).
#include "r2_internal.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>

using namespace r2;

namespace r2host {

int fail(r2_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}
int hip_fail(r2_ctx* ctx, hipError_t e, const char* what) {
  if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? R2_ERR_OOM : R2_ERR_HIP;
}
int ensure_scratch(r2_ctx* ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return R2_OK;
  if (ctx->scratch) {
    R2_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  R2_HIP(ctx, hipMalloc(&ctx->scratch, bytes));
  ctx->scratch_bytes = bytes;
  return R2_OK;
}

hipError_t pool_get(r2_ctx* ctx, int slot, size_t bytes, void** out) {
  *out = nullptr;
  if (slot < 0 || slot >= r2_ctx::kPoolSlots) return hipErrorInvalidValue;
  bytes = bytes ? bytes : 1;
  if (ctx->pool_bytes[slot] < bytes) {
    if (ctx->pool[slot]) {
      const hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return e;
      (void)hipFree(ctx->pool[slot]);
      ctx->pool[slot] = nullptr;
      ctx->pool_bytes[slot] = 0;
    }
    // grow by 1/8 to absorb small size changes between calls
    const size_t want = bytes + bytes / 8;
    const hipError_t e = hipMalloc(&ctx->pool[slot], want);
    if (e != hipSuccess) return e;
    ctx->pool_bytes[slot] = want;
  }
  *out = ctx->pool[slot];
  return hipSuccess;
}

// getRotation (functions.hpp:128-142)
m44 get_rotation(v3 n) {
  const float cs = dot(n, mk(0.f, 1.f, 0.f));
  if (cs == 1.f) return identity44();
  if (cs == -1.f) {
    m44 r = identity44();
    for (int i = 0; i < 3; ++i) r.c[i][i] = -1.f;
    return r;
  }
  const float angle = acosf(dot(n, mk(0.f, 1.f, 0.f)));
  const v3 ax = normalize(cross(n, mk(0.f, 1.f, 0.f)));
  const float c = cosf(angle), s = sinf(angle);
  const v3 tmp = ax * (1.f - c);
  m44 r = identity44();
  r.c[0][0] = c + tmp.x * ax.x;
  r.c[0][1] = tmp.x * ax.y + s * ax.z;
  r.c[0][2] = tmp.x * ax.z - s * ax.y;
  r.c[1][0] = tmp.y * ax.x - s * ax.z;
  r.c[1][1] = c + tmp.y * ax.y;
  r.c[1][2] = tmp.y * ax.z + s * ax.x;
  r.c[2][0] = tmp.z * ax.x + s * ax.y;
  r.c[2][1] = tmp.z * ax.y - s * ax.x;
  r.c[2][2] = c + tmp.z * ax.z;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.c[i][j] = roundf(r.c[i][j]);
  return r;
}

// ------------------------------------------------------ instance hierarchy --
// Median-split hierarchy over the instances' world boxes (the stand-in for the
// OptiX instance acceleration structure, scene.cu:188-256): nodes in depth-first
// order with the index of the node after each subtree, leaves of at most 4
// instances.  The split is on the longest axis of the box centres, instances
// ordered by centre then index, so the build is deterministic.
static void bvh_build(const std::vector<DInstance>& inst, std::vector<DBvhNode>& nodes, std::vector<int32_t>& idx) {
  const int n = (int)inst.size();
  idx.resize(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  nodes.clear();
  nodes.reserve(2 * (size_t)n);
  struct Build {
    const std::vector<DInstance>& inst;
    std::vector<DBvhNode>& nodes;
    std::vector<int32_t>& idx;
    void run(int lo, int hi) {
      const int me = (int)nodes.size();
      nodes.push_back(DBvhNode{});
      DBvhNode nd{};
      float clo[3], chi[3];
      for (int a = 0; a < 3; ++a) {
        nd.lo[a] = clo[a] = INFINITY;
        nd.hi[a] = chi[a] = -INFINITY;
      }
      for (int i = lo; i < hi; ++i) {
        const DInstance& in = inst[idx[i]];
        for (int a = 0; a < 3; ++a) {
          nd.lo[a] = std::min(nd.lo[a], in.bb_min[a]);
          nd.hi[a] = std::max(nd.hi[a], in.bb_max[a]);
          const float c = 0.5f * (in.bb_min[a] + in.bb_max[a]);
          clo[a] = std::min(clo[a], c);
          chi[a] = std::max(chi[a], c);
        }
      }
      if (hi - lo <= 4) {
        nd.leaf = (lo << 3) | (hi - lo);
        nd.skip = me + 1;
        nodes[me] = nd;
        return;
      }
      int ax = 0;
      for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
      const int mid = (lo + hi) / 2;
      std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int32_t x, int32_t y) {
        const float cx = inst[x].bb_min[ax] + inst[x].bb_max[ax], cy = inst[y].bb_min[ax] + inst[y].bb_max[ax];
        return cx < cy || (cx == cy && x < y);
      });
      nd.leaf = -1;
      run(lo, mid);
      run(mid, hi);
      nd.skip = (int)nodes.size();
      nodes[me] = nd;
    }
  } b{inst, nodes, idx};
  if (n > 0) b.run(0, n);
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t HostScene::device_bytes() const {
  return align256(sizeof(DScene)) + align256(inst.size() * sizeof(DInstance)) +
         align256(bvh.size() * sizeof(DBvhNode)) + align256(bvh_idx.size() * sizeof(int32_t));
}

void HostScene::pack(void* dev, std::vector<unsigned char>& out) const {
  out.assign(device_bytes(), 0);
  const size_t o_inst = align256(sizeof(DScene));
  const size_t o_bvh = o_inst + align256(inst.size() * sizeof(DInstance));
  const size_t o_idx = o_bvh + align256(bvh.size() * sizeof(DBvhNode));
  DScene d = s;
  char* base = (char*)dev;
  d.inst = (const DInstance*)(base + o_inst);
  d.bvh = bvh.empty() ? nullptr : (const DBvhNode*)(base + o_bvh);
  d.bvh_idx = bvh_idx.empty() ? nullptr : (const int32_t*)(base + o_idx);
  memcpy(out.data(), &d, sizeof(DScene));
  if (!inst.empty()) memcpy(out.data() + o_inst, inst.data(), inst.size() * sizeof(DInstance));
  if (!bvh.empty()) memcpy(out.data() + o_bvh, bvh.data(), bvh.size() * sizeof(DBvhNode));
  if (!bvh_idx.empty()) memcpy(out.data() + o_idx, bvh_idx.data(), bvh_idx.size() * sizeof(int32_t));
}

size_t scene_bytes_for(int n_inst) {
  HostScene h;
  h.inst.resize(n_inst);
  if (n_inst > kLinearInstances) {
    h.bvh.resize(2 * (size_t)n_inst);
    h.bvh_idx.resize(n_inst);
  }
  return h.device_bytes();
}

int upload_scene(r2_ctx* ctx, const HostScene& hs, bool may_alloc) {
  const size_t need = hs.device_bytes();
  if (need > ctx->scene_cap) {
    if (!may_alloc)
      return fail(ctx, R2_ERR_UNSUPPORTED,
                  "render: the scene's " + std::to_string(hs.inst.size()) + " instances exceed the context's " +
                      "preallocated scene table (" + std::to_string(kCtxSceneInstances) + " instances)");
    R2_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->d_scene);
    ctx->d_scene = nullptr;
    ctx->scene_cap = 0;
    R2_HIP(ctx, hipMalloc((void**)&ctx->d_scene, need));
    ctx->scene_cap = need;
  }
  // the image stays in the context until the next upload, past the stream's copy
  hs.pack(ctx->d_scene, ctx->scene_img);
  R2_HIP(ctx, hipMemcpyAsync(ctx->d_scene, ctx->scene_img.data(), need, hipMemcpyHostToDevice, ctx->stream));
  return R2_OK;
}

int build_scene(r2_ctx* ctx, const r2_render_params* p, HostScene* hs) {
  if (!p || p->width <= 0 || p->height <= 0 || p->spp <= 0 || p->max_depth < 0)
    return fail(ctx, R2_ERR_INVALID, "render: bad frame size / spp / max_depth");
  if ((int64_t)p->width * p->height > (1ll << 28)) return fail(ctx, R2_ERR_UNSUPPORTED, "render: frame too large");
  if (p->n_instances < 0 || p->n_instances > kMaxInstances)
    return fail(ctx, R2_ERR_UNSUPPORTED, "render: at most " + std::to_string(kMaxInstances) + " medium instances");
  if (p->n_instances > 0 && !p->instances) return fail(ctx, R2_ERR_INVALID, "render: instances missing");
  if (!p->env) return fail(ctx, R2_ERR_INVALID, "render: an environment map is required (config Scene.Envmap)");
  hs->inst.clear();
  hs->bvh.clear();
  hs->bvh_idx.clear();
  DScene& s = hs->s;
  memset(&s, 0, sizeof(s));
  s.W = p->width; s.H = p->height; s.spp = p->spp; s.max_depth = p->max_depth;
  s.seed = p->seed; s.stream = p->stream; s.sample_begin = p->sample_begin; s.flags = p->flags;
  s.env = p->env->dev;
  // CCamera (camera.cu:20-37): glm::lookAt, then glm::inverse of the mat4x3
  s.pixel_size = 1e-5f;
  const float fov = p->fov_deg * 0.01745329251994329576923690768489f;
  s.near_dist = ((float)p->width * s.pixel_size) / (2.f * tanf(fov / 2.f));
  const v3 eye = mk(p->cam_pos[0], p->cam_pos[1], p->cam_pos[2]);
  const v3 ctr = mk(p->cam_look_at[0], p->cam_look_at[1], p->cam_look_at[2]);
  const v3 up = mk(p->cam_up[0], p->cam_up[1], p->cam_up[2]);
  const v3 f = normalize(ctr - eye);
  const v3 sv = normalize(cross(f, up));
  const v3 u = cross(sv, f);
  m44 M = identity44();
  M.c[0][0] = sv.x; M.c[1][0] = sv.y; M.c[2][0] = sv.z;
  M.c[0][1] = u.x; M.c[1][1] = u.y; M.c[2][1] = u.z;
  M.c[0][2] = -f.x; M.c[1][2] = -f.y; M.c[2][2] = -f.z;
  M.c[3][0] = -dot(sv, eye); M.c[3][1] = -dot(u, eye); M.c[3][2] = dot(f, eye);
  s.view_to_world = to43(inverse44(to44(to43(M))));
  // instances (CHostSceneobject(medium, pos, orientation, scaling), sceneobject.cu:50-65);
  // instances of one grid share its device view
  s.n_inst = p->n_instances;
  hs->inst.resize(p->n_instances);
  hs->precision = -2;
  for (int k = 0; k < p->n_instances; ++k) {
    const r2_instance& id = p->instances[k];
    if (!id.grid) return fail(ctx, R2_ERR_INVALID, "render: instance without grid");
    int gidx = 0;
    while (gidx < s.n_grid && s.grid[gidx].leaf_table != id.grid->dev.leaf_table) ++gidx;
    if (gidx == s.n_grid) {
      if (s.n_grid == kMaxGrids)
        return fail(ctx, R2_ERR_UNSUPPORTED, "render: at most " + std::to_string(kMaxGrids) + " distinct grids");
      s.grid[s.n_grid++] = id.grid->dev;
      const int q = id.grid->dev.precision == 1 ? 1 : 0;
      hs->precision = hs->precision == -2 ? q : (hs->precision == q ? q : -1);
    }
    DInstance& in = hs->inst[k];
    in.grid = gidx;
    m44 T = identity44();
    T.c[3][0] = id.pos[0]; T.c[3][1] = id.pos[1]; T.c[3][2] = id.pos[2];
    const m44 R = get_rotation(mk(id.orientation[0], id.orientation[1], id.orientation[2]));
    m44 S = identity44();
    S.c[0][0] = id.scaling[0]; S.c[1][1] = id.scaling[1]; S.c[2][2] = id.scaling[2];
    in.m2w = to43(mul44(mul44(T, R), S));
    in.w2m = inverse43(in.m2w);
    const r2_grid_info& gi = id.grid->info;
    const v3 a = mul_point(in.m2w, mk(gi.bbox_min[0], gi.bbox_min[1], gi.bbox_min[2]));
    const v3 b = mul_point(in.m2w, mk(gi.bbox_max[0], gi.bbox_max[1], gi.bbox_max[2]));
    in.bb_min[0] = fminf(a.x, b.x); in.bb_min[1] = fminf(a.y, b.y); in.bb_min[2] = fminf(a.z, b.z);
    in.bb_max[0] = fmaxf(a.x, b.x); in.bb_max[1] = fmaxf(a.y, b.y); in.bb_max[2] = fmaxf(a.z, b.z);
    // getFilterRenderRatio (sceneobject.cu:108-114)
    const float vsr = (in.bb_max[0] - in.bb_min[0]) / (float)gi.size[0];
    in.ratio = id.grid->voxel_size_filtering / vsr;
  }
  if (hs->precision == -2) hs->precision = 0;
  if (p->n_instances > kLinearInstances) {
    bvh_build(hs->inst, hs->bvh, hs->bvh_idx);
    s.n_bvh = (int32_t)hs->bvh.size();
  }
  // analytic surfaces (CHostSceneobject(shape, material ...), sceneobject.cu:28-48)
  if (p->n_surfaces < 0 || p->n_surfaces > kMaxSurfaces)
    return fail(ctx, R2_ERR_UNSUPPORTED, "render: at most 8 surfaces");
  if (p->n_surfaces > 0 && !p->surfaces) return fail(ctx, R2_ERR_INVALID, "render: surfaces missing");
  s.n_surf = p->n_surfaces;
  for (int j = 0; j < p->n_surfaces; ++j) {
    const r2_surface& sd = p->surfaces[j];
    if (sd.shape != R2_SHAPE_CIRCLE && sd.shape != R2_SHAPE_SPHERE)
      return fail(ctx, R2_ERR_INVALID, "render: unknown surface shape");
    DSurface& sf = s.surf[j];
    sf.shape = sd.shape;
    for (int a = 0; a < 3; ++a) {
      sf.pos[a] = sd.pos[a];
      sf.normal[a] = sd.normal[a];
      sf.diffuse[a] = sd.diffuse[a];
      sf.specular[a] = sd.specular[a];
    }
    sf.radius = sd.radius;
    // CShape (shape.hpp:49-57): translate(worldPos) * getRotation(normal), and its inverse
    m44 T = identity44();
    T.c[3][0] = sd.pos[0]; T.c[3][1] = sd.pos[1]; T.c[3][2] = sd.pos[2];
    sf.m2w = to43(mul44(T, get_rotation(mk(sd.normal[0], sd.normal[1], sd.normal[2]))));
    sf.w2m = inverse43(sf.m2w);
    // COrenNayarBRDF(roughness) (oren_nayar_brdf.hpp:31-35)
    const float sig2 = sd.diffuse_roughness * sd.diffuse_roughness;
    sf.on_a = 1.f - (sig2 / (2.f * (sig2 + 0.33f)));
    sf.on_b = 0.45f * sig2 / (sig2 + 0.09f);
    sf.alpha_x = sd.alpha_x; sf.alpha_y = sd.alpha_y;
    sf.eta_i = sd.eta_i; sf.eta_t = sd.eta_t;
  }
  return R2_OK;
}

}  // namespace r2host

using namespace r2host;

extern "C" {

int r2_version(void) { return R2_VERSION; }

int r2_create(r2_ctx** out, int device) {
  if (!out) return R2_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return R2_ERR_HIP;
  if (device < 0 || device >= n) return R2_ERR_INVALID;
  r2_ctx* c = new (std::nothrow) r2_ctx();
  if (!c) return R2_ERR_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->main_mark, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void**)&c->d_scene, scene_bytes_for(kCtxSceneInstances)) != hipSuccess ||
      hipMalloc((void**)&c->d_counter, 64) != hipSuccess) {
    r2_destroy(c);
    return R2_ERR_HIP;
  }
  c->scene_cap = scene_bytes_for(kCtxSceneInstances);
  // the address of r2_render_noalloc's scratch array on this device (kept per
  // device; a failure is reported by that entry point, the context stays usable)
  (void)noalloc_scratch_init(c);
  c->err.clear();
  *out = c;
  return R2_OK;
}

int r2_destroy(r2_ctx* c) {
  if (!c) return R2_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->d_scene) (void)hipFree(c->d_scene);
  if (c->d_counter) (void)hipFree(c->d_counter);
  for (int i = 0; i < r2_ctx::kPoolSlots; ++i)
    if (c->pool[i]) (void)hipFree(c->pool[i]);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->side_done) (void)hipEventDestroy(c->side_done);
  if (c->main_mark) (void)hipEventDestroy(c->main_mark);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return R2_OK;
}

const char* r2_last_error(const r2_ctx* c) { return c ? c->err.c_str() : "null context"; }
void* r2_stream(r2_ctx* c) { return c ? (void*)c->stream : nullptr; }

int r2_grid_upload(r2_ctx* ctx, const r2_grid_desc* d, r2_grid** out) {
  if (!ctx || !d || !out) return R2_ERR_INVALID;
  *out = nullptr;
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !d->records) return fail(ctx, R2_ERR_INVALID, "grid: bad size / records");
  if (d->nx > 8192 || d->ny > 8192 || d->nz > 8192) return fail(ctx, R2_ERR_UNSUPPORTED, "grid: at most 8192 voxels per axis");
  if (d->precision != R2_PREC_FP32 && d->precision != R2_PREC_PACKED) return fail(ctx, R2_ERR_INVALID, "grid: bad precision");
  (void)hipSetDevice(ctx->device);
  r2_grid* g = new (std::nothrow) r2_grid();
  if (!g) return R2_ERR_OOM;
  const int rc = grid_build(ctx, d, g);
  if (rc) {
    r2_grid_free(ctx, g);
    return rc;
  }
  *out = g;
  return R2_OK;
}

int r2_grid_free(r2_ctx* ctx, r2_grid* g) {
  if (!g) return R2_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  for (void* p : g->allocs)
    if (p) (void)hipFree(p);
  delete g;
  return R2_OK;
}

int r2_grid_get_info(const r2_grid* g, r2_grid_info* out) {
  if (!g || !out) return R2_ERR_INVALID;
  *out = g->info;
  return R2_OK;
}

int r2_grid_download_bricks(r2_ctx* ctx, const r2_grid* g, uint32_t* range, uint8_t* u8, uint32_t* mips) {
  if (!ctx || !g) return R2_ERR_INVALID;
  const int* nb = g->info.n_bricks;
  const size_t nbt = (size_t)nb[0] * nb[1] * nb[2];
  hipStream_t st = ctx->stream;
  if (range) R2_HIP(ctx, hipMemcpyAsync(range, g->range0, nbt * 4, hipMemcpyDeviceToHost, st));
  if (mips) {
    size_t off = 0;
    for (int l = 0; l < 3; ++l) {
      const size_t cnt = (size_t)(nb[0] >> (l + 1)) * (nb[1] >> (l + 1)) * (nb[2] >> (l + 1));
      R2_HIP(ctx, hipMemcpyAsync(mips + off, g->mip[l], cnt * 4, hipMemcpyDeviceToHost, st));
      off += cnt;
    }
  }
  if (u8) {
    std::vector<uint2> bricks(nbt);
    const size_t nne = (size_t)std::max<int64_t>(g->info.n_nonempty_bricks, 1);
    std::vector<uint8_t> atlas(nne * 512);
    R2_HIP(ctx, hipMemcpyAsync(bricks.data(), g->dev.bricks, nbt * sizeof(uint2), hipMemcpyDeviceToHost, st));
    R2_HIP(ctx, hipMemcpyAsync(atlas.data(), g->dev.atlas, nne * 512, hipMemcpyDeviceToHost, st));
    R2_HIP(ctx, hipStreamSynchronize(st));
    const size_t ex = (size_t)nb[0] * 8, ey = (size_t)nb[1] * 8, ez = (size_t)nb[2] * 8;
    memset(u8, 0, ex * ey * ez);
    for (size_t bi = 0; bi < nbt; ++bi) {
      if (bricks[bi].y == 0xffffffffu) continue;
      const size_t bx = bi % nb[0], by = (bi / nb[0]) % nb[1], bz = bi / ((size_t)nb[0] * nb[1]);
      const uint8_t* src = atlas.data() + (size_t)bricks[bi].y * 512;
      for (int v = 0; v < 512; ++v) {
        const size_t x = bx * 8 + (v & 7), y = by * 8 + ((v >> 3) & 7), z = bz * 8 + (v >> 6);
        u8[(z * ey + y) * ex + x] = src[v];
      }
    }
  }
  R2_HIP(ctx, hipStreamSynchronize(st));
  return R2_OK;
}

int r2_grid_lookup(r2_ctx* ctx, const r2_grid* g, int32_t n, const int32_t* coords, uint8_t* out) {
  if (!ctx || !g || n < 0 || (n > 0 && (!coords || !out))) return R2_ERR_INVALID;
  if (n == 0) return R2_OK;
  return grid_lookup(ctx, g, n, coords, out);
}

int r2_grid_sample(r2_ctx* ctx, const r2_grid* g, int32_t n, const float* pos, float* out) {
  if (!ctx || !g || n < 0 || (n > 0 && (!pos || !out))) return R2_ERR_INVALID;
  if (n == 0) return R2_OK;
  return grid_sample(ctx, g, n, pos, out);
}

int r2_scene_intersect(r2_ctx* ctx, const r2_render_params* p, int32_t n, const float* rays7, float* out4) {
  if (!ctx || !p || n < 0 || (n > 0 && (!rays7 || !out4))) return R2_ERR_INVALID;
  if (n == 0) return R2_OK;
  (void)hipSetDevice(ctx->device);
  HostScene s;
  const int rc = build_scene(ctx, p, &s);
  if (rc) return rc;
  return scene_intersect_probe(ctx, s, n, rays7, out4);
}

int r2_dda_cells(r2_ctx* ctx, int32_t n, const float* rays7, float sigma_t, int32_t collide_every, int32_t max_cells,
                 int32_t* cells, int32_t* counts) {
  if (!ctx || n < 0 || max_cells <= 0 || !(sigma_t > 0.f) || collide_every < 0 ||
      (n > 0 && (!rays7 || !cells || !counts)))
    return R2_ERR_INVALID;
  if ((int64_t)n * max_cells > (1ll << 28)) return fail(ctx, R2_ERR_UNSUPPORTED, "dda_cells: at most 2^28 cells");
  if (n == 0) return R2_OK;
  return dda_cells(ctx, n, rays7, sigma_t, collide_every, max_cells, cells, counts);
}

int r2_env_upload(r2_ctx* ctx, const float* rgb, int32_t W, int32_t H, r2_env** out) {
  if (!ctx || !rgb || !out || W < 2 || H < 2) return fail(ctx, R2_ERR_INVALID, "env: need at least 2x2 texels");
  *out = nullptr;
  const size_t n = (size_t)W * H;
  std::vector<float> row_func(n), row_cdf((size_t)(W + 1) * H), row_int(H), marg_func(H), marg_cdf(H + 1);
  // radiancePerPixel (environmentmap.cu:16-29)
  for (int j = 0; j < H; ++j) {
    const double st = sin(3.14159265358979323846 * (double)((float)j + 0.5f) / (double)(float)H);
    for (int i = 0; i < W; ++i) {
      const size_t p = (size_t)j * W + i;
      const float rad = rgb[3 * p] + rgb[3 * p + 1] + rgb[3 * p + 2];
      row_func[p] = (float)((double)rad * st);
    }
  }
  // CDistribution1D ctor (distribution_1d.cu:6-39)
  auto build_1d = [](const float* f, int nn, float* cdf) -> float {
    cdf[0] = 0.f;
    for (int i = 1; i <= nn; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / (float)nn;
    const float fi = cdf[nn];
    if (fi == 0.f) {
      for (int i = 1; i <= nn; ++i) cdf[i] = (float)i / (float)nn;
    } else {
      for (int i = 1; i <= nn; ++i) cdf[i] /= fi;
    }
    return fi;
  };
  // CDistribution2D::sumRows (distribution_2d.hpp:59-71): pushes at i % W == 0
  {
    float sum = 0.f;
    int k = 0;
    for (size_t i = 0; i < n; ++i) {
      sum += row_func[i];
      if (i % (size_t)W == 0) {
        marg_func[k++] = sum;
        sum = 0.f;
      }
    }
  }
  const float marg_int = build_1d(marg_func.data(), H, marg_cdf.data());
  for (int j = 0; j < H; ++j) row_int[j] = build_1d(row_func.data() + (size_t)j * W, W, row_cdf.data() + (size_t)j * (W + 1));
  // guide tables: g = smallest power of two >= the number of intervals, so that
  // u * g is exact and a guided search brackets about one interval
  auto pow2_at_least = [](int v) {
    int g = 1;
    while (g < v && g < (1 << 16)) g <<= 1;
    return g;
  };
  auto build_guide = [](const float* cdf, int size, int g, int32_t* out) {
    int cnt = 0;  // number of cdf values <= j / g (the CDF is non-decreasing)
    for (int j = 0; j <= g; ++j) {
      const float v = (float)j / (float)g;
      while (cnt < size && cdf[cnt] <= v) ++cnt;
      out[j] = cnt;
    }
  };
  const int marg_g = pow2_at_least(H), row_g = pow2_at_least(W);
  std::vector<int32_t> marg_guide(marg_g + 1), row_guide((size_t)H * (row_g + 1));
  build_guide(marg_cdf.data(), H + 1, marg_g, marg_guide.data());
  for (int j = 0; j < H; ++j)
    build_guide(row_cdf.data() + (size_t)j * (W + 1), W + 1, row_g, row_guide.data() + (size_t)j * (row_g + 1));
  r2_env* e = new (std::nothrow) r2_env();
  if (!e) return R2_ERR_OOM;
  auto up = [&](const void* src, size_t bytes) -> void* {
    void* d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
    e->allocs.push_back(d);
    if (hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
  };
  e->dev.w = W;
  e->dev.h = H;
  e->dev.tex = (const float*)up(rgb, n * 12);
  e->dev.row_func = (const float*)up(row_func.data(), n * 4);
  e->dev.row_cdf = (const float*)up(row_cdf.data(), row_cdf.size() * 4);
  e->dev.row_int = (const float*)up(row_int.data(), H * 4);
  e->dev.marg_func = (const float*)up(marg_func.data(), H * 4);
  e->dev.marg_cdf = (const float*)up(marg_cdf.data(), (H + 1) * 4);
  e->dev.marg_guide = (const int32_t*)up(marg_guide.data(), marg_guide.size() * 4);
  e->dev.row_guide = (const int32_t*)up(row_guide.data(), row_guide.size() * 4);
  e->dev.marg_g = marg_g;
  e->dev.row_g = row_g;
  e->dev.marg_int = marg_int;
  if (!e->dev.tex || !e->dev.row_func || !e->dev.row_cdf || !e->dev.row_int || !e->dev.marg_func || !e->dev.marg_cdf ||
      !e->dev.marg_guide || !e->dev.row_guide) {
    r2_env_free(ctx, e);
    return fail(ctx, R2_ERR_HIP, "env: device upload failed");
  }
  *out = e;
  return R2_OK;
}

int r2_env_free(r2_ctx* ctx, r2_env* e) {
  if (!e) return R2_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  for (void* p : e->allocs) (void)hipFree(p);
  delete e;
  return R2_OK;
}

static int render_common(r2_ctx* ctx, const r2_render_params* p, float* out, int32_t out_on_device,
                         const RenderOpts& o) {
  if (!ctx || !p || !out) return R2_ERR_INVALID;
  const auto t0 = std::chrono::steady_clock::now();
  (void)hipSetDevice(ctx->device);
  HostScene s;
  int rc = build_scene(ctx, p, &s);
  if (rc) return rc;
  const size_t bytes = (size_t)p->width * p->height * 3 * sizeof(float);
  float* d_out = out;
  void* tmp = nullptr;
  if (!out_on_device) {
    R2_HIP(ctx, hipMalloc(&tmp, bytes));
    d_out = (float*)tmp;
  }
  rc = launch_render_ex(ctx, s, d_out, (p->flags & R2_FLAG_FP64_ACCUM) != 0, o);
  if ((rc == R2_OK || rc == R2_ERR_CANCELLED) && !out_on_device) {
    hipError_t e = hipMemcpy(out, d_out, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "hipMemcpy(framebuffer)");
  }
  if (tmp) (void)hipFree(tmp);
  ctx->stats.last_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int r2_render(r2_ctx* ctx, const r2_render_params* p, float* out, int32_t out_on_device) {
  return render_common(ctx, p, out, out_on_device, RenderOpts{});
}

int r2_render_progressive(r2_ctx* ctx, const r2_render_params* p, float* out, int32_t out_on_device,
                          r2_progress_fn progress, void* user) {
  if (!progress) return R2_ERR_INVALID;
  RenderOpts o;
  o.progress = progress;
  o.user = user;
  return render_common(ctx, p, out, out_on_device, o);
}

int r2_render_noalloc(r2_ctx* ctx, const r2_render_params* p, float* out) {
  if (!ctx || !p || !out) return R2_ERR_INVALID;
  const auto t0 = std::chrono::steady_clock::now();
  (void)hipSetDevice(ctx->device);
  void* scr = noalloc_scratch(ctx);
  if (!scr) return fail(ctx, R2_ERR_HIP, "render_noalloc: the scratch array has no address on this device");
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, out) != hipSuccess ||
      (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged)) {
    (void)hipGetLastError();
    return fail(ctx, R2_ERR_INVALID, "render_noalloc: out_rgb must be device memory");
  }
  HostScene s;
  int rc = build_scene(ctx, p, &s);
  if (rc) return rc;
  const bool fp64 = (p->flags & R2_FLAG_FP64_ACCUM) != 0;
  const uint64_t per_chunk = (uint64_t)p->width * p->height * 3 * (fp64 ? 8 : 4);
  if (per_chunk > kNoAllocScratchBytes)
    return fail(ctx, R2_ERR_UNSUPPORTED,
                "render_noalloc: one sample chunk of a " + std::to_string(p->width) + "x" + std::to_string(p->height) +
                    " frame needs " + std::to_string(per_chunk) + " bytes of chunk sums, the scratch holds " +
                    std::to_string(kNoAllocScratchBytes) + " (at most " +
                    std::to_string(kNoAllocScratchBytes / (3 * (fp64 ? 8 : 4))) + " pixels)");
  RenderOpts o;
  o.partial = scr;
  o.partial_bytes = kNoAllocScratchBytes;
  o.scene_alloc = false;
  rc = launch_render_ex(ctx, s, out, fp64, o);
  ctx->stats.last_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int r2_render_paths(r2_ctx* ctx, const r2_render_params* p, int32_t n, const int32_t* pix, const int32_t* smp,
                    float* out, uint32_t* counters) {
  if (!ctx || !p || n < 0 || !pix || !smp || !out) return R2_ERR_INVALID;
  if (n == 0) return R2_OK;
  (void)hipSetDevice(ctx->device);
  HostScene s;
  int rc = build_scene(ctx, p, &s);
  if (rc) return rc;
  for (int i = 0; i < n; ++i)
    if (pix[i] < 0 || pix[i] >= p->width * p->height) return fail(ctx, R2_ERR_INVALID, "render_paths: pixel out of range");
  void *dp = nullptr, *ds = nullptr, *dout = nullptr, *dc = nullptr;
  struct F {
    void** q[4];
    ~F() { for (auto x : q) if (*x) (void)hipFree(*x); }
  } f{{&dp, &ds, &dout, &dc}};
  R2_HIP(ctx, hipMalloc(&dp, (size_t)n * 4));
  R2_HIP(ctx, hipMalloc(&ds, (size_t)n * 4));
  R2_HIP(ctx, hipMalloc(&dout, (size_t)n * 12));
  if (counters) R2_HIP(ctx, hipMalloc(&dc, (size_t)n * kNumCnt * 4));
  R2_HIP(ctx, hipMemcpy(dp, pix, (size_t)n * 4, hipMemcpyHostToDevice));
  R2_HIP(ctx, hipMemcpy(ds, smp, (size_t)n * 4, hipMemcpyHostToDevice));
  rc = launch_paths(ctx, s, n, (const int32_t*)dp, (const int32_t*)ds, (float*)dout, (uint32_t*)dc);
  if (rc) return rc;
  R2_HIP(ctx, hipMemcpy(out, dout, (size_t)n * 12, hipMemcpyDeviceToHost));
  if (counters) R2_HIP(ctx, hipMemcpy(counters, dc, (size_t)n * kNumCnt * 4, hipMemcpyDeviceToHost));
  return R2_OK;
}

int r2_get_stats(const r2_ctx* ctx, r2_stats* out) {
  if (!ctx || !out) return R2_ERR_INVALID;
  *out = ctx->stats;
  return R2_OK;
}

}  // extern "C"
