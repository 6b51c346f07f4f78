// Mesh -> SGGX voxel filter on gfx950 (CMeshFilter, raytracer/include/filtering/mesh_filter.hpp:21-266,
// launched per voxel by __raygen__filtering, raytracer/src/backend/shaders.optix.cu:258-284).
//
// Per LOD grid:
//   1. k_tri_prep   per triangle: world AABB -> conservative range of voxels whose
//                   ray region (the clip sphere, or the voxel plus one ray length)
//                   can reach it, split into 64-voxel tasks.
//   2. k_pairs      one wave per task, one voxel per lane: exact point/triangle
//                   distance against the voxel's reach radius; passing (voxel, tri)
//                   pairs are counted (with the task's lane mask), then emitted
//                   from the masks as 64-bit keys voxel<<32 | tri.
//   3. radix sort + run-length encode: sorted candidate list per active voxel.
//   4. k_filter     one wave per active voxel, the 64 lanes are the voxel's rays.
//                   Every lane tests every candidate triangle (uniform loop; the
//                   triangle is a wave-uniform load), keeps the lexicographic
//                   minimum of (t, triangle index), and the per-ray sums are
//                   reduced in ray order through LDS, so the record equals the
//                   sequential loop of the reference bit for bit.
// A voxel with no candidate cannot be hit by any of its rays: its record is the
// no-hit record, written for all voxels up front by k_fill_empty.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#include "r2_internal.hpp"
#include "r2_codec.hpp"
#include "r2_detmath.hpp"

using namespace r2;

struct r2_mesh {
  int32_t n_vertices = 0, n_triangles = 0, n_materials = 0;
  float model_min[3] = {0, 0, 0}, model_max[3] = {0, 0, 0};
  m43 m2w{};               // scene object modelToWorld = T(pos) * getRotation(orientation) * S(scaling)
  float4* tri = nullptr;   // 3 per triangle: world-space vertices (w unused)
  int32_t* idx = nullptr;  // 3 per triangle
  float* nrm = nullptr;    // 3 per vertex, model space
  int32_t* mat = nullptr;  // material id per triangle
  float* matf = nullptr;   // 8 per material: diffuse rgb, specular rgb, roughness, ior
  // textured materials: texture coordinates, model-space positions (tangent
  // frames of normal-mapped hits), an RGBA texel pool with (first texel, width,
  // height) per texture, and 4 texture indices per material (diffuse, glossy,
  // normal map, alpha mask; -1 none)
  float2* uv = nullptr;
  float* pos = nullptr;
  float4* texels = nullptr;
  int4* tex = nullptr;
  int4* mtex = nullptr;
  bool any_alpha = false, any_normal = false;
  std::vector<void*> allocs;
};

namespace {

constexpr int kRow = 68;  // LDS row of one value over a batch's hits: 16-byte aligned, rows 4 banks apart
constexpr int kWaves = 4;  // waves per 256-thread block
constexpr int kVals = 17;  // per-ray sums: S6, diffuse3, specular3, normal3, ior, distance
constexpr uint32_t kTag = 0xF17Eu;

struct FConst {
  m43 i2m, m2w, w2m, i2w;  // indexToModel, modelToWorld, worldToModel, indexToWorld (mat4x3)
  m43 mesh_m2w;            // mesh object modelToWorld (hit normals)
  m44 rs;                  // getRotation((1,0,0)) of the clip sphere
  int nx, ny, nz;
  float diagonal, tmax_fixed, radius;
  float reach;             // conservative distance from the voxel centre a ray can reach
  float bbmin[3], inv_vs[3], range_margin;
  int n_samples, iterations, clip;
  float sigma_t, alpha;
  uint32_t seed, lod;
  int n_materials;
  // textured materials: uv null without texture coordinates, mtex null without
  // material textures
  const float2* uv;
  const float* pos;
  const float4* texels;
  const int4* tex;
  const int4* mtex;
  int alpha_on, normal_on;  // some material has an alpha mask / a normal map
};

__device__ __forceinline__ float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ v3 f4v(float4 a) { return mk(a.x, a.y, a.z); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// voxel centre: modelToWorld * (indexToModel * (v + 0.5))  (mesh_filter.hpp:47)
__device__ __forceinline__ v3 voxel_center(const FConst& c, int ix, int iy, int iz) {
  const v3 vm = mul_point(c.i2m, mk((float)ix + 0.5f, (float)iy + 0.5f, (float)iz + 0.5f));
  return mul_point(c.m2w, vm);
}

// ---------------------------------------------------------------- mesh ----
__global__ void k_world_tris(const float* __restrict__ pos, const int32_t* __restrict__ idx, m43 m2w, int n,
                             float4* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  for (int k = 0; k < 3; ++k) {
    const int v = idx[3 * t + k];
    const v3 p = mul_point(m2w, mk(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]));
    out[3 * t + k] = make_float4(p.x, p.y, p.z, 0.f);
  }
}

// -------------------------------------------------------------- binning ---
// squared distance from p to triangle (a, b, c): closest point by Voronoi
// regions (Ericson, Real-Time Collision Detection, 5.1.5), relative to a
__device__ float dist2_point_triangle(v3 p, v3 a, v3 b, v3 c) {
  const v3 ab = b - a, ac = c - a, ap = p - a;
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.f && d2 <= 0.f) return dot(ap, ap);
  const v3 bp = p - b;
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.f && d4 <= d3) return dot(bp, bp);
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
    const float v = d1 / (d1 - d3);
    const v3 q = ap - ab * v;
    return dot(q, q);
  }
  const v3 cp = p - c;
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.f && d5 <= d6) return dot(cp, cp);
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
    const float w = d2 / (d2 - d6);
    const v3 q = ap - ac * w;
    return dot(q, q);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    const v3 q = bp - (c - b) * w;
    return dot(q, q);
  }
  const float den = 1.f / (va + vb + vc);
  const float v = vb * den, w = vc * den;
  const v3 q = ap - ab * v - ac * w;
  return dot(q, q);
}

struct TriRange {
  int lo[3];
  int dim[3];
};

__global__ void k_tri_prep(const float4* __restrict__ tri, int n, FConst c, TriRange* __restrict__ rng,
                           unsigned long long* __restrict__ tasks) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float4 a = tri[3 * t], b = tri[3 * t + 1], d = tri[3 * t + 2];
  const float mn[3] = {fminf(a.x, fminf(b.x, d.x)), fminf(a.y, fminf(b.y, d.y)), fminf(a.z, fminf(b.z, d.z))};
  const float mx[3] = {fmaxf(a.x, fmaxf(b.x, d.x)), fmaxf(a.y, fmaxf(b.y, d.y)), fmaxf(a.z, fmaxf(b.z, d.z))};
  const int nv[3] = {c.nx, c.ny, c.nz};
  TriRange r;
  unsigned long long vol = 1;
  for (int k = 0; k < 3; ++k) {
    // voxel centres lie at bbmin + (i + 0.5) * vs; keep those within reach of the AABB
    const float flo = (mn[k] - c.reach - c.bbmin[k]) * c.inv_vs[k] - 0.5f - c.range_margin;
    const float fhi = (mx[k] + c.reach - c.bbmin[k]) * c.inv_vs[k] - 0.5f + c.range_margin;
    int lo = (flo > -1.f) ? (int)floorf(flo) : -1;
    int hi = (fhi < (float)nv[k]) ? (int)ceilf(fhi) : nv[k];
    lo = max(lo, 0);
    hi = min(hi, nv[k] - 1);
    if (!(flo == flo) || !(fhi == fhi) || hi < lo) {  // NaN vertex or out of the grid
      lo = 0;
      hi = -1;
    }
    r.lo[k] = lo;
    r.dim[k] = hi - lo + 1;
    vol *= (unsigned long long)max(r.dim[k], 0);
  }
  rng[t] = r;
  tasks[t] = (vol + 63ull) >> 6;
}

// task -> triangle map: mark each triangle's first task, then a max-scan fills the rest
__global__ void k_task_mark(const unsigned long long* __restrict__ tasks, const unsigned long long* __restrict__ begin,
                            int n, uint32_t* __restrict__ task_tri) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || tasks[t] == 0) return;
  task_tri[begin[t]] = (uint32_t)t;
}

// voxel of lane `local` of a triangle's voxel range (x fastest), in 32-bit
// arithmetic when the range allows it
__device__ __forceinline__ void range_voxel(const TriRange& r, unsigned long long local, int& ix, int& iy, int& iz) {
  const unsigned long long dxy = (unsigned long long)r.dim[0] * r.dim[1];
  if (dxy * (unsigned long long)r.dim[2] <= 0xFFFFFFFFull) {
    const uint32_t l = (uint32_t)local, d0 = (uint32_t)r.dim[0], d01 = (uint32_t)dxy;
    const uint32_t z = l / d01, rem = l - z * d01, y = rem / d0;
    iz = r.lo[2] + (int)z;
    iy = r.lo[1] + (int)y;
    ix = r.lo[0] + (int)(rem - y * d0);
  } else {
    const unsigned long long rem = local % dxy;
    iz = r.lo[2] + (int)(local / dxy);
    iy = r.lo[1] + (int)(rem / r.dim[0]);
    ix = r.lo[0] + (int)(rem % r.dim[0]);
  }
}

// 64-voxel tasks.  A block takes 256 consecutive tasks: its threads first load
// the tasks' metadata (triangle, voxel range, first task of the triangle, and the
// triangle or the stored mask) into LDS, one task each, so that the dependent
// loads of all 256 tasks are in flight together.
//   kEmit = false: the exact point/triangle test of every (task, voxel) pair.  The
//     block's pairs are packed densely onto its lanes through a prefix sum of the
//     tasks' voxel counts (a lane finds its task by binary search).  About nine in
//     ten pairs of a triangle's range are out of reach, so each pair first takes a
//     conservative sphere test: the distance to the triangle is at least the
//     distance to its centroid minus the largest vertex distance, and a pair is
//     dropped only when that bound exceeds the reach by a margin (0.1 % relative
//     plus 1e-5 of the coordinates' magnitude) far above the rounding of either
//     test, so no pair the exact test would accept is dropped.  The survivors
//     queue up per wave in LDS and take the exact test 64 at a time; a passing
//     pair sets its bit in the task's mask with an LDS atomic OR.  Each task's
//     mask and pair count then go to task_mask / task_cnt.
//   kEmit = true: one wave per task, one bit per lane: the pairs of the stored
//     mask are written from task_off[task] on, in lane order (no atomics: the
//     scan of the counts places them), without testing again.
constexpr int kTasksPerBlock = 256;
template <bool kEmit>
__global__ __launch_bounds__(256) void k_pairs(const float4* __restrict__ tri, FConst c,
                                               const TriRange* __restrict__ rng,
                                               const unsigned long long* __restrict__ task_begin,
                                               const uint32_t* __restrict__ task_tri, unsigned long long n_tasks,
                                               uint32_t* __restrict__ task_cnt,
                                               unsigned long long* __restrict__ task_mask,
                                               const uint32_t* __restrict__ task_off,
                                               unsigned long long* __restrict__ keys) {
  __shared__ uint32_t s_t[kTasksPerBlock];
  __shared__ TriRange s_r[kTasksPerBlock];
  __shared__ unsigned long long s_beg[kTasksPerBlock];
  __shared__ unsigned long long s_mask[kTasksPerBlock];
  __shared__ uint32_t s_off[kTasksPerBlock];  // emit: task_off; test: exclusive prefix of the voxel counts
  __shared__ uint32_t s_wsum[4];
  __shared__ float4 s_tri[kEmit ? 1 : 3 * kTasksPerBlock];
  __shared__ float4 s_sph[kEmit ? 1 : kTasksPerBlock];             // centroid, radius
  __shared__ uint32_t s_q[kEmit ? 1 : 4][kEmit ? 1 : 128];          // queued pairs: task << 8 | voxel
  __shared__ float4 s_qc[kEmit ? 1 : 4][kEmit ? 1 : 128];           // their voxel centres
  const int i = threadIdx.x, lane = i & 63;
  const int wave = __builtin_amdgcn_readfirstlane(i >> 6);
  const unsigned long long base = (unsigned long long)blockIdx.x * kTasksPerBlock;
  const unsigned long long mine = base + (unsigned long long)i;
  const int nk = (int)(n_tasks - base < (unsigned long long)kTasksPerBlock ? n_tasks - base : kTasksPerBlock);
  uint32_t act = 0;  // voxels of this thread's task
  if (mine < n_tasks) {
    const uint32_t t = task_tri[mine];
    const TriRange r = rng[t];
    const unsigned long long beg = task_begin[t];
    s_t[i] = t;
    s_r[i] = r;
    s_beg[i] = beg;
    if (kEmit) {
      s_mask[i] = task_mask[mine];
      s_off[i] = task_off[mine];
    } else {
      const float4 ta = tri[3 * t], tb = tri[3 * t + 1], tc = tri[3 * t + 2];
      s_tri[3 * i] = ta;
      s_tri[3 * i + 1] = tb;
      s_tri[3 * i + 2] = tc;
      const v3 va = f4v(ta), vb = f4v(tb), vc = f4v(tc);
      const v3 cen = (va + vb + vc) * (1.f / 3.f);
      const float r2 = fmaxf(dot(va - cen, va - cen), fmaxf(dot(vb - cen, vb - cen), dot(vc - cen, vc - cen)));
      s_sph[i] = make_float4(cen.x, cen.y, cen.z, sqrtf(r2));
      s_mask[i] = 0ull;
      const unsigned long long vol = (unsigned long long)r.dim[0] * r.dim[1] * r.dim[2];
      const unsigned long long first = (mine - beg) * 64ull;
      act = vol > first ? (uint32_t)(vol - first < 64ull ? vol - first : 64ull) : 0u;
    }
  }
  if (kEmit) {
    __syncthreads();
    for (int k = wave; k < nk; k += (int)(blockDim.x >> 6)) {
      const unsigned long long ball = s_mask[k];
      if ((ball >> lane) & 1ull) {
        int ix, iy, iz;
        range_voxel(s_r[k], (base + (unsigned long long)k - s_beg[k]) * 64ull + (unsigned long long)lane, ix, iy, iz);
        const uint32_t vid = (uint32_t)ix + (uint32_t)c.nx * ((uint32_t)iy + (uint32_t)c.ny * (uint32_t)iz);
        const unsigned long long slot =
            (unsigned long long)s_off[k] + (unsigned long long)__popcll(ball & ((1ull << lane) - 1ull));
        keys[slot] = ((unsigned long long)vid << 32) | (unsigned long long)s_t[k];
      }
    }
    return;
  }
  // exclusive prefix of the voxel counts over the block: within each wave, then
  // across the four waves
  uint32_t incl = act;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = (uint32_t)__shfl_up((int)incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == 63) s_wsum[wave] = incl;
  __syncthreads();
  uint32_t wbase = 0, total = 0;
  for (int w = 0; w < 4; ++w) {
    if (w < wave) wbase += s_wsum[w];
    total += s_wsum[w];
  }
  s_off[i] = wbase + incl - act;
  __syncthreads();
  // the block's (task, voxel) pairs, 64 per wave step: sphere test, then the
  // survivors through the per-wave queue to the exact test
  uint32_t* q = s_q[kEmit ? 0 : wave];
  float4* qc = s_qc[kEmit ? 0 : wave];
  int qn = 0;  // queued pairs (wave-uniform)
  auto exact = [&](int n) {
    if (lane < n) {
      const uint32_t e = q[lane];
      const int k = (int)(e >> 8), j = (int)(e & 63u);
      const float4 ctr = qc[lane];
      if (dist2_point_triangle(mk(ctr.x, ctr.y, ctr.z), f4v(s_tri[3 * k]), f4v(s_tri[3 * k + 1]),
                               f4v(s_tri[3 * k + 2])) <= c.reach * c.reach)
        atomicOr(&s_mask[k], 1ull << j);
    }
  };
  for (uint32_t p0 = (uint32_t)wave * 64u; p0 < total; p0 += 256u) {
    const uint32_t p = p0 + (uint32_t)lane;
    bool cand = false;
    int k = 0;
    uint32_t j = 0;
    v3 ctr = mk(0.f, 0.f, 0.f);
    if (p < total) {
      int lo = 0, hi = nk - 1;  // last task with prefix <= p
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_off[mid] <= p) lo = mid; else hi = mid - 1;
      }
      k = lo;
      j = p - s_off[k];
      int ix, iy, iz;
      range_voxel(s_r[k], (base + (unsigned long long)k - s_beg[k]) * 64ull + (unsigned long long)j, ix, iy, iz);
      ctr = voxel_center(c, ix, iy, iz);
      const float4 sp = s_sph[k];
      const v3 dv = ctr - mk(sp.x, sp.y, sp.z);
      const float mag = fmaxf(fabsf(ctr.x), fmaxf(fabsf(ctr.y), fabsf(ctr.z)));
      cand = !(sqrtf(dot(dv, dv)) > (sp.w + c.reach) * 1.001f + 1e-5f * mag);  // NaN: keep
    }
    const unsigned long long b = __ballot(cand);
    if (cand) {
      const int slot = qn + __popcll(b & ((1ull << lane) - 1ull));
      q[slot] = ((uint32_t)k << 8) | j;
      qc[slot] = make_float4(ctr.x, ctr.y, ctr.z, 0.f);
    }
    qn += __popcll(b);
    if (qn >= 64) {
      wave_sync();
      exact(64);
      // move the rest (fewer than 64) to the front of the queue
      uint32_t e = 0;
      float4 ec = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lane < qn - 64) {
        e = q[64 + lane];
        ec = qc[64 + lane];
      }
      wave_sync();
      if (lane < qn - 64) {
        q[lane] = e;
        qc[lane] = ec;
      }
      qn -= 64;
      wave_sync();
    }
  }
  if (qn > 0) {
    wave_sync();
    exact(qn);
  }
  __syncthreads();
  if (mine < n_tasks) {
    task_mask[mine] = s_mask[i];
    task_cnt[mine] = (uint32_t)__popcll(s_mask[i]);
  }
}

__global__ void k_split(const unsigned long long* __restrict__ keys, unsigned long long n, uint32_t* __restrict__ vox,
                        uint32_t* __restrict__ tri) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = keys[i];
  vox[i] = (uint32_t)(k >> 32);
  tri[i] = (uint32_t)k;
}

// no-hit record: density 0, S 0, colours 0, normal 0 (NaN -> 0), ior half(1)
__global__ void k_fill_empty(uint4* __restrict__ rec, unsigned long long nvox) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  rec[2 * i] = make_uint4(0u, 0u, 0u, 0u);
  rec[2 * i + 1] = make_uint4(0u, 0u, 0u, 0x3C00u);
}

// -------------------------------------------------------------- filter ----
// uniforms 5i .. 5i+4 of the voxel's Philox stream (ray i draws 3 for the
// origin, 2 for the direction, generateRay mesh_filter.hpp:260-266)
__device__ __forceinline__ void ray_uniforms(uint32_t seed, uint32_t lod, uint32_t vid, int i, float u[5]) {
  const uint32_t k0 = 5u * (uint32_t)i;
  const uint32_t b0 = k0 >> 2;
  uint32_t p0[4] = {vid, 0u, b0, kTag};
  uint32_t p1[4] = {vid, 0u, b0 + 1u, kTag};
  philox4x32_10(p0, seed, lod);
  philox4x32_10(p1, seed, lod);
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const uint32_t k = k0 + (uint32_t)j;
    // values (not pointers) are selected, so both blocks stay in registers
    const uint32_t s = k & 3u;
    const uint32_t x0 = s == 0 ? p0[0] : (s == 1 ? p0[1] : (s == 2 ? p0[2] : p0[3]));
    const uint32_t x1 = s == 0 ? p1[0] : (s == 1 ? p1[1] : (s == 2 ? p1[2] : p1[3]));
    const uint32_t x = ((k >> 2) == b0) ? x0 : x1;
    u[j] = (float)(x >> 8) * 5.9604644775390625e-08f;
  }
}

// Sphere::intersect of a world-space ray (sphere.hpp:33-90); 0 on a miss
__device__ float sphere_exit(const m43& w2m_s, float radius, v3 o, v3 d, float ray_tmax) {
  const v3 mo = mul_point(w2m_s, o);
  const v3 md = normalize(mul_dir(w2m_s, d));
  const float a = dot(md, md);
  const float b = 2.0f * dot(md, mo);
  const float cc = dot(mo, mo) - radius * radius;
  const float disc = b * b - 4.f * a * cc;
  if (disc == 0.f) {
    const float t = (-b + sqrtf(disc)) / (2.f * a);
    if (t > 0.f && t <= ray_tmax) return t;
  } else if (disc > 0.f) {
    const float sq = sqrtf(disc);
    const float den = 1.f / (2.0f * a);
    const float t1 = (-b + sq) * den, t2 = (-b - sq) * den;
    const float lo = gmin(t1, t2), hi = gmax(t1, t2);
    if (hi > 0.f) {
      if (lo > 0.f && lo <= ray_tmax) return lo;
      if (lo <= 0.f && hi <= ray_tmax) return hi;
    }
  }
  return 0.f;
}

// watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013)
struct WRay {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
__device__ __forceinline__ WRay wt_setup(v3 d) {
  WRay w;
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  w.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  w.kx = w.kz + 1 == 3 ? 0 : w.kz + 1;
  w.ky = w.kx + 1 == 3 ? 0 : w.kx + 1;
  if (comp(d, w.kz) < 0.f) {
    const int t = w.kx;
    w.kx = w.ky;
    w.ky = t;
  }
  w.Sx = comp(d, w.kx) / comp(d, w.kz);
  w.Sy = comp(d, w.ky) / comp(d, w.kz);
  w.Sz = 1.f / comp(d, w.kz);
  return w;
}
__device__ __forceinline__ bool wt_hit(const WRay& w, v3 o, v3 v0, v3 v1, v3 v2, float tmax, float* t, float* b1,
                                       float* b2) {
  const v3 A = v0 - o, B = v1 - o, Cc = v2 - o;
  const float Ax = comp(A, w.kx) - w.Sx * comp(A, w.kz), Ay = comp(A, w.ky) - w.Sy * comp(A, w.kz);
  const float Bx = comp(B, w.kx) - w.Sx * comp(B, w.kz), By = comp(B, w.ky) - w.Sy * comp(B, w.kz);
  const float Cx = comp(Cc, w.kx) - w.Sx * comp(Cc, w.kz), Cy = comp(Cc, w.ky) - w.Sy * comp(Cc, w.kz);
  float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  const float det = U + V + W;
  if (det == 0.f) return false;
  const float Az = w.Sz * comp(A, w.kz), Bz = w.Sz * comp(B, w.kz), Cz = w.Sz * comp(Cc, w.kz);
  const float T = U * Az + V * Bz + W * Cz;
  if (det < 0.f && (T >= 0.f || T < tmax * det)) return false;
  if (det > 0.f && (T <= 0.f || T > tmax * det)) return false;
  const float rd = 1.f / det;
  *t = T * rd;
  *b1 = V * rd;
  *b2 = W * rd;
  return true;
}

// ------------------------------------------------------------ textures ---
// CTexture::operator() (texture.hpp:55-58): tex2D<float4> with linear filtering,
// wrap addressing and normalised coordinates (texture.cu:166-189), texel centres
// at (i + 0.5) / width; full-precision weights (CUDA's are 8-bit fractions).
// The blend is two nested linear interpolations, so that a constant texture
// returns its value exactly, as tex2D's fixed-point weights (which sum to one)
// do: the normal map's zero test depends on it.  info = (first texel, width,
// height).
__device__ __forceinline__ float4 tex_fetch(const float4* __restrict__ texels, int4 info, float x, float y) {
  const int W = info.y, H = info.z;
  x = x - floorf(x);
  y = y - floorf(y);
  const float tx = x * (float)W - 0.5f, ty = y * (float)H - 0.5f;
  const float fx = floorf(tx), fy = floorf(ty);
  const float a = tx - fx, b = ty - fy;
  int i0 = (int)fx, j0 = (int)fy;
  int i1 = i0 + 1, j1 = j0 + 1;
  i0 = ((i0 % W) + W) % W;
  i1 = ((i1 % W) + W) % W;
  j0 = ((j0 % H) + H) % H;
  j1 = ((j1 % H) + H) % H;
  const float4* t = texels + info.x;
  const float4 p00 = t[j0 * W + i0], p10 = t[j0 * W + i1], p01 = t[j1 * W + i0], p11 = t[j1 * W + i1];
  auto blend = [a, b](float v00, float v10, float v01, float v11) {
    const float top = v00 + a * (v10 - v00);
    const float bot = v01 + a * (v11 - v01);
    return top + b * (bot - top);
  };
  return make_float4(blend(p00.x, p10.x, p01.x, p11.x), blend(p00.y, p10.y, p01.y, p11.y),
                     blend(p00.z, p10.z, p01.z, p11.z), blend(p00.w, p10.w, p01.w, p11.w));
}
// barycentric texture coordinate: (1 - b1 - b2) TC0 + b1 TC1 + b2 TC2 (shaders.optix.cu:61-66)
__device__ __forceinline__ float2 tri_tc(const FConst& c, const int32_t* __restrict__ tidx, int t, float b1,
                                         float b2) {
  const float2 U0 = c.uv[tidx[3 * t]], U1 = c.uv[tidx[3 * t + 1]], U2 = c.uv[tidx[3 * t + 2]];
  const float w0 = 1.f - b1 - b2;
  return make_float2((w0 * U0.x + b1 * U1.x) + b2 * U2.x, (w0 * U0.y + b1 * U1.y) + b2 * U2.y);
}
// __anyhit__mesh (shaders.optix.cu:132-152) with CMaterial::opaque (material.hpp:130-137):
// a hit on a textured mesh is ignored where the material's alpha mask is not > 0
__device__ __forceinline__ bool hit_opaque(const FConst& c, const int32_t* __restrict__ tidx,
                                           const int32_t* __restrict__ tmat, int t, float b1, float b2) {
  int m = tmat ? tmat[t] : 0;
  m = (m >= 0 && m < c.n_materials) ? m : 0;
  const int at = c.mtex[m].w;
  if (at < 0) return true;
  const float2 tc = tri_tc(c, tidx, t, b1, b2);
  return tex_fetch(c.texels, c.tex[at], tc.x, tc.y).x > 0.f;
}
// __closesthit__ch for a mesh (shaders.optix.cu:48-114): the smooth normal; where
// the material has a normal map and the triangle a tangent frame,
// CMaterial::normalmap(fromTBN(T, -B, N), tc) (material.hpp:143-156,
// objects.hpp:60-67); then modelToWorld and normalize
__device__ __forceinline__ v3 hit_normal(const FConst& c, const int32_t* __restrict__ tidx,
                                         const float* __restrict__ nrm, int t, float b1, float b2, int nt, float2 tc) {
  const int i0 = tidx[3 * t], i1 = tidx[3 * t + 1], i2 = tidx[3 * t + 2];
  const v3 N0 = mk(nrm[3 * i0], nrm[3 * i0 + 1], nrm[3 * i0 + 2]);
  const v3 N1 = mk(nrm[3 * i1], nrm[3 * i1 + 1], nrm[3 * i1 + 2]);
  const v3 N2 = mk(nrm[3 * i2], nrm[3 * i2 + 1], nrm[3 * i2 + 2]);
  const float w0 = 1.f - b1 - b2;
  const v3 gn = normalize((w0 * N0 + b1 * N1) + b2 * N2);
  v3 n = gn;
  if (nt >= 0) {
    const float* P = c.pos;
    const v3 P0 = mk(P[3 * i0], P[3 * i0 + 1], P[3 * i0 + 2]);
    const v3 P1 = mk(P[3 * i1], P[3 * i1 + 1], P[3 * i1 + 2]);
    const v3 P2 = mk(P[3 * i2], P[3 * i2 + 1], P[3 * i2 + 2]);
    const v3 Ph = (w0 * P0 + b1 * P1) + b2 * P2;
    const v3 e1 = P1 - Ph, e2 = P2 - Ph;
    const float2 U1 = c.uv[i1], U2 = c.uv[i2];
    const float d1x = U1.x - tc.x, d1y = U1.y - tc.y, d2x = U2.x - tc.x, d2y = U2.y - tc.y;
    const float den = d1x * d2y - d2x * d1y;
    if (den != 0.f) {
      const float f = 1.f / den;
      const v3 T = mk(f * (d2y * e1.x - d1y * e2.x), f * (d2y * e1.y - d1y * e2.y), f * (d2y * e1.z - d1y * e2.z));
      const v3 B = mk(f * (-d2x * e1.x + d1x * e2.x), f * (-d2x * e1.y + d1x * e2.y), f * (-d2x * e1.z + d1x * e2.z));
      const bool tz = T.x == 0.f && T.y == 0.f && T.z == 0.f, bz = B.x == 0.f && B.y == 0.f && B.z == 0.f;
      if (!tz && !bz) {
        const v3 Tn = normalize(T), Bn = normalize(mk(-B.x, -B.y, -B.z));
        const float4 v = tex_fetch(c.texels, c.tex[nt], tc.x, tc.y);
        const v3 m = mk(v.x * 2.f - 1.f, v.y * 2.f - 1.f, v.z * 2.f - 1.f);
        if (!(m.x == 0.f && m.y == 0.f && m.z == 0.f)) {
          const v3 mn = normalize(m);
          n = (mn.x * Tn + mn.y * Bn) + mn.z * gn;
        }
      }
    }
  }
  return normalize(mul_dir(c.mesh_m2w, n));
}

// one wave per work voxel.  direct: work[w] indexes act/off/cnt; otherwise work
// holds voxel ids, looked up in the sorted active list, and records go to out[w].
// kTex: the mesh has texture coordinates or material textures (alpha test,
// textured colours, normal maps); without, the kernel carries none of that code.
template <bool kTex>
__global__ __launch_bounds__(256) void k_filter(FConst c, const uint32_t* __restrict__ work, uint32_t n_work,
                                                int direct, const uint32_t* __restrict__ act,
                                                const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                                                uint32_t n_act, const uint32_t* __restrict__ cand,
                                                const float4* __restrict__ tri, const int32_t* __restrict__ tidx,
                                                const float* __restrict__ nrm, const int32_t* __restrict__ tmat,
                                                const float* __restrict__ matf, uint4* __restrict__ out,
                                                int ts_cap) {
  extern __shared__ __align__(16) float lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWaves + (uint32_t)wave));
  if (w >= n_work) return;
  // [kVals][kRow] per-hit values of a ray batch, value-major, hits compacted in
  // ray order (later: the exp terms of the density iteration)
  float* V = lds + wave * (kVals * kRow + ts_cap);
  float* TS = V + kVals * kRow;  // [ts_cap] clip distances of the hits, in ray order

  uint32_t vid, beg = 0, nc = 0;
  size_t orec;
  if (direct) {
    vid = act[w];
    beg = off[w];
    nc = cnt[w];
    orec = vid;
  } else {
    vid = work[w];
    orec = w;
    uint32_t lo = 0, hi = n_act;  // first active voxel >= vid
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (act[mid] < vid) lo = mid + 1; else hi = mid;
    }
    if (lo < n_act && act[lo] == vid) {
      beg = off[lo];
      nc = cnt[lo];
    }
  }
  vid = (uint32_t)__builtin_amdgcn_readfirstlane((int)vid);
  beg = (uint32_t)__builtin_amdgcn_readfirstlane((int)beg);
  nc = (uint32_t)__builtin_amdgcn_readfirstlane((int)nc);
  const int ix = (int)(vid % (uint32_t)c.nx), iy = (int)((vid / (uint32_t)c.nx) % (uint32_t)c.ny),
            iz = (int)(vid / ((uint32_t)c.nx * (uint32_t)c.ny));
  const v3 ctr = voxel_center(c, ix, iy, iz);
  // clip sphere: model = translate(ctr) * getRotation((1,0,0))
  m44 Ts = identity44();
  Ts.c[3][0] = ctr.x;
  Ts.c[3][1] = ctr.y;
  Ts.c[3][2] = ctr.z;
  const m43 m2w_s = to43(mul44(Ts, c.rs));
  const m43 w2m_s = to43(inverse44(to44(m2w_s)));

  float acc = 0.f;  // lane q < kVals: running sum of quantity q over the hit rays, in ray order
  uint32_t hits = 0;
  for (int r0 = 0; r0 < c.n_samples; r0 += 64) {
    const int i = r0 + lane;
    const bool live = i < c.n_samples;
    float u[5];
    ray_uniforms(c.seed, c.lod, vid, i, u);
    const v3 o = mul_point(c.i2w, mk((float)ix + u[0], (float)iy + u[1], (float)iz + u[2]));
    const float z = 1.f - 2.f * u[3];
    const float rr = sqrtf(gmax(0.f, 1.f - z * z));
    const float phi = (float)(2.0 * 3.14159265358979323846 * (double)u[4]);
    float sphi, cphi;
    det_sincosf(phi, &sphi, &cphi);
    const v3 d = mk(rr * cphi, rr * sphi, z);
    const float cur_tmax = c.clip ? sphere_exit(w2m_s, c.radius, o, d, 1e12f) : c.tmax_fixed;
    const WRay wr = wt_setup(d);
    int best = -1;
    float bt = 0.f, bb1 = 0.f, bb2 = 0.f;
    for (uint32_t k = 0; k < nc; ++k) {
      const uint32_t t = cand[beg + k];
      const v3 p0 = f4v(tri[3 * t]), p1 = f4v(tri[3 * t + 1]), p2 = f4v(tri[3 * t + 2]);
      float th, b1, b2;
      if (live && wt_hit(wr, o, p0, p1, p2, cur_tmax, &th, &b1, &b2) &&
          (!kTex || !c.alpha_on || hit_opaque(c, tidx, tmat, (int)t, b1, b2))) {
        if (best < 0 || th < bt || (th == bt && (int)t < best)) {
          best = (int)t;
          bt = th;
          bb1 = b1;
          bb2 = b2;
        }
      }
    }
    const bool hit = best >= 0;
    float val[kVals];
#pragma unroll
    for (int q = 0; q < kVals; ++q) val[q] = 0.f;
    if (hit) {
      // closest-hit: the hit's texture coordinate (0 without texture coordinates) and
      // normal (shaders.optix.cu:48-114)
      int m = tmat ? tmat[best] : 0;
      m = (m >= 0 && m < c.n_materials) ? m : 0;
      const float* mf = matf + 8 * m;
      const int4 mt = kTex && c.mtex ? c.mtex[m] : make_int4(-1, -1, -1, -1);
      float2 tc = make_float2(0.f, 0.f);
      if (kTex && c.uv) tc = tri_tc(c, tidx, best, bb1, bb2);
      const v3 n = hit_normal(c, tidx, nrm, best, bb1, bb2, kTex && c.normal_on ? mt.z : -1, tc);
      const float sg = mf[6];
      // buildS(n, sigma) = n n^T + sigma (I - n n^T)  (sggx_phase_function.hpp:95-105)
      const float xx = n.x * n.x, xy = n.x * n.y, xz = n.x * n.z, yy = n.y * n.y, yz = n.y * n.z, zz = n.z * n.z;
      val[0] = xx + sg * (yy + zz);
      val[1] = yy + sg * (xx + zz);
      val[2] = zz + sg * (xx + yy);
      val[3] = xy + sg * (-xy);
      val[4] = xz + sg * (-xz);
      val[5] = yz + sg * (-yz);
      // CMaterial::diffuse / glossy (material.hpp:111-127): the texture at tc, else the constant
      float4 dc = make_float4(mf[0], mf[1], mf[2], 0.f), gc = make_float4(mf[3], mf[4], mf[5], 0.f);
      if (kTex && mt.x >= 0) dc = tex_fetch(c.texels, c.tex[mt.x], tc.x, tc.y);
      if (kTex && mt.y >= 0) gc = tex_fetch(c.texels, c.tex[mt.y], tc.x, tc.y);
      val[6] = dc.x;
      val[7] = dc.y;
      val[8] = dc.z;
      val[9] = gc.x;
      val[10] = gc.y;
      val[11] = gc.z;
      val[12] = n.x;
      val[13] = n.y;
      val[14] = n.z;
      val[15] = mf[7];
      val[16] = cur_tmax;
    }
    // the hit rays' values, compacted by hit rank (ray order) into rows of one
    // value each; lane q < kVals then sums row q in order, four hits per load
    const unsigned long long hb = __ballot(hit);
    const uint32_t rank = (uint32_t)__popcll(hb & ((1ull << lane) - 1ull));
    if (hit) {
#pragma unroll
      for (int q = 0; q < kVals; ++q) V[q * kRow + rank] = val[q];
      TS[hits + rank] = cur_tmax;
    }
    wave_sync();
    const uint32_t nh = (uint32_t)__popcll(hb);
    if (lane < kVals) {
      const float* row = V + lane * kRow;
      const float4* row4 = reinterpret_cast<const float4*>(row);
      uint32_t h = 0;
      for (; h + 4u <= nh; h += 4u) {
        const float4 v = row4[h >> 2];
        acc += v.x;
        acc += v.y;
        acc += v.z;
        acc += v.w;
      }
      for (; h < nh; ++h) acc += row[h];
    }
    hits += (uint32_t)__popcll(hb);
    wave_sync();
  }
  float sum[kVals];
#pragma unroll
  for (int q = 0; q < kVals; ++q) sum[q] = __shfl(acc, q);

  uint32_t wout[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0x3C00u};
  if (hits > 0) {
    // estimateDensity (mesh_filter.hpp:182-204, 216-226)
    const float avg = sum[16] / (float)hits;
    const float P = (float)hits / (float)c.n_samples;
    float dens = -det_logf(1.f - P) / (avg * c.sigma_t);
    const float na = c.alpha / c.sigma_t;
    // with at most 64 rays, lane h keeps the clip distance of hit h in a register,
    // computes its exp term and stores it to the wave's LDS area (free after the
    // ray sums; 16-byte aligned); the hit-order sum reads four terms per load
    const bool lane_sums = c.clip && c.n_samples <= 64;
    const float ts_l = (lane_sums && (uint32_t)lane < hits) ? TS[lane] : 0.f;
    for (int it = 0; it < c.iterations; ++it) {
      float Pv;
      if (lane_sums) {
        V[lane] = (uint32_t)lane < hits ? det_expf(-dens * c.sigma_t * ts_l) : 0.f;
        wave_sync();
        const float4* V4 = reinterpret_cast<const float4*>(V);
        float e = 0.f;
        uint32_t h = 0;
        for (; h + 4u <= hits; h += 4u) {
          const float4 q = V4[h >> 2];
          e += q.x;
          e += q.y;
          e += q.z;
          e += q.w;
        }
        for (; h < hits; ++h) e += V[h];
        wave_sync();
        Pv = 1.f - e / (float)hits;
      } else if (c.clip) {
        for (uint32_t h = (uint32_t)lane; h < hits; h += 64) V[h] = det_expf(-dens * c.sigma_t * TS[h]);
        wave_sync();
        float e = 0.f;
        for (uint32_t h = 0; h < hits; ++h) e += V[h];
        wave_sync();
        Pv = 1.f - e / (float)hits;
      } else {
        Pv = 1.f - det_expf(-dens * c.sigma_t * c.tmax_fixed);
      }
      // each iteration is a function of dens alone: once it returns dens unchanged
      // (bit for bit), every later one would too, so the remaining ones are skipped
      // (the result is that of all iterations; dens is wave-uniform)
      const float nd = dens - na * (Pv - P);
      if (nd == dens) break;
      dens = nd;
    }
    const float ih = 1.f / (float)hits;
    const float S6[6] = {sum[0] * ih, sum[1] * ih, sum[2] * ih, sum[3] * ih, sum[4] * ih, sum[5] * ih};
    const float dif[3] = {sum[6] * ih, sum[7] * ih, sum[8] * ih};
    const float spec[3] = {sum[9] * ih, sum[10] * ih, sum[11] * ih};
    const v3 nm = normalize(mul_dir(c.w2m, mk(sum[12], sum[13], sum[14]) * ih));
    const float nv[3] = {nm.x, nm.y, nm.z};
    encode_record(dens, S6, dif, spec, nv, sum[15] * ih, wout);
  }
  if (lane == 0) {
    out[2 * orec] = make_uint4(wout[0], wout[1], wout[2], wout[3]);
    out[2 * orec + 1] = make_uint4(wout[4], wout[5], wout[6], wout[7]);
  }
}

// ----------------------------------------------------------------- host ---
#define R2F_HIP(expr, what)                           \
  do {                                                \
    const hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return r2host::hip_fail(ctx, e_, what); \
  } while (0)

// setupGrid (openvdb_backend.cpp:47-106) on the host, same arithmetic as glm
int setup_grid(const r2_mesh* m, float voxel_size, r2_filter_grid* g) {
  if (!(voxel_size > 0.f)) return R2_ERR_INVALID;
  const m44 m2w = to44(m->m2w);
  const m43 w2m = to43(inverse44(m2w));
  const v3 a = mul_point(m->m2w, mk(m->model_min[0], m->model_min[1], m->model_min[2]));
  const v3 b = mul_point(m->m2w, mk(m->model_max[0], m->model_max[1], m->model_max[2]));
  const v3 wmin = mk(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z));
  const v3 wmax = mk(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z));
  const v3 ext = wmax - wmin;
  const float fn[3] = {ext.x / voxel_size, ext.y / voxel_size, ext.z / voxel_size};
  int nv[3];
  for (int i = 0; i < 3; ++i) {
    const float v = ceilf(fn[i] - 0.001f);
    if (!(v >= 1.f) || v > 65536.f) return R2_ERR_INVALID;
    nv[i] = (int)v;
    g->num_voxels[i] = nv[i];
  }
  m44 T = identity44();
  T.c[3][0] = wmin.x;
  T.c[3][1] = wmin.y;
  T.c[3][2] = wmin.z;
  m44 S1 = identity44();
  S1.c[0][0] = ext.x;
  S1.c[1][1] = ext.y;
  S1.c[2][2] = ext.z;
  m44 S2 = identity44();
  S2.c[0][0] = 1.f / (float)nv[0];
  S2.c[1][1] = 1.f / (float)nv[1];
  S2.c[2][2] = 1.f / (float)nv[2];
  const m44 i2m = mul44(mul44(mul44(to44(w2m), T), S1), S2);
  const m44 m2i = inverse44(i2m);
  const m44 m2w4 = inverse44(to44(w2m));
  const m44 i2w = mul44(m2w4, i2m);
  // glm mat4 * vec4 adds the column products pairwise: (c0 x + c1 y) + (c2 z + c3 w)
  const float n3[3] = {(float)nv[0], (float)nv[1], (float)nv[2]};
  float c0[3], c1[3];
  for (int r = 0; r < 3; ++r) {
    c0[r] = (i2w.c[0][r] * 0.f + i2w.c[1][r] * 0.f) + (i2w.c[2][r] * 0.f + i2w.c[3][r] * 1.f);
    c1[r] = (i2w.c[0][r] * n3[0] + i2w.c[1][r] * n3[1]) + (i2w.c[2][r] * n3[2] + i2w.c[3][r] * 1.f);
  }
  const m43 i2m3 = to43(i2m), m2i3 = to43(m2i), m2w3 = to43(m2w4);
  memcpy(g->index_to_model, i2m3.m, sizeof(float) * 12);
  memcpy(g->model_to_index, m2i3.m, sizeof(float) * 12);
  memcpy(g->model_to_world, m2w3.m, sizeof(float) * 12);
  memcpy(g->world_to_model, w2m.m, sizeof(float) * 12);
  for (int r = 0; r < 3; ++r) {
    g->world_bb_min[r] = gmin(c0[r], c1[r]);
    g->world_bb_max[r] = gmax(c0[r], c1[r]);
  }
  return R2_OK;
}

FConst make_consts(const r2_mesh* m, const r2_filter_grid& g, const r2_filter_params& p) {
  FConst c{};
  memcpy(c.i2m.m, g.index_to_model, sizeof(float) * 12);
  memcpy(c.m2w.m, g.model_to_world, sizeof(float) * 12);
  memcpy(c.w2m.m, g.world_to_model, sizeof(float) * 12);
  c.i2w = to43(mul44(to44(c.m2w), to44(c.i2m)));
  c.mesh_m2w = m->m2w;
  c.rs = r2host::get_rotation(mk(1.f, 0.f, 0.f));
  c.nx = g.num_voxels[0];
  c.ny = g.num_voxels[1];
  c.nz = g.num_voxels[2];
  // m_voxelSize = (worldBB.max - worldBB.min) / numVoxels (mesh_filter.hpp:48)
  const v3 vs = mk((g.world_bb_max[0] - g.world_bb_min[0]) / (float)c.nx,
                   (g.world_bb_max[1] - g.world_bb_min[1]) / (float)c.ny,
                   (g.world_bb_max[2] - g.world_bb_min[2]) / (float)c.nz);
  c.diagonal = length(vs);
  c.tmax_fixed = vs.x;
  c.radius = 1.1f * c.diagonal / 2.f;
  // a ray starts inside the voxel (within diagonal/2 of the centre) and ends on
  // the clip sphere, or after tmax_fixed when rays are not clipped
  const float r = p.clip_rays ? c.radius : 0.5f * c.diagonal + c.tmax_fixed;
  float scale = 0.f;
  for (int k = 0; k < 3; ++k) scale = fmaxf(scale, fmaxf(fabsf(g.world_bb_min[k]), fabsf(g.world_bb_max[k])));
  c.reach = r * 1.001f + 1e-5f * scale;
  const float vsa[3] = {vs.x, vs.y, vs.z};
  float minvs = vsa[0];
  for (int k = 0; k < 3; ++k) {
    c.bbmin[k] = g.world_bb_min[k];
    c.inv_vs[k] = 1.f / vsa[k];
    minvs = fminf(minvs, vsa[k]);
  }
  c.range_margin = 0.05f + 1e-5f * scale / minvs;
  c.n_samples = p.samples_per_voxel;
  c.iterations = p.estimation_iterations;
  c.clip = p.clip_rays ? 1 : 0;
  c.sigma_t = p.sigma_t;
  c.alpha = p.alpha;
  c.seed = p.seed;
  c.lod = (uint32_t)p.lod;
  c.n_materials = m->n_materials;
  c.uv = m->uv;
  c.pos = m->pos;
  c.texels = m->texels;
  c.tex = m->tex;
  c.mtex = m->mtex;
  c.alpha_on = m->uv && m->any_alpha ? 1 : 0;
  c.normal_on = m->uv && m->any_normal ? 1 : 0;
  return c;
}

struct Binning {
  uint32_t* act = nullptr;
  uint32_t* off = nullptr;
  uint32_t* cnt = nullptr;
  uint32_t* cand = nullptr;
  uint32_t n_act = 0;
  unsigned long long n_pairs = 0, n_tasks = 0;
};

// context pool slots of the binning buffers
enum {
  kSlotRange = 0, kSlotTasks, kSlotBegin, kSlotCounter, kSlotTaskMark, kSlotTaskTri, kSlotKeys, kSlotKeys2,
  kSlotVox, kSlotCand, kSlotAct, kSlotCnt, kSlotOff, kSlotTemp, kSlotRuns, kSlotRecords, kSlotIds, kSlotTaskCnt,
  kSlotTaskOff, kSlotTaskMask
};

template <class T>
hipError_t pooled(r2_ctx* ctx, int slot, size_t count, T** out) {
  void* p = nullptr;
  const hipError_t e = r2host::pool_get(ctx, slot, sizeof(T) * count, &p);
  *out = (T*)p;
  return e;
}

// candidate lists per active voxel (sorted by voxel id, then triangle index)
int bin_triangles(r2_ctx* ctx, const r2_mesh* m, const FConst& c, Binning* B) {
  hipStream_t s = ctx->stream;
  const int n = m->n_triangles;
  TriRange* rng = nullptr;
  unsigned long long *tasks = nullptr, *begin = nullptr;
  R2F_HIP(pooled(ctx, kSlotRange, n, &rng), "filter: alloc ranges");
  R2F_HIP(pooled(ctx, kSlotTasks, n, &tasks), "filter: alloc tasks");
  R2F_HIP(pooled(ctx, kSlotBegin, n, &begin), "filter: alloc task offsets");
  hipLaunchKernelGGL(k_tri_prep, dim3((n + 255) / 256), dim3(256), 0, s, m->tri, n, c, rng, tasks);
  R2F_HIP(hipGetLastError(), "filter: k_tri_prep");
  void* tmp = nullptr;
  size_t tb = 0;
  R2F_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, tasks, begin, n, s), "filter: scan size");
  R2F_HIP(r2host::pool_get(ctx, kSlotTemp, tb, &tmp), "filter: alloc scan temp");
  R2F_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, tasks, begin, n, s), "filter: scan tasks");
  unsigned long long hb[2] = {0, 0};
  R2F_HIP(hipMemcpyAsync(&hb[0], begin + (n - 1), sizeof(unsigned long long), hipMemcpyDeviceToHost, s), "filter: d2h");
  R2F_HIP(hipMemcpyAsync(&hb[1], tasks + (n - 1), sizeof(unsigned long long), hipMemcpyDeviceToHost, s), "filter: d2h");
  R2F_HIP(hipStreamSynchronize(s), "filter: sync");
  B->n_tasks = hb[0] + hb[1];
  if (B->n_tasks == 0) return R2_OK;
  if (B->n_tasks > 0x7FFFFFFFull) return r2host::fail(ctx, R2_ERR_UNSUPPORTED, "filter: triangle/voxel overlap too large");
  const int nt = (int)B->n_tasks;
  uint32_t *mark = nullptr, *task_tri = nullptr;
  R2F_HIP(pooled(ctx, kSlotTaskMark, nt, &mark), "filter: alloc task map");
  R2F_HIP(pooled(ctx, kSlotTaskTri, nt, &task_tri), "filter: alloc task map");
  R2F_HIP(hipMemsetAsync(mark, 0, sizeof(uint32_t) * nt, s), "filter: clear task map");
  hipLaunchKernelGGL(k_task_mark, dim3((n + 255) / 256), dim3(256), 0, s, tasks, begin, n, mark);
  R2F_HIP(hipGetLastError(), "filter: k_task_mark");
  R2F_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, tb, mark, task_tri, hipcub::Max(), nt, s), "filter: scan size");
  R2F_HIP(r2host::pool_get(ctx, kSlotTemp, tb, &tmp), "filter: alloc scan temp");
  R2F_HIP(hipcub::DeviceScan::InclusiveScan(tmp, tb, mark, task_tri, hipcub::Max(), nt, s), "filter: task map");
  const dim3 grid((unsigned)((B->n_tasks + kTasksPerBlock - 1) / kTasksPerBlock));
  uint32_t *task_cnt = nullptr, *task_off = nullptr;
  unsigned long long* task_mask = nullptr;
  R2F_HIP(pooled(ctx, kSlotTaskCnt, nt, &task_cnt), "filter: alloc task counts");
  R2F_HIP(pooled(ctx, kSlotTaskOff, nt, &task_off), "filter: alloc task offsets");
  R2F_HIP(pooled(ctx, kSlotTaskMask, nt, &task_mask), "filter: alloc task masks");
  hipLaunchKernelGGL(k_pairs<false>, grid, dim3(256), 0, s, m->tri, c, rng, begin, task_tri, B->n_tasks, task_cnt,
                     task_mask, (const uint32_t*)nullptr, (unsigned long long*)nullptr);
  R2F_HIP(hipGetLastError(), "filter: k_pairs count");
  R2F_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, task_cnt, task_off, nt, s), "filter: scan size");
  R2F_HIP(r2host::pool_get(ctx, kSlotTemp, tb, &tmp), "filter: alloc scan temp");
  R2F_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, task_cnt, task_off, nt, s), "filter: scan pair counts");
  uint32_t last[2] = {0, 0};
  R2F_HIP(hipMemcpyAsync(&last[0], task_off + (nt - 1), 4, hipMemcpyDeviceToHost, s), "filter: d2h");
  R2F_HIP(hipMemcpyAsync(&last[1], task_cnt + (nt - 1), 4, hipMemcpyDeviceToHost, s), "filter: d2h");
  R2F_HIP(hipStreamSynchronize(s), "filter: sync");
  // each task holds at most 64 pairs, so the u32 offsets are exact while the total fits
  if ((unsigned long long)nt * 64ull > 0xFFFFFFFFull)
    return r2host::fail(ctx, R2_ERR_UNSUPPORTED, "filter: too many binning tasks");
  const unsigned long long np = (unsigned long long)last[0] + last[1];
  B->n_pairs = np;
  if (np == 0) return R2_OK;
  if (np > 0x7FFFFFFFull) return r2host::fail(ctx, R2_ERR_UNSUPPORTED, "filter: more than 2^31 voxel/triangle pairs");
  unsigned long long *keys = nullptr, *keys2 = nullptr;
  R2F_HIP(pooled(ctx, kSlotKeys, np, &keys), "filter: alloc keys");
  R2F_HIP(pooled(ctx, kSlotKeys2, np, &keys2), "filter: alloc keys");
  hipLaunchKernelGGL(k_pairs<true>, grid, dim3(256), 0, s, m->tri, c, rng, begin, task_tri, B->n_tasks, task_cnt,
                     task_mask, (const uint32_t*)task_off, keys);
  R2F_HIP(hipGetLastError(), "filter: k_pairs emit");
  const unsigned long long nvox = (unsigned long long)c.nx * c.ny * c.nz;
  int vbits = 1;
  while ((1ull << vbits) < nvox) ++vbits;
  const int npi = (int)np;
  size_t sb = 0;
  // the keys arrive in task order, i.e. ascending triangle; the radix sort is
  // stable, so sorting the voxel bits alone gives the same (voxel, triangle) order
  R2F_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, sb, keys, keys2, npi, 32, 32 + vbits, s), "filter: sort size");
  R2F_HIP(r2host::pool_get(ctx, kSlotTemp, sb, &tmp), "filter: alloc sort temp");
  R2F_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, sb, keys, keys2, npi, 32, 32 + vbits, s), "filter: sort");
  uint32_t *vox = nullptr, *runs = nullptr;
  R2F_HIP(pooled(ctx, kSlotVox, np, &vox), "filter: alloc voxels");
  R2F_HIP(pooled(ctx, kSlotCand, np, &B->cand), "filter: alloc candidates");
  hipLaunchKernelGGL(k_split, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, keys2, np, vox, B->cand);
  R2F_HIP(hipGetLastError(), "filter: k_split");
  R2F_HIP(pooled(ctx, kSlotAct, np, &B->act), "filter: alloc active");
  R2F_HIP(pooled(ctx, kSlotCnt, np, &B->cnt), "filter: alloc counts");
  R2F_HIP(pooled(ctx, kSlotOff, np, &B->off), "filter: alloc offsets");
  R2F_HIP(pooled(ctx, kSlotRuns, 2, &runs), "filter: alloc run count");
  size_t rb = 0;
  R2F_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, rb, vox, B->act, B->cnt, runs, npi, s), "filter: rle size");
  R2F_HIP(r2host::pool_get(ctx, kSlotTemp, rb, &tmp), "filter: alloc rle temp");
  R2F_HIP(hipcub::DeviceRunLengthEncode::Encode(tmp, rb, vox, B->act, B->cnt, runs, npi, s), "filter: rle");
  uint32_t nruns = 0;
  R2F_HIP(hipMemcpyAsync(&nruns, runs, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "filter: d2h");
  R2F_HIP(hipStreamSynchronize(s), "filter: sync");
  B->n_act = nruns;
  size_t ob = 0;
  R2F_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, ob, B->cnt, B->off, (int)nruns, s), "filter: scan size");
  R2F_HIP(r2host::pool_get(ctx, kSlotTemp, ob, &tmp), "filter: alloc scan temp");
  R2F_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, ob, B->cnt, B->off, (int)nruns, s), "filter: scan counts");
  return R2_OK;
}

int launch_filter(r2_ctx* ctx, const r2_mesh* m, const FConst& c, const Binning& B, const uint32_t* work,
                  uint32_t n_work, bool direct, uint4* out, float* kernel_ms) {
  if (n_work == 0) {
    *kernel_ms = 0.f;
    return R2_OK;
  }
  const int ts_cap = ((c.n_samples + 63) / 64) * 64;
  const size_t lds = sizeof(float) * kWaves * (kVals * kRow + ts_cap);
  const unsigned blocks = (n_work + kWaves - 1) / kWaves;
  R2F_HIP(hipEventRecord(ctx->ev0, ctx->stream), "filter: event");
  if (c.uv || c.mtex)
    hipLaunchKernelGGL(k_filter<true>, dim3(blocks), dim3(256), lds, ctx->stream, c, work, n_work, direct ? 1 : 0,
                       B.act, B.off, B.cnt, B.n_act, B.cand, m->tri, m->idx, m->nrm, m->mat, m->matf, out, ts_cap);
  else
    hipLaunchKernelGGL(k_filter<false>, dim3(blocks), dim3(256), lds, ctx->stream, c, work, n_work, direct ? 1 : 0,
                       B.act, B.off, B.cnt, B.n_act, B.cand, m->tri, m->idx, m->nrm, m->mat, m->matf, out, ts_cap);
  R2F_HIP(hipGetLastError(), "filter: k_filter");
  R2F_HIP(hipEventRecord(ctx->ev1, ctx->stream), "filter: event");
  R2F_HIP(hipEventSynchronize(ctx->ev1), "filter: sync");
  R2F_HIP(hipEventElapsedTime(kernel_ms, ctx->ev0, ctx->ev1), "filter: elapsed");
  return R2_OK;
}

int check_params(r2_ctx* ctx, const r2_mesh* m, const r2_filter_params* p) {
  if (!m || !p) return r2host::fail(ctx, R2_ERR_INVALID, "filter: null mesh / params");
  if (p->samples_per_voxel <= 0 || p->samples_per_voxel > 1000)
    return r2host::fail(ctx, R2_ERR_INVALID, "filter: samples_per_voxel must be in [1, 1000]");
  if (p->estimation_iterations < 0) return r2host::fail(ctx, R2_ERR_INVALID, "filter: negative iterations");
  if (!(p->sigma_t > 0.f)) return r2host::fail(ctx, R2_ERR_INVALID, "filter: sigma_t must be > 0");
  return R2_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI --
extern "C" int r2_mesh_upload(r2_ctx* ctx, const r2_mesh_desc* d, r2_mesh** out) {
  if (!ctx || !d || !out) return R2_ERR_INVALID;
  *out = nullptr;
  if (d->n_vertices <= 0 || d->n_triangles <= 0 || !d->positions || !d->normals || !d->indices)
    return r2host::fail(ctx, R2_ERR_INVALID, "mesh: empty or missing arrays");
  if (d->n_materials <= 0 || !d->mat_diffuse || !d->mat_specular || !d->mat_roughness || !d->mat_ior)
    return r2host::fail(ctx, R2_ERR_INVALID, "mesh: missing material table");
  for (int64_t i = 0; i < 3ll * d->n_triangles; ++i)
    if (d->indices[i] < 0 || d->indices[i] >= d->n_vertices)
      return r2host::fail(ctx, R2_ERR_INVALID, "mesh: vertex index out of range");
  if (d->material)
    for (int64_t t = 0; t < d->n_triangles; ++t)
      if (d->material[t] < 0 || d->material[t] >= d->n_materials)
        return r2host::fail(ctx, R2_ERR_INVALID, "mesh: material id out of range");
  if (d->n_textures < 0 || (d->n_textures > 0 && !d->textures))
    return r2host::fail(ctx, R2_ERR_INVALID, "mesh: missing texture table");
  int64_t n_texels = 0;
  for (int i = 0; i < d->n_textures; ++i) {
    const r2_texture_desc& t = d->textures[i];
    if (t.width <= 0 || t.height <= 0 || !t.rgba) return r2host::fail(ctx, R2_ERR_INVALID, "mesh: empty texture");
    n_texels += (int64_t)t.width * t.height;
  }
  if (n_texels >= (1ll << 31)) return r2host::fail(ctx, R2_ERR_UNSUPPORTED, "mesh: more than 2^31 texels");
  bool any_alpha = false, any_normal = false;
  if (d->mat_textures)
    for (int64_t i = 0; i < 4ll * d->n_materials; ++i) {
      if (d->mat_textures[i] < -1 || d->mat_textures[i] >= d->n_textures)
        return r2host::fail(ctx, R2_ERR_INVALID, "mesh: texture index out of range");
      any_normal = any_normal || (i % 4 == 2 && d->mat_textures[i] >= 0);
      any_alpha = any_alpha || (i % 4 == 3 && d->mat_textures[i] >= 0);
    }
  r2_mesh* m = new r2_mesh();
  m->any_alpha = any_alpha;
  m->any_normal = any_normal;
  m->n_vertices = d->n_vertices;
  m->n_triangles = d->n_triangles;
  m->n_materials = d->n_materials;
  // mesh AABB over the vertices (model space)
  for (int k = 0; k < 3; ++k) {
    m->model_min[k] = FLT_MAX;
    m->model_max[k] = -FLT_MAX;
  }
  for (int64_t v = 0; v < d->n_vertices; ++v)
    for (int k = 0; k < 3; ++k) {
      m->model_min[k] = gmin(d->positions[3 * v + k], m->model_min[k]);
      m->model_max[k] = gmax(d->positions[3 * v + k], m->model_max[k]);
    }
  m44 T = identity44();
  T.c[3][0] = d->pos[0];
  T.c[3][1] = d->pos[1];
  T.c[3][2] = d->pos[2];
  const m44 R = r2host::get_rotation(mk(d->orientation[0], d->orientation[1], d->orientation[2]));
  m44 S = identity44();
  S.c[0][0] = d->scaling[0];
  S.c[1][1] = d->scaling[1];
  S.c[2][2] = d->scaling[2];
  m->m2w = to43(mul44(mul44(T, R), S));
  std::vector<float> matf((size_t)8 * d->n_materials);
  for (int i = 0; i < d->n_materials; ++i) {
    for (int k = 0; k < 3; ++k) {
      matf[8 * i + k] = d->mat_diffuse[3 * i + k];
      matf[8 * i + 3 + k] = d->mat_specular[3 * i + k];
    }
    matf[8 * i + 6] = d->mat_roughness[i];
    matf[8 * i + 7] = d->mat_ior[i];
  }
  auto bail = [&](hipError_t e, const char* what) {
    for (void* q : m->allocs) (void)hipFree(q);
    delete m;
    return r2host::hip_fail(ctx, e, what);
  };
  auto dalloc = [&](void** p, size_t bytes) {
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) m->allocs.push_back(*p);
    return e;
  };
  float* dpos = nullptr;
  hipError_t e;
  const size_t nt = (size_t)d->n_triangles, nv = (size_t)d->n_vertices;
  if ((e = dalloc((void**)&dpos, sizeof(float) * 3 * nv)) != hipSuccess) return bail(e, "mesh: alloc");
  if ((e = dalloc((void**)&m->nrm, sizeof(float) * 3 * nv)) != hipSuccess) return bail(e, "mesh: alloc");
  if ((e = dalloc((void**)&m->idx, sizeof(int32_t) * 3 * nt)) != hipSuccess) return bail(e, "mesh: alloc");
  if ((e = dalloc((void**)&m->mat, sizeof(int32_t) * nt)) != hipSuccess) return bail(e, "mesh: alloc");
  if ((e = dalloc((void**)&m->matf, sizeof(float) * matf.size())) != hipSuccess) return bail(e, "mesh: alloc");
  if ((e = dalloc((void**)&m->tri, sizeof(float4) * 3 * nt)) != hipSuccess) return bail(e, "mesh: alloc");
  hipStream_t s = ctx->stream;
  if ((e = hipMemcpyAsync(dpos, d->positions, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, s)) != hipSuccess)
    return bail(e, "mesh: copy");
  if ((e = hipMemcpyAsync(m->nrm, d->normals, sizeof(float) * 3 * nv, hipMemcpyHostToDevice, s)) != hipSuccess)
    return bail(e, "mesh: copy");
  if ((e = hipMemcpyAsync(m->idx, d->indices, sizeof(int32_t) * 3 * nt, hipMemcpyHostToDevice, s)) != hipSuccess)
    return bail(e, "mesh: copy");
  if (d->material) {
    if ((e = hipMemcpyAsync(m->mat, d->material, sizeof(int32_t) * nt, hipMemcpyHostToDevice, s)) != hipSuccess)
      return bail(e, "mesh: copy");
  } else if ((e = hipMemsetAsync(m->mat, 0, sizeof(int32_t) * nt, s)) != hipSuccess) {
    return bail(e, "mesh: clear");
  }
  if ((e = hipMemcpyAsync(m->matf, matf.data(), sizeof(float) * matf.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
    return bail(e, "mesh: copy");
  // textured materials: texture coordinates, and the model-space positions that
  // normal-mapped hits build their tangent frames from
  const bool keep_pos = d->uvs && any_normal;
  if (d->uvs) {
    if ((e = dalloc((void**)&m->uv, sizeof(float2) * nv)) != hipSuccess) return bail(e, "mesh: alloc");
    if ((e = hipMemcpyAsync(m->uv, d->uvs, sizeof(float2) * nv, hipMemcpyHostToDevice, s)) != hipSuccess)
      return bail(e, "mesh: copy");
  }
  std::vector<int4> tex_tab((size_t)std::max(d->n_textures, 1));
  if (d->n_textures > 0) {
    if ((e = dalloc((void**)&m->texels, sizeof(float4) * (size_t)n_texels)) != hipSuccess) return bail(e, "mesh: alloc");
    if ((e = dalloc((void**)&m->tex, sizeof(int4) * (size_t)d->n_textures)) != hipSuccess) return bail(e, "mesh: alloc");
    int64_t first = 0;
    for (int i = 0; i < d->n_textures; ++i) {
      const r2_texture_desc& t = d->textures[i];
      const size_t n = (size_t)t.width * t.height;
      if ((e = hipMemcpyAsync(m->texels + first, t.rgba, sizeof(float4) * n, hipMemcpyHostToDevice, s)) != hipSuccess)
        return bail(e, "mesh: copy");
      tex_tab[i] = make_int4((int)first, t.width, t.height, 0);
      first += (int64_t)n;
    }
    if ((e = hipMemcpyAsync(m->tex, tex_tab.data(), sizeof(int4) * d->n_textures, hipMemcpyHostToDevice, s)) !=
        hipSuccess)
      return bail(e, "mesh: copy");
  }
  if (d->mat_textures) {
    if ((e = dalloc((void**)&m->mtex, sizeof(int4) * (size_t)d->n_materials)) != hipSuccess) return bail(e, "mesh: alloc");
    if ((e = hipMemcpyAsync(m->mtex, d->mat_textures, sizeof(int4) * d->n_materials, hipMemcpyHostToDevice, s)) !=
        hipSuccess)
      return bail(e, "mesh: copy");
  }
  hipLaunchKernelGGL(k_world_tris, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, dpos, m->idx, m->m2w,
                     (int)nt, m->tri);
  if ((e = hipGetLastError()) != hipSuccess) return bail(e, "mesh: k_world_tris");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return bail(e, "mesh: sync");
  if (keep_pos) {
    m->pos = dpos;
  } else {
    // positions are only needed for the world-space triangles
    (void)hipFree(dpos);
    m->allocs.erase(std::find(m->allocs.begin(), m->allocs.end(), (void*)dpos));
  }
  *out = m;
  return R2_OK;
}

extern "C" int r2_mesh_free(r2_ctx* ctx, r2_mesh* m) {
  (void)ctx;
  if (!m) return R2_OK;
  for (void* q : m->allocs) (void)hipFree(q);
  delete m;
  return R2_OK;
}

extern "C" int r2_filter_setup(const r2_mesh* m, float voxel_size, r2_filter_grid* out) {
  if (!m || !out) return R2_ERR_INVALID;
  return setup_grid(m, voxel_size, out);
}

static int filter_prepare(r2_ctx* ctx, const r2_mesh* m, const r2_filter_params* p, r2_filter_grid* g, FConst* c) {
  int rc = check_params(ctx, m, p);
  if (rc) return rc;
  rc = setup_grid(m, p->voxel_size, g);
  if (rc) return r2host::fail(ctx, rc, "filter: bad voxel size (grid set-up failed)");
  const unsigned long long nvox = (unsigned long long)g->num_voxels[0] * g->num_voxels[1] * g->num_voxels[2];
  if (nvox >= (1ull << 32)) return r2host::fail(ctx, R2_ERR_UNSUPPORTED, "filter: more than 2^32 voxels");
  *c = make_consts(m, *g, *p);
  return R2_OK;
}

static int filter_common(r2_ctx* ctx, const r2_mesh* m, const r2_filter_params* p, r2_filter_grid* g, FConst* c,
                         Binning* B) {
  const int rc = filter_prepare(ctx, m, p, g, c);
  return rc ? rc : bin_triangles(ctx, m, *c, B);
}

extern "C" int r2_filter_lod(r2_ctx* ctx, const r2_mesh* m, const r2_filter_params* p, void* out_records,
                             int32_t out_on_device, r2_filter_grid* grid_out) {
  if (!ctx || !out_records) return R2_ERR_INVALID;
  const auto t0 = std::chrono::steady_clock::now();
  r2_filter_grid g{};
  FConst c{};
  Binning B;
  int rc = filter_prepare(ctx, m, p, &g, &c);
  if (rc) return rc;
  const unsigned long long nvox = (unsigned long long)c.nx * c.ny * c.nz;
  uint4* rec = nullptr;
  if (out_on_device) {
    rec = (uint4*)out_records;
  } else {
    R2F_HIP(pooled(ctx, kSlotRecords, 2ull * nvox, &rec), "filter: alloc records");
  }
  // the no-hit fill of all records streams on the side stream while the binning
  // (latency-bound) runs on the main one; the main stream waits for it before
  // the filter writes the active voxels.  The fill itself first waits for the
  // work already queued on the main stream, which may still read the output.
  R2F_HIP(hipEventRecord(ctx->main_mark, ctx->stream), "filter: mark main stream");
  R2F_HIP(hipStreamWaitEvent(ctx->side, ctx->main_mark, 0), "filter: order fill after main stream");
  hipLaunchKernelGGL(k_fill_empty, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, ctx->side, rec, nvox);
  // from here on the fill may be in flight: every exit joins the side stream
  auto fail_joined = [&](int code) {
    (void)hipStreamSynchronize(ctx->side);
    return code;
  };
  hipError_t he = hipGetLastError();
  if (he == hipSuccess) he = hipEventRecord(ctx->side_done, ctx->side);
  if (he != hipSuccess) return fail_joined(r2host::hip_fail(ctx, he, "filter: k_fill_empty"));
  rc = bin_triangles(ctx, m, c, &B);
  if (rc) return fail_joined(rc);
  he = hipStreamWaitEvent(ctx->stream, ctx->side_done, 0);
  if (he != hipSuccess) return fail_joined(r2host::hip_fail(ctx, he, "filter: join fill"));
  float kms = 0.f;
  rc = launch_filter(ctx, m, c, B, nullptr, B.n_act, true, rec, &kms);
  if (rc) return fail_joined(rc);
  if (!out_on_device) {
    he = hipMemcpyAsync(out_records, rec, 32ull * nvox, hipMemcpyDeviceToHost, ctx->stream);
    if (he != hipSuccess) return fail_joined(r2host::hip_fail(ctx, he, "filter: d2h records"));
  }
  he = hipStreamSynchronize(ctx->stream);
  if (he != hipSuccess) return fail_joined(r2host::hip_fail(ctx, he, "filter: sync"));
  if (grid_out) *grid_out = g;
  const auto t1 = std::chrono::steady_clock::now();
  ctx->fstats.active_voxels = B.n_act;
  ctx->fstats.pairs = (int64_t)B.n_pairs;
  ctx->fstats.voxels = (int64_t)nvox;
  ctx->fstats.rays = (int64_t)B.n_act * p->samples_per_voxel;
  ctx->fstats.filter_kernel_ms = kms;
  ctx->fstats.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return R2_OK;
}

extern "C" int r2_filter_voxels(r2_ctx* ctx, const r2_mesh* m, const r2_filter_params* p, int32_t n,
                                const uint32_t* voxel_ids, void* out_records) {
  if (!ctx || n < 0 || (n > 0 && (!voxel_ids || !out_records))) return R2_ERR_INVALID;
  r2_filter_grid g{};
  FConst c{};
  Binning B;
  int rc = filter_common(ctx, m, p, &g, &c, &B);
  if (rc) return rc;
  if (n == 0) return R2_OK;
  const unsigned long long nvox = (unsigned long long)c.nx * c.ny * c.nz;
  for (int i = 0; i < n; ++i)
    if (voxel_ids[i] >= nvox) return r2host::fail(ctx, R2_ERR_INVALID, "filter: voxel id out of range");
  uint32_t* dids = nullptr;
  uint4* drec = nullptr;
  R2F_HIP(pooled(ctx, kSlotIds, (size_t)n, &dids), "filter: alloc ids");
  R2F_HIP(pooled(ctx, kSlotRecords, 2ull * n, &drec), "filter: alloc records");
  R2F_HIP(hipMemcpyAsync(dids, voxel_ids, sizeof(uint32_t) * n, hipMemcpyHostToDevice, ctx->stream), "filter: h2d");
  float kms = 0.f;
  rc = launch_filter(ctx, m, c, B, dids, (uint32_t)n, false, drec, &kms);
  if (rc) return rc;
  R2F_HIP(hipMemcpyAsync(out_records, drec, 32ull * n, hipMemcpyDeviceToHost, ctx->stream), "filter: d2h");
  R2F_HIP(hipStreamSynchronize(ctx->stream), "filter: sync");
  return R2_OK;
}

extern "C" int r2_filter_get_stats(const r2_ctx* ctx, r2_filter_stats* out) {
  if (!ctx || !out) return R2_ERR_INVALID;
  *out = ctx->fstats;
  return R2_OK;
}
