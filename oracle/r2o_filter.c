/*
 * Oracle (test infrastructure only): CPU restatement of the mesh -> SGGX filter of
 * FrederikBoehm/Renderer2 (FP32).  Never linked by the product library.
 *
 * Restates:
 *   COpenvdbBackend::init / setupGrid (raytracer/src/filtering/openvdb_backend.cpp:47-106)
 *   CSceneobject model/world transforms and AABBs (raytracer/src/scene/sceneobject.cu:90-97,
 *     raytracer/include/intersect/aabb.hpp:9-13, raytracer/include/utility/functions.hpp:128-142)
 *   CMeshFilter ctor / run / estimateDensity / estimatePhitVolumeGT(2) / generateRay
 *     (raytracer/include/filtering/mesh_filter.hpp:21-122,182-226,260-266)
 *   __raygen__filtering voxel id = x + y*nx + z*nx*ny (raytracer/src/backend/shaders.optix.cu:258-284)
 *   CSGGXMicroflakeDistribution::buildS (raytracer/include/medium/sggx_phase_function.hpp:95-105)
 *   SFilteredDataCompact(const SFilteredData&) (raytracer/include/filtering/filtered_data.hpp:91-107)
 *   Sphere::intersect + CRay::transform (raytracer/include/shapes/sphere.hpp:33-90,
 *     raytracer/include/intersect/ray.hpp:41-47), CShape model = translate * getRotation
 *     (raytracer/include/shapes/shape.hpp:49-57)
 *   CSampler::uniformSampleSphere (raytracer/include/sampling/sampler.hpp:86-94)
 *   __closesthit__ch mesh branch without texture coordinates: barycentric smooth normal,
 *     normalize, modelToWorld, normalize; ior of the material
 *     (raytracer/src/backend/shaders.optix.cu:48-58,106-121)
 *   optixTrace interval [0, tmax] (raytracer/include/scene/device_scene_impl.hpp:15-27)
 *
 * Replacements (parity with the reference unpinned at the last-ulp level):
 *   - OptiX 7.4 hardware triangle intersection -> the watertight ray/triangle test of
 *     Woop, Benthin and Wald (JCGT 2(1), 2013) on world-space triangles.  The closest hit
 *     is the lexicographic minimum of (t, triangle index) over all triangles, so it does
 *     not depend on the acceleration structure (a median-split BVH here).
 *   - cuRAND XORWOW per voxel -> Philox4x32-10 keyed (seed, lod), counter
 *     (voxel id lo, voxel id hi, block, 0xF17E); uniforms are (u32 >> 8) * 2^-24.
 *   - CUDA cosf/sinf/expf/logf -> the same functions evaluated in double from their
 *     Taylor / atanh series and rounded once to float (det_* below).
 * Materials are constant per triangle (no textures, so no normal map and no alpha test).
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "r2o.h"

typedef struct { float x, y, z; } f3;
static inline f3 F3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add(f3 a, f3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 scl(f3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
static inline f3 sscl(float s, f3 a) { return F3(s * a.x, s * a.y, s * a.z); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len(f3 a) { return sqrtf(dot(a, a)); }
/* glm::normalize = v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt */
static inline f3 nrm(f3 a) { return scl(a, 1.f / sqrtf(dot(a, a))); }
static inline f3 crs(f3 x, f3 y) { return F3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y); }
/* glm scalar min/max: (b < a) ? b : a and (a < b) ? b : a */
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }
static inline float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* deterministic sin/cos/exp/log: r2o_common.c */
#define det_sincosf r2o_det_sincosf
#define det_expf r2o_det_expf
#define det_logf r2o_det_logf

/* ------------------------------------------------------------ matrices -- */
typedef struct { float c[4][4]; } m44;
static m44 ident(void) {
  m44 r;
  memset(&r, 0, sizeof(r));
  for (int i = 0; i < 4; ++i) r.c[i][i] = 1.f;
  return r;
}
/* glm mat4 * mat4: Result[i] = A0*B[i][0] + A1*B[i][1] + A2*B[i][2] + A3*B[i][3] */
static m44 mul(const m44* a, const m44* b) {
  m44 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r.c[i][j] = a->c[0][j] * b->c[i][0] + a->c[1][j] * b->c[i][1] + a->c[2][j] * b->c[i][2] + a->c[3][j] * b->c[i][3];
  return r;
}
/* glm mat4 inverse by 2x2 cofactors */
static m44 inv44(const m44* mm) {
  const float(*m)[4] = mm->c;
  const float c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3], c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
  const float c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3], c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
  const float c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3], c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
  const float c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2], c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2];
  const float c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2], c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
  const float c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3], c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
  const float c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2], c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2];
  const float c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2], c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
  const float c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1], c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
  const float f3v[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
  const float v0[4] = {m[1][0], m[0][0], m[0][0], m[0][0]}, v1[4] = {m[1][1], m[0][1], m[0][1], m[0][1]};
  const float v2[4] = {m[1][2], m[0][2], m[0][2], m[0][2]}, v3[4] = {m[1][3], m[0][3], m[0][3], m[0][3]};
  const float sa[4] = {1.f, -1.f, 1.f, -1.f}, sb[4] = {-1.f, 1.f, -1.f, 1.f};
  m44 inv;
  for (int k = 0; k < 4; ++k) {
    inv.c[0][k] = (v1[k] * f0[k] - v2[k] * f1[k] + v3[k] * f2[k]) * sa[k];
    inv.c[1][k] = (v0[k] * f0[k] - v2[k] * f3v[k] + v3[k] * f4[k]) * sb[k];
    inv.c[2][k] = (v0[k] * f1[k] - v1[k] * f3v[k] + v3[k] * f5[k]) * sa[k];
    inv.c[3][k] = (v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k]) * sb[k];
  }
  const float d0 = m[0][0] * inv.c[0][0], d1 = m[0][1] * inv.c[1][0], d2 = m[0][2] * inv.c[2][0], d3 = m[0][3] * inv.c[3][0];
  const float o = 1.f / ((d0 + d1) + (d2 + d3));
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) inv.c[i][j] = inv.c[i][j] * o;
  return inv;
}
/* glm::mat4(glm::mat4x3(m)): bottom row (0,0,0,1) */
static m44 trunc43(const m44* a) {
  m44 r = *a;
  for (int i = 0; i < 4; ++i) r.c[i][3] = (i == 3) ? 1.f : 0.f;
  return r;
}
/* glm mat4x3 * vec4, left to right */
static f3 xpoint(const m44* m, f3 v) {
  return F3(m->c[0][0] * v.x + m->c[1][0] * v.y + m->c[2][0] * v.z + m->c[3][0] * 1.f,
            m->c[0][1] * v.x + m->c[1][1] * v.y + m->c[2][1] * v.z + m->c[3][1] * 1.f,
            m->c[0][2] * v.x + m->c[1][2] * v.y + m->c[2][2] * v.z + m->c[3][2] * 1.f);
}
static f3 xdir(const m44* m, f3 v) {
  return F3(m->c[0][0] * v.x + m->c[1][0] * v.y + m->c[2][0] * v.z + m->c[3][0] * 0.f,
            m->c[0][1] * v.x + m->c[1][1] * v.y + m->c[2][1] * v.z + m->c[3][1] * 0.f,
            m->c[0][2] * v.x + m->c[1][2] * v.y + m->c[2][2] * v.z + m->c[3][2] * 0.f);
}
/* glm mat4 * vec4: (m0 x + m1 y) + (m2 z + m3 w) */
static f3 xpoint44(const m44* m, f3 v) {
  return F3((m->c[0][0] * v.x + m->c[1][0] * v.y) + (m->c[2][0] * v.z + m->c[3][0] * 1.f),
            (m->c[0][1] * v.x + m->c[1][1] * v.y) + (m->c[2][1] * v.z + m->c[3][1] * 1.f),
            (m->c[0][2] * v.x + m->c[1][2] * v.y) + (m->c[2][2] * v.z + m->c[3][2] * 1.f));
}
static void to12(const m44* m, float* o) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) o[3 * i + j] = m->c[i][j];
}
static m44 from12(const float* o) {
  m44 m = ident();
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) m.c[i][j] = o[3 * i + j];
  return m;
}

/* getRotation (functions.hpp:128-142): glm::rotate(I, acos(n.y), normalize(n x y)), rounded */
static m44 get_rotation(f3 n) {
  const float cs = dot(n, F3(0.f, 1.f, 0.f));
  if (cs == 1.f) return ident();
  if (cs == -1.f) {
    m44 r = ident();
    for (int i = 0; i < 3; ++i) r.c[i][i] = -1.f;
    return r;
  }
  const float ang = acosf(dot(n, F3(0.f, 1.f, 0.f)));
  const f3 ax = nrm(crs(n, F3(0.f, 1.f, 0.f)));
  const float c = cosf(ang), s = sinf(ang);
  const f3 t = scl(ax, 1.f - c);
  m44 r = ident();
  r.c[0][0] = c + t.x * ax.x; r.c[0][1] = t.x * ax.y + s * ax.z; r.c[0][2] = t.x * ax.z - s * ax.y;
  r.c[1][0] = t.y * ax.x - s * ax.z; r.c[1][1] = c + t.y * ax.y; r.c[1][2] = t.y * ax.z + s * ax.x;
  r.c[2][0] = t.z * ax.x + s * ax.y; r.c[2][1] = t.z * ax.y - s * ax.x; r.c[2][2] = c + t.z * ax.z;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.c[i][j] = roundf(r.c[i][j]);
  return r;
}

/* sceneobject modelToWorld = translate(pos) * getRotation(orientation) * scale(scaling) */
static m44 model_to_world_of(const r2o_mesh_desc* d) {
  m44 T = ident();
  T.c[3][0] = d->pos[0]; T.c[3][1] = d->pos[1]; T.c[3][2] = d->pos[2];
  const m44 R = get_rotation(F3(d->orientation[0], d->orientation[1], d->orientation[2]));
  m44 S = ident();
  S.c[0][0] = d->scaling[0]; S.c[1][1] = d->scaling[1]; S.c[2][2] = d->scaling[2];
  const m44 TR = mul(&T, &R);
  const m44 M = mul(&TR, &S);
  return trunc43(&M);
}

/* COpenvdbBackend::init + setupGrid for one mesh object and one voxel size */
int r2o_filter_setup(const r2o_mesh_desc* d, float voxel_size, r2o_filter_grid* g) {
  if (d->n_vertices <= 0 || !(voxel_size > 0.f)) return -1;
  f3 mn = F3(FLT_MAX, FLT_MAX, FLT_MAX), mx = F3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
  for (int v = 0; v < d->n_vertices; ++v) {
    const f3 p = F3(d->positions[3 * v], d->positions[3 * v + 1], d->positions[3 * v + 2]);
    mn = F3(gmin(p.x, mn.x), gmin(p.y, mn.y), gmin(p.z, mn.z));
    mx = F3(gmax(p.x, mx.x), gmax(p.y, mx.y), gmax(p.z, mx.z));
  }
  const m44 m2w = model_to_world_of(d);
  const m44 w2m_full = inv44(&m2w);
  const m44 w2m = trunc43(&w2m_full);
  /* world AABB = SAABB::transform(model AABB, modelToWorld): both corners, min/max */
  const f3 a = xpoint(&m2w, mn), b = xpoint(&m2w, mx);
  const f3 wmin = F3(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z));
  const f3 wmax = F3(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z));
  const f3 ext = sub(wmax, wmin);
  const float fn[3] = {ext.x / voxel_size, ext.y / voxel_size, ext.z / voxel_size};
  int nv[3];
  for (int i = 0; i < 3; ++i) {
    nv[i] = (int)ceilf(fn[i] - 0.001f);
    if (nv[i] < 1) return -1;
    g->nv[i] = nv[i];
  }
  /* indexToModel = worldToModel * translate(minWorld) * scale(extent) * scale(1 / numVoxels) */
  m44 T = ident();
  T.c[3][0] = wmin.x; T.c[3][1] = wmin.y; T.c[3][2] = wmin.z;
  m44 S1 = ident();
  S1.c[0][0] = ext.x; S1.c[1][1] = ext.y; S1.c[2][2] = ext.z;
  m44 S2 = ident();
  S2.c[0][0] = 1.f / (float)nv[0]; S2.c[1][1] = 1.f / (float)nv[1]; S2.c[2][2] = 1.f / (float)nv[2];
  const m44 A = mul(&w2m, &T);
  const m44 B = mul(&A, &S1);
  const m44 i2m = mul(&B, &S2);
  const m44 m2i = inv44(&i2m);
  const m44 m2w4 = inv44(&w2m);
  const m44 i2w = mul(&m2w4, &i2m);
  const f3 c0 = xpoint44(&i2w, F3(0.f, 0.f, 0.f)), c1 = xpoint44(&i2w, F3((float)nv[0], (float)nv[1], (float)nv[2]));
  to12(&i2m, g->index_to_model);
  to12(&m2i, g->model_to_index);
  to12(&m2w4, g->model_to_world);
  to12(&w2m, g->world_to_model);
  g->world_bb_min[0] = gmin(c0.x, c1.x); g->world_bb_min[1] = gmin(c0.y, c1.y); g->world_bb_min[2] = gmin(c0.z, c1.z);
  g->world_bb_max[0] = gmax(c0.x, c1.x); g->world_bb_max[1] = gmax(c0.y, c1.y); g->world_bb_max[2] = gmax(c0.z, c1.z);
  return 0;
}

/* --------------------------------------------------------------- BVH ---- */
typedef struct { f3 mn, mx; int a, count; /* count > 0: leaf of tris order[a..a+count); else children a, -count */ } bnode;

struct r2o_filter_ctx {
  const r2o_mesh_desc* mesh;
  m44 m2w;
  int n_tris;
  f3 *p0, *p1, *p2; /* world-space triangle vertices */
  int* order;
  bnode* nodes;
  int n_nodes;
  float* keys;
};

static void tri_bounds(const r2o_filter_ctx* b, int t, f3* mn, f3* mx) {
  const f3 a = b->p0[t], c = b->p1[t], d = b->p2[t];
  *mn = F3(fminf(a.x, fminf(c.x, d.x)), fminf(a.y, fminf(c.y, d.y)), fminf(a.z, fminf(c.z, d.z)));
  *mx = F3(fmaxf(a.x, fmaxf(c.x, d.x)), fmaxf(a.y, fmaxf(c.y, d.y)), fmaxf(a.z, fmaxf(c.z, d.z)));
}

static const float* g_sort_keys;
static int cmp_key(const void* a, const void* b) {
  const int ia = *(const int*)a, ib = *(const int*)b;
  const float ka = g_sort_keys[ia], kb = g_sort_keys[ib];
  return (ka < kb) ? -1 : (ka > kb ? 1 : (ia - ib));
}

static int build_rec(r2o_filter_ctx* b, int first, int count) {
  const int id = b->n_nodes++;
  f3 mn = F3(FLT_MAX, FLT_MAX, FLT_MAX), mx = F3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
  f3 cmn = mn, cmx = mx;
  for (int i = first; i < first + count; ++i) {
    f3 a, c;
    tri_bounds(b, b->order[i], &a, &c);
    mn = F3(fminf(mn.x, a.x), fminf(mn.y, a.y), fminf(mn.z, a.z));
    mx = F3(fmaxf(mx.x, c.x), fmaxf(mx.y, c.y), fmaxf(mx.z, c.z));
    const f3 ce = scl(add(a, c), 0.5f);
    cmn = F3(fminf(cmn.x, ce.x), fminf(cmn.y, ce.y), fminf(cmn.z, ce.z));
    cmx = F3(fmaxf(cmx.x, ce.x), fmaxf(cmx.y, ce.y), fmaxf(cmx.z, ce.z));
  }
  /* pad the box so that float slab tests stay conservative */
  const f3 pad = F3(fabsf(mn.x) * 1e-6f + 1e-30f, fabsf(mn.y) * 1e-6f + 1e-30f, fabsf(mn.z) * 1e-6f + 1e-30f);
  const f3 padx = F3(fabsf(mx.x) * 1e-6f + 1e-30f, fabsf(mx.y) * 1e-6f + 1e-30f, fabsf(mx.z) * 1e-6f + 1e-30f);
  b->nodes[id].mn = sub(mn, pad);
  b->nodes[id].mx = add(mx, padx);
  if (count <= 4) {
    b->nodes[id].a = first;
    b->nodes[id].count = count;
    return id;
  }
  const f3 ce = sub(cmx, cmn);
  const int ax = (ce.x >= ce.y && ce.x >= ce.z) ? 0 : (ce.y >= ce.z ? 1 : 2);
  for (int i = first; i < first + count; ++i) {
    f3 a, c;
    tri_bounds(b, b->order[i], &a, &c);
    b->keys[b->order[i]] = 0.5f * (comp(a, ax) + comp(c, ax));
  }
  g_sort_keys = b->keys;
  qsort(b->order + first, count, sizeof(int), cmp_key);
  const int half = count / 2;
  const int l = build_rec(b, first, half);
  const int r = build_rec(b, first + half, count - half);
  b->nodes[id].a = l;
  b->nodes[id].count = -r;
  return id;
}

r2o_filter_ctx* r2o_filter_new(const r2o_mesh_desc* md) {
  if (md->n_triangles <= 0) return NULL;
  r2o_filter_ctx* b = (r2o_filter_ctx*)calloc(1, sizeof(r2o_filter_ctx));
  b->mesh = md;
  b->n_tris = md->n_triangles;
  b->p0 = (f3*)malloc(sizeof(f3) * b->n_tris);
  b->p1 = (f3*)malloc(sizeof(f3) * b->n_tris);
  b->p2 = (f3*)malloc(sizeof(f3) * b->n_tris);
  b->order = (int*)malloc(sizeof(int) * b->n_tris);
  b->nodes = (bnode*)malloc(sizeof(bnode) * 2 * (size_t)(b->n_tris + 1));
  b->keys = (float*)malloc(sizeof(float) * b->n_tris);
  b->m2w = model_to_world_of(md);
  for (int t = 0; t < b->n_tris; ++t) {
    const int32_t* id = md->indices + 3 * t;
    const float* P = md->positions;
    b->p0[t] = xpoint(&b->m2w, F3(P[3 * id[0]], P[3 * id[0] + 1], P[3 * id[0] + 2]));
    b->p1[t] = xpoint(&b->m2w, F3(P[3 * id[1]], P[3 * id[1] + 1], P[3 * id[1] + 2]));
    b->p2[t] = xpoint(&b->m2w, F3(P[3 * id[2]], P[3 * id[2] + 1], P[3 * id[2] + 2]));
    b->order[t] = t;
  }
  build_rec(b, 0, b->n_tris);
  return b;
}

void r2o_filter_free(r2o_filter_ctx* b) {
  if (!b) return;
  free(b->p0); free(b->p1); free(b->p2); free(b->order); free(b->nodes); free(b->keys);
  free(b);
}

/* watertight ray / triangle (Woop, Benthin, Wald 2013) */
typedef struct { int kx, ky, kz; float Sx, Sy, Sz; } wt_ray;
static wt_ray wt_setup(f3 d) {
  wt_ray w;
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  w.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  w.kx = w.kz + 1 == 3 ? 0 : w.kz + 1;
  w.ky = w.kx + 1 == 3 ? 0 : w.kx + 1;
  if (comp(d, w.kz) < 0.f) { const int t = w.kx; w.kx = w.ky; w.ky = t; }
  w.Sx = comp(d, w.kx) / comp(d, w.kz);
  w.Sy = comp(d, w.ky) / comp(d, w.kz);
  w.Sz = 1.f / comp(d, w.kz);
  return w;
}
/* 1 and (t, b1, b2) when the ray hits the triangle with 0 < t <= tmax */
static int wt_hit(const wt_ray* w, f3 o, f3 v0, f3 v1, f3 v2, float tmax, float* t, float* b1, float* b2) {
  const f3 A = sub(v0, o), B = sub(v1, o), Cc = sub(v2, o);
  const float Ax = comp(A, w->kx) - w->Sx * comp(A, w->kz), Ay = comp(A, w->ky) - w->Sy * comp(A, w->kz);
  const float Bx = comp(B, w->kx) - w->Sx * comp(B, w->kz), By = comp(B, w->ky) - w->Sy * comp(B, w->kz);
  const float Cx = comp(Cc, w->kx) - w->Sx * comp(Cc, w->kz), Cy = comp(Cc, w->ky) - w->Sy * comp(Cc, w->kz);
  float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return 0;
  const float det = U + V + W;
  if (det == 0.f) return 0;
  const float Az = w->Sz * comp(A, w->kz), Bz = w->Sz * comp(B, w->kz), Cz = w->Sz * comp(Cc, w->kz);
  const float T = U * Az + V * Bz + W * Cz;
  if (det < 0.f && (T >= 0.f || T < tmax * det)) return 0;
  if (det > 0.f && (T <= 0.f || T > tmax * det)) return 0;
  const float rd = 1.f / det;
  *t = T * rd;
  *b1 = V * rd;
  *b2 = W * rd;
  return 1;
}

static int box_hit(f3 mn, f3 mx, f3 o, f3 inv, float tmax) {
  float t0 = 0.f, t1 = tmax;
  const float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z}, oo[3] = {o.x, o.y, o.z}, iv[3] = {inv.x, inv.y, inv.z};
  for (int a = 0; a < 3; ++a) {
    float ta = (lo[a] - oo[a]) * iv[a], tb = (hi[a] - oo[a]) * iv[a];
    if (ta > tb) { const float t = ta; ta = tb; tb = t; }
    t0 = ta > t0 ? ta : t0;
    t1 = tb < t1 ? tb : t1;
    if (t0 > t1) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------ textures --- */
/* CTexture::operator() (texture.hpp:55-58): tex2D<float4> with linear filtering,
 * wrap addressing and normalised coordinates (texture.cu:166-189), texel centres
 * at (i + 0.5) / width; full-precision weights (CUDA's are 8-bit fractions).
 * The blend is two nested linear interpolations, so that a constant texture
 * returns its value exactly, as tex2D's fixed-point weights (which sum to one)
 * do: the normal map's zero test depends on it. */
static void tex_fetch(const r2o_texture* t, float x, float y, float out[4]) {
  const int W = t->width, H = t->height;
  x = x - floorf(x);
  y = y - floorf(y);
  const float tx = x * (float)W - 0.5f, ty = y * (float)H - 0.5f;
  const float fx = floorf(tx), fy = floorf(ty);
  const float a = tx - fx, b = ty - fy;
  int i0 = (int)fx, j0 = (int)fy;
  int i1 = i0 + 1, j1 = j0 + 1;
  i0 = ((i0 % W) + W) % W; i1 = ((i1 % W) + W) % W;
  j0 = ((j0 % H) + H) % H; j1 = ((j1 % H) + H) % H;
  const float* p00 = t->rgba + ((size_t)j0 * W + i0) * 4;
  const float* p10 = t->rgba + ((size_t)j0 * W + i1) * 4;
  const float* p01 = t->rgba + ((size_t)j1 * W + i0) * 4;
  const float* p11 = t->rgba + ((size_t)j1 * W + i1) * 4;
  for (int c = 0; c < 4; ++c) {
    const float top = p00[c] + a * (p10[c] - p00[c]);
    const float bot = p01[c] + a * (p11[c] - p01[c]);
    out[c] = top + b * (bot - top);
  }
}
enum { TEX_DIFFUSE = 0, TEX_GLOSSY = 1, TEX_NORMAL = 2, TEX_ALPHA = 3 };
/* texture of a material slot, or NULL (CMaterial's m_albedoTexture, m_glossyTexture,
 * m_normalTexture, m_alphaTexture, material.hpp:54-57) */
static const r2o_texture* mat_tex(const r2o_mesh_desc* md, int mat, int slot) {
  if (!md->mat_textures) return NULL;
  const int k = md->mat_textures[4 * mat + slot];
  return (k >= 0 && k < md->n_textures) ? &md->textures[k] : NULL;
}
static int tri_material(const r2o_mesh_desc* md, int t) { return md->material ? md->material[t] : 0; }
/* barycentric texture coordinate: (1 - b1 - b2) TC0 + b1 TC1 + b2 TC2 (shaders.optix.cu:61-66) */
static void tri_tc(const r2o_mesh_desc* md, int t, float b1, float b2, float tc[2]) {
  const int32_t* id = md->indices + 3 * t;
  const float* U = md->uvs;
  const float w0 = 1.f - b1 - b2;
  for (int c = 0; c < 2; ++c) tc[c] = (w0 * U[2 * id[0] + c] + b1 * U[2 * id[1] + c]) + b2 * U[2 * id[2] + c];
}
/* __anyhit__mesh (shaders.optix.cu:132-152) with CMaterial::opaque (material.hpp:130-137):
 * a hit on a mesh with texture coordinates is ignored where the material's alpha
 * mask is not > 0 */
static int hit_opaque(const r2o_mesh_desc* md, int t, float b1, float b2) {
  if (!md->uvs) return 1;
  const r2o_texture* at = mat_tex(md, tri_material(md, t), TEX_ALPHA);
  if (!at) return 1;
  float tc[2], v[4];
  tri_tc(md, t, b1, b2, tc);
  tex_fetch(at, tc[0], tc[1], v);
  return v[0] > 0.f;
}

/* closest opaque hit in (0, tmax]: lexicographic minimum of (t, triangle index) */
static int closest_hit(const r2o_filter_ctx* b, f3 o, f3 d, float tmax, int* tri, float* tb1, float* tb2, float* tt) {
  const wt_ray w = wt_setup(d);
  const f3 inv = F3(1.f / d.x, 1.f / d.y, 1.f / d.z);
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  int best = -1;
  float bt = tmax;
  while (sp) {
    const bnode* nd = &b->nodes[stack[--sp]];
    if (!box_hit(nd->mn, nd->mx, o, inv, bt * 1.0001f)) continue;
    if (nd->count > 0) {
      for (int i = nd->a; i < nd->a + nd->count; ++i) {
        const int t = b->order[i];
        float th, b1, b2;
        if (wt_hit(&w, o, b->p0[t], b->p1[t], b->p2[t], tmax, &th, &b1, &b2) && hit_opaque(b->mesh, t, b1, b2)) {
          if (best < 0 || th < bt || (th == bt && t < best)) {
            bt = th; best = t; *tb1 = b1; *tb2 = b2;
          }
        }
      }
    } else {
      stack[sp++] = -nd->count;
      stack[sp++] = nd->a;
    }
  }
  *tri = best;
  *tt = bt;
  return best >= 0;
}

/* ------------------------------------------------------------ filter ---- */
static float uni(r2o_rng* s) { return (float)(r2o_rng_next_u32(s) >> 8) * 5.9604644775390625e-08f; }

/* SFilteredDataCompact(const SFilteredData&): glm::clamp, float -> u16/u8 truncation, NaN -> 0 */
static uint32_t u16c(float x) {
  const float c = gmin(gmax(x, 0.f), 65535.f);
  return (c != c) ? 0u : (uint32_t)(uint16_t)c;
}
static uint32_t u8c(float x) {
  const float c = gmin(gmax(x, 0.f), 255.f);
  return (c != c) ? 0u : (uint32_t)(uint8_t)c;
}
static void encode(float dens, const float S[6], f3 dif, f3 spec, f3 n, float ior, uint8_t* out) {
  const float M = 65535.f, M8 = 255.f;
  uint32_t w[8];
  const uint32_t sx = u16c(M * sqrtf(S[0])), sy = u16c(M * sqrtf(S[1])), sz = u16c(M * sqrtf(S[2]));
  const uint32_t rxy = u16c(M * 0.5f * (S[3] / sqrtf(S[0] * S[1]) + 1.f));
  const uint32_t rxz = u16c(M * 0.5f * (S[4] / sqrtf(S[0] * S[2]) + 1.f));
  const uint32_t ryz = u16c(M * 0.5f * (S[5] / sqrtf(S[1] * S[2]) + 1.f));
  const f3 nn = nrm(n);
  memcpy(&w[0], &dens, 4);
  w[1] = sx | (sy << 16);
  w[2] = sz | (rxy << 16);
  w[3] = rxz | (ryz << 16);
  w[4] = u8c(M8 * dif.x) | (u8c(M8 * dif.y) << 8) | (u8c(M8 * dif.z) << 16) | (u8c(M8 * spec.x) << 24);
  w[5] = u8c(M8 * spec.y) | (u8c(M8 * spec.z) << 8) | (u16c(M * 0.5f * (nn.x + 1.f)) << 16);
  w[6] = u16c(M * 0.5f * (nn.y + 1.f)) | (u16c(M * 0.5f * (nn.z + 1.f)) << 16);
  w[7] = r2o_half_from_float_rne(ior);
  memcpy(out, w, 32);
}

/* Sphere::intersect for a ray given in world space; returns the hit t or 0 on a miss */
static float sphere_exit(const m44* w2m_s, float radius, f3 o, f3 d, float ray_tmax) {
  const f3 mo = xpoint(w2m_s, o);
  const f3 md = nrm(xdir(w2m_s, d));
  const float a = dot(md, md);
  const float b = 2.0f * dot(md, mo);
  const float c = dot(mo, mo) - radius * radius;
  const float disc = b * b - 4.f * a * c;
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

typedef struct {
  m44 i2m, m2w, w2m, i2w, Rs;
  int nx, ny, nz;
  float diagonal, tmax_fixed;
} fconst;

static void fconst_init(fconst* c, const r2o_filter_grid* g) {
  c->i2m = from12(g->index_to_model);
  c->m2w = from12(g->model_to_world);
  c->w2m = from12(g->world_to_model);
  /* indexToWorld = mat4(modelToWorld) * mat4(indexToModel), as mat4x3 */
  const m44 i2w_full = mul(&c->m2w, &c->i2m);
  c->i2w = trunc43(&i2w_full);
  c->nx = g->nv[0]; c->ny = g->nv[1]; c->nz = g->nv[2];
  /* m_voxelSize = (worldBB.max - worldBB.min) / numVoxels */
  const f3 vsz = F3((g->world_bb_max[0] - g->world_bb_min[0]) / (float)c->nx,
                    (g->world_bb_max[1] - g->world_bb_min[1]) / (float)c->ny,
                    (g->world_bb_max[2] - g->world_bb_min[2]) / (float)c->nz);
  c->diagonal = len(vsz);
  c->tmax_fixed = vsz.x;
  c->Rs = get_rotation(F3(1.f, 0.f, 0.f));
}

typedef struct { float* o; float* d; float* tmax; int32_t* tri; float* t; } fprobe;

/* __closesthit__ch for a mesh (shaders.optix.cu:48-114): the smooth normal from the
 * barycentrics; with texture coordinates, the hit's tc and, where the material has a
 * normal map and the triangle a tangent frame, CMaterial::normalmap(fromTBN(T, -B, N),
 * tc) (material.hpp:143-156, objects.hpp:60-67); then modelToWorld and normalize.
 * Without a normal map normalmap returns N, so the frame is only built with one. */
static f3 hit_normal(const r2o_filter_ctx* b, int tri, float b1, float b2, int mat, float tc[2]) {
  const r2o_mesh_desc* md = b->mesh;
  const int32_t* idx = md->indices + 3 * tri;
  const float* Nn = md->normals;
  const f3 N0 = F3(Nn[3 * idx[0]], Nn[3 * idx[0] + 1], Nn[3 * idx[0] + 2]);
  const f3 N1 = F3(Nn[3 * idx[1]], Nn[3 * idx[1] + 1], Nn[3 * idx[1] + 2]);
  const f3 N2 = F3(Nn[3 * idx[2]], Nn[3 * idx[2] + 1], Nn[3 * idx[2] + 2]);
  const float w0 = 1.f - b1 - b2;
  const f3 gn = nrm(add(add(sscl(w0, N0), sscl(b1, N1)), sscl(b2, N2)));
  f3 n = gn;
  if (md->uvs) {
    tri_tc(md, tri, b1, b2, tc);
    const r2o_texture* nt = mat_tex(md, mat, TEX_NORMAL);
    if (nt) {
      const float* P = md->positions;
      const float* U = md->uvs;
      const f3 P0 = F3(P[3 * idx[0]], P[3 * idx[0] + 1], P[3 * idx[0] + 2]);
      const f3 P1 = F3(P[3 * idx[1]], P[3 * idx[1] + 1], P[3 * idx[1] + 2]);
      const f3 P2 = F3(P[3 * idx[2]], P[3 * idx[2] + 1], P[3 * idx[2] + 2]);
      const f3 Ph = add(add(sscl(w0, P0), sscl(b1, P1)), sscl(b2, P2));
      const f3 e1 = sub(P1, Ph), e2 = sub(P2, Ph);
      const float d1x = U[2 * idx[1]] - tc[0], d1y = U[2 * idx[1] + 1] - tc[1];
      const float d2x = U[2 * idx[2]] - tc[0], d2y = U[2 * idx[2] + 1] - tc[1];
      const float den = d1x * d2y - d2x * d1y;
      if (den != 0.f) {
        const float f = 1.f / den;
        const f3 T = F3(f * (d2y * e1.x - d1y * e2.x), f * (d2y * e1.y - d1y * e2.y), f * (d2y * e1.z - d1y * e2.z));
        const f3 B = F3(f * (-d2x * e1.x + d1x * e2.x), f * (-d2x * e1.y + d1x * e2.y), f * (-d2x * e1.z + d1x * e2.z));
        const int tz = T.x == 0.f && T.y == 0.f && T.z == 0.f, bz = B.x == 0.f && B.y == 0.f && B.z == 0.f;
        if (!tz && !bz) {
          const f3 Tn = nrm(T), Bn = nrm(F3(-B.x, -B.y, -B.z));
          float v[4];
          tex_fetch(nt, tc[0], tc[1], v);
          const f3 m = F3(v[0] * 2.f - 1.f, v[1] * 2.f - 1.f, v[2] * 2.f - 1.f);
          if (!(m.x == 0.f && m.y == 0.f && m.z == 0.f)) {
            const f3 mn = nrm(m);
            n = add(add(sscl(mn.x, Tn), sscl(mn.y, Bn)), sscl(mn.z, gn));
          }
        }
      }
    }
  }
  return nrm(xdir(&b->m2w, n));
}

/* CMaterial::diffuse / glossy (material.hpp:111-127): the texture at tc, else the constant */
static f3 mat_color(const r2o_mesh_desc* md, int mat, int slot, const float* constant, const float tc[2]) {
  const r2o_texture* t = mat_tex(md, mat, slot);
  if (!t) return F3(constant[3 * mat], constant[3 * mat + 1], constant[3 * mat + 2]);
  float v[4];
  tex_fetch(t, tc[0], tc[1], v);
  return F3(v[0], v[1], v[2]);
}

/* CMeshFilter::run for one voxel id; returns the hit count */
static uint32_t filter_voxel(const r2o_filter_ctx* b, const fconst* fc, const r2o_filter_params* fp, int64_t vid,
                             uint8_t* o8, const fprobe* pr) {
  const r2o_mesh_desc* md = b->mesh;
  const int nx = fc->nx, ny = fc->ny;
  const int ix = (int)(vid % nx), iy = (int)((vid / nx) % ny), iz = (int)(vid / ((int64_t)nx * ny));
  r2o_rng rng;
  r2o_rng_init(&rng, fp->seed, (uint32_t)fp->lod, (uint32_t)vid, (uint32_t)((uint64_t)vid >> 32), 0xF17Eu);
  /* m_voxelCenter = modelToWorld * (indexToModel * (v + 0.5)) */
  const f3 vc_model = xpoint(&fc->i2m, F3((float)ix + 0.5f, (float)iy + 0.5f, (float)iz + 0.5f));
  const f3 ctr = xpoint(&fc->m2w, vc_model);
  /* Sphere(ctr, 1.1 * diagonal / 2, (1,0,0)): model = translate(ctr) * getRotation((1,0,0)) */
  m44 Ts = ident();
  Ts.c[3][0] = ctr.x; Ts.c[3][1] = ctr.y; Ts.c[3][2] = ctr.z;
  const m44 m2w_s_full = mul(&Ts, &fc->Rs);
  const m44 m2w_s = trunc43(&m2w_s_full);
  const m44 w2m_s_full = inv44(&m2w_s);
  const m44 w2m_s = trunc43(&w2m_s_full);
  const float radius = 1.1f * fc->diagonal / 2.f;
  const int N = fp->samples_per_voxel;
  uint32_t hits = 0;
  float S[6] = {0, 0, 0, 0, 0, 0};
  f3 dif = F3(0, 0, 0), spec = F3(0, 0, 0), nsum = F3(0, 0, 0);
  float dist = 0.f, ior = 0.f;
  float rayTs[1000];
  for (int i = 0; i < N; ++i) {
    /* generateRay: origin = voxel + (u, u, u) in index space; direction uniform on the sphere */
    const float u0 = uni(&rng), u1 = uni(&rng), u2 = uni(&rng);
    const f3 o = xpoint(&fc->i2w, F3((float)ix + u0, (float)iy + u1, (float)iz + u2));
    const float su1 = uni(&rng), su2 = uni(&rng);
    const float z = 1.f - 2.f * su1;
    const float r = sqrtf(gmax(0.f, 1.f - z * z));
    const float phi = (float)(2.0 * 3.14159265358979323846 * (double)su2);
    float sphi, cphi;
    det_sincosf(phi, &sphi, &cphi);
    const f3 d = F3(r * cphi, r * sphi, z);
    const float cur_tmax = fp->clip_rays ? sphere_exit(&w2m_s, radius, o, d, 1e12f) : fc->tmax_fixed;
    int tri;
    float b1, b2, th;
    const int hit = closest_hit(b, o, d, cur_tmax, &tri, &b1, &b2, &th);
    if (pr) {
      pr->o[3 * i] = o.x; pr->o[3 * i + 1] = o.y; pr->o[3 * i + 2] = o.z;
      pr->d[3 * i] = d.x; pr->d[3 * i + 1] = d.y; pr->d[3 * i + 2] = d.z;
      pr->tmax[i] = cur_tmax;
      pr->tri[i] = hit ? tri : -1;
      pr->t[i] = hit ? th : 0.f;
    }
    if (hit) {
      rayTs[hits] = cur_tmax;
      ++hits;
      const int mat = tri_material(md, tri);
      float tc[2] = {0.f, 0.f};
      const f3 n = hit_normal(b, tri, b1, b2, mat, tc);
      const float sg = md->mat_roughness[mat];
      /* buildS(n, sigma) = n n^T + sigma (I - n n^T) */
      const float xx = n.x * n.x, xy = n.x * n.y, xz = n.x * n.z, yy = n.y * n.y, yz = n.y * n.z, zz = n.z * n.z;
      S[0] += xx + sg * (yy + zz);
      S[1] += yy + sg * (xx + zz);
      S[2] += zz + sg * (xx + yy);
      S[3] += xy + sg * (-xy);
      S[4] += xz + sg * (-xz);
      S[5] += yz + sg * (-yz);
      dif = add(dif, mat_color(md, mat, TEX_DIFFUSE, md->mat_diffuse, tc));
      spec = add(spec, mat_color(md, mat, TEX_GLOSSY, md->mat_specular, tc));
      dist += cur_tmax;
      nsum = add(nsum, n);
      ior += md->mat_ior[mat];
    }
  }
  if (hits > 0) {
    const float avg = dist / (float)hits;
    /* estimateDensity */
    const float P = (float)hits / (float)N;
    float dens = -det_logf(1.f - P) / (avg * fp->sigma_t);
    const float na = fp->alpha / fp->sigma_t;
    for (int it = 0; it < fp->estimation_iterations; ++it) {
      float Pv;
      if (fp->clip_rays) {
        float acc = 0.f;
        for (uint32_t h = 0; h < hits; ++h) acc += det_expf(-dens * fp->sigma_t * rayTs[h]);
        Pv = 1.f - acc / (float)hits;
      } else {
        Pv = 1.f - det_expf(-dens * fp->sigma_t * fc->tmax_fixed);
      }
      dens = dens - na * (Pv - P);
    }
    const float ih = 1.f / (float)hits;
    float Sm[6];
    for (int q = 0; q < 6; ++q) Sm[q] = S[q] * ih;
    const f3 nm = nrm(xdir(&fc->w2m, scl(nsum, ih)));
    encode(dens, Sm, scl(dif, ih), scl(spec, ih), nm, ior * ih, o8);
  } else {
    const float Z[6] = {0, 0, 0, 0, 0, 0};
    encode(0.f, Z, F3(0, 0, 0), F3(0, 0, 0), F3(0, 0, 0), 1.f, o8);
  }
  return hits;
}

int r2o_filter_run(const r2o_filter_ctx* b, const r2o_filter_grid* g, const r2o_filter_params* fp, int64_t n_ids,
                   const int64_t* ids, uint8_t* out) {
  if (!b || fp->samples_per_voxel > 1000 || fp->samples_per_voxel <= 0) return -1;
  fconst fc;
  fconst_init(&fc, g);
  const int64_t total = ids ? n_ids : (int64_t)fc.nx * fc.ny * fc.nz;
  const int nthreads = fp->threads > 0 ? fp->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t k = 0; k < total; ++k) filter_voxel(b, &fc, fp, ids ? ids[k] : k, out + 32 * k, NULL);
  return 0;
}

/* probe: every ray of one voxel (origin, direction, ray tmax, hit triangle or -1, hit t) */
int r2o_filter_probe(const r2o_filter_ctx* b, const r2o_filter_grid* g, const r2o_filter_params* fp, int64_t vid,
                     uint8_t* rec, float* o, float* d, float* tmax, int32_t* tri, float* t) {
  if (!b || fp->samples_per_voxel > 1000 || fp->samples_per_voxel <= 0) return -1;
  fconst fc;
  fconst_init(&fc, g);
  const fprobe pr = {o, d, tmax, tri, t};
  return (int)filter_voxel(b, &fc, fp, vid, rec, &pr);
}

/* probe: one texture fetch (for tests of the texture restatement) */
void r2o_tex_fetch(const r2o_texture* t, float x, float y, float out[4]) { tex_fetch(t, x, y, out); }

/* probe: closest hit of one world-space ray (for tests of the triangle query) */
int r2o_filter_closest_hit(const r2o_filter_ctx* b, const float o[3], const float d[3], float tmax, float out[3]) {
  int tri;
  float b1, b2, t;
  if (!closest_hit(b, F3(o[0], o[1], o[2]), F3(d[0], d[1], d[2]), tmax, &tri, &b1, &b2, &t)) return -1;
  out[0] = t; out[1] = b1; out[2] = b2;
  return tri;
}
