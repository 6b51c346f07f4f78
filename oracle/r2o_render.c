/*
 * Oracle (test infrastructure only): CPU restatement of the volume integrator of
 * FrederikBoehm/Renderer2, compiled twice -- R2O_DOUBLE=0 gives the FP32 form
 * (the measured CPU baseline), R2O_DOUBLE=1 the FP64 form (the reference
 * image oracle).  Every function cites the reference file:line it restates.
 * The product library never links this file.
 *
 * Reference quirks kept on purpose (they change converged images):
 *   pixel jitter uses the horizontal sample on both axes (pixel_sampler.hpp:37-40);
 *   the camera ray's entry segment is tracked twice (path_integrator_impl.hpp:144-191);
 *   SGGX VNDF is sampled about the voxel normal (sggx_phase_function.hpp:116-121);
 *   the diffuse SGGX lobe is evaluated stochastically (sggx_phase_function.hpp:186-189);
 *   the env le() pdf is the unnormalised bilinear Distribution2D::pdf and the
 *   marginal is built from shifted row sums (distribution_2d.hpp:59-84);
 *   trDDA is effectively binary (nvdb_medium_impl.hpp:279);
 *   the DDA advances dt/sigma_t per step with mip hysteresis (nvdb_medium_impl.hpp:13-15,148-185).
 *
 * Deviations (stated):
 *   RNG: Philox4x32-10 keyed (seed, stream), counter (pixel, sample, block, 0);
 *     uniforms are (x >> 8) * 2^-24 in [0, 1).
 *   Constants that the reference evaluates in double through M_PI are evaluated
 *     in the working precision.
 *   Env texture bilinear weights use full precision (CUDA hardware uses 8-bit
 *     fractions, texture.cu:182-186).
 *   The hard-coded 50-bounce loop is `max_depth` (a render parameter), and every
 *   loop carries an iteration cap (never reached on valid input).
 *   Vector constructor arguments that draw random numbers are evaluated left to
 *   right (C++ leaves the order unspecified).
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "r2o.h"

#if R2O_DOUBLE
typedef double R;
#define SFX(n) n##_f64
#define RLOG log
#define REXP exp
#define RSQRT sqrt
#define RSIN sin
#define RCOS cos
#define RTAN tan
#define RACOS acos
#define RATAN2 atan2
#define RFLOOR floor
#define RCEIL ceil
#define RROUND round
#define RABS fabs
#define RNEXTUP(x) nextafter((x), INFINITY)
#define RNEXTDN(x) nextafter((x), -INFINITY)
typedef r2o_scene_f64 scene_t;
#else
typedef float R;
#define SFX(n) n##_f32
#define RLOG logf
#define REXP expf
#define RSQRT sqrtf
#define RSIN sinf
#define RCOS cosf
#define RTAN tanf
#define RACOS acosf
#define RATAN2 atan2f
#define RFLOOR floorf
#define RCEIL ceilf
#define RROUND roundf
#define RABS fabsf
#define RNEXTUP(x) nextafterf((x), INFINITY)
#define RNEXTDN(x) nextafterf((x), -INFINITY)
typedef r2o_scene_f32 scene_t;
#endif

#define C(x) ((R)(x))
#define PI_R C(3.14159265358979323846)

#define LOOP_CAP 65536
#define DEFAULT_TMAX C(1e12)
#define RAY_OFFSET C(1e-4)

/* ------------------------------------------------------------ vec / mat -- */
typedef struct { R x, y, z; } v3;
typedef struct { R c[4][3]; } m43; /* glm::mat4x3, column major */
typedef struct { R c[4][4]; } m44; /* glm::mat4 */
typedef struct { R c[3][3]; } m33; /* glm::mat3 */

static inline v3 V(R x, R y, R z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, R s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 sscale(R s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline v3 vdivs(v3 a, R s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline R vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline R vlen(v3 a) { return RSQRT(vdot(a, a)); }
static inline v3 vnorm(v3 a) { return vscale(a, C(1) / RSQRT(vdot(a, a))); }
static inline v3 vcross(v3 x, v3 y) {
  return V(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
/* glm::min / glm::max on scalars: (y < x) ? y : x and (x < y) ? y : x */
static inline R gmin(R x, R y) { return (y < x) ? y : x; }
static inline R gmax(R x, R y) { return (x < y) ? y : x; }
static inline R vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

static inline v3 m43_point(const m43* m, v3 v) {
  return V(m->c[0][0] * v.x + m->c[1][0] * v.y + m->c[2][0] * v.z + m->c[3][0] * C(1),
           m->c[0][1] * v.x + m->c[1][1] * v.y + m->c[2][1] * v.z + m->c[3][1] * C(1),
           m->c[0][2] * v.x + m->c[1][2] * v.y + m->c[2][2] * v.z + m->c[3][2] * C(1));
}
static inline v3 m43_dir(const m43* m, v3 v) {
  return V(m->c[0][0] * v.x + m->c[1][0] * v.y + m->c[2][0] * v.z + m->c[3][0] * C(0),
           m->c[0][1] * v.x + m->c[1][1] * v.y + m->c[2][1] * v.z + m->c[3][1] * C(0),
           m->c[0][2] * v.x + m->c[1][2] * v.y + m->c[2][2] * v.z + m->c[3][2] * C(0));
}
static inline v3 m33_mul(const m33* m, v3 v) {
  return V(m->c[0][0] * v.x + m->c[1][0] * v.y + m->c[2][0] * v.z,
           m->c[0][1] * v.x + m->c[1][1] * v.y + m->c[2][1] * v.z,
           m->c[0][2] * v.x + m->c[1][2] * v.y + m->c[2][2] * v.z);
}

static m44 m44_from43(const m43* a) {
  m44 r;
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 3; ++j) r.c[i][j] = a->c[i][j];
    r.c[i][3] = (i == 3) ? C(1) : C(0);
  }
  return r;
}
static m43 m43_from44(const m44* a) {
  m43 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) r.c[i][j] = a->c[i][j];
  return r;
}
/* glm mat4 * mat4: column i = sum_k a[k] * b[i][k], summed left to right */
static m44 m44_mul(const m44* a, const m44* b) {
  m44 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r.c[i][j] = a->c[0][j] * b->c[i][0] + a->c[1][j] * b->c[i][1] + a->c[2][j] * b->c[i][2] + a->c[3][j] * b->c[i][3];
  return r;
}
/* General 4x4 inverse through 2x2 cofactors (the scheme of glm 0.9.9.8
 * detail::compute_inverse<4,4>, used by glm::inverse at every call site). */
static m44 m44_inverse(const m44* mm) {
  const R(*m)[4] = mm->c;
  const R c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
  const R c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
  const R c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3];
  const R c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
  const R c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3];
  const R c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
  const R c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
  const R c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2];
  const R c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  const R c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
  const R c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3];
  const R c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
  const R c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
  const R c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2];
  const R c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  const R c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
  const R c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1];
  const R c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  const R f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
  const R f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
  const R v0[4] = {m[1][0], m[0][0], m[0][0], m[0][0]};
  const R v1[4] = {m[1][1], m[0][1], m[0][1], m[0][1]};
  const R v2[4] = {m[1][2], m[0][2], m[0][2], m[0][2]};
  const R v3v[4] = {m[1][3], m[0][3], m[0][3], m[0][3]};
  m44 inv;
  const R sa[4] = {C(1), C(-1), C(1), C(-1)}, sb[4] = {C(-1), C(1), C(-1), C(1)};
  for (int k = 0; k < 4; ++k) {
    inv.c[0][k] = (v1[k] * f0[k] - v2[k] * f1[k] + v3v[k] * f2[k]) * sa[k];
    inv.c[1][k] = (v0[k] * f0[k] - v2[k] * f3[k] + v3v[k] * f4[k]) * sb[k];
    inv.c[2][k] = (v0[k] * f1[k] - v1[k] * f3[k] + v3v[k] * f5[k]) * sa[k];
    inv.c[3][k] = (v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k]) * sb[k];
  }
  const R d0 = m[0][0] * inv.c[0][0], d1 = m[0][1] * inv.c[1][0], d2 = m[0][2] * inv.c[2][0], d3 = m[0][3] * inv.c[3][0];
  const R det = (d0 + d1) + (d2 + d3);
  const R one_over = C(1) / det;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) inv.c[i][j] = inv.c[i][j] * one_over;
  return inv;
}
static m44 m44_identity(void) {
  m44 r;
  memset(&r, 0, sizeof(r));
  for (int i = 0; i < 4; ++i) r.c[i][i] = C(1);
  return r;
}

/* 3x3 inverse by adjugate (glm detail::compute_inverse<3,3>), used by
 * CSGGXMicroflakeDistribution (sggx_phase_function.hpp:89-93). */
static m33 m33_inverse(const m33* mm) {
  const R(*m)[3] = mm->c;
  const R det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
  const R o = C(1) / det;
  m33 r;
  r.c[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * o;
  r.c[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * o;
  r.c[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * o;
  r.c[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * o;
  r.c[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * o;
  r.c[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * o;
  r.c[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * o;
  r.c[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * o;
  r.c[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * o;
  return r;
}
static R m33_det(const m33* mm) {
  const R(*m)[3] = mm->c;
  return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
         m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
         m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}

/* ------------------------------------------------------------------ RNG -- */
static inline R uni(r2o_rng* s) {
  const uint32_t x = r2o_rng_next_u32(s);
  return (R)(x >> 8) * C(5.9604644775390625e-08);
}

/* The random streams of one path.  Philox4x32-10 keyed (seed, stream) with
 * counter (pixel, sample, block, sub): every vertex of the path draws from
 * streams of its own, sub = 4 k + role, where vertex k = 0 is the camera and
 * k >= 1 the k-th scatter event:
 *   role 0: the vertex's own draws, in the reference's order (the camera
 *           jitter; at a scatter: direct()'s light sample, its phase
 *           evaluation and phase sample, Li's continuation sample, Russian
 *           roulette);
 *   role 1: the transmittance walk of direct()'s light-sample shadow ray;
 *   role 2: the transmittance walk of its phase-sample shadow ray;
 *   role 3: the free-flight walk of the ray that leaves the vertex (bounding
 *           box pass-throughs included) up to the next scatter event.
 * The reference draws all of these from one cuRAND sequence per pixel
 * (sampler.hpp:11-37, raytracer.cu:33-39).  cuRAND is not reproduced, so which
 * uniform feeds which draw is this restatement's own choice; the estimator is
 * the reference's.  Streams per ray let a GPU shade a scatter event once and
 * then trace its three rays one after another. */
typedef struct { uint32_t k0, k1, c0, c1; } path_key;
static inline void stream_init(r2o_rng* r, const path_key* pk, uint32_t k, uint32_t role) {
  r2o_rng_init(r, pk->k0, pk->k1, pk->c0, pk->c1, 4u * k + role);
}
/* intersectTr (device_scene_impl.hpp:93-114) loops until a miss, and Li's
 * bounce loop does not count box pass-throughs; both are capped here: 64
 * segments (iterations beyond 4 max_depth) plus 4 per instance beyond the
 * fourth.  A segment or pass-through enters or leaves an instance box, so a
 * walk that crosses every box once stays below the cap. */
static inline int inst_extra(int n_inst) { return n_inst > 4 ? 4 * (n_inst - 4) : 0; }
#define SHADOW_SEG_CAP(s) (64 + inst_extra((s)->n_inst))

/* ------------------------------------------------------------------ ray -- */
typedef struct { v3 o, d; R tmax; int medium; } ray_t;

static inline ray_t mkray(v3 o, v3 d, R tmax, int medium) {
  ray_t r; r.o = o; r.d = d; r.tmax = tmax; r.medium = medium; return r;
}

/* CRay::transform (ray.hpp:41-47) */
static ray_t ray_transform(const ray_t* r, const m43* M) {
  const v3 o = m43_point(M, r->o);
  const v3 e = m43_point(M, vadd(r->o, sscale(r->tmax, r->d)));
  const v3 d = vnorm(m43_dir(M, r->d));
  return mkray(o, d, vlen(vsub(e, o)), r->medium);
}
/* CRay::transform2 (ray.hpp:49-55): keeps the ray parameter */
static ray_t ray_transform2(const ray_t* r, const m43* M) {
  const v3 o = m43_point(M, r->o);
  const v3 e = m43_point(M, vadd(r->o, sscale(r->tmax, r->d)));
  const v3 d = m43_dir(M, r->d);
  return mkray(o, d, vlen(vsub(e, o)) / vlen(d), r->medium);
}
/* nextFloatUp / nextFloatDown (functions.hpp:94-126) */
static inline R next_up(R v) {
  if (isinf(v) && v > 0) return v;
  if (v == C(0)) return C(0); /* `v == -0.f` also holds for +0: zero is not bumped */
  return RNEXTUP(v);
}
static inline R next_dn(R v) {
  if (isinf(v) && v < 0) return v;
  if (v == C(0)) return -C(0);
  return RNEXTDN(v);
}
/* CRay::offsetRayOrigin (ray.hpp:71-84) */
static v3 offset_point(v3 p, v3 dir) {
  const R d = RABS(dir.x) * RAY_OFFSET + RABS(dir.y) * RAY_OFFSET + RABS(dir.z) * RAY_OFFSET;
  const v3 off = vscale(dir, d);
  v3 po = vadd(p, off);
  po.x = off.x > 0 ? next_up(po.x) : (off.x < 0 ? next_dn(po.x) : po.x);
  po.y = off.y > 0 ? next_up(po.y) : (off.y < 0 ? next_dn(po.y) : po.y);
  po.z = off.z > 0 ? next_up(po.z) : (off.z < 0 ? next_dn(po.z) : po.z);
  return po;
}
/* CRay::offsetRayOrigin(offsetDir) member (ray.hpp:86-91) */
static void ray_offset(ray_t* r, v3 dir) {
  const v3 end = vadd(r->o, sscale(r->tmax, r->d));
  r->o = offset_point(r->o, dir);
  r->tmax = vlen(vsub(end, r->o));
}
/* CRay::spawnRay (ray.hpp:62-69) */
static ray_t spawn_ray(v3 start, v3 end, int medium) {
  v3 dir = vsub(end, start);
  const v3 off = offset_point(start, vnorm(vsub(end, start)));
  const v3 nd = vsub(end, off);
  const R t = vlen(nd);
  dir = vdivs(dir, t);
  return mkray(off, dir, t, medium);
}

/* ------------------------------------------------------------- record --- */
typedef struct {
  R density;
  R S[6]; /* xx yy zz xy xz yz */
  v3 diffuse, specular, normal;
  R ior;
} rec_t;

/* SFilteredData(const SFilteredDataCompact&) (filtered_data.hpp:63-89) */
static rec_t decode_record(const uint8_t* p) {
  rec_t r;
  float dens;
  uint16_t sg[3], rr[3], nn[3], ior;
  memcpy(&dens, p + 0, 4);
  memcpy(sg, p + 4, 6);
  memcpy(rr, p + 10, 6);
  memcpy(nn, p + 22, 6);
  memcpy(&ior, p + 28, 2);
  const R inv16 = C(1) / C(65535);
  const R inv8 = C(1) / C(255);
  r.density = (R)dens;
  const R sx = (R)sg[0] * inv16, sy = (R)sg[1] * inv16, sz = (R)sg[2] * inv16;
  const R rxy = C(2) * ((R)rr[0] * inv16) - C(1);
  const R rxz = C(2) * ((R)rr[1] * inv16) - C(1);
  const R ryz = C(2) * ((R)rr[2] * inv16) - C(1);
  r.S[0] = sx * sx; r.S[1] = sy * sy; r.S[2] = sz * sz;
  r.S[3] = rxy * sx * sy; r.S[4] = rxz * sx * sz; r.S[5] = ryz * sy * sz;
  r.diffuse = V((R)p[16] * inv8, (R)p[17] * inv8, (R)p[18] * inv8);
  r.specular = V((R)p[19] * inv8, (R)p[20] * inv8, (R)p[21] * inv8);
  r.normal = vnorm(V(C(2) * (R)nn[0] * inv16 - C(1), C(2) * (R)nn[1] * inv16 - C(1), C(2) * (R)nn[2] * inv16 - C(1)));
  r.ior = (R)r2o_half_to_float(ior);
  return r;
}

/* operator*(float, SFilteredData) (filtered_data.hpp:115-132) */
static rec_t rec_scale(R v, const rec_t* a) {
  rec_t o;
  o.density = v * a->density;
  for (int i = 0; i < 6; ++i) o.S[i] = v * a->S[i];
  o.diffuse = sscale(v, a->diffuse);
  o.specular = sscale(v, a->specular);
  o.normal = sscale(v, vnorm(a->normal));
  o.ior = v * a->ior;
  return o;
}
/* operator+(SFilteredData, SFilteredData) (filtered_data.hpp:134-151) */
static rec_t rec_add(const rec_t* a, const rec_t* b) {
  rec_t o;
  o.density = a->density + b->density;
  for (int i = 0; i < 6; ++i) o.S[i] = a->S[i] + b->S[i];
  o.diffuse = vadd(a->diffuse, b->diffuse);
  o.specular = vadd(a->specular, b->specular);
  o.normal = vnorm(vadd(a->normal, b->normal));
  o.ior = a->ior + b->ior;
  return o;
}
/* interpolate (functions.hpp:66-72) */
static rec_t rec_lerp(R t, const rec_t* a, const rec_t* b) {
  const rec_t x = rec_scale(C(1) - t, a);
  const rec_t y = rec_scale(t, b);
  return rec_add(&x, &y);
}

/* ------------------------------------------------------------- medium --- */
typedef struct {
  int nx, ny, nz;
  const uint8_t* records;
  uint8_t* active;
  uint8_t background[32];
  r2o_bricks bricks;
  int ibb_min[3];
  int size[3];
  m43 index_to_model, model_to_index;
  m43 model_to_scaled, scaled_to_model;
  v3 bb_min, bb_max; /* model-space world bbox of the grid */
  v3 sigma_a, sigma_s;
  R sigma_t;
  R inv_max_density;
  R voxel_size_filtering;
} medium_t;

typedef struct {
  int medium;
  m43 model_to_world, world_to_model;
  v3 bb_min, bb_max; /* world AABB (SAABB::transform of the model bbox) */
  R ratio;           /* filterRenderRatio */
} instance_t;

/* CCircle / Sphere with the CMaterial of CHostSceneobject (sceneobject.cu:28-48) */
typedef struct {
  int shape;
  v3 pos, normal;
  R radius;
  m43 model_to_world, world_to_model; /* translate(pos) * getRotation(normal) (shape.hpp:49-57) */
  v3 diffuse, specular;
  R on_a, on_b;                       /* COrenNayarBRDF(roughness) (oren_nayar_brdf.hpp:31-35) */
  R alpha_x, alpha_y, eta_i, eta_t;   /* CMicrofacetBRDF */
} surface_t;

struct SFX(r2o_scene) {
  int n_media;
  medium_t media[R2O_MAX_GRIDS];
  int n_inst;
  instance_t* inst; /* n_inst */
  int n_surf;
  surface_t surf[R2O_MAX_SURFACES];
  /* env map */
  int env_w, env_h;
  R* tex;       /* env_h * env_w * 3 */
  R* row_func;  /* env_h * env_w */
  R* row_cdf;   /* env_h * (env_w + 1) */
  R* row_int;   /* env_h */
  R* marg_func; /* env_h */
  R* marg_cdf;  /* env_h + 1 */
  R marg_int;
};

/* NanoVDB-read of one voxel: active -> stored record, else background
 * (ReadAccessor::getValue at nvdb_medium_impl.hpp:72-78). */
static const uint8_t* voxel_ptr(const medium_t* m, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= m->nx || y >= m->ny || z >= m->nz) return m->background;
  const size_t i = ((size_t)z * m->ny + y) * m->nx + x;
  return m->active[i] ? m->records + 32 * i : m->background;
}
static rec_t get_value(const medium_t* m, int x, int y, int z, uint64_t* cnt) {
  if (cnt) cnt[R2O_CNT_VOX] += 1;
  return decode_record(voxel_ptr(m, x, y, z));
}

/* CNVDBMedium::filteredData (nvdb_medium_impl.hpp:47-64); p in normalised [0,1]^3 */
static rec_t filtered_data(const medium_t* m, v3 p, uint64_t* cnt) {
  const v3 ps = V(p.x * (R)m->size[0] + (R)m->ibb_min[0] - C(0.5), p.y * (R)m->size[1] + (R)m->ibb_min[1] - C(0.5),
                  p.z * (R)m->size[2] + (R)m->ibb_min[2] - C(0.5));
  const R fx = RFLOOR(ps.x), fy = RFLOOR(ps.y), fz = RFLOOR(ps.z);
  const int x = (int)fx, y = (int)fy, z = (int)fz;
  const v3 d = V(ps.x - (R)x, ps.y - (R)y, ps.z - (R)z);
  rec_t a, b;
  a = get_value(m, x, y, z, cnt); b = get_value(m, x + 1, y, z, cnt);
  const rec_t d00 = rec_lerp(d.x, &a, &b);
  a = get_value(m, x, y + 1, z, cnt); b = get_value(m, x + 1, y + 1, z, cnt);
  const rec_t d10 = rec_lerp(d.x, &a, &b);
  a = get_value(m, x, y, z + 1, cnt); b = get_value(m, x + 1, y, z + 1, cnt);
  const rec_t d01 = rec_lerp(d.x, &a, &b);
  a = get_value(m, x, y + 1, z + 1, cnt); b = get_value(m, x + 1, y + 1, z + 1, cnt);
  const rec_t d11 = rec_lerp(d.x, &a, &b);
  const rec_t d0 = rec_lerp(d.y, &d00, &d10);
  const rec_t d1 = rec_lerp(d.y, &d01, &d11);
  return rec_lerp(d.z, &d0, &d1);
}

/* CDeviceBrickGrid::lookupMajorant (device_grid_brick_impl.hpp:29-33) */
static R lookup_majorant(const medium_t* m, v3 ipos, int mip, uint64_t* cnt) {
  if (cnt) cnt[R2O_CNT_MAJ] += 1;
  const int ix = (int)RFLOOR(ipos.x), iy = (int)RFLOOR(ipos.y), iz = (int)RFLOOR(ipos.z);
  return (R)r2o_brick_majorant(&m->bricks, ix >> 3, iy >> 3, iz >> 3, mip);
}
/* CDeviceBrickGrid::lookupDensity (device_grid_brick_impl.hpp:9-27) */
static R lookup_density(const medium_t* m, v3 ipos, v3 u, uint64_t* cnt) {
  if (cnt) cnt[R2O_CNT_BRICK] += 1;
  const v3 p = V(ipos.x + u.x - C(0.5), ipos.y + u.y - C(0.5), ipos.z + u.z - C(0.5));
  return (R)r2o_brick_density(&m->bricks, (int)RFLOOR(p.x), (int)RFLOOR(p.y), (int)RFLOOR(p.z));
}

/* CNVDBMedium::stepDDA (nvdb_medium_impl.hpp:241-246) */
static R step_dda(v3 pos, v3 inv_dir, int mip) {
  const R dim = (R)(8 << mip);
  const R ox = inv_dir.x >= C(0) ? dim + C(0.5) : C(-0.5);
  const R oy = inv_dir.y >= C(0) ? dim + C(0.5) : C(-0.5);
  const R oz = inv_dir.z >= C(0) ? dim + C(0.5) : C(-0.5);
  const R inv_dim = C(1) / dim;
  const R tx = (RFLOOR(pos.x * inv_dim) * dim + ox - pos.x) * inv_dir.x;
  const R ty = (RFLOOR(pos.y * inv_dim) * dim + oy - pos.y) * inv_dir.y;
  const R tz = (RFLOOR(pos.z * inv_dim) * dim + oz - pos.z) * inv_dir.z;
  /* glm::min(a, glm::min(b, c)): min(x, y) = (y < x) ? y : x */
  const R myz = (tz < ty) ? tz : ty;
  return (myz < tx) ? myz : tx;
}

/* CFresnel::evaluate (fresnel.hpp:28-54) */
static R fresnel_ie(R cos_i, R eta_i_in, R eta_t_in) {
  cos_i = cos_i < C(-1) ? C(-1) : (cos_i > C(1) ? C(1) : cos_i);
  R eta_i = eta_i_in, eta_t = eta_t_in;
  if (!(cos_i > C(0))) {
    eta_i = eta_t_in; eta_t = eta_i_in;
    cos_i = RABS(cos_i);
  }
  const R sin_i = RSQRT(gmax(C(0), C(1) - cos_i * cos_i));
  const R sin_t = eta_i / eta_t * sin_i;
  if (sin_t >= C(1)) return C(1);
  const R cos_t = RSQRT(gmax(C(0), C(1) - sin_t * sin_t));
  const R rpar = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
  const R rper = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
  return (rpar * rpar + rper * rper) / C(2);
}
/* the SGGX phase's Fresnel: etaI = 1, etaT = voxel ior */
static R fresnel(R cos_i, R eta_t_in) { return fresnel_ie(cos_i, C(1), eta_t_in); }

/* Scattering event recorded by a medium sample (SInteraction of a medium hit). */
typedef struct {
  int valid;
  int inst;
  v3 pos, normal;
  R S[6];
  R ior;
} mhit_t;

/* Shared tail of sampleInternal / sampleDDA at a real collision
 * (nvdb_medium_impl.hpp:120-131, 169-180). `ray_model` is the collision ray in
 * model space (rayWorldNew). */
static v3 collision_albedo(const medium_t* m, const rec_t* fd, const ray_t* ray_model, mhit_t* mi) {
  const v3 pos = vadd(ray_model->o, sscale(ray_model->tmax, ray_model->d));
  const v3 n = vnorm(fd->normal);
  mi->valid = 1;
  mi->pos = pos;
  mi->normal = n;
  for (int i = 0; i < 6; ++i) mi->S[i] = fd->S[i];
  mi->ior = fd->ior;
  const R w = fresnel(RABS(vdot(vnorm(vneg(ray_model->d)), n)), fd->ior);
  /* (diffuse * (1 - w) + specular * w) * sigma_s / sigma_t */
  const v3 b = vadd(vscale(fd->diffuse, C(1) - w), vscale(fd->specular, w));
  return vdivs(vmul(b, m->sigma_s), m->sigma_t);
}

/* CNVDBMedium::sampleInternal (nvdb_medium_impl.hpp:94-137), Vec4d accessor */
static v3 sample_internal(const medium_t* m, ray_t* ray_model, r2o_rng* rng, R ratio, mhit_t* mi, uint64_t* cnt) {
  ray_t ray = ray_transform2(ray_model, &m->model_to_index);
  R t = C(0);
  const R inv_max = C(1) / (ratio / m->inv_max_density);
  for (int it = 0; it < LOOP_CAP; ++it) {
    if (cnt) cnt[R2O_CNT_FLIGHT] += 1;
    t -= RLOG(C(1) - uni(rng)) * inv_max / m->sigma_t;
    if (t >= ray.tmax) break;
    const rec_t fd = filtered_data(m, vadd(ray.o, sscale(t, ray.d)), cnt);
    if (ratio * fd.density * inv_max > uni(rng)) {
      ray.tmax = t;
      ray_t rw = ray_transform2(&ray, &m->index_to_model);
      ray_model->tmax = rw.tmax;
      return collision_albedo(m, &fd, &rw, mi);
    }
  }
  return V(C(1), C(1), C(1));
}

/* CNVDBMedium::sampleDDA (nvdb_medium_impl.hpp:139-187) */
static v3 sample_dda(const medium_t* m, ray_t* ray_model, r2o_rng* rng, R ratio, mhit_t* mi, uint64_t* cnt) {
  ray_t iray = ray_transform2(ray_model, &m->model_to_scaled);
  const v3 ri = V(C(1) / iray.d.x, C(1) / iray.d.y, C(1) / iray.d.z);
  R t = C(0);
  R tau = -RLOG(C(1) - uni(rng));
  R mip = C(3);
  for (int it = 0; it < LOOP_CAP && t < iray.tmax; ++it) {
    if (cnt) cnt[R2O_CNT_DDA] += 1;
    const v3 cur = vadd(iray.o, sscale(t, iray.d));
    const int imip = (int)RROUND(mip);
    const R maj = ratio * lookup_majorant(m, cur, imip, cnt);
    const R dt = step_dda(cur, ri, imip);
    const R prev = t;
    t += dt / m->sigma_t;
    if (t == prev) break;
    tau -= maj * dt;
    mip = gmin(mip + C(0.25), C(3));
    if (tau > C(0)) continue;
    t += tau / (maj * m->sigma_t);
    if (t >= iray.tmax) break;
    const R u0 = uni(rng), u1 = uni(rng), u2 = uni(rng);
    const R d = lookup_density(m, vadd(iray.o, sscale(t, iray.d)), V(u0, u1, u2), cnt);
    if (uni(rng) * maj < d * ratio) {
      iray.tmax = t;
      const v3 sp = vadd(iray.o, sscale(t, iray.d));
      const v3 pos = V((sp.x - (R)m->ibb_min[0]) / (R)m->size[0], (sp.y - (R)m->ibb_min[1]) / (R)m->size[1],
                       (sp.z - (R)m->ibb_min[2]) / (R)m->size[2]);
      const rec_t fd = filtered_data(m, pos, cnt);
      ray_t rw = ray_transform2(&iray, &m->scaled_to_model);
      ray_model->tmax = rw.tmax;
      return collision_albedo(m, &fd, &rw, mi);
    }
    tau = -RLOG(C(1) - uni(rng));
    mip = gmax(C(0), mip - C(2));
  }
  return V(C(1), C(1), C(1));
}

/* CNVDBMedium::trInternal (nvdb_medium_impl.hpp:203-239) */
static R tr_internal(const medium_t* m, const ray_t* ray_model, r2o_rng* rng, R ratio, uint64_t* cnt) {
  const ray_t ray = ray_transform2(ray_model, &m->model_to_index);
  R tr = C(1), t = C(0);
  const R inv_max = C(1) / (ratio / m->inv_max_density);
  for (int it = 0; it < LOOP_CAP; ++it) {
    if (cnt) cnt[R2O_CNT_FLIGHT] += 1;
    t -= RLOG(C(1) - uni(rng)) * inv_max / m->sigma_t;
    if (t >= ray.tmax) break;
    const R d = filtered_data(m, vadd(ray.o, sscale(t, ray.d)), cnt).density;
    tr *= C(1) - gmax(C(0), ratio * d * inv_max);
    if (tr < C(1)) {
      const R p = C(1) - tr;
      if (uni(rng) < p) return C(0);
      tr /= C(1) - p;
    }
  }
  return tr;
}

/* CNVDBMedium::trDDA (nvdb_medium_impl.hpp:248-294) */
static R tr_dda(const medium_t* m, const ray_t* ray_model, r2o_rng* rng, R ratio, uint64_t* cnt) {
  const ray_t iray = ray_transform2(ray_model, &m->model_to_scaled);
  const v3 ri = V(C(1) / iray.d.x, C(1) / iray.d.y, C(1) / iray.d.z);
  R tr = C(1), t = C(0);
  const R inv_max = C(1) / (ratio / m->inv_max_density);
  R tau = -RLOG(C(1) - uni(rng));
  R mip = C(3);
  const R vol_maj = C(1) / inv_max;
  for (int it = 0; it < LOOP_CAP && t < iray.tmax; ++it) {
    if (cnt) cnt[R2O_CNT_DDA] += 1;
    const v3 cur = vadd(iray.o, sscale(t, iray.d));
    const int imip = (int)RROUND(mip);
    const R maj = ratio * lookup_majorant(m, cur, imip, cnt);
    const R dt = step_dda(cur, ri, imip);
    const R prev = t;
    t += dt / m->sigma_t;
    if (t == prev) break;
    tau -= maj * dt;
    mip = gmin(mip + C(0.25), C(3));
    if (tau > C(0)) continue;
    t += tau / (maj * m->sigma_t);
    if (t >= iray.tmax) break;
    const R u0 = uni(rng), u1 = uni(rng), u2 = uni(rng);
    const R d = lookup_density(m, vadd(iray.o, sscale(t, iray.d)), V(u0, u1, u2), cnt);
    if (uni(rng) * maj < d * ratio) {
      tr *= gmax(C(0), C(1) - vol_maj / maj);
      if (tr < C(0.1)) {
        const R prob = C(1) - tr;
        if (uni(rng) < prob) return C(0);
        tr /= C(1) - prob;
      }
    }
    tau = -RLOG(C(1) - uni(rng));
    mip = gmax(C(0), mip - C(2));
  }
  return tr;
}

/* CNVDBMedium::moveToVoxelBorder (nvdb_medium_impl.hpp:296-320); ray in model space */
static ray_t medium_move_to_voxel_border(const medium_t* m, const ray_t* ray) {
  const ray_t ri = ray_transform(ray, &m->model_to_index);
  const R size[3] = {(R)m->size[0], (R)m->size[1], (R)m->size[2]};
  const R ibb[3] = {(R)m->ibb_min[0], (R)m->ibb_min[1], (R)m->ibb_min[2]};
  const R o[3] = {ri.o.x, ri.o.y, ri.o.z};
  const R d[3] = {ri.d.x, ri.d.y, ri.d.z};
  const R gamma = (R)((3.f * FLT_EPSILON) / (1.f - 3.f * FLT_EPSILON));
  R t1 = ri.tmax;
  for (int i = 0; i < 3; ++i) {
    const R ip = o[i] * size[i] + ibb[i];
    const R vmin = (RFLOOR(ip) - ibb[i]) / size[i];
    const R vmax = (RCEIL(ip) - ibb[i]) / size[i];
    const R inv = C(1) / d[i];
    R tn = (vmin - o[i]) * inv;
    R tf = (vmax - o[i]) * inv;
    if (tn > tf) { const R tmp = tn; tn = tf; tf = tmp; }
    tf *= C(1) + C(2) * gamma;
    t1 = tf < t1 ? tf : t1;
  }
  t1 *= (C(1) - RAY_OFFSET);
  const ray_t moved = mkray(vadd(ri.o, sscale(t1, ri.d)), ri.d, ri.tmax - t1, ri.medium);
  ray_t r = ray_transform(&moved, &m->index_to_model);
  ray_offset(&r, vneg(ray->d));
  return r;
}

/* ---------------------------------------------------------------- SGGX --- */
typedef struct { m33 S, invS; v3 n; } sggx_t;

static sggx_t sggx_make(const R S6[6], v3 n) {
  sggx_t g;
  g.S.c[0][0] = S6[0]; g.S.c[0][1] = S6[3]; g.S.c[0][2] = S6[4];
  g.S.c[1][0] = S6[3]; g.S.c[1][1] = S6[1]; g.S.c[1][2] = S6[5];
  g.S.c[2][0] = S6[4]; g.S.c[2][1] = S6[5]; g.S.c[2][2] = S6[2];
  g.invS = m33_inverse(&g.S);
  g.n = n;
  return g;
}
/* D (sggx_phase_function.hpp:107-110) */
static R sggx_D(const sggx_t* g, v3 w) {
  const R v = vdot(w, m33_mul(&g->invS, w));
  return C(1) / (PI_R * RSQRT(RABS(m33_det(&g->S))) * v * v);
}
/* sigma (sggx_phase_function.hpp:112-114) */
static R sggx_sigma(const sggx_t* g, v3 w) {
  return RSQRT(gmax(vdot(w, m33_mul(&g->S, w)), C(0)));
}
/* CCoordinateFrame::fromNormal2, Frisvad (objects.hpp:42-58) */
static void frame_frisvad(v3 N, v3* T, v3* B) {
  if (N.z < C(-0.9999999)) {
    *T = V(C(0), C(-1), C(0));
    *B = V(C(-1), C(0), C(0));
  } else {
    const R a = C(1) / (C(1) + N.z);
    const R b = -N.x * N.y * a;
    *T = vnorm(V(C(1) - N.x * N.x * a, b, -N.x));
    *B = vnorm(V(b, C(1) - N.y * N.y * a, -N.y));
  }
}
/* CCoordinateFrame::fromNormal (objects.hpp:28-40), tangent-to-world columns */
static void frame_basic(v3 N, v3* T, v3* B) {
  if (RABS(N.x) > RABS(N.y))
    *T = vdivs(V(-N.z, C(0), N.x), RSQRT(N.x * N.x + N.z * N.z));
  else
    *T = vdivs(V(C(0), N.z, -N.y), RSQRT(N.y * N.y + N.z * N.z));
  *B = vnorm(vcross(N, *T));
}
/* sampleVNDF (sggx_phase_function.hpp:116-145), about the voxel normal */
static v3 sggx_sample_vndf(const sggx_t* g, R ux, R uy) {
  v3 wj, wk;
  frame_frisvad(g->n, &wj, &wk); /* w_k = B, w_j = T */
  const v3 wi = g->n;
  const v3 Swk = m33_mul(&g->S, wk), Swj = m33_mul(&g->S, wj), Swi = m33_mul(&g->S, wi);
  const R Skk = vdot(wk, Swk), Skj = vdot(wk, Swj), Ski = vdot(wk, Swi);
  const R Sjj = vdot(wj, Swj), Sji = vdot(wj, Swi), Sii = vdot(wi, Swi);
  m33 P;
  P.c[0][0] = Skk; P.c[0][1] = Skj; P.c[0][2] = Ski;
  P.c[1][0] = Skj; P.c[1][1] = Sjj; P.c[1][2] = Sji;
  P.c[2][0] = Ski; P.c[2][1] = Sji; P.c[2][2] = Sii;
  const R tmp = RSQRT(Sjj * Sii - Sji * Sji);
  const v3 Mk = V(RSQRT(RABS(m33_det(&P))) / tmp, C(0), C(0));
  const R sc = C(1) / RSQRT(Sii);
  const v3 Mj = vscale(V(-(Ski * Sji - Skj * Sii) / tmp, tmp, C(0)), sc);
  const v3 Mi = vscale(V(Ski, Sji, Sii), sc);
  const R sux = RSQRT(ux);
  const R u = sux * RCOS(C(2) * PI_R * uy);
  const R v = sux * RSIN(C(2) * PI_R * uy);
  const R w = RSQRT(C(1) - u * u - v * v);
  v3 wm = vadd(vadd(sscale(u, Mk), sscale(v, Mj)), sscale(w, Mi));
  wm = vnorm(wm);
  /* mat3(w_k, w_j, n) * wm */
  const v3 r = V(wk.x * wm.x + wj.x * wm.y + wi.x * wm.z, wk.y * wm.x + wj.y * wm.y + wi.y * wm.z,
                 wk.z * wm.x + wj.z * wm.y + wi.z * wm.z);
  return vnorm(r);
}
/* specular p (sggx_phase_function.hpp:168-171) */
static R sggx_spec_p(const sggx_t* g, v3 wo, v3 wi) {
  const v3 wh = vnorm(vadd(wi, wo));
  return sggx_D(g, wh) / (C(4) * sggx_sigma(g, wo));
}
/* concentricSampleDisk + cosineSampleHemisphere (sampler.hpp:52-80) */
static v3 cosine_hemisphere(r2o_rng* rng) {
  const R r1 = uni(rng), r2 = uni(rng);
  const R ox = C(2) * r1 - C(1), oy = C(2) * r2 - C(1);
  v3 d;
  if (ox == C(0) && oy == C(0)) {
    d = V(C(0), C(0), C(0));
  } else {
    R theta, r;
    if (RABS(ox) > RABS(oy)) {
      r = ox;
      theta = (PI_R / C(4)) * (oy / ox);
    } else {
      r = oy;
      theta = PI_R / C(2) - (PI_R / C(4)) * (ox / oy);
    }
    d = sscale(r, V(RCOS(theta), RSIN(theta), C(0)));
  }
  const R z = RSQRT(gmax(C(0), C(1) - d.x * d.x - d.y * d.y));
  return V(d.x, d.y, z);
}
/* CSGGXPhaseFunction::p (sggx_phase_function.hpp:214-218) */
static R phase_p(const R S6[6], v3 n, R ior, v3 wo, v3 wi, r2o_rng* rng) {
  const sggx_t g = sggx_make(S6, n);
  const R w = fresnel(RABS(vdot(vnorm(wo), n)), ior);
  const R ux = uni(rng), uy = uni(rng);
  const v3 wm = sggx_sample_vndf(&g, ux, uy);
  const R pd = gmax(vdot(wi, wm), C(0)) / PI_R;
  return pd * (C(1) - w) + sggx_spec_p(&g, wo, wi) * w;
}
/* CSGGXPhaseFunction::sampleP (sggx_phase_function.hpp:220-229) */
static R phase_sample(const R S6[6], v3 n, R ior, v3 wo, v3* wi, r2o_rng* rng) {
  const sggx_t g = sggx_make(S6, n);
  const R pf = fresnel(RABS(vdot(vnorm(wo), n)), ior);
  if (uni(rng) < pf) {
    const R ux = uni(rng), uy = uni(rng);
    const v3 wm = sggx_sample_vndf(&g, ux, uy);
    /* glm::reflect(-wo, wm) = I - N * dot(N, I) * 2 */
    const v3 I = vneg(wo);
    *wi = vsub(I, vscale(vscale(wm, vdot(wm, I)), C(2)));
    return sggx_spec_p(&g, wo, *wi);
  }
  const R ux = uni(rng), uy = uni(rng);
  const v3 wn = sggx_sample_vndf(&g, ux, uy);
  v3 T, B;
  frame_basic(wn, &T, &B);
  const v3 wm = cosine_hemisphere(rng);
  const v3 ww = V(T.x * wm.x + B.x * wm.y + wn.x * wm.z, T.y * wm.x + B.y * wm.y + wn.y * wm.z,
                  T.z * wm.x + B.z * wm.y + wn.z * wm.z);
  *wi = vnorm(ww);
  return gmax(vdot(*wi, wn), C(0)) / PI_R;
}

/* ------------------------------------------------------------------ env --- */
/* findInterval (functions.hpp:10-27) with predicate cdf[i] <= u */
static int find_interval(const R* cdf, int size, R u) {
  int first = 0, len = size;
  while (len > 0) {
    const int half = len >> 1, mid = first + half;
    if (cdf[mid] <= u) { first = mid + 1; len -= half + 1; }
    else len = half;
  }
  int r = first - 1;
  return r < 0 ? 0 : (r > size - 2 ? size - 2 : r);
}
/* CDistribution1D::CDistribution1D (distribution_1d.cu:6-39) */
static R build_1d(const R* f, int n, R* cdf) {
  cdf[0] = C(0);
  for (int i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / (R)n;
  const R fi = cdf[n];
  if (fi == C(0)) {
    for (int i = 1; i <= n; ++i) cdf[i] = (R)i / (R)n;
  } else {
    for (int i = 1; i <= n; ++i) cdf[i] /= fi;
  }
  return fi;
}
/* CDistribution1D::sampleContinuous (distribution_1d.hpp:54-75) */
static R sample_1d(const R* f, const R* cdf, R fint, int n, R u, R* pdf, int* off) {
  const int o = find_interval(cdf, n + 1, u);
  if (off) *off = o;
  R du = u - cdf[o];
  if ((cdf[o + 1] - cdf[o]) > C(0)) du /= (cdf[o + 1] - cdf[o]);
  if (pdf) *pdf = f[o] / fint;
  return ((R)o + du) / (R)n;
}

/* Env texel fetch with bilinear filtering, wrap addressing and normalised
 * coordinates (texture.cu:166-189; CTexture::operator(), texture.hpp:56-59). */
static v3 env_tex(const scene_t* s, R x, R y) {
  const int W = s->env_w, H = s->env_h;
  x = x - RFLOOR(x);
  y = y - RFLOOR(y);
  const R tx = x * (R)W - C(0.5), ty = y * (R)H - C(0.5);
  const R fx = RFLOOR(tx), fy = RFLOOR(ty);
  const R a = tx - fx, b = ty - fy;
  int i0 = (int)fx, j0 = (int)fy;
  int i1 = i0 + 1, j1 = j0 + 1;
  i0 = ((i0 % W) + W) % W; i1 = ((i1 % W) + W) % W;
  j0 = ((j0 % H) + H) % H; j1 = ((j1 % H) + H) % H;
  const R* t = s->tex;
  v3 r;
  R* rr = &r.x;
  for (int c = 0; c < 3; ++c) {
    const R t00 = t[((size_t)j0 * W + i0) * 3 + c], t10 = t[((size_t)j0 * W + i1) * 3 + c];
    const R t01 = t[((size_t)j1 * W + i0) * 3 + c], t11 = t[((size_t)j1 * W + i1) * 3 + c];
    rr[c] = (C(1) - a) * (C(1) - b) * t00 + a * (C(1) - b) * t10 + (C(1) - a) * b * t01 + a * b * t11;
  }
  return r;
}

/* CEnvironmentMap::sample (environmentmap.hpp:28-42) with
 * CDistribution2D::sampleContinuous (distribution_2d.hpp:50-57) */
static v3 env_sample(const scene_t* s, r2o_rng* rng, v3* dir, R* pdf, uint64_t* cnt) {
  if (cnt) cnt[R2O_CNT_ENV] += 1;
  R p1, p0;
  int v;
  const R y = sample_1d(s->marg_func, s->marg_cdf, s->marg_int, s->env_h, uni(rng), &p1, &v);
  const R x = sample_1d(s->row_func + (size_t)v * s->env_w, s->row_cdf + (size_t)v * (s->env_w + 1), s->row_int[v],
                        s->env_w, uni(rng), &p0, NULL);
  *pdf = p0 * p1;
  const R phi = x * C(2) * PI_R;
  const R theta = y * PI_R;
  const R st = RSIN(theta);
  if (st == C(0)) *pdf = C(0);
  else *pdf /= (st * C(2) * PI_R * PI_R);
  *dir = V(RSIN(phi) * st, RCOS(theta), RCOS(phi) * st);
  return env_tex(s, x, y);
}

/* CDistribution2D::pdf (distribution_2d.hpp:73-84): unnormalised bilinear func */
static R dist2d_pdf(const scene_t* s, R px, R py) {
  const int W = s->env_w, H = s->env_h;
  int lr = (int)(py * (R)(H - 2));
  lr = lr < 0 ? 0 : (lr > H - 2 ? H - 2 : lr);
  const int ur = lr + 1;
  int lc = (int)(px * (R)(W - 2));
  lc = lc < 0 ? 0 : (lc > W - 2 ? W - 2 : lc);
  const int uc = lc + 1;
  const R* f = s->row_func;
  const R sx = px * ((R)W - C(2)), sy = py * ((R)H - C(2));
  const R lo = f[(size_t)lr * W + lc] * ((R)uc - sx) + f[(size_t)lr * W + uc] * (sx - (R)lc);
  const R up = f[(size_t)ur * W + lc] * ((R)uc - sx) + f[(size_t)ur * W + uc] * (sx - (R)lc);
  return lo * ((R)ur - sy) + up * (sy - (R)lr);
}

/* CEnvironmentMap::le (environmentmap.hpp:44-55) */
static v3 env_le(const scene_t* s, v3 d, R* pdf, uint64_t* cnt) {
  if (cnt) cnt[R2O_CNT_ENV] += 1;
  const R yc = d.y < C(-1) ? C(-1) : (d.y > C(1) ? C(1) : d.y);
  const R theta = RACOS(yc);
  const R p = RATAN2(d.x, d.z);
  const R phi = p < C(0) ? (p + C(2) * PI_R) : p;
  const R x = phi / (C(2) * PI_R);
  const R y = theta / PI_R;
  const R st = RSIN(theta);
  *pdf = st == C(0) ? C(0) : dist2d_pdf(s, x, y) / (C(2) * PI_R * PI_R * st);
  return env_tex(s, x, y);
}

/* ------------------------------------------------------ surface BSDFs --- */
/* The reference evaluates M_PI, M_1_PI and a few literals in double inside
 * float code; DMUL keeps those products in double and rounds once, as the
 * float build does. */
#define DPI 3.14159265358979323846
#define D1PI 0.318309886183790671538
static inline R dmul(double c, R x) { return (R)(c * (double)x); }

/* brdf_functions.hpp:7-55 and sameHemisphere (functions.hpp:58-60) */
static inline R cos2_theta(v3 w) { return w.z * w.z; }
static inline R sin2_theta(v3 w) { return gmax(C(0), C(1) - cos2_theta(w)); }
static inline R sin_theta(v3 w) { return RSQRT(sin2_theta(w)); }
static inline R tan_theta(v3 w) { return sin_theta(w) / w.z; }
static inline R tan2_theta(v3 w) { return sin2_theta(w) / cos2_theta(w); }
static inline R clamp11(R x) { return x < C(-1) ? C(-1) : (x > C(1) ? C(1) : x); }
static inline R cos_phi(v3 w) {
  const R st = sin_theta(w);
  return st == C(0) ? C(1) : clamp11(w.x / st);
}
static inline R sin_phi(v3 w) {
  const R st = sin_theta(w);
  return st == C(0) ? C(0) : clamp11(w.y / st);
}
static inline int same_hemisphere(v3 a, v3 b) { return a.z * b.z > C(0); }

/* COrenNayarBRDF::f (oren_nayar_brdf.hpp:42-67) */
static R on_f(const surface_t* m, v3 wo, v3 wi) {
  const R st_i = sin_theta(wi), st_o = sin_theta(wo);
  R max_cos = C(0);
  if ((double)st_i > 1.0e-4 && (double)st_o > 1.0e-4) {
    const R sp_i = sin_phi(wi), cp_i = cos_phi(wi), sp_o = sin_phi(wo), cp_o = cos_phi(wo);
    max_cos = gmax(C(0), cp_i * cp_o + sp_i * sp_o);
  }
  R sin_a, tan_b;
  if (RABS(wi.z) > RABS(wo.z)) {
    sin_a = st_o;
    tan_b = st_i / RABS(wi.z);
  } else {
    sin_a = st_i;
    tan_b = st_o / RABS(wo.z);
  }
  return dmul(D1PI, m->on_a + m->on_b * max_cos * sin_a * tan_b);
}
/* COrenNayarBRDF::pdf (oren_nayar_brdf.hpp:78-80), cosineHemispherePdf (sampler.hpp:82-84) */
static R on_pdf(v3 wo, v3 wi) { return same_hemisphere(wo, wi) ? dmul(D1PI, RABS(wi.z)) : C(0); }
/* COrenNayarBRDF::sampleF (oren_nayar_brdf.hpp:69-76) */
static R on_sample(const surface_t* m, v3 wo, v3* wi, r2o_rng* rng, R* pdf) {
  *wi = cosine_hemisphere(rng);
  if (wo.z < C(0)) wi->z = -wi->z;
  *pdf = on_pdf(wo, *wi);
  return on_f(m, wo, *wi);
}

/* CMicrofacetDistribution, Trowbridge-Reitz (microfacet_distribution.hpp:40-140) */
static R ggx_D(const surface_t* m, v3 h) {
  const R t2 = tan2_theta(h);
  if (isinf(t2)) return C(0);
  const R c2 = cos2_theta(h);
  const R c4 = c2 * c2;
  const R cp = cos_phi(h), sp = sin_phi(h);
  const R e = (cp * cp / (m->alpha_x * m->alpha_x) + sp * sp / (m->alpha_y * m->alpha_y)) * t2;
  const R e1 = C(1) + e;
  return (R)(1.0 / ((((DPI * (double)m->alpha_x) * (double)m->alpha_y) * (double)c4) * (double)e1 * (double)e1));
}
static R ggx_lambda(const surface_t* m, v3 w) {
  const R at = RABS(tan_theta(w));
  if (isinf(at)) return C(0);
  const R cp = cos_phi(w), sp = sin_phi(w);
  const R alpha = RSQRT(cp * cp * m->alpha_x * m->alpha_x + sp * sp * m->alpha_y * m->alpha_y);
  const R a2t2 = (alpha * at) * (alpha * at);
  return (C(-1) + RSQRT(C(1) + a2t2)) / C(2);
}
static R ggx_G(const surface_t* m, v3 wo, v3 wi) { return C(1) / (C(1) + ggx_lambda(m, wo) + ggx_lambda(m, wi)); }
static R ggx_G1(const surface_t* m, v3 w) { return C(1) / (C(1) + ggx_lambda(m, w)); }
/* sample11: visible-slope sampling of the unit-roughness distribution
 * (Heitz and d'Eon 2014, with pbrt-v3's rational fit of the y slope) */
static void ggx_sample11(R ct, R u1, R u2, R* sx, R* sy) {
  if ((double)ct > 0.9999) {
    const R r = RSQRT(u1 / (C(1) - u1));
    const R phi = dmul(6.28318530718, u2);
    *sx = r * RCOS(phi);
    *sy = r * RSIN(phi);
    return;
  }
  const R st = RSQRT(gmax(C(0), C(1) - ct * ct));
  const R tt = st / ct;
  const R a = C(1) / tt;
  const R g1 = C(2) / (C(1) + RSQRT(C(1) + C(1) / (a * a)));
  const R A = C(2) * u1 / g1 - C(1);
  R tmp = C(1) / (A * A - C(1));
  if ((double)tmp > 1e10) tmp = C(1e10);
  const R B = tt;
  const R D = RSQRT(gmax(B * B * tmp * tmp - (A * A - B * B) * tmp, C(0)));
  const R x1 = B * tmp - D, x2 = B * tmp + D;
  *sx = (A < C(0) || x2 > C(1) / tt) ? x1 : x2;
  R S;
  if (u2 > C(0.5)) {
    S = C(1);
    u2 = C(2) * (u2 - C(0.5));
  } else {
    S = C(-1);
    u2 = C(2) * (C(0.5) - u2);
  }
  const R z = (u2 * (u2 * (u2 * C(0.27385f) - C(0.73369f)) + C(0.46341f))) /
              (u2 * (u2 * (u2 * C(0.093073f) + C(0.309420f)) - C(1)) + C(0.597999f));
  *sy = S * z * RSQRT(C(1) + *sx * *sx);
}
static v3 ggx_sample_wi(const surface_t* m, v3 wi, R u1, R u2) {
  const v3 ws = vnorm(V(m->alpha_x * wi.x, m->alpha_y * wi.y, wi.z));
  R sx, sy;
  ggx_sample11(ws.z, u1, u2, &sx, &sy);
  const R cp = cos_phi(ws), sp = sin_phi(ws);
  const R tmp = cp * sx - sp * sy;
  sy = sp * sx + cp * sy;
  sx = tmp;
  sx = m->alpha_x * sx;
  sy = m->alpha_y * sy;
  return vnorm(V(-sx, -sy, C(1)));
}
static v3 ggx_sample_h(const surface_t* m, v3 wo, R u1, R u2) {
  const int flip = wo.z < C(0);
  v3 h = ggx_sample_wi(m, flip ? vneg(wo) : wo, u1, u2);
  if (flip) h = vneg(h);
  return h;
}
static R ggx_pdf(const surface_t* m, v3 wo, v3 wh) {
  return ggx_D(m, wh) * ggx_G1(m, wo) * RABS(vdot(wo, wh)) / RABS(wo.z);
}

/* CMicrofacetBRDF (microfacet_brdf.hpp:57-100) */
static R mf_f(const surface_t* m, v3 wo, v3 wi) {
  const R co = RABS(wo.z), ci = RABS(wi.z);
  v3 h = vadd(wi, wo);
  if (ci == C(0) || co == C(0)) return C(0);
  if (h.x == C(0) && h.y == C(0) && h.z == C(0)) return C(0);
  h = vnorm(h);
  /* faceforward(h, +z) */
  const v3 hf = h.z < C(0) ? vneg(h) : h;
  const R F = fresnel_ie(vdot(wi, hf), m->eta_i, m->eta_t);
  return ggx_D(m, h) * ggx_G(m, wo, wi) * F / (C(4) * ci * co);
}
/* returns f; leaves *wi and *pdf untouched on the early exits, as the reference does */
static R mf_sample(const surface_t* m, v3 wo, v3* wi, r2o_rng* rng, R* pdf) {
  if (wo.z == C(0)) return C(0);
  const R u1 = uni(rng), u2 = uni(rng);
  const v3 h = ggx_sample_h(m, wo, u1, u2);
  const R dh = vdot(wo, h);
  if (dh < C(0)) return C(0);
  /* glm::reflect(-wo, h) = I - N * dot(N, I) * 2 */
  const v3 I = vneg(wo);
  const R d = vdot(h, I);
  *wi = V(I.x - h.x * d * C(2), I.y - h.y * d * C(2), I.z - h.z * d * C(2));
  if (!same_hemisphere(wo, *wi)) return C(0);
  *pdf = ggx_pdf(m, wo, h) / (C(4) * dh);
  return mf_f(m, wo, *wi);
}
static R mf_pdf(const surface_t* m, v3 wo, v3 wi) {
  if (!same_hemisphere(wo, wi)) return C(0);
  const v3 h = vnorm(vadd(wo, wi));
  return ggx_pdf(m, wo, h) / (C(4) * vdot(wo, h));
}

/* CMaterial::f / sampleF / pdf (material.hpp:67-100): the two lobes mixed by
 * the Fresnel weight of |cos theta_o| */
static v3 mat_f(const surface_t* m, v3 wo, v3 wi) {
  const R fd = on_f(m, wo, wi), fg = mf_f(m, wo, wi);
  const R w = fresnel_ie(RABS(wo.z), m->eta_i, m->eta_t);
  return vadd(vscale(vscale(m->diffuse, fd), C(1) - w), vscale(vscale(m->specular, fg), w));
}
static v3 mat_sample(const surface_t* m, v3 wo, v3* wi, r2o_rng* rng, R* pdf) {
  const R p = fresnel_ie(RABS(wo.z), m->eta_i, m->eta_t);
  if (uni(rng) < p) return vscale(m->specular, mf_sample(m, wo, wi, rng, pdf));
  return vscale(m->diffuse, on_sample(m, wo, wi, rng, pdf));
}
static R mat_pdf(const surface_t* m, v3 wo, v3 wi) {
  const R w = fresnel_ie(RABS(wo.z), m->eta_i, m->eta_t);
  return on_pdf(wo, wi) * (C(1) - w) + mf_pdf(m, wo, wi) * w;
}

/* CCoordinateFrame::fromNormal (objects.hpp:28-40): tangent-to-world columns
 * (T, B, N) and its glm::inverse */
typedef struct { m33 t2w, w2t; } sframe_t;
static sframe_t surface_frame(v3 n) {
  v3 T, B;
  frame_basic(n, &T, &B);
  sframe_t f;
  f.t2w.c[0][0] = T.x; f.t2w.c[0][1] = T.y; f.t2w.c[0][2] = T.z;
  f.t2w.c[1][0] = B.x; f.t2w.c[1][1] = B.y; f.t2w.c[1][2] = B.z;
  f.t2w.c[2][0] = n.x; f.t2w.c[2][1] = n.y; f.t2w.c[2][2] = n.z;
  f.w2t = m33_inverse(&f.t2w);
  return f;
}

/* -------------------------------------------------------------- scene ---- */
typedef struct { int hit; R t; v3 pos; int medium; int surf; v3 normal; } hit_t;

/* CCircle::intersect (circle.hpp:88-113): the disk in the plane through pos
 * with normal m_normal; t in (0, tmax] */
static int circle_hit(const surface_t* sf, const ray_t* ray, R tmax, R* t_out, v3* n_out) {
  const R den = vdot(ray->d, sf->normal);
  if (den == C(0)) return 0;
  const R t = vdot(vsub(sf->pos, ray->o), sf->normal) / den;
  if (!(t > C(0) && t <= tmax)) return 0;
  const v3 ip = vadd(ray->o, sscale(t, ray->d));
  if (!(vlen(vsub(ip, sf->pos)) < sf->radius)) return 0;
  *t_out = t;
  *n_out = sf->normal;
  return 1;
}
/* Sphere::intersect (sphere.hpp:33-90): quadratic in model space (the ray
 * through CRay::transform, ray.hpp:41-47), t compared against the world ray's
 * tmax; the far root when the origin is inside */
static int sphere_hit(const surface_t* sf, const ray_t* ray, R tmax, R* t_out, v3* n_out) {
  const v3 om = m43_point(&sf->world_to_model, ray->o);
  const v3 dm = vnorm(m43_dir(&sf->world_to_model, ray->d));
  const R a = vdot(dm, dm);
  const R b = C(2) * vdot(dm, om);
  const R c = vdot(om, om) - sf->radius * sf->radius;
  const R disc = b * b - C(4) * a * c;
  R t;
  if (disc == C(0)) {
    t = (-b + RSQRT(disc)) / (C(2) * a);
    if (!(t > C(0) && t <= tmax)) return 0;
  } else if (disc > C(0)) {
    const R sq = RSQRT(disc);
    const R den = C(1) / (C(2) * a);
    const R t1 = (-b + sq) * den, t2 = (-b - sq) * den;
    const R mn = gmin(t1, t2), mx = gmax(t1, t2);
    if (!(mx > C(0))) return 0;
    if (mn > C(0) && mn <= tmax) t = mn;
    else if (mn <= C(0) && mx <= tmax) t = mx;
    else return 0;
  } else {
    return 0;
  }
  *t_out = t;
  *n_out = vnorm(m43_dir(&sf->model_to_world, vadd(om, sscale(t, dm))));
  return 1;
}

/* __intersection__volume (shaders.optix.cu:192-236) with nanovdb::Ray::intersects
 * (Williams et al. 2005 slab test, NanoVDB Ray.h) and BBox::isInside; the
 * closest accepted hit over all instances wins (OptiX closest-hit). */
static hit_t scene_intersect(const scene_t* s, ray_t* ray) {
  hit_t h;
  h.hit = 0; h.t = C(0); h.medium = -1; h.surf = -1; h.pos = V(0, 0, 0); h.normal = V(0, 0, 0);
  R best = ray->tmax;
  for (int k = 0; k < s->n_inst; ++k) {
    const instance_t* in = &s->inst[k];
    const R o[3] = {ray->o.x, ray->o.y, ray->o.z};
    const R inv[3] = {C(1) / ray->d.x, C(1) / ray->d.y, C(1) / ray->d.z};
    const R bb[2][3] = {{in->bb_min.x, in->bb_min.y, in->bb_min.z}, {in->bb_max.x, in->bb_max.y, in->bb_max.z}};
    const int sg[3] = {inv[0] < C(0), inv[1] < C(0), inv[2] < C(0)};
    const R span0 = C(0), span1 = DEFAULT_TMAX;
    R t0 = (bb[sg[0]][0] - o[0]) * inv[0];
    R t2 = (bb[1 - sg[1]][1] - o[1]) * inv[1];
    if (t0 > t2) continue;
    R t1 = (bb[1 - sg[0]][0] - o[0]) * inv[0];
    R t3 = (bb[sg[1]][1] - o[1]) * inv[1];
    if (t3 > t1) continue;
    if (t3 > t0) t0 = t3;
    if (t2 < t1) t1 = t2;
    t3 = (bb[sg[2]][2] - o[2]) * inv[2];
    if (t3 > t1) continue;
    t2 = (bb[1 - sg[2]][2] - o[2]) * inv[2];
    if (t0 > t2) continue;
    if (t3 > t0) t0 = t3;
    if (span1 < t0) continue;
    if (t2 < t1) t1 = t2;
    if (span0 > t1) continue;
    if (span0 > t0) t0 = span0;
    if (span1 < t1) t1 = span1;
    if (!(t1 < DEFAULT_TMAX)) continue;
    const int inside = o[0] > bb[0][0] && o[1] > bb[0][1] && o[2] > bb[0][2] && o[0] < bb[1][0] &&
                       o[1] < bb[1][1] && o[2] < bb[1][2];
    const R t = inside ? t1 : t0;
    if (!(t >= C(0) && t <= best)) continue;
    if (h.hit && !(t < h.t)) continue;
    h.hit = 1; h.t = t; h.medium = k;
  }
  /* __intersection__surface (shaders.optix.cu:170-189): the shape reports t
   * and its normal; the closest-hit program normalises the normal (:117) */
  for (int j = 0; j < s->n_surf; ++j) {
    const surface_t* sf = &s->surf[j];
    R t;
    v3 n;
    const R tmax = h.hit ? h.t : best;
    const int ok = sf->shape == R2O_SHAPE_CIRCLE ? circle_hit(sf, ray, tmax, &t, &n) : sphere_hit(sf, ray, tmax, &t, &n);
    if (!ok || (h.hit && !(t < h.t))) continue;
    h.hit = 1; h.t = t; h.medium = -1; h.surf = j; h.normal = n;
  }
  if (h.hit) {
    if (h.surf >= 0) h.normal = vnorm(h.normal);
    h.pos = V(ray->o.x + h.t * ray->d.x, ray->o.y + h.t * ray->d.y, ray->o.z + h.t * ray->d.z);
    ray->tmax = h.t; /* CDeviceScene::intersect mutates m_t_max (device_scene_impl.hpp:28-30) */
  }
  return h;
}

/* CMediumInstance::sample (medium_instance_impl.hpp:7-18) */
static v3 inst_sample(const scene_t* s, int k, ray_t* ray_world, r2o_rng* rng, int use_dda, mhit_t* mi,
                      uint64_t* cnt) {
  const instance_t* in = &s->inst[k];
  const medium_t* m = &s->media[in->medium];
  ray_t rm = ray_transform2(ray_world, &in->world_to_model);
  mhit_t h;
  h.valid = 0;
  const v3 alb = use_dda ? sample_dda(m, &rm, rng, in->ratio, &h, cnt) : sample_internal(m, &rm, rng, in->ratio, &h, cnt);
  const ray_t rwn = ray_transform2(&rm, &in->model_to_world);
  ray_world->tmax = rwn.tmax;
  mi->valid = h.valid;
  if (h.valid) {
    mi->inst = k;
    mi->pos = m43_point(&in->model_to_world, h.pos);
    mi->normal = vnorm(m43_dir(&in->model_to_world, h.normal));
    for (int i = 0; i < 6; ++i) mi->S[i] = h.S[i];
    mi->ior = h.ior;
    if (cnt) cnt[R2O_CNT_SCATTER] += 1;
  }
  return alb;
}
/* CMediumInstance::tr (medium_instance_impl.hpp:20-23) */
static R inst_tr(const scene_t* s, int k, const ray_t* ray_world, r2o_rng* rng, int use_dda, uint64_t* cnt) {
  const instance_t* in = &s->inst[k];
  const medium_t* m = &s->media[in->medium];
  const ray_t rm = ray_transform2(ray_world, &in->world_to_model);
  return use_dda ? tr_dda(m, &rm, rng, in->ratio, cnt) : tr_internal(m, &rm, rng, in->ratio, cnt);
}
/* CMediumInstance::moveToVoxelBorder (medium_instance_impl.hpp:38-41) */
static ray_t inst_move_to_voxel_border(const scene_t* s, int k, const ray_t* ray) {
  const instance_t* in = &s->inst[k];
  const ray_t rm = ray_transform(ray, &in->world_to_model);
  const ray_t moved = medium_move_to_voxel_border(&s->media[in->medium], &rm);
  ray_t r = ray_transform(&moved, &in->model_to_world);
  ray_offset(&r, vneg(ray->d));
  return r;
}

/* CDeviceScene::intersectTr (device_scene_impl.hpp:93-114); returns hit flag,
 * writes the accumulated transmittance. */
static int intersect_tr(const scene_t* s, const ray_t* ray, r2o_rng* rng, int use_dda, R* Tr, uint64_t* cnt) {
  *Tr = C(1);
  v3 p0 = ray->o;
  const v3 p1 = vadd(p0, sscale(ray->tmax, ray->d));
  int cur = ray->medium;
  for (int it = 0; it < SHADOW_SEG_CAP(s); ++it) {
    ray_t r = spawn_ray(p0, p1, cur);
    const hit_t h = scene_intersect(s, &r);
    /* a surface ends the walk; its (zero) emission zeroes the shadow term, so
     * the transmittance up to it is not tracked */
    if (h.hit && h.surf >= 0) return 1;
    if (h.hit && r.medium >= 0 && (h.medium < 0 || h.medium == cur)) *Tr *= inst_tr(s, r.medium, &r, rng, use_dda, cnt);
    if (!h.hit) return 0;
    p0 = h.pos;
    cur = r.medium < 0 ? h.medium : -1;
  }
  return 1;
}

/* direct (path_integrator_impl.hpp:23-118), medium interaction only, at scatter
 * vertex k: the vertex's own draws come from `rng`, the two shadow walks from
 * the path's role-1 and role-2 streams of vertex k */
static v3 direct(const scene_t* s, const mhit_t* mi, v3 wo, r2o_rng* rng, const path_key* pk, uint32_t k, int use_dda,
                 uint64_t* cnt) {
  v3 L = V(0, 0, 0);
  {
    v3 ldir;
    R light_pdf = C(0);
    const v3 Le = env_sample(s, rng, &ldir, &light_pdf, cnt);
    if (light_pdf > C(0)) {
      ldir = vnorm(ldir);
      ray_t rl = mkray(mi->pos, ldir, DEFAULT_TMAX, mi->inst);
      ray_offset(&rl, mi->normal);
      rl = inst_move_to_voxel_border(s, mi->inst, &rl);
      R trs;
      r2o_rng walk;
      stream_init(&walk, pk, k, 1);
      const int hit = intersect_tr(s, &rl, &walk, use_dda, &trs, cnt);
      const R p = phase_p(mi->S, mi->normal, mi->ior, wo, ldir, rng);
      const R mis = (C(1) * light_pdf) / (C(1) * light_pdf + C(1) * p);
      const R tr = hit ? C(0) : trs;
      L = vadd(L, vdivs(vscale(vscale(Le, mis * p), tr), light_pdf));
    }
  }
  {
    v3 wi;
    const R p = phase_sample(mi->S, mi->normal, mi->ior, wo, &wi, rng);
    if (p > C(0)) {
      wi = vnorm(wi);
      ray_t rb = mkray(mi->pos, wi, DEFAULT_TMAX, mi->inst);
      ray_offset(&rb, mi->normal);
      rb = inst_move_to_voxel_border(s, mi->inst, &rb);
      R trs;
      r2o_rng walk;
      stream_init(&walk, pk, k, 2);
      const int hit = intersect_tr(s, &rb, &walk, use_dda, &trs, cnt);
      R light_pdf;
      const v3 Le = env_le(s, wi, &light_pdf, cnt);
      const R mis = (C(1) * p) / (C(1) * p + C(1) * light_pdf);
      const R tr = hit ? C(0) : trs;
      L = vadd(L, vdivs(vscale(vscale(Le, mis * p), tr), p));
    }
  }
  return L;
}

/* direct (path_integrator_impl.hpp:23-118) at a surface vertex k: light
 * sample and BSDF sample in the tangent frame of the shading normal, the
 * shadow walks start in the medium the incoming ray travelled in */
static v3 direct_surface(const scene_t* s, const surface_t* sf, const hit_t* si, int cur_medium, v3 wo, r2o_rng* rng,
                         const path_key* pk, uint32_t k, int use_dda, uint64_t* cnt) {
  v3 L = V(0, 0, 0);
  const sframe_t fr = surface_frame(si->normal);
  const v3 wot = m33_mul(&fr.w2t, wo);
  {
    v3 ldir;
    R light_pdf = C(0);
    const v3 Le = env_sample(s, rng, &ldir, &light_pdf, cnt);
    if (light_pdf > C(0)) {
      ldir = vnorm(ldir);
      const v3 lt = vnorm(m33_mul(&fr.w2t, ldir));
      ray_t rl = mkray(si->pos, ldir, DEFAULT_TMAX, cur_medium);
      ray_offset(&rl, si->normal);
      R trs;
      r2o_rng walk;
      stream_init(&walk, pk, k, 1);
      const int hit = intersect_tr(s, &rl, &walk, use_dda, &trs, cnt);
      const v3 f = vscale(mat_f(sf, wot, lt), RABS(vdot(si->normal, rl.d)));
      const R sp = mat_pdf(sf, wot, lt);
      const R mis = (C(1) * light_pdf) / (C(1) * light_pdf + C(1) * sp);
      const R tr = hit ? C(0) : trs;
      L = vadd(L, vdivs(vscale(vmul(vscale(f, mis), Le), tr), light_pdf));
    }
  }
  {
    R sp = C(0);
    v3 wit = V(0, 0, 0);
    v3 f = mat_sample(sf, wot, &wit, rng, &sp);
    v3 wi = vnorm(m33_mul(&fr.t2w, wit));
    f = vscale(f, RABS(vdot(si->normal, wi)));
    if (sp > C(0)) {
      wi = vnorm(wi);
      ray_t rb = mkray(si->pos, wi, DEFAULT_TMAX, cur_medium);
      ray_offset(&rb, si->normal);
      R trs;
      r2o_rng walk;
      stream_init(&walk, pk, k, 2);
      const int hit = intersect_tr(s, &rb, &walk, use_dda, &trs, cnt);
      R light_pdf;
      const v3 Le = env_le(s, wi, &light_pdf, cnt);
      const R mis = (C(1) * sp) / (C(1) * sp + C(1) * light_pdf);
      const R tr = hit ? C(0) : trs;
      L = vadd(L, vdivs(vscale(vmul(vscale(f, mis), Le), tr), sp));
    }
  }
  return L;
}

/* --------------------------------------------------------------- camera -- */
typedef struct { m43 view_to_world; R pixel_size, near; int W, H; } camera_t;

/* CCamera::CCamera (camera.cu:20-32): glm::lookAt then glm::inverse */
static camera_t make_camera(const r2o_render_params* p) {
  camera_t c;
  c.W = p->width; c.H = p->height;
  c.pixel_size = C(1e-5);
  const R fov = (R)p->fov_deg * C(0.01745329251994329576923690768489);
  c.near = ((R)p->width * c.pixel_size) / (C(2) * RTAN(fov / C(2)));
  const v3 eye = V(p->cam_pos[0], p->cam_pos[1], p->cam_pos[2]);
  const v3 ctr = V(p->cam_look_at[0], p->cam_look_at[1], p->cam_look_at[2]);
  const v3 up = V(p->cam_up[0], p->cam_up[1], p->cam_up[2]);
  const v3 f = vnorm(vsub(ctr, eye));
  const v3 sv = vnorm(vcross(f, up));
  const v3 u = vcross(sv, f);
  m44 M = m44_identity();
  M.c[0][0] = sv.x; M.c[1][0] = sv.y; M.c[2][0] = sv.z;
  M.c[0][1] = u.x; M.c[1][1] = u.y; M.c[2][1] = u.z;
  M.c[0][2] = -f.x; M.c[1][2] = -f.y; M.c[2][2] = -f.z;
  M.c[3][0] = -vdot(sv, eye); M.c[3][1] = -vdot(u, eye); M.c[3][2] = vdot(f, eye);
  /* m_worldToView is stored as mat4x3, then inverted as mat4 */
  const m43 w2v = m43_from44(&M);
  const m44 w2v4 = m44_from43(&w2v);
  const m44 inv = m44_inverse(&w2v4);
  c.view_to_world = m43_from44(&inv);
  return c;
}
/* CPixelSampler::samplePixel (pixel_sampler.hpp:36-47) */
static ray_t camera_ray(const camera_t* c, int px, int py, r2o_rng* rng) {
  const R rh = uni(rng);
  const R h = ((R)(px - c->W / 2) + rh) * c->pixel_size;
  (void)uni(rng); /* randomVertical is drawn but unused */
  const R v = ((R)(py - c->H / 2) + rh) * c->pixel_size;
  const v3 d = vnorm(V(h, v, -c->near));
  const ray_t vr = mkray(V(0, 0, 0), d, DEFAULT_TMAX, -1);
  return ray_transform(&vr, &c->view_to_world);
}

/* ---------------------------------------------------------- integrator --- */
static int any_zero(v3 a) { return a.x == C(0) || a.y == C(0) || a.z == C(0); }

/* CPathIntegrator::Li (path_integrator_impl.hpp:120-247): media and analytic surfaces */
static v3 Li(const scene_t* s, const camera_t* cam, int px, int py, int spp, int max_depth, int use_dda,
             const path_key* pk, uint64_t* cnt) {
  v3 L = V(0, 0, 0), thr = V(C(1), C(1), C(1));
  r2o_rng vtx, walk; /* the current vertex's own stream, the walk leaving it */
  uint32_t k = 0;
  stream_init(&vtx, pk, 0, 0);
  ray_t ray = camera_ray(cam, px, py, &vtx);
  stream_init(&walk, pk, 0, 3);
  r2o_rng* rng = &walk;
  int iters = 0;
  for (int b = 0; b < max_depth; ++b) {
    if (++iters > 4 * max_depth + 64 + inst_extra(s->n_inst)) break;
    const hit_t si = scene_intersect(s, &ray);
    mhit_t mi;
    mi.valid = 0;
    if (si.medium >= 0) {
      if (ray.medium >= 0) {
        if (si.medium != ray.medium) break;
        thr = vmul(thr, inst_sample(s, ray.medium, &ray, rng, use_dda, &mi, cnt));
      } else {
        ray_t mr = mkray(si.pos, vnorm(ray.d), DEFAULT_TMAX, si.medium);
        ray_offset(&mr, ray.d);
        const hit_t se = scene_intersect(s, &mr);
        if (se.hit) {
          if (se.medium >= 0 && si.medium != se.medium) break;
          thr = vmul(thr, inst_sample(s, mr.medium, &mr, rng, use_dda, &mi, cnt));
        }
      }
      if (any_zero(thr)) break;
    }
    if (mi.valid) {
      ++k;
      stream_init(&vtx, pk, k, 0);
      const v3 wo = vneg(ray.d);
      L = vadd(L, vmul(thr, direct(s, &mi, wo, &vtx, pk, k, use_dda, cnt)));
      v3 wi;
      (void)phase_sample(mi.S, mi.normal, mi.ior, wo, &wi, &vtx);
      ray_t nr = mkray(mi.pos, vnorm(wi), DEFAULT_TMAX, mi.inst);
      ray_offset(&nr, wi);
      ray = inst_move_to_voxel_border(s, mi.inst, &nr);
      stream_init(&walk, pk, k, 3);
    } else {
      if (b == 0 && !si.hit) {
        R p;
        L = vadd(L, vmul(env_le(s, ray.d, &p, cnt), thr));
      }
      if (!si.hit) break;
      if (si.surf < 0) {
        ray_t nr = mkray(si.pos, vnorm(ray.d), DEFAULT_TMAX, ray.medium < 0 ? si.medium : -1);
        ray_offset(&nr, ray.d);
        ray = nr;
        --b;
        continue;
      }
      /* surface vertex (path_integrator_impl.hpp:194-223) */
      const surface_t* sf = &s->surf[si.surf];
      ++k;
      stream_init(&vtx, pk, k, 0);
      L = vadd(L, vmul(direct_surface(s, sf, &si, ray.medium, vneg(ray.d), &vtx, pk, k, use_dda, cnt), thr));
      const sframe_t fr = surface_frame(si.normal);
      const v3 wt = vnorm(m33_mul(&fr.w2t, ray.d)); /* ray.transform(worldToTangent) */
      v3 wi = V(0, 0, 0);
      R pdf = C(0);
      const v3 f = mat_sample(sf, vneg(wt), &wi, &vtx, &pdf);
      if (pdf == C(0)) break;
      ray_t nr = mkray(si.pos, vnorm(m33_mul(&fr.t2w, wi)), DEFAULT_TMAX, ray.medium);
      ray_offset(&nr, si.normal);
      const R cosine = RABS(vdot(si.normal, nr.d));
      thr = vmul(thr, vdivs(vscale(f, cosine), pdf));
      ray = nr;
      stream_init(&walk, pk, k, 3);
    }
    if (b > 3) {
      const R p = C(1) - (thr.x + thr.y + thr.z) / C(3);
      if (uni(&vtx) <= p) break;
      thr = vdivs(thr, C(1) - p);
    }
  }
  if (isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z)) return V(0, 0, 0);
  return vdivs(L, (R)spp);
}

/* ------------------------------------------------------ scene building --- */
static m43 m43_from_cols(v3 c0, v3 c1, v3 c2, v3 c3) {
  m43 m;
  m.c[0][0] = c0.x; m.c[0][1] = c0.y; m.c[0][2] = c0.z;
  m.c[1][0] = c1.x; m.c[1][1] = c1.y; m.c[1][2] = c1.z;
  m.c[2][0] = c2.x; m.c[2][1] = c2.y; m.c[2][2] = c2.z;
  m.c[3][0] = c3.x; m.c[3][1] = c3.y; m.c[3][2] = c3.z;
  return m;
}

/* getRotation (functions.hpp:128-142) */
static m44 get_rotation(v3 n) {
  const R cs = vdot(n, V(0, 1, 0));
  if (cs == C(1)) return m44_identity();
  if (cs == C(-1)) {
    m44 r = m44_identity();
    for (int i = 0; i < 3; ++i) r.c[i][i] = C(-1);
    return r;
  }
  const R angle = RACOS(vdot(n, V(0, 1, 0)));
  const v3 ax = vnorm(vcross(n, V(0, 1, 0)));
  /* glm::rotate(mat4(1), angle, axis) */
  const R c = RCOS(angle), sn = RSIN(angle);
  const v3 tmp = vscale(ax, C(1) - c);
  m44 r = m44_identity();
  r.c[0][0] = c + tmp.x * ax.x;
  r.c[0][1] = tmp.x * ax.y + sn * ax.z;
  r.c[0][2] = tmp.x * ax.z - sn * ax.y;
  r.c[1][0] = tmp.y * ax.x - sn * ax.z;
  r.c[1][1] = c + tmp.y * ax.y;
  r.c[1][2] = tmp.y * ax.z + sn * ax.x;
  r.c[2][0] = tmp.z * ax.x + sn * ax.y;
  r.c[2][1] = tmp.z * ax.y - sn * ax.x;
  r.c[2][2] = c + tmp.z * ax.z;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.c[i][j] = RROUND(r.c[i][j]);
  return r;
}

static void medium_init(medium_t* m, const r2o_grid_desc* g) {
  memset(m, 0, sizeof(*m));
  m->nx = g->nx; m->ny = g->ny; m->nz = g->nz;
  m->records = g->records;
  const size_t nv = (size_t)g->nx * g->ny * g->nz;
  float* dens = (float*)malloc(nv * sizeof(float));
  m->active = (uint8_t*)malloc(nv);
  for (size_t i = 0; i < nv; ++i) memcpy(&dens[i], g->records + 32 * i, 4);
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  float maxd = 0.f;
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        const size_t i = ((size_t)z * g->ny + y) * g->nx + x;
        m->active[i] = (uint8_t)r2o_voxel_active(dens, g->nx, g->ny, g->nz, x, y, z);
        if (m->active[i]) {
          lo[0] = x < lo[0] ? x : lo[0]; lo[1] = y < lo[1] ? y : lo[1]; lo[2] = z < lo[2] ? z : lo[2];
          hi[0] = x > hi[0] ? x : hi[0]; hi[1] = y > hi[1] ? y : hi[1]; hi[2] = z > hi[2] ? z : hi[2];
        }
      }
  /* getMaxValue (nvdb_medium.cu:308-323): max over the active index bbox, reading through the accessor */
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const size_t i = ((size_t)z * g->ny + y) * g->nx + x;
        const float d = m->active[i] ? dens[i] : 0.f;
        maxd = fmaxf(d, maxd);
      }
  r2o_brick_build(dens, g->nx, g->ny, g->nz, &m->bricks);
  free(dens);
  /* background record (openvdb_backend.cpp:30-43): all zero except ior = 1 */
  memset(m->background, 0, 32);
  const uint16_t one = 0x3C00u;
  memcpy(m->background + 28, &one, 2);

  for (int a = 0; a < 3; ++a) {
    m->ibb_min[a] = lo[a];
    m->size[a] = hi[a] - lo[a] + 1;
  }
  /* getIndexToModelTransformation (nvdb_medium.cu:325-335) */
  m44 nano = m44_identity();
  const float* mf = g->map_mat;
  nano.c[0][0] = mf[0]; nano.c[0][1] = mf[3]; nano.c[0][2] = mf[6];
  nano.c[1][0] = mf[1]; nano.c[1][1] = mf[4]; nano.c[1][2] = mf[7];
  nano.c[2][0] = mf[2]; nano.c[2][1] = mf[5]; nano.c[2][2] = mf[8];
  nano.c[3][0] = g->map_vec[0]; nano.c[3][1] = g->map_vec[1]; nano.c[3][2] = g->map_vec[2];
  m44 i2n = m44_identity();
  i2n.c[0][0] = (R)(float)m->size[0]; i2n.c[1][1] = (R)(float)m->size[1]; i2n.c[2][2] = (R)(float)m->size[2];
  i2n.c[3][0] = (R)m->ibb_min[0]; i2n.c[3][1] = (R)m->ibb_min[1]; i2n.c[3][2] = (R)m->ibb_min[2];
  const m44 i2m = m44_mul(&nano, &i2n);
  m->index_to_model = m43_from44(&i2m);
  const m44 i2m4 = m44_from43(&m->index_to_model);
  const m44 m2i = m44_inverse(&i2m4);
  m->model_to_index = m43_from44(&m2i);
  const m43 i2s = m43_from_cols(V((R)m->size[0], 0, 0), V(0, (R)m->size[1], 0), V(0, 0, (R)m->size[2]),
                                V((R)m->ibb_min[0], (R)m->ibb_min[1], (R)m->ibb_min[2]));
  const m44 i2s4 = m44_from43(&i2s);
  const m44 s2i4 = m44_inverse(&i2s4);
  const m43 s2i = m43_from44(&s2i4);
  const m44 m2i4 = m44_from43(&m->model_to_index);
  const m44 m2s = m44_mul(&i2s4, &m2i4);
  m->model_to_scaled = m43_from44(&m2s);
  const m44 s2i44 = m44_from43(&s2i);
  const m44 s2m = m44_mul(&i2m4, &s2i44);
  m->scaled_to_model = m43_from44(&s2m);
  /* world bbox: the index bbox [min, max + 1] through the map */
  const v3 c0 = m43_point(&m->index_to_model, V(0, 0, 0));
  const v3 c1 = m43_point(&m->index_to_model, V(C(1), C(1), C(1)));
  m->bb_min = V(fmin(c0.x, c1.x), fmin(c0.y, c1.y), fmin(c0.z, c1.z));
  m->bb_max = V(fmax(c0.x, c1.x), fmax(c0.y, c1.y), fmax(c0.z, c1.z));
  m->sigma_a = V(g->sigma_a[0], g->sigma_a[1], g->sigma_a[2]);
  m->sigma_s = V(g->sigma_s[0], g->sigma_s[1], g->sigma_s[2]);
  m->sigma_t = gmax(m->sigma_a.x + m->sigma_s.x, gmax(m->sigma_a.y + m->sigma_s.y, m->sigma_a.z + m->sigma_s.z));
  m->inv_max_density = (R)(1.f / maxd);
  m->voxel_size_filtering = g->voxel_size_filtering;
}

/* CShape(worldPos, normal) (shape.hpp:49-57) and CHostSceneobject's material
 * (sceneobject.cu:28-39) */
static void surface_init(surface_t* sf, const r2o_surface_desc* sd) {
  sf->shape = sd->shape;
  sf->pos = V(sd->pos[0], sd->pos[1], sd->pos[2]);
  sf->normal = V(sd->normal[0], sd->normal[1], sd->normal[2]);
  sf->radius = sd->radius;
  m44 T = m44_identity();
  T.c[3][0] = sd->pos[0]; T.c[3][1] = sd->pos[1]; T.c[3][2] = sd->pos[2];
  const m44 Rm = get_rotation(sf->normal);
  const m44 M = m44_mul(&T, &Rm);
  sf->model_to_world = m43_from44(&M);
  const m44 M4 = m44_from43(&sf->model_to_world);
  const m44 Mi = m44_inverse(&M4);
  sf->world_to_model = m43_from44(&Mi);
  sf->diffuse = V(sd->diffuse[0], sd->diffuse[1], sd->diffuse[2]);
  sf->specular = V(sd->specular[0], sd->specular[1], sd->specular[2]);
  const R sig2 = (R)sd->diffuse_roughness * (R)sd->diffuse_roughness;
  sf->on_a = C(1) - (sig2 / (C(2) * (sig2 + C(0.33f))));
  sf->on_b = C(0.45f) * sig2 / (sig2 + C(0.09f));
  sf->alpha_x = sd->alpha_x; sf->alpha_y = sd->alpha_y;
  sf->eta_i = sd->eta_i; sf->eta_t = sd->eta_t;
}

scene_t* SFX(r2o_scene_new)(const r2o_scene_desc* d) {
  scene_t* s = (scene_t*)calloc(1, sizeof(scene_t));
  s->n_media = d->n_grids;
  for (int i = 0; i < d->n_grids; ++i) medium_init(&s->media[i], &d->grids[i]);
  s->n_inst = d->n_instances;
  s->inst = (instance_t*)calloc(d->n_instances > 0 ? (size_t)d->n_instances : 1, sizeof(instance_t));
  for (int k = 0; k < d->n_instances; ++k) {
    const r2o_instance_desc* id = &d->instances[k];
    instance_t* in = &s->inst[k];
    in->medium = id->grid;
    const medium_t* m = &s->media[id->grid];
    /* getModelToWorldTransform (sceneobject.cu:375-377): translate * getRotation * scale */
    m44 T = m44_identity();
    T.c[3][0] = id->pos[0]; T.c[3][1] = id->pos[1]; T.c[3][2] = id->pos[2];
    const m44 Rm = get_rotation(V(id->orientation[0], id->orientation[1], id->orientation[2]));
    m44 Sm = m44_identity();
    Sm.c[0][0] = id->scaling[0]; Sm.c[1][1] = id->scaling[1]; Sm.c[2][2] = id->scaling[2];
    const m44 TR = m44_mul(&T, &Rm);
    const m44 M = m44_mul(&TR, &Sm);
    in->model_to_world = m43_from44(&M);
    const m44 M4 = m44_from43(&in->model_to_world);
    const m44 Mi = m44_inverse(&M4);
    in->world_to_model = m43_from44(&Mi);
    /* SAABB::transform (aabb.hpp:9-13) */
    const v3 a = m43_point(&in->model_to_world, m->bb_min), b = m43_point(&in->model_to_world, m->bb_max);
    in->bb_min = V(fmin(a.x, b.x), fmin(a.y, b.y), fmin(a.z, b.z));
    in->bb_max = V(fmax(a.x, b.x), fmax(a.y, b.y), fmax(a.z, b.z));
    /* getFilterRenderRatio (sceneobject.cu:108-114) */
    const R vsr = (in->bb_max.x - in->bb_min.x) / (R)m->size[0];
    in->ratio = m->voxel_size_filtering / vsr;
  }
  s->n_surf = d->n_surfaces;
  for (int j = 0; j < d->n_surfaces; ++j) surface_init(&s->surf[j], &d->surfaces[j]);
  /* env map + CDistribution2D (environmentmap.cu:11-29, distribution_2d.cu:7-14) */
  const int W = d->env_w, H = d->env_h;
  s->env_w = W; s->env_h = H;
  s->tex = (R*)malloc(sizeof(R) * 3 * (size_t)W * H);
  s->row_func = (R*)malloc(sizeof(R) * (size_t)W * H);
  s->row_cdf = (R*)malloc(sizeof(R) * (size_t)(W + 1) * H);
  s->row_int = (R*)malloc(sizeof(R) * H);
  s->marg_func = (R*)malloc(sizeof(R) * H);
  s->marg_cdf = (R*)malloc(sizeof(R) * (H + 1));
  for (size_t i = 0; i < (size_t)W * H * 3; ++i) s->tex[i] = d->env_rgb[i];
  for (int j = 0; j < H; ++j) {
    const double st = sin(3.14159265358979323846 * (double)((float)j + 0.5f) / (double)(float)H);
    for (int i = 0; i < W; ++i) {
      const size_t p = (size_t)j * W + i;
      const R rad = s->tex[3 * p] + s->tex[3 * p + 1] + s->tex[3 * p + 2];
      s->row_func[p] = (R)((double)rad * st);
    }
  }
  /* sumRows quirk (distribution_2d.hpp:59-71): push at i % W == 0 */
  {
    R sum = C(0);
    int k = 0;
    for (size_t i = 0; i < (size_t)W * H; ++i) {
      sum += s->row_func[i];
      if (i % (size_t)W == 0) { s->marg_func[k++] = sum; sum = C(0); }
    }
  }
  s->marg_int = build_1d(s->marg_func, H, s->marg_cdf);
  for (int j = 0; j < H; ++j) s->row_int[j] = build_1d(s->row_func + (size_t)j * W, W, s->row_cdf + (size_t)j * (W + 1));
  return s;
}

void SFX(r2o_scene_free)(scene_t* s) {
  if (!s) return;
  for (int i = 0; i < s->n_media; ++i) {
    free(s->media[i].active);
    r2o_brick_free(&s->media[i].bricks);
  }
  free(s->tex); free(s->row_func); free(s->row_cdf); free(s->row_int); free(s->marg_func); free(s->marg_cdf);
  free(s->inst);
  free(s);
}

/* ------------------------------------------------------------- render ---- */
/* Raytracer::renderFrame (raytracer.cu:161-190) + __raygen__rg (shaders.optix.cu:240-256):
 * data[px] += L / spp for sample = 0..spp-1.  `out` holds W*H*3 values (bottom row first). */
void SFX(r2o_render)(const scene_t* s, const r2o_render_params* p, R* out, uint64_t* counters) {
  const camera_t cam = make_camera(p);
  const int W = p->width, H = p->height;
  for (size_t i = 0; i < (size_t)W * H * 3; ++i) out[i] = C(0);
  uint64_t tot[R2O_NCNT] = {0};
#pragma omp parallel num_threads(p->threads > 0 ? p->threads : omp_get_max_threads())
  {
    uint64_t loc[R2O_NCNT] = {0};
#pragma omp for schedule(dynamic, 1)
    for (int y = 0; y < H; ++y) {
      for (int x = 0; x < W; ++x) {
        R acc[3] = {C(0), C(0), C(0)};
        for (int sidx = 0; sidx < p->spp; ++sidx) {
          const path_key pk = {p->seed, p->stream, (uint32_t)(y * W + x), (uint32_t)(p->sample_begin + sidx)};
          const v3 L = Li(s, &cam, x, y, p->spp, p->max_depth, p->use_brick_grid, &pk, counters ? loc : NULL);
          acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
        }
        R* o = out + ((size_t)y * W + x) * 3;
        o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2];
      }
    }
    if (counters) {
#pragma omp critical
      for (int k = 0; k < R2O_NCNT; ++k) tot[k] += loc[k];
    }
  }
  if (counters)
    for (int k = 0; k < R2O_NCNT; ++k) counters[k] = tot[k];
}

/* One path per (pixel, sample) pair listed, with its own counters
 * (out: n*3 radiance values already divided by spp; cnt: n*R2O_NCNT or NULL). */
void SFX(r2o_render_paths)(const scene_t* s, const r2o_render_params* p, int n, const int32_t* pix,
                           const int32_t* smp, R* out, uint64_t* cnt) {
  const camera_t cam = make_camera(p);
  const int W = p->width;
#pragma omp parallel for schedule(dynamic, 16) num_threads(p->threads > 0 ? p->threads : omp_get_max_threads())
  for (int i = 0; i < n; ++i) {
    const int x = pix[i] % W, y = pix[i] / W;
    const path_key pk = {p->seed, p->stream, (uint32_t)pix[i], (uint32_t)smp[i]};
    uint64_t* c = cnt ? cnt + (size_t)i * R2O_NCNT : NULL;
    if (c) memset(c, 0, sizeof(uint64_t) * R2O_NCNT);
    const v3 L = Li(s, &cam, x, y, p->spp, p->max_depth, p->use_brick_grid, &pk, c);
    out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
  }
}

/* --------------------------------------------------- unit-level probes --- */
void SFX(r2o_camera_view_to_world)(const r2o_render_params* p, R* out12) {
  const camera_t c = make_camera(p);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) out12[3 * i + j] = c.view_to_world.c[i][j];
}
void SFX(r2o_camera_ray)(const r2o_render_params* p, int px, int py, uint32_t sample, R* out7) {
  const camera_t c = make_camera(p);
  r2o_rng rng;
  r2o_rng_init(&rng, p->seed, p->stream, (uint32_t)(py * p->width + px), sample, 0u);
  const ray_t r = camera_ray(&c, px, py, &rng);
  out7[0] = r.o.x; out7[1] = r.o.y; out7[2] = r.o.z;
  out7[3] = r.d.x; out7[4] = r.d.y; out7[5] = r.d.z; out7[6] = r.tmax;
}
/* medium transforms: index_to_model, model_to_index, model_to_scaled, scaled_to_model (4 x 12),
 * then ibb_min(3), size(3), inv_max_density, world bbox of instance 0 (6), ratio of instance 0 */
void SFX(r2o_scene_info)(const scene_t* s, int medium, R* out) {
  const medium_t* m = &s->media[medium];
  const m43* ms[4] = {&m->index_to_model, &m->model_to_index, &m->model_to_scaled, &m->scaled_to_model};
  int k = 0;
  for (int q = 0; q < 4; ++q)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j) out[k++] = ms[q]->c[i][j];
  for (int a = 0; a < 3; ++a) out[k++] = (R)m->ibb_min[a];
  for (int a = 0; a < 3; ++a) out[k++] = (R)m->size[a];
  out[k++] = m->inv_max_density;
  out[k++] = s->inst[0].bb_min.x; out[k++] = s->inst[0].bb_min.y; out[k++] = s->inst[0].bb_min.z;
  out[k++] = s->inst[0].bb_max.x; out[k++] = s->inst[0].bb_max.y; out[k++] = s->inst[0].bb_max.z;
  out[k++] = s->inst[0].ratio;
}
/* instance k: model -> world and world -> model 4x3 (column major, 12 each), world
 * bbox min and max, filterRenderRatio (31 values) */
void SFX(r2o_instance_info)(const scene_t* s, int k, R* out) {
  const instance_t* in = &s->inst[k];
  int q = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) out[q++] = in->model_to_world.c[i][j];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) out[q++] = in->world_to_model.c[i][j];
  out[q++] = in->bb_min.x; out[q++] = in->bb_min.y; out[q++] = in->bb_min.z;
  out[q++] = in->bb_max.x; out[q++] = in->bb_max.y; out[q++] = in->bb_max.z;
  out[q++] = in->ratio;
}
/* trilinear record at normalised p: density, S[6], diffuse, specular, normal, ior (17 values) */
void SFX(r2o_filtered_data)(const scene_t* s, int medium, const R* p3, R* out17) {
  const rec_t r = filtered_data(&s->media[medium], V(p3[0], p3[1], p3[2]), NULL);
  out17[0] = r.density;
  for (int i = 0; i < 6; ++i) out17[1 + i] = r.S[i];
  out17[7] = r.diffuse.x; out17[8] = r.diffuse.y; out17[9] = r.diffuse.z;
  out17[10] = r.specular.x; out17[11] = r.specular.y; out17[12] = r.specular.z;
  const v3 n = vnorm(r.normal);
  out17[13] = n.x; out17[14] = n.y; out17[15] = n.z;
  out17[16] = r.ior;
}
void SFX(r2o_decode_record)(const uint8_t* rec, R* out17) {
  const rec_t r = decode_record(rec);
  out17[0] = r.density;
  for (int i = 0; i < 6; ++i) out17[1 + i] = r.S[i];
  out17[7] = r.diffuse.x; out17[8] = r.diffuse.y; out17[9] = r.diffuse.z;
  out17[10] = r.specular.x; out17[11] = r.specular.y; out17[12] = r.specular.z;
  out17[13] = r.normal.x; out17[14] = r.normal.y; out17[15] = r.normal.z;
  out17[16] = r.ior;
}
/* DDA cell walk of one free-flight traversal (the loop of sampleDDA,
 * nvdb_medium_impl.hpp:139-187) without its density decisions: records (brick x, y,
 * z, mip) of every majorant lookup.  collide_every = k > 0 applies the mip drop of a
 * tentative collision (mip - 2, floored at 0, nvdb_medium_impl.hpp:184) after every
 * k-th step, so that the finer mips are walked too; 0 never collides.
 * ray: origin(3), dir(3), tmax in voxel ("scaled index") space. */
int SFX(r2o_dda_cells)(const scene_t* s, int medium, const R* ray7, R sigma_t, int collide_every, int max_cells,
                       int32_t* cells) {
  const medium_t* m = &s->media[medium];
  const v3 o = V(ray7[0], ray7[1], ray7[2]), d = V(ray7[3], ray7[4], ray7[5]);
  const R tmax = ray7[6];
  const v3 ri = V(C(1) / d.x, C(1) / d.y, C(1) / d.z);
  R t = C(0), mip = C(3);
  int n = 0;
  (void)m;
  for (int it = 0; it < LOOP_CAP && t < tmax && n < max_cells; ++it) {
    const v3 cur = vadd(o, sscale(t, d));
    const int imip = (int)RROUND(mip);
    const int ix = (int)RFLOOR(cur.x), iy = (int)RFLOOR(cur.y), iz = (int)RFLOOR(cur.z);
    cells[4 * n + 0] = ix >> 3; cells[4 * n + 1] = iy >> 3; cells[4 * n + 2] = iz >> 3; cells[4 * n + 3] = imip;
    ++n;
    const R dt = step_dda(cur, ri, imip);
    const R prev = t;
    t += dt / sigma_t;
    if (t == prev) break;
    mip = gmin(mip + C(0.25), C(3));
    if (collide_every > 0 && n % collide_every == 0) mip = gmax(C(0), mip - C(2));
  }
  return n;
}
R SFX(r2o_fresnel)(R cos_i, R ior) { return fresnel(cos_i, ior); }
R SFX(r2o_sggx_D)(const R* S6, const R* w) {
  const sggx_t g = sggx_make(S6, V(0, 1, 0));
  return sggx_D(&g, V(w[0], w[1], w[2]));
}
R SFX(r2o_sggx_sigma)(const R* S6, const R* w) {
  const sggx_t g = sggx_make(S6, V(0, 1, 0));
  return sggx_sigma(&g, V(w[0], w[1], w[2]));
}
void SFX(r2o_sggx_vndf)(const R* S6, const R* n, R ux, R uy, R* out3) {
  const sggx_t g = sggx_make(S6, V(n[0], n[1], n[2]));
  const v3 r = sggx_sample_vndf(&g, ux, uy);
  out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
/* phase sample / eval with a fixed RNG stream (key0, counter c0) */
R SFX(r2o_phase_sample)(const R* S6, const R* n, R ior, const R* wo, uint32_t key, uint32_t c0, R* wi3) {
  r2o_rng rng;
  r2o_rng_init(&rng, key, 0, c0, 0, 0);
  v3 wi;
  const R p = phase_sample(S6, V(n[0], n[1], n[2]), ior, V(wo[0], wo[1], wo[2]), &wi, &rng);
  wi3[0] = wi.x; wi3[1] = wi.y; wi3[2] = wi.z;
  return p;
}
R SFX(r2o_phase_p)(const R* S6, const R* n, R ior, const R* wo, const R* wi, uint32_t key, uint32_t c0) {
  r2o_rng rng;
  r2o_rng_init(&rng, key, 0, c0, 0, 0);
  return phase_p(S6, V(n[0], n[1], n[2]), ior, V(wo[0], wo[1], wo[2]), V(wi[0], wi[1], wi[2]), &rng);
}
/* env: sample with a fixed RNG stream -> radiance(3), dir(3), pdf */
void SFX(r2o_env_sample)(const scene_t* s, uint32_t key, uint32_t c0, R* out7) {
  r2o_rng rng;
  r2o_rng_init(&rng, key, 0, c0, 0, 0);
  v3 d;
  R pdf;
  const v3 L = env_sample(s, &rng, &d, &pdf, NULL);
  out7[0] = L.x; out7[1] = L.y; out7[2] = L.z; out7[3] = d.x; out7[4] = d.y; out7[5] = d.z; out7[6] = pdf;
}
void SFX(r2o_env_le)(const scene_t* s, const R* d, R* out4) {
  R pdf;
  const v3 L = env_le(s, V(d[0], d[1], d[2]), &pdf, NULL);
  out4[0] = L.x; out4[1] = L.y; out4[2] = L.z; out4[3] = pdf;
}
void SFX(r2o_env_tables)(const scene_t* s, R* marg_cdf, R* row_cdf, R* ints) {
  for (int j = 0; j <= s->env_h; ++j) marg_cdf[j] = s->marg_cdf[j];
  for (size_t i = 0; i < (size_t)(s->env_w + 1) * s->env_h; ++i) row_cdf[i] = s->row_cdf[i];
  ints[0] = s->marg_int;
  for (int j = 0; j < s->env_h; ++j) ints[1 + j] = s->row_int[j];
}
/* transmittance of one world-space segment through instance 0 (repeat with
 * different counters to estimate the mean) */
R SFX(r2o_tr_segment)(const scene_t* s, const R* ray7, int use_dda, uint32_t key, uint32_t c0) {
  r2o_rng rng;
  r2o_rng_init(&rng, key, 0, c0, 0, 0);
  const ray_t r = mkray(V(ray7[0], ray7[1], ray7[2]), V(ray7[3], ray7[4], ray7[5]), ray7[6], 0);
  return inst_tr(s, 0, &r, &rng, use_dda, NULL);
}
/* free flight through instance 0: returns 1 if it collided (and the world hit t in *t) */
int SFX(r2o_sample_segment)(const scene_t* s, const R* ray7, int use_dda, uint32_t key, uint32_t c0, R* t_out) {
  r2o_rng rng;
  r2o_rng_init(&rng, key, 0, c0, 0, 0);
  ray_t r = mkray(V(ray7[0], ray7[1], ray7[2]), V(ray7[3], ray7[4], ray7[5]), ray7[6], 0);
  mhit_t mi;
  mi.valid = 0;
  (void)inst_sample(s, 0, &r, &rng, use_dda, &mi, NULL);
  *t_out = r.tmax;
  return mi.valid;
}
void SFX(r2o_move_to_voxel_border)(const scene_t* s, const R* ray7, R* out7) {
  const ray_t r = mkray(V(ray7[0], ray7[1], ray7[2]), V(ray7[3], ray7[4], ray7[5]), ray7[6], 0);
  const ray_t m = inst_move_to_voxel_border(s, 0, &r);
  out7[0] = m.o.x; out7[1] = m.o.y; out7[2] = m.o.z; out7[3] = m.d.x; out7[4] = m.d.y; out7[5] = m.d.z; out7[6] = m.tmax;
}
/* closest-hit query: returns hit flag; out: t, pos(3), medium */
/* out: t, hit position, instance (-1: none), surface (-1: none) */
int SFX(r2o_intersect)(const scene_t* s, const R* ray7, R* out6) {
  ray_t r = mkray(V(ray7[0], ray7[1], ray7[2]), V(ray7[3], ray7[4], ray7[5]), ray7[6], -1);
  const hit_t h = scene_intersect(s, &r);
  out6[0] = h.t; out6[1] = h.pos.x; out6[2] = h.pos.y; out6[3] = h.pos.z; out6[4] = (R)h.medium;
  out6[5] = (R)h.surf;
  return h.hit;
}
/* direct() at a medium interaction of instance `inst` at world position pos with
 * voxel data (S6, n, ior), seen from wo: mean and variance per channel over `count`
 * draws with counters (key, c0 + i), accumulated in double */
void SFX(r2o_direct_mean)(const scene_t* s, int inst, const R* pos, const R* S6, const R* n, R ior, const R* wo,
                          int use_dda, uint32_t key, uint32_t c0, int64_t count, double* mean3, double* var3) {
  mhit_t mi;
  mi.valid = 1;
  mi.inst = inst;
  mi.pos = V(pos[0], pos[1], pos[2]);
  mi.normal = V(n[0], n[1], n[2]);
  for (int k = 0; k < 6; ++k) mi.S[k] = S6[k];
  mi.ior = ior;
  double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
  for (int64_t i = 0; i < count; ++i) {
    /* the first scatter vertex of a path keyed (key, 0) at pixel c0 + i, sample 0 */
    const path_key pk = {key, 0, c0 + (uint32_t)i, 0};
    r2o_rng rng;
    stream_init(&rng, &pk, 1, 0);
    const v3 L = direct(s, &mi, V(wo[0], wo[1], wo[2]), &rng, &pk, 1, use_dda, NULL);
    const double l[3] = {(double)L.x, (double)L.y, (double)L.z};
    for (int k = 0; k < 3; ++k) {
      sum[k] += l[k];
      sq[k] += l[k] * l[k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    mean3[k] = sum[k] / (double)count;
    var3[k] = sq[k] / (double)count - mean3[k] * mean3[k];
  }
}
/* phase_p averaged over the VNDF draws of counters (key, c0 .. c0 + n_c - 1), for each
 * of n_wi directions */
void SFX(r2o_phase_p_mean)(const R* S6, const R* n, R ior, const R* wo, const R* wi, int64_t n_wi, uint32_t key,
                           uint32_t c0, int n_c, double* out) {
  for (int64_t j = 0; j < n_wi; ++j) {
    double acc = 0;
    for (int c = 0; c < n_c; ++c) {
      r2o_rng rng;
      r2o_rng_init(&rng, key, 0, c0 + (uint32_t)c, 0, 0);
      acc += (double)phase_p(S6, V(n[0], n[1], n[2]), ior, V(wo[0], wo[1], wo[2]),
                             V(wi[3 * j], wi[3 * j + 1], wi[3 * j + 2]), &rng);
    }
    out[j] = acc / n_c;
  }
}

/* ---- surface probes (BSDF lobes, shapes) ---- */
/* out: mat_f rgb, mat_pdf, on_f, on_pdf, mf_f, mf_pdf, Fresnel weight of |wo.z| */
void SFX(r2o_bsdf_eval)(const r2o_surface_desc* sd, const R* wo, const R* wi, R* out9) {
  surface_t sf;
  surface_init(&sf, sd);
  const v3 o = V(wo[0], wo[1], wo[2]), i = V(wi[0], wi[1], wi[2]);
  const v3 f = mat_f(&sf, o, i);
  out9[0] = f.x; out9[1] = f.y; out9[2] = f.z;
  out9[3] = mat_pdf(&sf, o, i);
  out9[4] = on_f(&sf, o, i);
  out9[5] = on_pdf(o, i);
  out9[6] = mf_f(&sf, o, i);
  out9[7] = mf_pdf(&sf, o, i);
  out9[8] = fresnel_ie(RABS(o.z), sf.eta_i, sf.eta_t);
}
/* out: D(h), G1(wo), G(wo, wi), VNDF pdf(wo, h) */
void SFX(r2o_ggx_eval)(const r2o_surface_desc* sd, const R* wo, const R* wi, const R* h, R* out4) {
  surface_t sf;
  surface_init(&sf, sd);
  const v3 o = V(wo[0], wo[1], wo[2]), i = V(wi[0], wi[1], wi[2]), hh = V(h[0], h[1], h[2]);
  out4[0] = ggx_D(&sf, hh);
  out4[1] = ggx_G1(&sf, o);
  out4[2] = ggx_G(&sf, o, i);
  out4[3] = ggx_pdf(&sf, o, hh);
}
void SFX(r2o_ggx_sample_h)(const r2o_surface_desc* sd, const R* wo, R u1, R u2, R* out3) {
  surface_t sf;
  surface_init(&sf, sd);
  const v3 h = ggx_sample_h(&sf, V(wo[0], wo[1], wo[2]), u1, u2);
  out3[0] = h.x; out3[1] = h.y; out3[2] = h.z;
}
/* CMaterial::sampleF with its own stream; out: wi, f rgb, pdf (pdf 0 and wi 0
 * when the sampled lobe rejected the draw) */
void SFX(r2o_bsdf_sample)(const r2o_surface_desc* sd, const R* wo, uint32_t key, uint32_t c0, R* out7) {
  surface_t sf;
  surface_init(&sf, sd);
  r2o_rng rng;
  r2o_rng_init(&rng, key, 0, c0, 0, 0);
  v3 wi = V(0, 0, 0);
  R pdf = C(0);
  const v3 f = mat_sample(&sf, V(wo[0], wo[1], wo[2]), &wi, &rng, &pdf);
  out7[0] = wi.x; out7[1] = wi.y; out7[2] = wi.z;
  out7[3] = f.x; out7[4] = f.y; out7[5] = f.z;
  out7[6] = pdf;
}
R SFX(r2o_fresnel_ie)(R cos_i, R eta_i, R eta_t) { return fresnel_ie(cos_i, eta_i, eta_t); }
/* shape test of one surface against a ray (o, d, tmax); out: t, normal
 * (normalised as the closest-hit program does); returns the hit flag */
int SFX(r2o_surface_hit)(const r2o_surface_desc* sd, const R* ray7, R* out4) {
  surface_t sf;
  surface_init(&sf, sd);
  const ray_t r = mkray(V(ray7[0], ray7[1], ray7[2]), V(ray7[3], ray7[4], ray7[5]), ray7[6], -1);
  R t;
  v3 n;
  const int ok = sf.shape == R2O_SHAPE_CIRCLE ? circle_hit(&sf, &r, r.tmax, &t, &n) : sphere_hit(&sf, &r, r.tmax, &t, &n);
  if (ok) {
    n = vnorm(n);
    out4[0] = t; out4[1] = n.x; out4[2] = n.y; out4[3] = n.z;
  }
  return ok;
}
