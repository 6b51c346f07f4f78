/*
 * Oracle (test infrastructure only) -- precision-independent pieces of the CPU
 * restatement of FrederikBoehm/Renderer2's volumetric hot path.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this code, and only as the checker / CPU baseline.  The product library never
 * links it.
 *
 * Contents and what pins each part (see tests/test_oracle_*.py):
 *   - Philox4x32-10 (Salmon et al., SC'11) replaces cuRAND XORWOW
 *     (raytracer/include/sampling/sampler.hpp:28).  Pinned by the Random123
 *     known-answer vectors.
 *   - float->half with round-to-nearest, ties away from zero: the algorithm of
 *     glm 0.9.9.8 detail::toFloat16, used by encode_range
 *     (raytracer/include/grid_brick/common.hpp:23-29).  Pinned by exhaustive
 *     comparison against numpy's RNE conversion away from ties, and by hand
 *     tie cases.
 *   - float->half RNE (CUDA __float2half, used for SFilteredDataCompact::ior,
 *     raytracer/include/filtering/filtered_data.hpp:44,106).  Pinned against
 *     numpy float16.
 *   - brick-grid build (raytracer/src/grid_brick/host_grid_brick.cpp:61-152):
 *     dilated min/max ranges, u8 atlas quantisation, 3 min/max mip levels.
 *     Pinned by brute-force invariants (range brackets every dilated voxel,
 *     decode error <= 1/510 of the range, mips are max/min of children).
 */
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include "r2o.h"

/* ------------------------------------------------------------------ RNG -- */

static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void r2o_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
  uint32_t k[2] = {key_in[0], key_in[1]};
  philox_round(c, k);
  for (int r = 1; r < 10; ++r) {
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
    philox_round(c, k);
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

void r2o_rng_init(r2o_rng* s, uint32_t key0, uint32_t key1, uint32_t c0, uint32_t c1, uint32_t c3) {
  s->key[0] = key0; s->key[1] = key1;
  s->ctr[0] = c0; s->ctr[1] = c1; s->ctr[2] = 0; s->ctr[3] = c3;
  s->n = 0; s->block = 0xFFFFFFFFu;
}

uint32_t r2o_rng_next_u32(r2o_rng* s) {
  const uint32_t blk = s->n >> 2;
  if (blk != s->block) {
    uint32_t c[4] = {s->ctr[0], s->ctr[1], blk, s->ctr[3]};
    r2o_philox4x32_10(c, s->key, s->buf);
    s->block = blk;
  }
  return s->buf[(s->n++) & 3u];
}

/* ----------------------------------------------------------------- half -- */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

uint16_t r2o_half_from_float_glm(float f) {
  const uint32_t i = f2u(f);
  const uint32_t s = (i >> 16) & 0x8000u;
  int e = (int)((i >> 23) & 0xffu) - (127 - 15);
  uint32_t m = i & 0x007fffffu;
  if (e <= 0) {
    if (e < -10) return (uint16_t)s;
    m = (m | 0x00800000u) >> (1 - e);
    if (m & 0x00001000u) m += 0x00002000u;
    return (uint16_t)(s | (m >> 13));
  }
  if (e == 0xff - (127 - 15)) {
    if (m == 0) return (uint16_t)(s | 0x7c00u);
    m >>= 13;
    return (uint16_t)(s | 0x7c00u | m | (m == 0));
  }
  if (m & 0x00001000u) {
    m += 0x00002000u;
    if (m & 0x00800000u) { m = 0; e += 1; }
  }
  if (e > 30) return (uint16_t)(s | 0x7c00u);
  return (uint16_t)(s | ((uint32_t)e << 10) | (m >> 13));
}

uint16_t r2o_half_from_float_rne(float f) {
  const uint32_t x = f2u(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to >= 65520 -> inf */
  if (ax < 0x33000001u) return (uint16_t)sign;              /* < 2^-25 (or ==) -> 0 */
  int e = (int)(ax >> 23);
  uint32_t m = (ax & 0x7fffffu) | 0x800000u;
  int shift;
  uint32_t base;
  if (e < 113) { /* half subnormal */
    shift = 126 - e; /* 14 - (e - 112) + ... */
    base = 0;
  } else {
    shift = 13;
    base = (uint32_t)(e - 112) << 10;
    m &= 0x7fffffu;
  }
  uint32_t q = m >> shift;
  const uint32_t rem = m & ((1u << shift) - 1u);
  const uint32_t half = 1u << (shift - 1);
  if (rem > half || (rem == half && (q & 1u))) q += 1;
  return (uint16_t)(sign | (base + q));
}

float r2o_half_to_float(uint16_t h) {
  const uint32_t s = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  if (e == 0) {
    if (m == 0) return u2f(s);
    /* subnormal: value = m * 2^-24 */
    float v = (float)m * 5.9604644775390625e-08f;
    return (s ? -v : v);
  }
  if (e == 31) return u2f(s | 0x7f800000u | (m << 13));
  return u2f(s | ((e + 112u) << 23) | (m << 13));
}

/* --------------------------------------------------------- brick codecs -- */

uint32_t r2o_encode_range(float lo, float hi) {
  return (uint32_t)r2o_half_from_float_glm(lo) | ((uint32_t)r2o_half_from_float_glm(hi) << 16);
}

void r2o_decode_range(uint32_t d, float* lo, float* hi) {
  *lo = r2o_half_to_float((uint16_t)(d & 0xFFFFu));
  *hi = r2o_half_to_float((uint16_t)(d >> 16));
}

uint8_t r2o_encode_voxel(float value, float lo, float hi) {
  float norm = (value - lo) / (hi - lo);
  norm = norm < 0.f ? 0.f : (norm > 1.f ? 1.f : norm); /* glm::clamp */
  return (uint8_t)roundf(255.f * norm);
}

float r2o_decode_voxel(uint8_t v, float lo, float hi) {
  return lo + (float)v * (1.f / 255.f) * (hi - lo);
}

/* --------------------------------------------------------- voxel access -- */

/* The dense density source with the reference's active-voxel rule
 * (raytracer/src/filtering/openvdb_backend.cpp:108-132): a voxel holds its value
 * iff density > 0 or it is one of the 8 corners; everything else (and everything
 * outside [0,n)) reads the background record (density 0). */
int r2o_voxel_active(const float* dens, int nx, int ny, int nz, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0;
  const float d = dens[((size_t)z * ny + y) * nx + x];
  if (d > 0.f) return 1;
  const int cx = (x == 0 || x == nx - 1), cy = (y == 0 || y == ny - 1), cz = (z == 0 || z == nz - 1);
  return cx && cy && cz;
}

static inline float density_at(const float* dens, int nx, int ny, int nz, int x, int y, int z) {
  return r2o_voxel_active(dens, nx, ny, nz, x, y, z) ? dens[((size_t)z * ny + y) * nx + x] : 0.f;
}

/* ---------------------------------------------------------- brick build -- */

/* Restates CHostBrickGrid::CHostBrickGrid (host_grid_brick.cpp:61-152).
 * Atlas slots are assigned in brick-linear order here; the reference assigns
 * them with a relaxed atomic under par_unseq.  Lookups do not depend on the
 * slot, so `u8` is kept as a dense per-voxel array (value only meaningful in
 * non-empty bricks). */
int r2o_brick_build(const float* dens, int nx, int ny, int nz, r2o_bricks* b) {
  memset(b, 0, sizeof(*b));
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    const int nb = (int)ceilf((float)n[a] / 8.f);
    b->nb[a] = (int)ceilf((float)nb / 8.f) * 8;
    if (b->nb[a] >= 1024) return -1; /* MAX_BRICKS, host_grid_brick.cpp:69-70 */
  }
  b->nx = nx; b->ny = ny; b->nz = nz;
  const size_t nbt = (size_t)b->nb[0] * b->nb[1] * b->nb[2];
  b->range = (uint32_t*)calloc(nbt, 4);
  b->nonempty = (uint8_t*)calloc(nbt, 1);
  const size_t nvx = (size_t)b->nb[0] * 8 * b->nb[1] * 8 * b->nb[2] * 8;
  b->u8 = (uint8_t*)calloc(nvx, 1);
  if (!b->range || !b->nonempty || !b->u8) return -2;
  const int ex = b->nb[0] * 8, ey = b->nb[1] * 8;
  for (int bz = 0; bz < b->nb[2]; ++bz)
    for (int by = 0; by < b->nb[1]; ++by)
      for (int bx = 0; bx < b->nb[0]; ++bx) {
        float lo = FLT_MAX, hi = -FLT_MAX;
        for (int z = -1; z < 9; ++z)
          for (int y = -1; y < 9; ++y)
            for (int x = -1; x < 9; ++x) {
              const float d = density_at(dens, nx, ny, nz, bx * 8 + x, by * 8 + y, bz * 8 + z);
              lo = fminf(lo, d);
              hi = fmaxf(hi, d);
            }
        const size_t bi = ((size_t)bz * b->nb[1] + by) * b->nb[0] + bx;
        b->range[bi] = r2o_encode_range(lo, hi);
        if (hi == lo) continue;
        b->nonempty[bi] = 1;
        b->n_nonempty++;
        float rlo, rhi;
        r2o_decode_range(b->range[bi], &rlo, &rhi);
        for (int z = 0; z < 8; ++z)
          for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
              const int vx = bx * 8 + x, vy = by * 8 + y, vz = bz * 8 + z;
              b->u8[((size_t)vz * ey + vy) * ex + vx] =
                  r2o_encode_voxel(density_at(dens, nx, ny, nz, vx, vy, vz), rlo, rhi);
            }
      }
  /* min/max mips (host_grid_brick.cpp:124-151) */
  for (int l = 0; l < 3; ++l) {
    const int sx = b->nb[0] >> (l + 1), sy = b->nb[1] >> (l + 1), sz = b->nb[2] >> (l + 1);
    const uint32_t* src = l == 0 ? b->range : b->mip[l - 1];
    const int px = b->nb[0] >> l, py = b->nb[1] >> l;
    b->mip[l] = (uint32_t*)calloc((size_t)sx * sy * sz, 4);
    if (!b->mip[l]) return -2;
    for (int z = 0; z < sz; ++z)
      for (int y = 0; y < sy; ++y)
        for (int x = 0; x < sx; ++x) {
          float lo = FLT_MAX, hi = -FLT_MAX;
          for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
              for (int dx = 0; dx < 2; ++dx) {
                float clo, chi;
                r2o_decode_range(src[((size_t)(2 * z + dz) * py + (2 * y + dy)) * px + (2 * x + dx)], &clo, &chi);
                lo = fminf(lo, clo);
                hi = fmaxf(hi, chi);
              }
          b->mip[l][((size_t)z * sy + y) * sx + x] = r2o_encode_range(lo, hi);
        }
  }
  return 0;
}

void r2o_brick_free(r2o_bricks* b) {
  free(b->range); free(b->nonempty); free(b->u8);
  for (int l = 0; l < 3; ++l) free(b->mip[l]);
  memset(b, 0, sizeof(*b));
}

/* Brick majorant (device_grid_brick_impl.hpp:29-33): border (= 0) outside the
 * brick grid, else range.y at mip level `mip`. */
float r2o_brick_majorant(const r2o_bricks* b, int bx, int by, int bz, int mip) {
  if (bx < 0 || by < 0 || bz < 0 || bx >= b->nb[0] || by >= b->nb[1] || bz >= b->nb[2]) return 0.f;
  float lo, hi;
  if (mip == 0) {
    r2o_decode_range(b->range[((size_t)bz * b->nb[1] + by) * b->nb[0] + bx], &lo, &hi);
  } else {
    const int sx = b->nb[0] >> mip, sy = b->nb[1] >> mip;
    r2o_decode_range(b->mip[mip - 1][((size_t)(bz >> mip) * sy + (by >> mip)) * sx + (bx >> mip)], &lo, &hi);
  }
  return hi;
}

/* Brick density at integer voxel coordinate (device_grid_brick_impl.hpp:17-27). */
float r2o_brick_density(const r2o_bricks* b, int ix, int iy, int iz) {
  const int bx = ix >> 3, by = iy >> 3, bz = iz >> 3;
  if (bx < 0 || by < 0 || bz < 0 || bx >= b->nb[0] || by >= b->nb[1] || bz >= b->nb[2]) return 0.f;
  const size_t bi = ((size_t)bz * b->nb[1] + by) * b->nb[0] + bx;
  float lo, hi;
  r2o_decode_range(b->range[bi], &lo, &hi);
  uint8_t v = 0;
  if (b->nonempty[bi]) {
    const int ex = b->nb[0] * 8, ey = b->nb[1] * 8;
    v = b->u8[((size_t)iz * ey + iy) * ex + ix];
  }
  return r2o_decode_voxel(v, lo, hi);
}

/* ------------------------------------------------ deterministic math ---- */
/* sin/cos: quadrant reduction by pi/2 (two-part constant) and Taylor series on
 * |r| <= pi/4 to degree 17/16, all in double, rounded once to float. */
void r2o_det_sincosf(float x, float* s, float* c) {
  const double xd = (double)x;
  const double q = rint(xd * 0.63661977236758134308);
  const double r = (xd - q * 1.5707963267948966) - q * 6.123233995736766e-17;
  const double r2 = r * r;
  double sp = 1.0 / 355687428096000.0;
  sp = sp * r2 - 1.0 / 1307674368000.0;
  sp = sp * r2 + 1.0 / 6227020800.0;
  sp = sp * r2 - 1.0 / 39916800.0;
  sp = sp * r2 + 1.0 / 362880.0;
  sp = sp * r2 - 1.0 / 5040.0;
  sp = sp * r2 + 1.0 / 120.0;
  sp = sp * r2 - 1.0 / 6.0;
  sp = sp * r2 + 1.0;
  const double sr = sp * r;
  double cp = 1.0 / 20922789888000.0;
  cp = cp * r2 - 1.0 / 87178291200.0;
  cp = cp * r2 + 1.0 / 479001600.0;
  cp = cp * r2 - 1.0 / 3628800.0;
  cp = cp * r2 + 1.0 / 40320.0;
  cp = cp * r2 - 1.0 / 720.0;
  cp = cp * r2 + 1.0 / 24.0;
  cp = cp * r2 - 1.0 / 2.0;
  const double cr = cp * r2 + 1.0;
  const int iq = ((int)q) & 3;
  const double so = iq == 0 ? sr : (iq == 1 ? cr : (iq == 2 ? -sr : -cr));
  const double co = iq == 0 ? cr : (iq == 1 ? -sr : (iq == 2 ? -cr : sr));
  *s = (float)so;
  *c = (float)co;
}

/* exp: n = rint(x / ln2), r = x - n ln2, Taylor series to degree 13, scaled by 2^n */
double r2o_det_exp_d(double xd) {
  const double n = rint(xd * 1.4426950408889634);
  const double r = xd - n * 0.6931471805599453;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)n);
}

float r2o_det_expf(float x) {
  if (x != x) return x;
  if (x < -104.f) return 0.f;
  if (x > 89.f) return INFINITY;
  return (float)r2o_det_exp_d((double)x);
}

/* log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) by its series */
double r2o_det_log_d(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752) { m = m * 2.0; e = e - 1; }
  const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  return 2.0 * s * p + (double)e * 0.6931471805599453;
}

float r2o_det_logf(float x) {
  if (x != x || x < 0.f) return NAN;
  if (x == 0.f) return -INFINITY;
  if (x == INFINITY) return x;
  return (float)r2o_det_log_d((double)x);
}

/* x^y = exp(y log x) in double, rounded once */
float r2o_det_powf(float x, float y) {
  if (x != x || y != y || x < 0.f) return NAN;
  if (x == 0.f) return y > 0.f ? 0.f : (y == 0.f ? 1.f : INFINITY);
  if (x == INFINITY) return y > 0.f ? INFINITY : 0.f;
  const double t = (double)y * r2o_det_log_d((double)x);
  if (t < -745.0) return 0.f;
  if (t > 709.0) return INFINITY;
  return (float)r2o_det_exp_d(t);
}

