// Host/device math for the volumetric hot path: float3 algebra, glm-style 4x3 /
// 4x4 / 3x3 matrices (column major), Philox4x32-10 counter RNG and the half-float
// codecs used by the voxel records and brick ranges.  Every operation keeps the
// evaluation order of the glm expression it replaces, so host-built transforms
// are bit-identical wherever they are recomputed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <float.h>

#define R2_HD __host__ __device__ inline

namespace r2 {

// Read-only launch data (a scene, the entry table of a batched launch) is read
// through the constant address space: wave-uniform fields then load with scalar
// instructions, and the device pointers read from it are known to address global
// memory, so loads through them are global loads instead of flat ones (which
// also wait on the LDS counter).  The empty asm keeps the optimiser from folding
// the address-space round trip away.  The pointer must be wave-uniform and the
// data must not change while the kernel runs.
template <typename T>
__device__ __forceinline__ const T& kconst(const T* p) {
  uint64_t v = (uint64_t)(uintptr_t)p;
  asm volatile("" : "+s"(v));
  return *(const T*)(const __attribute__((address_space(4))) T*)v;
}
// element i of such a table: the base is wave-uniform, the index may differ per
// lane (then the load is a vector load of the constant address space)
template <typename T>
__device__ __forceinline__ const T& kconst_at(const T* p, int i) {
  uint64_t v = (uint64_t)(uintptr_t)p;
  asm volatile("" : "+s"(v));
  return ((const T*)(const __attribute__((address_space(4))) T*)v)[i];
}

struct v3 { float x, y, z; };
R2_HD v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
R2_HD v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
R2_HD v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
R2_HD v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
R2_HD v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
R2_HD v3 operator*(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
R2_HD v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
R2_HD v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
R2_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
R2_HD float length(v3 a) { return sqrtf(dot(a, a)); }
// glm::normalize: v * inversesqrt(dot(v, v)) with inversesqrt = 1 / sqrt
R2_HD v3 normalize(v3 a) { return a * (1.f / sqrtf(dot(a, a))); }
R2_HD v3 cross(v3 x, v3 y) {
  return mk(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
// glm scalar min / max
R2_HD float gmin(float x, float y) { return (y < x) ? y : x; }
R2_HD float gmax(float x, float y) { return (x < y) ? y : x; }

// glm::mat4x3 (4 columns of vec3), stored as m[col*3 + row]
struct m43 { float m[12]; };
R2_HD v3 mul_point(const m43& M, v3 v) {
  const float* m = M.m;
  return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z + m[9] * 1.f,
            m[1] * v.x + m[4] * v.y + m[7] * v.z + m[10] * 1.f,
            m[2] * v.x + m[5] * v.y + m[8] * v.z + m[11] * 1.f);
}
R2_HD v3 mul_dir(const m43& M, v3 v) {
  const float* m = M.m;
  return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z + m[9] * 0.f,
            m[1] * v.x + m[4] * v.y + m[7] * v.z + m[10] * 0.f,
            m[2] * v.x + m[5] * v.y + m[8] * v.z + m[11] * 0.f);
}

// glm::mat4, stored as c[col][row]
struct m44 { float c[4][4]; };
R2_HD m44 identity44() {
  m44 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.c[i][j] = (i == j) ? 1.f : 0.f;
  return r;
}
R2_HD m44 to44(const m43& a) {
  m44 r;
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.m[i * 3 + j];
    r.c[i][3] = (i == 3) ? 1.f : 0.f;
  }
  return r;
}
R2_HD m43 to43(const m44& a) {
  m43 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) r.m[i * 3 + j] = a.c[i][j];
  return r;
}
R2_HD m44 mul44(const m44& a, const m44& b) {
  m44 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r.c[i][j] = a.c[0][j] * b.c[i][0] + a.c[1][j] * b.c[i][1] + a.c[2][j] * b.c[i][2] + a.c[3][j] * b.c[i][3];
  return r;
}
// General 4x4 inverse by 2x2 cofactors (the scheme glm::inverse uses for mat4).
R2_HD m44 inverse44(const m44& mm) {
  const float(*m)[4] = mm.c;
  const float c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
  const float c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
  const float c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3];
  const float c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
  const float c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3];
  const float c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
  const float c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
  const float c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2];
  const float c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  const float c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
  const float c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3];
  const float c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
  const float c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
  const float c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2];
  const float c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  const float c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
  const float c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1];
  const float c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
  const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
  const float v0[4] = {m[1][0], m[0][0], m[0][0], m[0][0]};
  const float v1[4] = {m[1][1], m[0][1], m[0][1], m[0][1]};
  const float v2[4] = {m[1][2], m[0][2], m[0][2], m[0][2]};
  const float v3v[4] = {m[1][3], m[0][3], m[0][3], m[0][3]};
  const float sa[4] = {1.f, -1.f, 1.f, -1.f}, sb[4] = {-1.f, 1.f, -1.f, 1.f};
  m44 inv;
  for (int k = 0; k < 4; ++k) {
    inv.c[0][k] = (v1[k] * f0[k] - v2[k] * f1[k] + v3v[k] * f2[k]) * sa[k];
    inv.c[1][k] = (v0[k] * f0[k] - v2[k] * f3[k] + v3v[k] * f4[k]) * sb[k];
    inv.c[2][k] = (v0[k] * f1[k] - v1[k] * f3[k] + v3v[k] * f5[k]) * sa[k];
    inv.c[3][k] = (v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k]) * sb[k];
  }
  const float d0 = m[0][0] * inv.c[0][0], d1 = m[0][1] * inv.c[1][0], d2 = m[0][2] * inv.c[2][0],
              d3 = m[0][3] * inv.c[3][0];
  const float det = (d0 + d1) + (d2 + d3);
  const float o = 1.f / det;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) inv.c[i][j] = inv.c[i][j] * o;
  return inv;
}
R2_HD m43 inverse43(const m43& a) { return to43(inverse44(to44(a))); }

// ------------------------------------------------------------------ RNG --
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): key = (seed, stream),
// counter = (c0, c1, block, c3).  Uniforms are (x >> 8) * 2^-24 in [0, 1).
R2_HD void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
  const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k[0];
  const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k[1];
  c[1] = (uint32_t)p1;
  c[3] = (uint32_t)p0;
  c[0] = n0;
  c[2] = n2;
}
R2_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  uint32_t k[2] = {k0, k1};
  philox_round(c, k);
#pragma unroll
  for (int r = 1; r < 10; ++r) {
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
    philox_round(c, k);
  }
}

struct Rng {
  uint32_t k0, k1, c0, c1, c3;
  uint32_t n;
  uint32_t buf[4];
  R2_HD void init(uint32_t key0, uint32_t key1, uint32_t ctr0, uint32_t ctr1, uint32_t ctr3) {
    k0 = key0; k1 = key1; c0 = ctr0; c1 = ctr1; c3 = ctr3; n = 0;
  }
  R2_HD uint32_t next_u32() {
    const uint32_t slot = n & 3u;
    if (slot == 0) {
      uint32_t c[4] = {c0, c1, n >> 2, c3};
      philox4x32_10(c, k0, k1);
      buf[0] = c[0]; buf[1] = c[1]; buf[2] = c[2]; buf[3] = c[3];
    }
    ++n;
    // register-friendly select instead of a dynamic index
    return slot == 0 ? buf[0] : (slot == 1 ? buf[1] : (slot == 2 ? buf[2] : buf[3]));
  }
  R2_HD float uni() { return (float)(next_u32() >> 8) * 5.9604644775390625e-08f; }
};

// ----------------------------------------------------------------- half --
R2_HD uint32_t f2u(float f) { union { float f; uint32_t u; } x; x.f = f; return x.u; }
R2_HD float u2f(uint32_t u) { union { float f; uint32_t u; } x; x.u = u; return x.f; }

// float -> half, round to nearest with ties away from zero (the algorithm of
// glm::detail::toFloat16, used by encode_range, grid_brick/common.hpp:23-25).
R2_HD uint16_t half_from_float_tie_away(float f) {
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
// float -> half, IEEE round to nearest even (the __half(float) conversion of
// SFilteredDataCompact::ior, filtered_data.hpp:106).
R2_HD uint16_t half_from_float_rne(float f) {
  const uint32_t x = f2u(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (ax < 0x33000001u) return (uint16_t)sign;
  const int e = (int)(ax >> 23);
  uint32_t m = (ax & 0x7fffffu) | 0x800000u;
  int shift;
  uint32_t base;
  if (e < 113) {
    shift = 126 - e;
    base = 0;
  } else {
    shift = 13;
    base = (uint32_t)(e - 112) << 10;
    m &= 0x7fffffu;
  }
  uint32_t q = m >> shift;
  const uint32_t rem = m & ((1u << shift) - 1u);
  const uint32_t hf = 1u << (shift - 1);
  if (rem > hf || (rem == hf && (q & 1u))) q += 1;
  return (uint16_t)(sign | (base + q));
}
R2_HD float half_to_float(uint16_t h) {
  const uint32_t s = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  const uint32_t m = h & 0x3ffu;
  if (e == 0) {
    if (m == 0) return u2f(s);
    const float v = (float)m * 5.9604644775390625e-08f;
    return s ? -v : v;
  }
  if (e == 31) return u2f(s | 0x7f800000u | (m << 13));
  return u2f(s | ((e + 112u) << 23) | (m << 13));
}
// brick-grid codecs (grid_brick/common.hpp:23-53)
R2_HD uint32_t encode_range(float lo, float hi) {
  return (uint32_t)half_from_float_tie_away(lo) | ((uint32_t)half_from_float_tie_away(hi) << 16);
}
R2_HD float range_lo(uint32_t d) { return half_to_float((uint16_t)(d & 0xFFFFu)); }
R2_HD float range_hi(uint32_t d) { return half_to_float((uint16_t)(d >> 16)); }
R2_HD uint8_t encode_voxel(float value, float lo, float hi) {
  float norm = (value - lo) / (hi - lo);
  norm = norm < 0.f ? 0.f : (norm > 1.f ? 1.f : norm);
  return (uint8_t)roundf(255.f * norm);
}
R2_HD float decode_voxel(uint8_t v, float lo, float hi) { return lo + (float)v * (1.f / 255.f) * (hi - lo); }

// bfloat16 (round to nearest even) for the packed record form
R2_HD uint16_t bf16_from_float(float f) {
  const uint32_t u = f2u(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  const uint32_t lsb = (u >> 16) & 1u;
  return (uint16_t)((u + 0x7fffu + lsb) >> 16);
}
R2_HD float bf16_to_float(uint16_t h) { return u2f((uint32_t)h << 16); }

}  // namespace r2
