// Device algorithms of the volume integrator: NanoVDB-layout voxel lookups,
// brick majorant DDA, delta / ratio tracking, SGGX microflake phase, environment
// map importance sampling and the path integrator.  Each function restates the
// reference routine it cites; arithmetic keeps the reference's evaluation order.
#pragma once
#include "r2_types.hpp"
#include "r2_packed.hpp"

namespace r2 {

struct Cnt {
  uint32_t c[kNumCnt];
};

template <bool kCount>
__device__ __forceinline__ void count(Cnt& c, int k, uint32_t v = 1) {
  if constexpr (kCount) c.c[k] += v;
}

struct Ray {
  v3 o, d;
  float tmax;
  int medium;  // instance index, -1 = none
};
__device__ __forceinline__ Ray mkray(v3 o, v3 d, float tmax, int medium) {
  Ray r; r.o = o; r.d = d; r.tmax = tmax; r.medium = medium; return r;
}
// CRay::transform (ray.hpp:41-47)
__device__ __forceinline__ Ray ray_transform(const Ray& r, const m43& M) {
  const v3 o = mul_point(M, r.o);
  const v3 e = mul_point(M, r.o + r.tmax * r.d);
  const v3 d = normalize(mul_dir(M, r.d));
  return mkray(o, d, length(e - o), r.medium);
}
// CRay::transform2 (ray.hpp:49-55)
__device__ __forceinline__ Ray ray_transform2(const Ray& r, const m43& M) {
  const v3 o = mul_point(M, r.o);
  const v3 e = mul_point(M, r.o + r.tmax * r.d);
  const v3 d = mul_dir(M, r.d);
  return mkray(o, d, length(e - o) / length(d), r.medium);
}
// nextFloatUp / nextFloatDown (functions.hpp:94-126): zero is not bumped
__device__ __forceinline__ float next_up(float v) {
  if (isinf(v) && v > 0.f) return v;
  if (v == 0.f) return 0.f;
  uint32_t u = f2u(v);
  u = (v >= 0.f) ? u + 1u : u - 1u;
  return u2f(u);
}
__device__ __forceinline__ float next_dn(float v) {
  if (isinf(v) && v < 0.f) return v;
  if (v == 0.f) return -0.f;
  uint32_t u = f2u(v);
  u = (v > 0.f) ? u - 1u : u + 1u;
  return u2f(u);
}
// CRay::offsetRayOrigin (ray.hpp:71-84)
__device__ __forceinline__ v3 offset_point(v3 p, v3 dir) {
  const float d = fabsf(dir.x) * kRayOffset + fabsf(dir.y) * kRayOffset + fabsf(dir.z) * kRayOffset;
  const v3 off = dir * d;
  v3 po = p + off;
  po.x = off.x > 0.f ? next_up(po.x) : (off.x < 0.f ? next_dn(po.x) : po.x);
  po.y = off.y > 0.f ? next_up(po.y) : (off.y < 0.f ? next_dn(po.y) : po.y);
  po.z = off.z > 0.f ? next_up(po.z) : (off.z < 0.f ? next_dn(po.z) : po.z);
  return po;
}
__device__ __forceinline__ void ray_offset(Ray& r, v3 dir) {
  const v3 end = r.o + r.tmax * r.d;
  r.o = offset_point(r.o, dir);
  r.tmax = length(end - r.o);
}
// CRay::spawnRay (ray.hpp:62-69)
__device__ __forceinline__ Ray spawn_ray(v3 start, v3 end, int medium) {
  v3 dir = end - start;
  const v3 off = offset_point(start, normalize(end - start));
  const float t = length(end - off);
  dir = dir / t;
  return mkray(off, dir, t, medium);
}

// ------------------------------------------------------------ records --
struct Rec {
  float density;
  float S[6];  // xx yy zz xy xz yz
  v3 diffuse, specular, normal;
  float ior;
};

// SFilteredData(const SFilteredDataCompact&) (filtered_data.hpp:63-89), from the
// eight 32-bit words of one 32-byte record.
__device__ __forceinline__ Rec decode_record(uint4 a, uint4 b) {
  Rec r;
  const float inv16 = 1.f / 65535.f;
  const float inv8 = 1.f / 255.f;
  r.density = u2f(a.x);
  const float sx = (float)(a.y & 0xffffu) * inv16, sy = (float)(a.y >> 16) * inv16;
  const float sz = (float)(a.z & 0xffffu) * inv16;
  const float rxy = 2.f * ((float)(a.z >> 16) * inv16) - 1.f;
  const float rxz = 2.f * ((float)(a.w & 0xffffu) * inv16) - 1.f;
  const float ryz = 2.f * ((float)(a.w >> 16) * inv16) - 1.f;
  r.S[0] = sx * sx; r.S[1] = sy * sy; r.S[2] = sz * sz;
  r.S[3] = rxy * sx * sy; r.S[4] = rxz * sx * sz; r.S[5] = ryz * sy * sz;
  r.diffuse = mk((float)(b.x & 0xffu) * inv8, (float)((b.x >> 8) & 0xffu) * inv8, (float)((b.x >> 16) & 0xffu) * inv8);
  r.specular = mk((float)(b.x >> 24) * inv8, (float)(b.y & 0xffu) * inv8, (float)((b.y >> 8) & 0xffu) * inv8);
  const float nx = (float)(b.y >> 16), ny = (float)(b.z & 0xffffu), nz = (float)(b.z >> 16);
  r.normal = normalize(mk(2.f * nx * inv16 - 1.f, 2.f * ny * inv16 - 1.f, 2.f * nz * inv16 - 1.f));
  r.ior = half_to_float((uint16_t)(b.w & 0xffffu));
  return r;
}

// Record precision of the lookups: R2_PREC_MODE 0 compiles the FP32-record path
// only, 1 the packed path only, 2 (default) both behind a runtime test.
#ifndef R2_PREC_MODE
#define R2_PREC_MODE 2
#endif
__device__ __forceinline__ bool grid_packed(const DGrid& g) {
#if R2_PREC_MODE == 0
  (void)g;
  return false;
#elif R2_PREC_MODE == 1
  (void)g;
  return true;
#else
  return g.precision == 1;
#endif
}

// packed record (r2_packed.hpp): FP16 density / S / ior, u8 colours, octahedral normal
__device__ __forceinline__ Rec decode_packed(uint4 h, uint2 c) {
  Rec r;
  r.density = half_to_float((uint16_t)(h.x & 0xffffu));
  r.S[0] = half_to_float((uint16_t)(h.x >> 16));
  r.S[1] = half_to_float((uint16_t)(h.y & 0xffffu));
  r.S[2] = half_to_float((uint16_t)(h.y >> 16));
  r.S[3] = half_to_float((uint16_t)(h.z & 0xffffu));
  r.S[4] = half_to_float((uint16_t)(h.z >> 16));
  r.S[5] = half_to_float((uint16_t)(h.w & 0xffffu));
  r.ior = half_to_float((uint16_t)(h.w >> 16));
  const float inv8 = 1.f / 255.f;
  r.diffuse = mk((float)(c.x & 0xffu) * inv8, (float)((c.x >> 8) & 0xffu) * inv8, (float)((c.x >> 16) & 0xffu) * inv8);
  r.specular = mk((float)(c.x >> 24) * inv8, (float)(c.y & 0xffu) * inv8, (float)((c.y >> 8) & 0xffu) * inv8);
  r.normal = oct_decode(c.y >> 16);
  return r;
}

// operator*(float, SFilteredData) and operator+ (filtered_data.hpp:115-151);
// interpolate(t, a, b) = (1 - t) * a + t * b (functions.hpp:66-72)
__device__ __forceinline__ Rec rec_lerp(float t, const Rec& a, const Rec& b) {
  const float wa = 1.f - t;
  Rec o;
  o.density = wa * a.density + t * b.density;
#pragma unroll
  for (int i = 0; i < 6; ++i) o.S[i] = wa * a.S[i] + t * b.S[i];
  o.diffuse = wa * a.diffuse + t * b.diffuse;
  o.specular = wa * a.specular + t * b.specular;
  o.normal = normalize(wa * normalize(a.normal) + t * normalize(b.normal));
  o.ior = wa * a.ior + t * b.ior;
  return o;
}

// index of entry (x, y, z) of a row-major 3-D table with rows of sx and planes of
// sx * sy entries, for the leaf, brick and mip tables.  The grid build allows
// fewer than 1024 bricks (and so leaves) per axis, so every operand stays below
// 2^20 and every product below 2^30: the full-rate 24-bit multiply is exact.
__device__ __forceinline__ uint32_t idx3(uint32_t x, uint32_t y, uint32_t z, uint32_t sx, uint32_t sy) {
  return __umul24(__umul24(z, sy) + y, sx) + x;
}

// leaf index of voxel (x, y, z), or -1 outside the grid or in an empty leaf
__device__ __forceinline__ int32_t leaf_of(const DGrid& g, int x, int y, int z) {
  if ((unsigned)x < (unsigned)g.nx && (unsigned)y < (unsigned)g.ny && (unsigned)z < (unsigned)g.nz)
    return g.leaf_table[idx3((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3, (uint32_t)g.nlx, (uint32_t)g.nly)];
  return -1;
}

// NanoVDB ReadAccessor<Vec4d>::getValue over the collapsed tree.
template <bool kCount>
__device__ __forceinline__ Rec get_value(const DGrid& g, int x, int y, int z, Cnt& cnt) {
  count<kCount>(cnt, kCntVox);
  int32_t leaf = -1;
  if ((unsigned)x < (unsigned)g.nx && (unsigned)y < (unsigned)g.ny && (unsigned)z < (unsigned)g.nz)
    leaf = g.leaf_table[idx3((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3, (uint32_t)g.nlx, (uint32_t)g.nly)];
  if (grid_packed(g)) {
    // background: the packed form of the FP32 background record (ior 1, normal
    // decoded from the zero code, normalize(-1,-1,-1), octahedral code 0xABAB)
    uint4 h = make_uint4(0u, 0u, 0u, 0x3C00u << 16);
    uint2 c = make_uint2(0u, 0xABABu << 16);
    if (leaf >= 0) {
      const size_t v = (size_t)leaf * kLeafVoxels + (((z & 7) << 6) | ((y & 7) << 3) | (x & 7));
      h = g.leaves_packed[v];
      c = g.leaves_cold[v];
    }
    return decode_packed(h, c);
  }
  uint4 a = make_uint4(0u, 0u, 0u, 0u), b = make_uint4(0u, 0u, 0u, 0x3C00u);  // background: ior = 1
  if (leaf >= 0) {
    const uint4* p = g.leaves + ((size_t)leaf * kLeafVoxels + (((z & 7) << 6) | ((y & 7) << 3) | (x & 7))) * 2;
    a = p[0];
    b = p[1];
  }
  return decode_record(a, b);
}

// CNVDBMedium::filteredData (nvdb_medium_impl.hpp:47-64), p normalised in [0,1]^3
template <bool kCount>
__device__ __forceinline__ Rec filtered_data(const DGrid& g, v3 p, Cnt& cnt) {
  const float psx = p.x * (float)g.size[0] + (float)g.ibb_min[0] - 0.5f;
  const float psy = p.y * (float)g.size[1] + (float)g.ibb_min[1] - 0.5f;
  const float psz = p.z * (float)g.size[2] + (float)g.ibb_min[2] - 0.5f;
  const int x = (int)floorf(psx), y = (int)floorf(psy), z = (int)floorf(psz);
  const float dx = psx - (float)x, dy = psy - (float)y, dz = psz - (float)z;
  if (grid_packed(g)) {
    // packed form: density, S, colours and ior as one weighted sum of the eight
    // corners; normals interpolated pairwise along x, y, z with renormalisation,
    // as operator*/operator+ do (filtered_data.hpp:115-151); corners are decoded
    // in pairs so that at most two decoded records are live
    Rec o;
    o.density = 0.f;
#pragma unroll
    for (int i = 0; i < 6; ++i) o.S[i] = 0.f;
    o.diffuse = o.specular = mk(0.f, 0.f, 0.f);
    o.ior = 0.f;
    auto acc = [&](float w, const Rec& r) {
      o.density += w * r.density;
#pragma unroll
      for (int i = 0; i < 6; ++i) o.S[i] += w * r.S[i];
      o.diffuse = o.diffuse + w * r.diffuse;
      o.specular = o.specular + w * r.specular;
      o.ior += w * r.ior;
    };
    auto nlerp = [](float t, v3 u, v3 v) { return normalize((1.f - t) * normalize(u) + t * normalize(v)); };
    // the eight leaf-table entries are loaded at once; the records of each
    // pair of corners are loaded when the pair is decoded (four more round
    // trips, with two corners' words live instead of eight)
    int32_t lf[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) lf[c] = leaf_of(g, x + (c & 1), y + ((c >> 1) & 1), z + (c >> 2));
#pragma unroll
    for (int c = 0; c < 8; ++c) count<kCount>(cnt, kCntVox);
    v3 e = mk(0.f, 0.f, 0.f), d0 = e, d1 = e;
#pragma unroll 1
    for (int pr = 0; pr < 4; ++pr) {
      const int cy = pr & 1, cz = pr >> 1;
      const int vy = y + cy, vz = z + cz;
      uint4 hw[2];
      uint2 cw[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        // background: the packed form of the FP32 background record (ior 1,
        // normal decoded from the zero code, normalize(-1,-1,-1): octahedral 0xABAB)
        hw[c] = make_uint4(0u, 0u, 0u, 0x3C00u << 16);
        cw[c] = make_uint2(0u, 0xABABu << 16);
        if (lf[c] >= 0) {
          const int vx = x + c;
          const size_t v = (size_t)lf[c] * kLeafVoxels + (((vz & 7) << 6) | ((vy & 7) << 3) | (vx & 7));
          hw[c] = g.leaves_packed[v];
          cw[c] = g.leaves_cold[v];
        }
      }
      // the next pair's leaf entries move down (constant indices keep them in registers)
#pragma unroll
      for (int c = 0; c < 6; ++c) lf[c] = lf[c + 2];
      const Rec ra = decode_packed(hw[0], cw[0]);
      acc((cy ? dy : 1.f - dy) * (cz ? dz : 1.f - dz) * (1.f - dx), ra);
      const Rec rb = decode_packed(hw[1], cw[1]);
      acc((cy ? dy : 1.f - dy) * (cz ? dz : 1.f - dz) * dx, rb);
      const v3 f = nlerp(dx, ra.normal, rb.normal);
      if (cy == 0) e = f;
      else if (cz == 0) d0 = nlerp(dy, e, f);
      else d1 = nlerp(dy, e, f);
    }
    o.normal = nlerp(dz, d0, d1);
    return o;
  }
  // FP32 form: the leaf-table entries of the eight corners, then all sixteen
  // record words are requested before any is used (two memory round trips
  // instead of eight); then the same arithmetic as d00, d10 -> d0, d01, d11 ->
  // d1, then z, with at most three decoded records live at a time
  uint4 ra[8], rb[8];
  {
    int32_t lf[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) lf[c] = leaf_of(g, x + (c & 1), y + ((c >> 1) & 1), z + (c >> 2));
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      ra[c] = make_uint4(0u, 0u, 0u, 0u);
      rb[c] = make_uint4(0u, 0u, 0u, 0x3C00u);  // background: ior = 1
      if (lf[c] >= 0) {
        const int vx = x + (c & 1), vy = y + ((c >> 1) & 1), vz = z + (c >> 2);
        const uint4* q = g.leaves + ((size_t)lf[c] * kLeafVoxels + (((vz & 7) << 6) | ((vy & 7) << 3) | (vx & 7))) * 2;
        ra[c] = q[0];
        rb[c] = q[1];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) count<kCount>(cnt, kCntVox);
  Rec a = decode_record(ra[0], rb[0]), b = decode_record(ra[1], rb[1]);
  Rec e = rec_lerp(dx, a, b);
  a = decode_record(ra[2], rb[2]);
  b = decode_record(ra[3], rb[3]);
  const Rec d0 = rec_lerp(dy, e, rec_lerp(dx, a, b));
  a = decode_record(ra[4], rb[4]);
  b = decode_record(ra[5], rb[5]);
  e = rec_lerp(dx, a, b);
  a = decode_record(ra[6], rb[6]);
  b = decode_record(ra[7], rb[7]);
  const Rec d1 = rec_lerp(dy, e, rec_lerp(dx, a, b));
  return rec_lerp(dz, d0, d1);
}

// CDeviceBrickGrid::lookupMajorant (device_grid_brick_impl.hpp:29-33)
// texel of brick (bx, by, bz) at mip level `mip` of the range texture.  Brick
// counts are multiples of 2^3 (host_grid_brick.cpp:64), so the point-sampled texel
// at the normalised coordinate (b + 0.5) / n of a level of extent n >> mip
// (tex3DLod, device_grid_brick_impl.hpp:29-33) is b >> mip.
__device__ __forceinline__ uint32_t mip_texel(const DGrid& g, int bx, int by, int bz, int mip) {
  return idx3((uint32_t)(bx >> mip), (uint32_t)(by >> mip), (uint32_t)(bz >> mip), (uint32_t)(g.nb[0] >> mip),
              (uint32_t)(g.nb[1] >> mip));
}
template <bool kCount>
__device__ __forceinline__ float lookup_majorant(const DGrid& g, v3 ipos, int mip, Cnt& cnt) {
  count<kCount>(cnt, kCntMaj);
  const int bx = ((int)floorf(ipos.x)) >> 3, by = ((int)floorf(ipos.y)) >> 3, bz = ((int)floorf(ipos.z)) >> 3;
  if ((unsigned)bx >= (unsigned)g.nb[0] || (unsigned)by >= (unsigned)g.nb[1] || (unsigned)bz >= (unsigned)g.nb[2])
    return 0.f;
  // The grid build keeps the four levels in one buffer, level l at (texels of
  // level 0) / 64 x {0, 64, 72, 73}[l]: the lane's level finds its table by
  // arithmetic (indexing g.maj with it would load the pointer from memory, one
  // more dependent load on every DDA step).  Level-0 texels are fewer than 2^30
  // (fewer than 1024 bricks per axis), so the 24-bit products are exact.
  const uint32_t n64 = (uint32_t)(g.nb[0] * g.nb[1] * g.nb[2]) >> 6;
  const uint32_t k = (0x49484000u >> (8 * mip)) & 0xffu;
  return g.maj[0][__umul24(n64, k) + mip_texel(g, bx, by, bz, mip)];
}
// CDeviceBrickGrid::lookupDensity (device_grid_brick_impl.hpp:9-27)
template <bool kCount>
__device__ __forceinline__ float lookup_density(const DGrid& g, v3 ipos, v3 u, Cnt& cnt) {
  count<kCount>(cnt, kCntBrick);
  const int ix = (int)floorf(ipos.x + u.x - 0.5f), iy = (int)floorf(ipos.y + u.y - 0.5f),
            iz = (int)floorf(ipos.z + u.z - 0.5f);
  const int bx = ix >> 3, by = iy >> 3, bz = iz >> 3;
  if ((unsigned)bx >= (unsigned)g.nb[0] || (unsigned)by >= (unsigned)g.nb[1] || (unsigned)bz >= (unsigned)g.nb[2])
    return 0.f;
  const uint2 br = g.bricks[idx3((uint32_t)bx, (uint32_t)by, (uint32_t)bz, (uint32_t)g.nb[0], (uint32_t)g.nb[1])];
  const float lo = range_lo(br.x), hi = range_hi(br.x);
  uint8_t v = 0;
  if (br.y != 0xffffffffu) v = g.atlas[(size_t)br.y * 512 + (((iz & 7) << 6) | ((iy & 7) << 3) | (ix & 7))];
  return decode_voxel(v, lo, hi);
}

// CNVDBMedium::stepDDA (nvdb_medium_impl.hpp:241-246)
__device__ __forceinline__ float step_dda(v3 pos, v3 inv_dir, int mip) {
  // dim = 8 << mip and its reciprocal, both powers of two: exact by ldexp, without
  // a conversion and a quarter-rate reciprocal
  const float dim = ldexpf(1.f, 3 + mip);
  const float ox = inv_dir.x >= 0.f ? dim + 0.5f : -0.5f;
  const float oy = inv_dir.y >= 0.f ? dim + 0.5f : -0.5f;
  const float oz = inv_dir.z >= 0.f ? dim + 0.5f : -0.5f;
  const float inv_dim = ldexpf(1.f, -3 - mip);
  const float tx = (floorf(pos.x * inv_dim) * dim + ox - pos.x) * inv_dir.x;
  const float ty = (floorf(pos.y * inv_dim) * dim + oy - pos.y) * inv_dir.y;
  const float tz = (floorf(pos.z * inv_dim) * dim + oz - pos.z) * inv_dir.z;
  const float myz = (tz < ty) ? tz : ty;
  return (myz < tx) ? myz : tx;
}

// CFresnel::evaluate (fresnel.hpp:28-54), eta_i = 1, eta_t = ior
__device__ __forceinline__ float fresnel(float cos_i, float ior) {
  cos_i = cos_i < -1.f ? -1.f : (cos_i > 1.f ? 1.f : cos_i);
  float eta_i = 1.f, eta_t = ior;
  if (!(cos_i > 0.f)) {
    eta_i = ior; eta_t = 1.f;
    cos_i = fabsf(cos_i);
  }
  const float sin_i = sqrtf(gmax(0.f, 1.f - cos_i * cos_i));
  const float sin_t = eta_i / eta_t * sin_i;
  if (sin_t >= 1.f) return 1.f;
  const float cos_t = sqrtf(gmax(0.f, 1.f - sin_t * sin_t));
  const float rpar = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
  const float rper = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
  return (rpar * rpar + rper * rper) / 2.f;
}

struct MHit {
  int valid;
  int inst;
  v3 pos, normal;
  float S[6];
  float ior;
};

// tail of sampleInternal / sampleDDA at a real collision (nvdb_medium_impl.hpp:120-131,169-180)
__device__ __forceinline__ v3 collision_albedo(const DGrid& g, const Rec& fd, const Ray& rm, MHit& mi) {
  mi.valid = 1;
  mi.pos = rm.o + rm.tmax * rm.d;
  const v3 n = normalize(fd.normal);
  mi.normal = n;
#pragma unroll
  for (int i = 0; i < 6; ++i) mi.S[i] = fd.S[i];
  mi.ior = fd.ior;
  const float w = fresnel(fabsf(dot(normalize(-rm.d), n)), fd.ior);
  const v3 b = fd.diffuse * (1.f - w) + fd.specular * w;
  return (b * mk(g.sigma_s[0], g.sigma_s[1], g.sigma_s[2])) / g.sigma_t;
}

// x / size[i] and x / sigma_t of a grid: with a power-of-two divisor the product
// with the exact reciprocal is the same number as the IEEE quotient, so the
// (wave-uniform) branch skips the division without changing any result
__device__ __forceinline__ float div_size(const DGrid& g, float x, int i) {
  return g.size_pow2 ? x * g.inv_size[i] : x / (float)g.size[i];
}
__device__ __forceinline__ float div_sigma_t(const DGrid& g, float x) {
  return g.sigma_t_pow2 ? x * g.inv_sigma_t : x / g.sigma_t;
}

// CNVDBMedium::moveToVoxelBorder (nvdb_medium_impl.hpp:296-320), ray in model space
__device__ __forceinline__ Ray medium_move_to_voxel_border(const DGrid& g, const Ray& ray) {
  const Ray ri = ray_transform(ray, g.m2i);
  const float gamma = (3.f * FLT_EPSILON) / (1.f - 3.f * FLT_EPSILON);
  float t1 = ri.tmax;
  const float o[3] = {ri.o.x, ri.o.y, ri.o.z};
  const float d[3] = {ri.d.x, ri.d.y, ri.d.z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float size = (float)g.size[i], ibb = (float)g.ibb_min[i];
    const float ip = o[i] * size + ibb;
    const float vmin = div_size(g, floorf(ip) - ibb, i);
    const float vmax = div_size(g, ceilf(ip) - ibb, i);
    const float inv = 1.f / d[i];
    float tn = (vmin - o[i]) * inv;
    float tf = (vmax - o[i]) * inv;
    if (tn > tf) { const float tmp = tn; tn = tf; tf = tmp; }
    tf *= 1.f + 2.f * gamma;
    t1 = tf < t1 ? tf : t1;
  }
  t1 *= (1.f - kRayOffset);
  const Ray moved = mkray(ri.o + t1 * ri.d, ri.d, ri.tmax - t1, ri.medium);
  Ray r = ray_transform(moved, g.i2m);
  ray_offset(r, -ray.d);
  return r;
}

// ----------------------------------------------------------------- SGGX --
struct Sggx {
  float S[3][3];
  float invS[3][3];
  v3 n;
};
__device__ __forceinline__ float det3(const float m[3][3]) {
  return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
         m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}
__device__ __forceinline__ v3 mul3(const float m[3][3], v3 v) {
  return mk(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
            m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}
// CSGGXMicroflakeDistribution ctor (sggx_phase_function.hpp:89-93): S and glm::inverse(S)
__device__ __forceinline__ Sggx sggx_make(const float S6[6], v3 n) {
  Sggx g;
  g.S[0][0] = S6[0]; g.S[0][1] = S6[3]; g.S[0][2] = S6[4];
  g.S[1][0] = S6[3]; g.S[1][1] = S6[1]; g.S[1][2] = S6[5];
  g.S[2][0] = S6[4]; g.S[2][1] = S6[5]; g.S[2][2] = S6[2];
  const float(*m)[3] = g.S;
  const float o = 1.f / det3(m);
  g.invS[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * o;
  g.invS[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * o;
  g.invS[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * o;
  g.invS[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * o;
  g.invS[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * o;
  g.invS[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * o;
  g.invS[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * o;
  g.invS[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * o;
  g.invS[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * o;
  g.n = n;
  return g;
}
constexpr float kPi = 3.14159265358979323846f;
// D (sggx_phase_function.hpp:107-110)
__device__ __forceinline__ float sggx_D(const Sggx& g, v3 w) {
  const float v = dot(w, mul3(g.invS, w));
  return 1.f / (kPi * sqrtf(fabsf(det3(g.S))) * v * v);
}
// sigma (sggx_phase_function.hpp:112-114)
__device__ __forceinline__ float sggx_sigma(const Sggx& g, v3 w) { return sqrtf(gmax(dot(w, mul3(g.S, w)), 0.f)); }
// CCoordinateFrame::fromNormal2 (objects.hpp:42-58)
__device__ __forceinline__ void frame_frisvad(v3 N, v3& T, v3& B) {
  if (N.z < -0.9999999f) {
    T = mk(0.f, -1.f, 0.f);
    B = mk(-1.f, 0.f, 0.f);
  } else {
    const float a = 1.f / (1.f + N.z);
    const float b = -N.x * N.y * a;
    T = normalize(mk(1.f - N.x * N.x * a, b, -N.x));
    B = normalize(mk(b, 1.f - N.y * N.y * a, -N.y));
  }
}
// CCoordinateFrame::fromNormal (objects.hpp:28-40)
__device__ __forceinline__ void frame_basic(v3 N, v3& T, v3& B) {
  if (fabsf(N.x) > fabsf(N.y))
    T = mk(-N.z, 0.f, N.x) / sqrtf(N.x * N.x + N.z * N.z);
  else
    T = mk(0.f, N.z, -N.y) / sqrtf(N.y * N.y + N.z * N.z);
  B = normalize(cross(N, T));
}
// sampleVNDF (sggx_phase_function.hpp:116-145), about the voxel normal
__device__ __forceinline__ v3 sggx_vndf(const Sggx& g, float ux, float uy) {
  v3 wj, wk;
  frame_frisvad(g.n, wj, wk);
  const v3 wi = g.n;
  const v3 Swk = mul3(g.S, wk), Swj = mul3(g.S, wj), Swi = mul3(g.S, wi);
  const float Skk = dot(wk, Swk), Skj = dot(wk, Swj), Ski = dot(wk, Swi);
  const float Sjj = dot(wj, Swj), Sji = dot(wj, Swi), Sii = dot(wi, Swi);
  float P[3][3];
  P[0][0] = Skk; P[0][1] = Skj; P[0][2] = Ski;
  P[1][0] = Skj; P[1][1] = Sjj; P[1][2] = Sji;
  P[2][0] = Ski; P[2][1] = Sji; P[2][2] = Sii;
  const float tmp = sqrtf(Sjj * Sii - Sji * Sji);
  const v3 Mk = mk(sqrtf(fabsf(det3(P))) / tmp, 0.f, 0.f);
  const float sc = 1.f / sqrtf(Sii);
  const v3 Mj = mk(-(Ski * Sji - Skj * Sii) / tmp, tmp, 0.f) * sc;
  const v3 Mi = mk(Ski, Sji, Sii) * sc;
  const float sux = sqrtf(ux);
  const float u = sux * cosf(2.f * kPi * uy);
  const float v = sux * sinf(2.f * kPi * uy);
  const float w = sqrtf(1.f - u * u - v * v);
  v3 wm = (u * Mk + v * Mj) + w * Mi;
  wm = normalize(wm);
  const v3 r = mk(wk.x * wm.x + wj.x * wm.y + wi.x * wm.z, wk.y * wm.x + wj.y * wm.y + wi.y * wm.z,
                  wk.z * wm.x + wj.z * wm.y + wi.z * wm.z);
  return normalize(r);
}
// CSGGXSpecularPhaseFunction::p (sggx_phase_function.hpp:168-171)
__device__ __forceinline__ float sggx_spec_p(const Sggx& g, v3 wo, v3 wi) {
  const v3 wh = normalize(wi + wo);
  return sggx_D(g, wh) / (4.f * sggx_sigma(g, wo));
}
// concentricSampleDisk + cosineSampleHemisphere (sampler.hpp:52-80)
template <typename R>
__device__ __forceinline__ v3 cosine_hemisphere(R& rng) {
  const float r1 = rng.uni(), r2 = rng.uni();
  const float ox = 2.f * r1 - 1.f, oy = 2.f * r2 - 1.f;
  v3 d = mk(0.f, 0.f, 0.f);
  if (!(ox == 0.f && oy == 0.f)) {
    float theta, r;
    if (fabsf(ox) > fabsf(oy)) {
      r = ox;
      theta = (kPi / 4.f) * (oy / ox);
    } else {
      r = oy;
      theta = kPi / 2.f - (kPi / 4.f) * (ox / oy);
    }
    d = r * mk(cosf(theta), sinf(theta), 0.f);
  }
  const float z = sqrtf(gmax(0.f, 1.f - d.x * d.x - d.y * d.y));
  return mk(d.x, d.y, z);
}
// CSGGXPhaseFunction::p (sggx_phase_function.hpp:214-218)
template <typename R>
__device__ __forceinline__ float phase_p(const float S6[6], v3 n, float ior, v3 wo, v3 wi, R& rng) {
  const Sggx g = sggx_make(S6, n);
  const float w = fresnel(fabsf(dot(normalize(wo), n)), ior);
  const float ux = rng.uni(), uy = rng.uni();
  const v3 wm = sggx_vndf(g, ux, uy);
  const float pd = gmax(dot(wi, wm), 0.f) / kPi;
  return pd * (1.f - w) + sggx_spec_p(g, wo, wi) * w;
}
// CSGGXPhaseFunction::sampleP (sggx_phase_function.hpp:220-229).  Both lobes
// start by drawing (ux, uy) and sampling the VNDF with them (the specular lobe's
// microflake normal, the diffuse lobe's frame normal), so that part runs once,
// before the lobe choice: same draws in the same order, same arithmetic, and a
// wavefront whose lanes choose different lobes does not run it twice.
template <typename R>
__device__ __forceinline__ float phase_sample(const float S6[6], v3 n, float ior, v3 wo, v3& wi, R& rng) {
  const Sggx g = sggx_make(S6, n);
  const float pf = fresnel(fabsf(dot(normalize(wo), n)), ior);
  const bool specular = rng.uni() < pf;
  const float ux = rng.uni(), uy = rng.uni();
  const v3 wn = sggx_vndf(g, ux, uy);
  if (specular) {
    const v3 I = -wo;
    wi = I - (wn * dot(wn, I)) * 2.f;
    return sggx_spec_p(g, wo, wi);
  }
  v3 T, B;
  frame_basic(wn, T, B);
  const v3 wm = cosine_hemisphere(rng);
  const v3 ww = mk(T.x * wm.x + B.x * wm.y + wn.x * wm.z, T.y * wm.x + B.y * wm.y + wn.y * wm.z,
                   T.z * wm.x + B.z * wm.y + wn.z * wm.z);
  wi = normalize(ww);
  return gmax(dot(wi, wn), 0.f) / kPi;
}

// ------------------------------------------------------------------ env --
// find_interval over the bracket a guide table gives: the CDF is non-decreasing,
// so the number of values <= u lies between the counts at j / g and (j + 1) / g
// for j = floor(u * g), and the same bisection over that bracket returns the
// same interval as over the whole table (u in [0, 1), g a power of two)
__device__ __forceinline__ int find_interval_guided(const float* cdf, int size, float u, const int32_t* guide, int g) {
  int j = (int)(u * (float)g);
  j = j < 0 ? 0 : (j > g - 1 ? g - 1 : j);
  int first = guide[j], len = guide[j + 1] - first;
  while (len > 0) {
    const int half = len >> 1, mid = first + half;
    if (cdf[mid] <= u) {
      first = mid + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  const int r = first - 1;
  return r < 0 ? 0 : (r > size - 2 ? size - 2 : r);
}
// CDistribution1D::sampleContinuous (distribution_1d.hpp:54-75)
__device__ __forceinline__ float sample_1d(const float* f, const float* cdf, float fint, int n, float u, float& pdf,
                                           int& off, const int32_t* guide, int g) {
  const int o = find_interval_guided(cdf, n + 1, u, guide, g);
  off = o;
  const float c0 = cdf[o], c1 = cdf[o + 1];
  float du = u - c0;
  if ((c1 - c0) > 0.f) du /= (c1 - c0);
  pdf = f[o] / fint;
  return ((float)o + du) / (float)n;
}
// bilinear, wrap-addressed, normalised-coordinate texel fetch (texture.cu:166-189)
__device__ __forceinline__ v3 env_tex(const DEnv& e, float x, float y) {
  const int W = e.w, H = e.h;
  x = x - floorf(x);
  y = y - floorf(y);
  const float tx = x * (float)W - 0.5f, ty = y * (float)H - 0.5f;
  const float fx = floorf(tx), fy = floorf(ty);
  const float a = tx - fx, b = ty - fy;
  int i0 = (int)fx, j0 = (int)fy;
  int i1 = i0 + 1, j1 = j0 + 1;
  i0 = ((i0 % W) + W) % W; i1 = ((i1 % W) + W) % W;
  j0 = ((j0 % H) + H) % H; j1 = ((j1 % H) + H) % H;
  const float* t = e.tex;
  const float* p00 = t + ((size_t)j0 * W + i0) * 3;
  const float* p10 = t + ((size_t)j0 * W + i1) * 3;
  const float* p01 = t + ((size_t)j1 * W + i0) * 3;
  const float* p11 = t + ((size_t)j1 * W + i1) * 3;
  float r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    r[c] = (1.f - a) * (1.f - b) * p00[c] + a * (1.f - b) * p10[c] + (1.f - a) * b * p01[c] + a * b * p11[c];
  return mk(r[0], r[1], r[2]);
}
// CEnvironmentMap::sample (environmentmap.hpp:28-42)
template <bool kCount, typename R>
__device__ __forceinline__ v3 env_sample(const DEnv& e, R& rng, v3& dir, float& pdf, Cnt& cnt) {
  count<kCount>(cnt, kCntEnv);
  float p1, p0;
  int v, dummy;
  const float y = sample_1d(e.marg_func, e.marg_cdf, e.marg_int, e.h, rng.uni(), p1, v, e.marg_guide, e.marg_g);
  const float x = sample_1d(e.row_func + (size_t)v * e.w, e.row_cdf + (size_t)v * (e.w + 1), e.row_int[v], e.w,
                            rng.uni(), p0, dummy, e.row_guide + (size_t)v * (e.row_g + 1), e.row_g);
  pdf = p0 * p1;
  const float phi = x * 2.f * kPi;
  const float theta = y * kPi;
  const float st = sinf(theta);
  if (st == 0.f) pdf = 0.f;
  else pdf /= (st * 2.f * kPi * kPi);
  dir = mk(sinf(phi) * st, cosf(theta), cosf(phi) * st);
  return env_tex(e, x, y);
}
// CDistribution2D::pdf (distribution_2d.hpp:73-84)
__device__ __forceinline__ float dist2d_pdf(const DEnv& e, float px, float py) {
  const int W = e.w, H = e.h;
  int lr = (int)(py * (float)(H - 2));
  lr = lr < 0 ? 0 : (lr > H - 2 ? H - 2 : lr);
  const int ur = lr + 1;
  int lc = (int)(px * (float)(W - 2));
  lc = lc < 0 ? 0 : (lc > W - 2 ? W - 2 : lc);
  const int uc = lc + 1;
  const float* f = e.row_func;
  const float sx = px * ((float)W - 2.f), sy = py * ((float)H - 2.f);
  const float lo = f[(size_t)lr * W + lc] * ((float)uc - sx) + f[(size_t)lr * W + uc] * (sx - (float)lc);
  const float up = f[(size_t)ur * W + lc] * ((float)uc - sx) + f[(size_t)ur * W + uc] * (sx - (float)lc);
  return lo * ((float)ur - sy) + up * (sy - (float)lr);
}
// CEnvironmentMap::le (environmentmap.hpp:44-55)
template <bool kCount>
__device__ __forceinline__ v3 env_le(const DEnv& e, v3 d, float& pdf, Cnt& cnt) {
  count<kCount>(cnt, kCntEnv);
  const float yc = d.y < -1.f ? -1.f : (d.y > 1.f ? 1.f : d.y);
  const float theta = acosf(yc);
  const float p = atan2f(d.x, d.z);
  const float phi = p < 0.f ? (p + 2.f * kPi) : p;
  const float x = phi / (2.f * kPi);
  const float y = theta / kPi;
  const float st = sinf(theta);
  pdf = st == 0.f ? 0.f : dist2d_pdf(e, x, y) / (2.f * kPi * kPi * st);
  return env_tex(e, x, y);
}

// ---------------------------------------------------------------- scene --
}  // namespace r2
#include "r2_surface.hpp"
namespace r2 {

struct Hit {
  int hit;
  float t;
  v3 pos;
  int medium;
  int surf;   // surface index, -1 for a box or a miss
  v3 normal;  // a surface's normalised normal
};
// instance k of the scene (the table follows the scene in device memory)
__device__ __forceinline__ const DInstance& inst_at(const DScene& s, int k) { return kconst_at(s.inst, k); }

// The slab test of nanovdb::Ray::intersects (Williams et al. 2005) over the
// world box [bmin, bmax] in the reference's order of comparisons, and
// BBox::isInside: the t that __intersection__volume reports (the exit when the
// origin is inside, else the entry), or false.
__device__ __forceinline__ bool instance_box_t(const float* bmin, const float* bmax, const float o[3],
                                               const float inv[3], int sg0, int sg1, int sg2, float& t_out) {
  const float span1 = kDefaultTmax;
  float t0 = ((sg0 ? bmax : bmin)[0] - o[0]) * inv[0];
  float t2 = ((sg1 ? bmin : bmax)[1] - o[1]) * inv[1];
  if (t0 > t2) return false;
  float t1 = ((sg0 ? bmin : bmax)[0] - o[0]) * inv[0];
  float t3 = ((sg1 ? bmax : bmin)[1] - o[1]) * inv[1];
  if (t3 > t1) return false;
  if (t3 > t0) t0 = t3;
  if (t2 < t1) t1 = t2;
  t3 = ((sg2 ? bmax : bmin)[2] - o[2]) * inv[2];
  if (t3 > t1) return false;
  t2 = ((sg2 ? bmin : bmax)[2] - o[2]) * inv[2];
  if (t0 > t2) return false;
  if (t3 > t0) t0 = t3;
  if (span1 < t0) return false;
  if (t2 < t1) t1 = t2;
  if (0.f > t1) return false;
  if (0.f > t0) t0 = 0.f;
  if (span1 < t1) t1 = span1;
  if (!(t1 < kDefaultTmax)) return false;
  const bool inside =
      o[0] > bmin[0] && o[1] > bmin[1] && o[2] > bmin[2] && o[0] < bmax[0] && o[1] < bmax[1] && o[2] < bmax[2];
  t_out = inside ? t1 : t0;
  return true;
}
// Culling test of a hierarchy node: the same rejections as instance_box_t on
// the node's box, which contains its instances' boxes.  Each rejection compares
// an entry against an exit, and a contained box has a later (or equal) entry
// and an earlier (or equal) exit under the same rounded arithmetic, so a node
// is rejected only when all its instances are (comparisons with NaN never
// reject).  The node is also skipped when its entry lies beyond `cur`: every
// instance inside then reports a larger t.
__device__ __forceinline__ bool node_may_hit(const DBvhNode& nd, const float o[3], const float inv[3], int sg0,
                                             int sg1, int sg2, float cur) {
  const float* bmin = nd.lo;
  const float* bmax = nd.hi;
  float t0 = ((sg0 ? bmax : bmin)[0] - o[0]) * inv[0];
  float t2 = ((sg1 ? bmin : bmax)[1] - o[1]) * inv[1];
  if (t0 > t2) return false;
  float t1 = ((sg0 ? bmin : bmax)[0] - o[0]) * inv[0];
  float t3 = ((sg1 ? bmax : bmin)[1] - o[1]) * inv[1];
  if (t3 > t1) return false;
  if (t3 > t0) t0 = t3;
  if (t2 < t1) t1 = t2;
  t3 = ((sg2 ? bmax : bmin)[2] - o[2]) * inv[2];
  if (t3 > t1) return false;
  t2 = ((sg2 ? bmin : bmax)[2] - o[2]) * inv[2];
  if (t0 > t2) return false;
  if (t3 > t0) t0 = t3;
  if (kDefaultTmax < t0) return false;
  if (t2 < t1) t1 = t2;
  if (0.f > t1) return false;
  return !(t0 > cur);
}

// __intersection__volume (shaders.optix.cu:192-236) over the instance boxes, then
// (kSurf) the analytic surfaces through __intersection__surface (:170-189);
// closest accepted hit wins, the lowest instance index among equal t (the order
// of a linear scan, which the hierarchy walk reproduces).
template <bool kSurf = false, bool kSingle = false>
__device__ __forceinline__ Hit scene_intersect(const DScene& s, Ray& ray) {
  Hit h;
  h.hit = 0; h.t = 0.f; h.medium = -1; h.pos = mk(0.f, 0.f, 0.f);
  h.surf = -1; h.normal = mk(0.f, 0.f, 0.f);
  const float best = ray.tmax;
  const float o[3] = {ray.o.x, ray.o.y, ray.o.z};
  const float inv[3] = {1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z};
  const int sg0 = inv[0] < 0.f, sg1 = inv[1] < 0.f, sg2 = inv[2] < 0.f;
  if (kSingle || s.n_bvh == 0) {
    const int n = kSingle ? 1 : s.n_inst;
    for (int k = 0; k < n; ++k) {
      const DInstance& in = inst_at(s, k);
      float t;
      if (!instance_box_t(in.bb_min, in.bb_max, o, inv, sg0, sg1, sg2, t)) continue;
      if (!(t >= 0.f && t <= best)) continue;
      if (h.hit && !(t < h.t)) continue;
      h.hit = 1; h.t = t; h.medium = k;
    }
  } else {
    for (int i = 0; i < s.n_bvh;) {
      const DBvhNode nd = kconst_at(s.bvh, i);
      if (!node_may_hit(nd, o, inv, sg0, sg1, sg2, h.hit ? h.t : best)) {
        i = nd.skip;
        continue;
      }
      if (nd.leaf >= 0) {
        const int first = nd.leaf >> 3, cnt = nd.leaf & 7;
        for (int j = 0; j < cnt; ++j) {
          const int k = kconst_at(s.bvh_idx, first + j);
          const DInstance& in = inst_at(s, k);
          float t;
          if (!instance_box_t(in.bb_min, in.bb_max, o, inv, sg0, sg1, sg2, t)) continue;
          if (!(t >= 0.f && t <= best)) continue;
          if (h.hit && !(t < h.t || (t == h.t && k < h.medium))) continue;
          h.hit = 1; h.t = t; h.medium = k;
        }
      }
      ++i;
    }
  }
  if constexpr (kSurf) {
    for (int j = 0; j < s.n_surf; ++j) {
      float t;
      v3 n;
      if (!surface_hit(s.surf[j], ray, h.hit ? h.t : best, t, n) || (h.hit && !(t < h.t))) continue;
      h.hit = 1; h.t = t; h.medium = -1; h.surf = j; h.normal = n;
    }
    if (h.surf >= 0) h.normal = normalize(h.normal);
  }
  if (h.hit) {
    h.pos = mk(ray.o.x + h.t * ray.d.x, ray.o.y + h.t * ray.d.y, ray.o.z + h.t * ray.d.z);
    ray.tmax = h.t;
  }
  return h;
}

// CMediumInstance::moveToVoxelBorder (medium_instance_impl.hpp:38-41)
__device__ __forceinline__ Ray inst_move_to_voxel_border(const DScene& s, int k, const Ray& ray) {
  const DInstance& in = inst_at(s, k);
  const Ray rm = ray_transform(ray, in.w2m);
  const Ray moved = medium_move_to_voxel_border(s.grid[in.grid], rm);
  Ray r = ray_transform(moved, in.m2w);
  ray_offset(r, -ray.d);
  return r;
}
// CPixelSampler::samplePixel (pixel_sampler.hpp:36-47)
template <typename R>
__device__ __forceinline__ Ray camera_ray(const DScene& s, int px, int py, R& rng) {
  const float rh = rng.uni();
  const float h = ((float)(px - s.W / 2) + rh) * s.pixel_size;
  (void)rng.uni();  // randomVertical: drawn, unused
  const float v = ((float)(py - s.H / 2) + rh) * s.pixel_size;
  const v3 d = normalize(mk(h, v, -s.near_dist));
  const Ray vr = mkray(mk(0.f, 0.f, 0.f), d, kDefaultTmax, -1);
  return ray_transform(vr, s.view_to_world);
}

}  // namespace r2
