// Seeded synthetic inputs of the benchmark workloads (host, multi-threaded):
//   kind 1: rho = scale * VN_k(x), k-octave value noise (lattice hash, smoothstep
//           trilinear, base 4 cells per extent, lacunarity 2, gain 0.5,
//           normalised to [0,1]); S = I; normal +y; colours 1; ior 1.5.
//   kind 2: as kind 1, plus fibre axis t ~ vMF(mu(x), kappa = 20) (Wood 1994)
//           around mu(x) = normalize(3-channel low-frequency noise - 0.5);
//           S = R diag(1,1,0.05) R^T = I - 0.95 t t^T; normal = unit vector
//           perpendicular to t.
//   env kind 0: constant white; kind 1: sky gradient + 0.5 degree sun scaled to
//           a 1e4:1 texel dynamic range.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>
#include "../../include/r2.h"
#include "r2_codec.hpp"

namespace {
using namespace r2;

inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
inline float lattice(uint32_t seed, int x, int y, int z) {
  uint32_t h = fmix32(seed * 0x9E3779B1u ^ fmix32((uint32_t)x * 0x85EBCA77u ^ fmix32((uint32_t)y * 0xC2B2AE3Du ^
                                                                                     fmix32((uint32_t)z * 0x27D4EB2Fu))));
  return (float)(h >> 8) * (1.f / 16777216.f);
}
inline float smooth(float f) { return f * f * (3.f - 2.f * f); }
float value_noise(uint32_t seed, float px, float py, float pz) {
  const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const float sx = smooth(px - fx), sy = smooth(py - fy), sz = smooth(pz - fz);
  float c[2][2][2];
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) c[dz][dy][dx] = lattice(seed, ix + dx, iy + dy, iz + dz);
  const float x00 = c[0][0][0] + sx * (c[0][0][1] - c[0][0][0]);
  const float x10 = c[0][1][0] + sx * (c[0][1][1] - c[0][1][0]);
  const float x01 = c[1][0][0] + sx * (c[1][0][1] - c[1][0][0]);
  const float x11 = c[1][1][0] + sx * (c[1][1][1] - c[1][1][0]);
  const float y0 = x00 + sy * (x10 - x00), y1 = x01 + sy * (x11 - x01);
  return y0 + sz * (y1 - y0);
}
float fbm(uint32_t seed, int octaves, float base, float x, float y, float z) {
  float sum = 0.f, norm = 0.f, amp = 1.f, freq = base;
  for (int o = 0; o < octaves; ++o) {
    sum += amp * value_noise(seed + (uint32_t)o * 1013u, x * freq, y * freq, z * freq);
    norm += amp;
    amp *= 0.5f;
    freq *= 2.f;
  }
  return sum / norm;
}

template <typename F>
void parallel_for(int64_t n, F&& f) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 32u));
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t a = (int64_t)t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=, &f]() {
      for (int64_t i = a; i < b; ++i) f(i);
    });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" int r2_synth_volume(int32_t kind, int32_t n, uint32_t seed, float scale, int32_t octaves, void* out) {
  if ((kind != 1 && kind != 2) || n <= 0 || n > 2048 || !out || octaves <= 0 || octaves > 12) return R2_ERR_INVALID;
  uint32_t* w = (uint32_t*)out;
  const float inv = 1.f / (float)n;
  parallel_for((int64_t)n * n, [&](int64_t zy) {
    const int z = (int)(zy / n), y = (int)(zy % n);
    for (int x = 0; x < n; ++x) {
      const size_t i = ((size_t)z * n + y) * n + x;
      const float px = ((float)x + 0.5f) * inv, py = ((float)y + 0.5f) * inv, pz = ((float)z + 0.5f) * inv;
      const float rho = scale * fbm(seed, octaves, 4.f, px, py, pz);
      float S6[6] = {1.f, 1.f, 1.f, 0.f, 0.f, 0.f};
      float nrm[3] = {0.f, 1.f, 0.f};
      const float one[3] = {1.f, 1.f, 1.f};
      if (kind == 2) {
        v3 mu = mk(fbm(seed + 100u, 1, 2.f, px, py, pz) - 0.5f, fbm(seed + 200u, 1, 2.f, px, py, pz) - 0.5f,
                   fbm(seed + 300u, 1, 2.f, px, py, pz) - 0.5f);
        const float ml = length(mu);
        mu = ml > 1e-6f ? mu / ml : mk(0.f, 1.f, 0.f);
        uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, 0x5F17u};
        philox4x32_10(c, seed, 0xF1B3u);
        const float u1 = (float)(c[0] >> 8) * (1.f / 16777216.f), u2 = (float)(c[1] >> 8) * (1.f / 16777216.f);
        const double kappa = 20.0;
        const double wv = 1.0 + std::log((double)u1 + (1.0 - (double)u1) * std::exp(-2.0 * kappa)) / kappa;
        const double sw = std::sqrt(std::max(0.0, 1.0 - wv * wv));
        const v3 e = std::fabs(mu.x) < 0.9f ? mk(1.f, 0.f, 0.f) : mk(0.f, 1.f, 0.f);
        const v3 b1 = normalize(cross(mu, e));
        const v3 b2 = cross(mu, b1);
        const double phi = 2.0 * 3.14159265358979323846 * (double)u2;
        v3 t = mk((float)(wv * mu.x + sw * (std::cos(phi) * b1.x + std::sin(phi) * b2.x)),
                  (float)(wv * mu.y + sw * (std::cos(phi) * b1.y + std::sin(phi) * b2.y)),
                  (float)(wv * mu.z + sw * (std::cos(phi) * b1.z + std::sin(phi) * b2.z)));
        t = normalize(t);
        S6[0] = 1.f - 0.95f * t.x * t.x;
        S6[1] = 1.f - 0.95f * t.y * t.y;
        S6[2] = 1.f - 0.95f * t.z * t.z;
        S6[3] = -0.95f * t.x * t.y;
        S6[4] = -0.95f * t.x * t.z;
        S6[5] = -0.95f * t.y * t.z;
        const float ax = std::fabs(t.x), ay = std::fabs(t.y), az = std::fabs(t.z);
        const v3 least = (ax <= ay && ax <= az) ? mk(1.f, 0.f, 0.f) : (ay <= az ? mk(0.f, 1.f, 0.f) : mk(0.f, 0.f, 1.f));
        const v3 nn = normalize(cross(t, least));
        nrm[0] = nn.x; nrm[1] = nn.y; nrm[2] = nn.z;
      }
      encode_record(rho, S6, one, one, nrm, 1.5f, w + 8 * i);
    }
  });
  return R2_OK;
}

extern "C" int r2_synth_env(int32_t kind, int32_t W, int32_t H, float* out) {
  if ((kind != 0 && kind != 1) || W < 2 || H < 2 || !out) return R2_ERR_INVALID;
  const size_t n = (size_t)W * H;
  if (kind == 0) {
    for (size_t i = 0; i < n * 3; ++i) out[i] = 1.f;
    return R2_OK;
  }
  const double PI = 3.14159265358979323846;
  const v3 sun = normalize(mk(0.4f, 0.7f, -0.6f));
  const double sun_cos = std::cos(0.25 * PI / 180.0);
  std::vector<double> cover(n, 0.0);
  std::vector<float> sky(3 * n);
  const int SS = 16;
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i) {
      const size_t p = (size_t)j * W + i;
      const double th = PI * ((double)j + 0.5) / H, ph = 2.0 * PI * ((double)i + 0.5) / W;
      const double dy = std::cos(th);
      float r, g, b;
      if (dy >= 0.0) {
        const float y = (float)dy;
        r = 1.0f * (1.f - y) + 0.25f * y;
        g = 0.85f * (1.f - y) + 0.45f * y;
        b = 0.7f * (1.f - y) + 1.0f * y;
      } else {
        r = 0.12f; g = 0.1f; b = 0.08f;
      }
      sky[3 * p] = r; sky[3 * p + 1] = g; sky[3 * p + 2] = b;
      (void)ph;
      int inside = 0;
      for (int sj = 0; sj < SS; ++sj)
        for (int si = 0; si < SS; ++si) {
          const double t2 = PI * ((double)j + (sj + 0.5) / SS) / H, p2 = 2.0 * PI * ((double)i + (si + 0.5) / SS) / W;
          const double dx = std::sin(p2) * std::sin(t2), dyy = std::cos(t2), dz = std::cos(p2) * std::sin(t2);
          if (dx * sun.x + dyy * sun.y + dz * sun.z >= sun_cos) ++inside;
        }
      cover[p] = (double)inside / (SS * SS);
    }
  double min_lum = 1e30, max_cover = 0.0;
  for (size_t p = 0; p < n; ++p) {
    min_lum = std::min(min_lum, (double)(sky[3 * p] + sky[3 * p + 1] + sky[3 * p + 2]) / 3.0);
    max_cover = std::max(max_cover, cover[p]);
  }
  const double L_sun = max_cover > 0 ? 1e4 * min_lum / max_cover : 0.0;
  for (size_t p = 0; p < n; ++p) {
    out[3 * p] = sky[3 * p] + (float)(cover[p] * L_sun * 1.0);
    out[3 * p + 1] = sky[3 * p + 1] + (float)(cover[p] * L_sun * 0.95);
    out[3 * p + 2] = sky[3 * p + 2] + (float)(cover[p] * L_sun * 0.85);
  }
  return R2_OK;
}

// Procedural mesh of the mesh->volume workload: a cube-sphere of 6*res^2 quads
// (2 triangles each), radius 1 displaced along the normal by 0.1*(fbm - 0.5),
// with smooth (area-weighted) vertex normals, plus n_quads randomly oriented
// square quads of side `quad_size` (2 triangles, flat normals) whose centres are
// uniform in the shell 1.05 < r < 1.3.  Triangle material ids: 0 for the
// sphere, 1 for the quads.  Returns the vertex / triangle counts when the output
// pointers are NULL.
extern "C" int r2_synth_mesh(int32_t res, int32_t n_quads, float quad_size, uint32_t seed, int32_t* n_vertices,
                             int32_t* n_triangles, float* positions, float* normals, int32_t* indices,
                             int32_t* material) {
  if (res < 1 || res > 2048 || n_quads < 0 || !n_vertices || !n_triangles) return R2_ERR_INVALID;
  const int64_t vs = (int64_t)6 * (res + 1) * (res + 1), ts = (int64_t)12 * res * res;
  const int64_t nv = vs + (int64_t)4 * n_quads, nt = ts + (int64_t)2 * n_quads;
  if (nv > INT32_MAX || nt > INT32_MAX) return R2_ERR_UNSUPPORTED;
  *n_vertices = (int32_t)nv;
  *n_triangles = (int32_t)nt;
  if (!positions) return R2_OK;
  if (!normals || !indices) return R2_ERR_INVALID;
  // faces: (axis, sign) -> vertex grid on the cube face, projected to the sphere
  const int R1 = res + 1;
  parallel_for((int64_t)6 * R1, [&](int64_t fr) {
    const int f = (int)(fr / R1), j = (int)(fr % R1);
    const int ax = f >> 1;
    const float sg = (f & 1) ? -1.f : 1.f;
    for (int i = 0; i < R1; ++i) {
      const float a = 2.f * (float)i / (float)res - 1.f, b = 2.f * (float)j / (float)res - 1.f;
      float c[3];
      c[ax] = sg;
      c[(ax + 1) % 3] = a * sg;
      c[(ax + 2) % 3] = b;
      const v3 d = normalize(mk(c[0], c[1], c[2]));
      const float r = 1.f + 0.1f * (fbm(seed + 77u, 4, 3.f, d.x + 2.f, d.y + 2.f, d.z + 2.f) - 0.5f);
      const int64_t v = ((int64_t)f * R1 + j) * R1 + i;
      positions[3 * v] = d.x * r;
      positions[3 * v + 1] = d.y * r;
      positions[3 * v + 2] = d.z * r;
    }
  });
  parallel_for((int64_t)6 * res, [&](int64_t fr) {
    const int f = (int)(fr / res), j = (int)(fr % res);
    for (int i = 0; i < res; ++i) {
      const int32_t v00 = (int32_t)(((int64_t)f * R1 + j) * R1 + i), v10 = v00 + 1, v01 = v00 + R1, v11 = v01 + 1;
      const int64_t t = ((int64_t)f * res + j) * res * 2 + 2 * i;
      indices[3 * t] = v00; indices[3 * t + 1] = v10; indices[3 * t + 2] = v11;
      indices[3 * t + 3] = v00; indices[3 * t + 4] = v11; indices[3 * t + 5] = v01;
      if (material) material[t] = material[t + 1] = 0;
    }
  });
  // smooth normals: area-weighted face normals accumulated per vertex (vertices
  // on cube edges are duplicated per face, so neighbouring faces share position
  // but not normal; the displacement field makes those seams sub-voxel)
  std::vector<float> acc((size_t)vs * 3, 0.f);
  for (int64_t t = 0; t < ts; ++t) {
    const int32_t* id = indices + 3 * t;
    const v3 p0 = mk(positions[3 * id[0]], positions[3 * id[0] + 1], positions[3 * id[0] + 2]);
    const v3 p1 = mk(positions[3 * id[1]], positions[3 * id[1] + 1], positions[3 * id[1] + 2]);
    const v3 p2 = mk(positions[3 * id[2]], positions[3 * id[2] + 1], positions[3 * id[2] + 2]);
    const v3 fn = cross(p1 - p0, p2 - p0);
    for (int k = 0; k < 3; ++k) {
      acc[3 * id[k]] += fn.x;
      acc[3 * id[k] + 1] += fn.y;
      acc[3 * id[k] + 2] += fn.z;
    }
  }
  for (int64_t v = 0; v < vs; ++v) {
    v3 n = mk(acc[3 * v], acc[3 * v + 1], acc[3 * v + 2]);
    const v3 p = mk(positions[3 * v], positions[3 * v + 1], positions[3 * v + 2]);
    if (dot(n, p) < 0.f) n = -n;  // outward
    n = normalize(n);
    normals[3 * v] = n.x;
    normals[3 * v + 1] = n.y;
    normals[3 * v + 2] = n.z;
  }
  // foliage quads
  parallel_for(n_quads, [&](int64_t q) {
    uint32_t c[4] = {(uint32_t)q, 0u, 0u, 0x0A1Du};
    philox4x32_10(c, seed, 0x9u);
    const float u[4] = {(float)(c[0] >> 8) * (1.f / 16777216.f), (float)(c[1] >> 8) * (1.f / 16777216.f),
                        (float)(c[2] >> 8) * (1.f / 16777216.f), (float)(c[3] >> 8) * (1.f / 16777216.f)};
    uint32_t c2[4] = {(uint32_t)q, 1u, 0u, 0x0A1Du};
    philox4x32_10(c2, seed, 0x9u);
    const float w[2] = {(float)(c2[0] >> 8) * (1.f / 16777216.f), (float)(c2[1] >> 8) * (1.f / 16777216.f)};
    // centre: uniform direction, radius uniform in volume of the shell
    const float z = 1.f - 2.f * u[0], rr = sqrtf(fmaxf(0.f, 1.f - z * z)), ph = 6.2831853f * u[1];
    const float r3a = 1.05f * 1.05f * 1.05f, r3b = 1.3f * 1.3f * 1.3f;
    const float rad = cbrtf(r3a + u[2] * (r3b - r3a));
    const v3 ctr = mk(rr * cosf(ph), rr * sinf(ph), z) * rad;
    // orientation: uniform normal, random spin
    const float nz = 1.f - 2.f * u[3], nr = sqrtf(fmaxf(0.f, 1.f - nz * nz)), nph = 6.2831853f * w[0];
    const v3 n = mk(nr * cosf(nph), nr * sinf(nph), nz);
    const v3 e = fabsf(n.x) < 0.9f ? mk(1.f, 0.f, 0.f) : mk(0.f, 1.f, 0.f);
    const v3 t0 = normalize(cross(n, e)), b0 = cross(n, t0);
    const float sp = 6.2831853f * w[1];
    const v3 tt = t0 * cosf(sp) + b0 * sinf(sp), bb = cross(n, tt);
    const float h = 0.5f * quad_size;
    const v3 corner[4] = {ctr - tt * h - bb * h, ctr + tt * h - bb * h, ctr + tt * h + bb * h, ctr - tt * h + bb * h};
    const int64_t v0 = vs + 4 * q;
    for (int k = 0; k < 4; ++k) {
      positions[3 * (v0 + k)] = corner[k].x;
      positions[3 * (v0 + k) + 1] = corner[k].y;
      positions[3 * (v0 + k) + 2] = corner[k].z;
      normals[3 * (v0 + k)] = n.x;
      normals[3 * (v0 + k) + 1] = n.y;
      normals[3 * (v0 + k) + 2] = n.z;
    }
    const int64_t t = ts + 2 * q;
    indices[3 * t] = (int32_t)v0; indices[3 * t + 1] = (int32_t)(v0 + 1); indices[3 * t + 2] = (int32_t)(v0 + 2);
    indices[3 * t + 3] = (int32_t)v0; indices[3 * t + 4] = (int32_t)(v0 + 2); indices[3 * t + 5] = (int32_t)(v0 + 3);
    if (material) material[t] = material[t + 1] = 1;
  });
  return R2_OK;
}
