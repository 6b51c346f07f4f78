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
