// Reduced-precision voxel records (R2_PREC_PACKED), 24 bytes per voxel in two
// arrays of the leaf pool:
//   hot  (uint4): x = density FP16 | S_xx FP16, y = S_yy FP16 | S_zz FP16,
//                 z = S_xy FP16 | S_xz FP16,   w = S_yz FP16 | ior FP16
//   cold (uint2): x = diffuse u8 r,g,b | specular u8 r,   y = specular u8 g,b | normal octahedral 8+8
// The six SGGX coefficients are the entries of S itself (linear in S, so
// interpolation needs no decode beyond the FP16 -> FP32 conversion), after an
// FP64 projection of S onto the positive semi-definite cone by a Jacobi
// eigen-decomposition: negative eigenvalues, which 16-bit quantisation of a
// near-singular S can produce, are raised to a floor of 2e-3 of the largest,
// which keeps S positive definite after the FP16 rounding of its entries.
#pragma once
#include "r2_math.hpp"

namespace r2 {

// 3x3 symmetric eigen-decomposition by cyclic Jacobi rotations (Golub & Van
// Loan, Matrix Computations, 8.5) in double; S6 = (xx, yy, zz, xy, xz, yz).
R2_HD void sggx_psd_fp64(const float S6[6], float out[6]) {
  double a[3][3] = {{S6[0], S6[3], S6[4]}, {S6[3], S6[1], S6[5]}, {S6[4], S6[5], S6[2]}};
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 12; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    const double scale = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
    if (off <= 1e-17 * scale || off == 0.0) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- A J
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^T A
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {  // V <- V J
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  double lam[3] = {a[0][0], a[1][1], a[2][2]};
  const double lmax = fmax(lam[0], fmax(lam[1], lam[2]));
  const double floor_v = lmax > 0.0 ? 2e-3 * lmax : 0.0;
  for (int i = 0; i < 3; ++i) lam[i] = fmax(lam[i], floor_v);
  double r[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = v[i][0] * lam[0] * v[j][0] + v[i][1] * lam[1] * v[j][1] + v[i][2] * lam[2] * v[j][2];
  out[0] = (float)r[0][0];
  out[1] = (float)r[1][1];
  out[2] = (float)r[2][2];
  out[3] = (float)r[0][1];
  out[4] = (float)r[0][2];
  out[5] = (float)r[1][2];
}

R2_HD uint32_t oct8(float v) {
  float c = v < -1.f ? -1.f : (v > 1.f ? 1.f : v);
  return (uint32_t)(int32_t)rintf(c * 127.f) & 0xffu;
}
// octahedral normal encoding, 8 bits per component (snorm)
R2_HD uint32_t oct_encode(float x, float y, float z) {
  const float l1 = fabsf(x) + fabsf(y) + fabsf(z);
  if (!(l1 > 0.f)) return 0u;
  float u = x / l1, w = y / l1;
  if (z < 0.f) {
    const float uu = (1.f - fabsf(w)) * (u >= 0.f ? 1.f : -1.f);
    const float ww = (1.f - fabsf(u)) * (w >= 0.f ? 1.f : -1.f);
    u = uu;
    w = ww;
  }
  return oct8(u) | (oct8(w) << 8);
}
R2_HD v3 oct_decode(uint32_t c) {
  float u = (float)(int8_t)(c & 0xffu) * (1.f / 127.f), w = (float)(int8_t)((c >> 8) & 0xffu) * (1.f / 127.f);
  u = u < -1.f ? -1.f : u;
  w = w < -1.f ? -1.f : w;
  const float z = 1.f - fabsf(u) - fabsf(w);
  if (z < 0.f) {
    const float uu = (1.f - fabsf(w)) * (u >= 0.f ? 1.f : -1.f);
    const float ww = (1.f - fabsf(u)) * (w >= 0.f ? 1.f : -1.f);
    u = uu;
    w = ww;
  }
  const v3 n = mk(u, w, z);
  const float l = sqrtf(dot(n, n));
  return l > 0.f ? n * (1.f / l) : mk(0.f, 0.f, 0.f);
}

// fields of one decoded record: density, S6, diffuse, specular, normal (unit or 0), ior
struct PackIn {
  float density, S[6], diffuse[3], specular[3], normal[3], ior;
};

R2_HD uint32_t u8q(float v) {
  const float c = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
  return (uint32_t)rintf(c * 255.f);
}

R2_HD void pack_record(const PackIn& in, uint4* hot, uint2* cold) {
  float S[6];
  sggx_psd_fp64(in.S, S);
  const uint32_t h0 = half_from_float_rne(in.density), h1 = half_from_float_rne(S[0]);
  const uint32_t h2 = half_from_float_rne(S[1]), h3 = half_from_float_rne(S[2]);
  const uint32_t h4 = half_from_float_rne(S[3]), h5 = half_from_float_rne(S[4]);
  const uint32_t h6 = half_from_float_rne(S[5]), h7 = half_from_float_rne(in.ior);
  *hot = make_uint4(h0 | (h1 << 16), h2 | (h3 << 16), h4 | (h5 << 16), h6 | (h7 << 16));
  const uint32_t dr = u8q(in.diffuse[0]), dg = u8q(in.diffuse[1]), db = u8q(in.diffuse[2]);
  const uint32_t sr = u8q(in.specular[0]), sg = u8q(in.specular[1]), sb = u8q(in.specular[2]);
  const uint32_t n = oct_encode(in.normal[0], in.normal[1], in.normal[2]);
  *cold = make_uint2(dr | (dg << 8) | (db << 16) | (sr << 24), sg | (sb << 8) | (n << 16));
}

}  // namespace r2
