// SFilteredDataCompact encoder (filtered_data.hpp:91-107): 32-byte voxel record
//   density f32 @0 | sigma_x,y,z u16 @4 | r_xy,xz,yz u16 @10 | diffuse u8x3 @16 |
//   specular u8x3 @19 | normal u16x3 @22 | ior half @28 | pad @30.
// Float -> integer conversions truncate after glm::clamp, as in the reference.
// A NaN input (undefined behaviour in the reference) encodes as 0.
#pragma once
#include "r2_math.hpp"

namespace r2 {

R2_HD float clampf_glm(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
R2_HD uint32_t to_u16(float x) {
  const float c = clampf_glm(x, 0.f, 65535.f);
  return (c != c) ? 0u : (uint32_t)(uint16_t)c;
}
R2_HD uint32_t to_u8(float x) {
  const float c = clampf_glm(x, 0.f, 255.f);
  return (c != c) ? 0u : (uint32_t)(uint8_t)c;
}

// S6 = (xx, yy, zz, xy, xz, yz); `normal` is normalised here (SFilteredData::n()).
R2_HD void encode_record(float density, const float S6[6], const float diffuse[3], const float specular[3],
                         const float normal[3], float ior, uint32_t w[8]) {
  const float M = 65535.f, M8 = 255.f;
  const uint32_t sx = to_u16(M * sqrtf(S6[0])), sy = to_u16(M * sqrtf(S6[1])), sz = to_u16(M * sqrtf(S6[2]));
  const uint32_t rxy = to_u16(M * 0.5f * (S6[3] / sqrtf(S6[0] * S6[1]) + 1.f));
  const uint32_t rxz = to_u16(M * 0.5f * (S6[4] / sqrtf(S6[0] * S6[2]) + 1.f));
  const uint32_t ryz = to_u16(M * 0.5f * (S6[5] / sqrtf(S6[1] * S6[2]) + 1.f));
  const uint32_t dr = to_u8(M8 * diffuse[0]), dg = to_u8(M8 * diffuse[1]), db = to_u8(M8 * diffuse[2]);
  const uint32_t sr = to_u8(M8 * specular[0]), sg = to_u8(M8 * specular[1]), sb = to_u8(M8 * specular[2]);
  const v3 n = normalize(mk(normal[0], normal[1], normal[2]));
  const uint32_t nx = to_u16(M * 0.5f * (n.x + 1.f)), ny = to_u16(M * 0.5f * (n.y + 1.f)),
                 nz = to_u16(M * 0.5f * (n.z + 1.f));
  const uint32_t h = half_from_float_rne(ior);
  w[0] = f2u(density);
  w[1] = sx | (sy << 16);
  w[2] = sz | (rxy << 16);
  w[3] = rxz | (ryz << 16);
  w[4] = dr | (dg << 8) | (db << 16) | (sr << 24);
  w[5] = sg | (sb << 8) | (nx << 16);
  w[6] = ny | (nz << 16);
  w[7] = h;
}

}  // namespace r2
