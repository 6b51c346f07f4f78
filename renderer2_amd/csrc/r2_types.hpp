// Device-side scene description shared by the render kernels and the host glue.
#pragma once
#include "r2_math.hpp"

namespace r2 {

// A scene holds up to kMaxGrids distinct grids and any number (up to
// kMaxInstances) of instances of them.  Up to kLinearInstances instances are
// tested by a linear scan; more go through a bounding-volume hierarchy over the
// instances' world boxes, which gives the same closest hit.
constexpr int kMaxGrids = 16;
constexpr int kMaxInstances = 1 << 20;
constexpr int kLinearInstances = 8;
constexpr int kLeafDim = 8;            // NanoVDB leaf node: 8^3 voxels
constexpr int kLeafVoxels = 512;
constexpr float kDefaultTmax = 1e12f;  // CRay::DEFAULT_TMAX (ray.hpp:12)
constexpr float kRayOffset = 1e-4f;    // CRay::OFFSET (ray.hpp:11)
constexpr int kLoopCap = 65536;        // safety cap on every traversal loop

// One filtered volume resident in HBM.
//  * records: NanoVDB-layout sparse tree collapsed to one dense leaf table
//    (leaf index per 8^3 node, -1 = background tile) + a pool of leaves, each
//    512 voxel records contiguous (x fastest) so a wavefront's trilinear
//    footprints fall into a few 32-byte-aligned lines.
//  * brick majorant grid (8^3 bricks = the same nodes): per brick an encoded half
//    (min, max) range and an atlas slot; a u8 atlas of 512 bytes per non-empty
//    brick; and the max of the range per mip level 0..3 as float.
struct DGrid {
  const int32_t* leaf_table;
  const uint4* leaves;          // 2 x uint4 (32 B) per voxel record (FP32 form)
  const uint4* leaves_packed;   // packed form: 16 B hot record per voxel (density, S, ior in FP16)
  const uint2* leaves_cold;     // packed form: 8 B cold record per voxel (u8 colours, octahedral normal)
  int32_t nx, ny, nz;
  int32_t nlx, nly, nlz;
  const uint2* bricks;          // x = encoded range, y = atlas slot (0xffffffff: empty)
  const uint8_t* atlas;
  const float* maj[4];
  int32_t nb[3];
  int32_t precision;
  m43 m2i, i2m, m2s, s2m;
  int32_t ibb_min[3], size[3];
  float sigma_s[3];
  float sigma_t;
  float inv_max_density;
  // reciprocals of the index extent and of sigma_t, used in place of a division
  // when the divisor is a power of two (then x * (1/d) and x / d are the same number)
  float inv_size[3];
  float inv_sigma_t;
  int32_t size_pow2, sigma_t_pow2;
};

struct DInstance {
  int32_t grid;
  m43 m2w, w2m;
  float bb_min[3], bb_max[3];
  float ratio;  // filterRenderRatio
};

struct DEnv {
  int32_t w, h;
  const float* tex;        // h * w * 3
  const float* row_func;   // h * w
  const float* row_cdf;    // h * (w + 1)
  const float* row_int;    // h
  const float* marg_func;  // h
  const float* marg_cdf;   // h + 1
  // guide tables of the CDF searches: entry j of a table of g + 1 entries is the
  // number of CDF values <= j / g, which brackets the search for any u in
  // [j / g, (j + 1) / g)
  const int32_t* marg_guide;  // marg_g + 1
  const int32_t* row_guide;   // h * (row_g + 1)
  int32_t marg_g, row_g;
  float marg_int;
};

// An analytic surface (CCircle / Sphere) with its CMaterial: shape 0 circle, 1
// sphere; transforms translate(pos) * getRotation(normal) and its inverse; the
// Oren-Nayar A, B of the diffuse roughness; the microfacet alphas and the
// Fresnel indices of the mixture weight.
constexpr int kMaxSurfaces = 8;
struct DSurface {
  int32_t shape;
  float pos[3], normal[3];
  float radius;
  m43 m2w, w2m;
  float diffuse[3], specular[3];
  float on_a, on_b;
  float alpha_x, alpha_y, eta_i, eta_t;
};

// Node of the instance hierarchy, in depth-first order: the box of its
// subtree, the index of the node after the subtree, and for a leaf
// (first << 3 | count) into DScene::bvh_idx (count 1..4); -1 for an inner node.
struct DBvhNode {
  float lo[3];
  int32_t skip;
  float hi[3];
  int32_t leaf;
};

struct DScene {
  int32_t n_inst;
  const DInstance* inst;     // n_inst instances (device memory, after the scene)
  int32_t n_bvh;             // 0: linear scan
  const DBvhNode* bvh;
  const int32_t* bvh_idx;
  int32_t n_grid;
  DGrid grid[kMaxGrids];
  int32_t n_surf;
  DSurface surf[kMaxSurfaces];
  DEnv env;
  m43 view_to_world;
  float pixel_size, near_dist;
  int32_t W, H, spp, max_depth;
  uint32_t seed, stream;
  int32_t sample_begin;
  uint32_t flags;
};

enum { kCntVox = 0, kCntMaj, kCntBrick, kCntEnv, kCntDda, kCntFlight, kCntScatter, kNumCnt = 8 };

}  // namespace r2
