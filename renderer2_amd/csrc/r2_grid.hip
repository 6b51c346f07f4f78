// Grid upload on the GPU: applies the reference's active-voxel rule, lays the
// records out as a NanoVDB-style leaf pool (8^3 leaves) behind a dense leaf
// table, finds the active index bbox and the max density, and builds the brick
// majorant grid (dilated half ranges, u8 atlas, 3 min/max mips).
//
// Restates: COpenvdbBackend::setValues (raytracer/src/filtering/openvdb_backend.cpp:108-132),
// CNVDBMedium::init / getMaxValue (raytracer/src/medium/nvdb_medium.cu:130-173,308-323),
// CHostBrickGrid::CHostBrickGrid (raytracer/src/grid_brick/host_grid_brick.cpp:61-152).
#include "r2_internal.hpp"
#include "r2_volume.hpp"
#include <algorithm>
#include <cstring>

namespace {

using namespace r2;

__device__ __forceinline__ bool is_corner(int x, int y, int z, int nx, int ny, int nz) {
  return (x == 0 || x == nx - 1) && (y == 0 || y == ny - 1) && (z == 0 || z == nz - 1);
}

// effective density (active rule) + leaf occupancy + bbox + max density
// order-preserving unsigned key of a float (for atomicMin over any sign)
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// wave, then block reduction of a thread's extrema and active box, published
// with one set of atomics per block (same-address atomics from every wavefront
// serialise in L2): max_bits[0] = max active density (float bits), max_bits[1] =
// order key of the minimum, bbox = active index box
__device__ __forceinline__ void publish_extrema(float lmax, float lmin, int lo[3], int hi[3], int* __restrict__ bbox,
                                                uint32_t* __restrict__ max_bits) {
  for (int off = 32; off > 0; off >>= 1) {
    lmax = fmaxf(lmax, __shfl_xor(lmax, off));
    lmin = fminf(lmin, __shfl_xor(lmin, off));
    for (int a = 0; a < 3; ++a) {
      lo[a] = min(lo[a], __shfl_xor(lo[a], off));
      hi[a] = max(hi[a], __shfl_xor(hi[a], off));
    }
  }
  __shared__ float s_max[4], s_min[4];
  __shared__ int s_lo[4][3], s_hi[4][3];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_max[wave] = lmax;
    s_min[wave] = lmin;
    for (int a = 0; a < 3; ++a) s_lo[wave][a] = lo[a], s_hi[wave][a] = hi[a];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 1; w < nw; ++w) {
      lmax = fmaxf(lmax, s_max[w]);
      lmin = fminf(lmin, s_min[w]);
      for (int a = 0; a < 3; ++a) lo[a] = min(lo[a], s_lo[w][a]), hi[a] = max(hi[a], s_hi[w][a]);
    }
    atomicMax(max_bits, __float_as_uint(fmaxf(lmax, 0.f)));
    atomicMin(max_bits + 1, float_key(lmin));
    for (int a = 0; a < 3; ++a) {
      atomicMin(&bbox[a], lo[a]);
      atomicMax(&bbox[3 + a], hi[a]);
    }
  }
}

// dense records: the active rule (density > 0, or a corner of the index box),
// leaf occupancy, and the extrema and active box over all voxels (inactive ones
// reading 0 for the minimum, the brick grid's minorant)
__global__ void k_active(const uint4* __restrict__ rec, int nx, int ny, int nz, uint8_t* __restrict__ active,
                         uint32_t* __restrict__ leaf_flag, int nlx, int nly, int* __restrict__ bbox,
                         uint32_t* __restrict__ max_bits) {
  const size_t n = (size_t)nx * ny * nz;
  const bool small = n <= 0xFFFFFFFFull;  // 32-bit index arithmetic (the common case)
  float lmax = 0.f, lmin = FLT_MAX;
  int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    int x, y, z;
    if (small) {
      const uint32_t i32 = (uint32_t)i, r = i32 / (uint32_t)nx;
      x = (int)(i32 - r * (uint32_t)nx);
      y = (int)(r % (uint32_t)ny);
      z = (int)(r / (uint32_t)ny);
    } else {
      x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((size_t)nx * ny));
    }
    const float d = __uint_as_float(rec[2 * i].x);
    const bool a = (d > 0.f) || is_corner(x, y, z, nx, ny, nz);
    active[i] = a ? 1 : 0;
    lmin = fminf(lmin, a ? d : 0.f);
    if (a) {
      leaf_flag[((size_t)(z >> 3) * nly + (y >> 3)) * nlx + (x >> 3)] = 1u;
      lmax = fmaxf(lmax, d);
      lo[0] = min(lo[0], x); lo[1] = min(lo[1], y); lo[2] = min(lo[2], z);
      hi[0] = max(hi[0], x); hi[1] = max(hi[1], y); hi[2] = max(hi[2], z);
    }
  }
  publish_extrema(lmax, lmin, lo, hi, bbox, max_bits);
}

// sparse records: leaf k of the caller's list becomes pool slot k; a voxel is
// active by its mask bit (or, without masks, by the dense rule), inactive voxels
// and voxels beyond the index extent hold the background record; extrema and
// active box over the voxels of the leaves
__global__ void k_sparse_leaves(const uint4* __restrict__ rec, const uint8_t* __restrict__ masks,
                                const int32_t* __restrict__ origins, int64_t n_leaves, int nx, int ny, int nz,
                                uint4* __restrict__ leaves, int* __restrict__ bbox, uint32_t* __restrict__ max_bits) {
  float lmax = 0.f, lmin = FLT_MAX;
  int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (int64_t li = blockIdx.x; li < n_leaves; li += gridDim.x) {
    const int ox = origins[3 * li], oy = origins[3 * li + 1], oz = origins[3 * li + 2];
    for (int v = threadIdx.x; v < kLeafVoxels; v += blockDim.x) {
      const int x = ox + (v & 7), y = oy + ((v >> 3) & 7), z = oz + (v >> 6);
      const size_t i = (size_t)li * kLeafVoxels + v;
      uint4 a = rec[2 * i], b = rec[2 * i + 1];
      const float d = __uint_as_float(a.x);
      const bool inside = x < nx && y < ny && z < nz;
      const bool act = inside && (masks ? ((masks[(size_t)li * 64 + (v >> 3)] >> (v & 7)) & 1u) != 0u
                                        : (d > 0.f || is_corner(x, y, z, nx, ny, nz)));
      if (inside) lmin = fminf(lmin, act ? d : 0.f);
      if (act) {
        lmax = fmaxf(lmax, d);
        lo[0] = min(lo[0], x); lo[1] = min(lo[1], y); lo[2] = min(lo[2], z);
        hi[0] = max(hi[0], x); hi[1] = max(hi[1], y); hi[2] = max(hi[2], z);
      } else {
        a = make_uint4(0u, 0u, 0u, 0u);
        b = make_uint4(0u, 0u, 0u, 0x3C00u);
      }
      leaves[2 * i] = a;
      leaves[2 * i + 1] = b;
    }
  }
  publish_extrema(lmax, lmin, lo, hi, bbox, max_bits);
}

// scatter records into the leaf pool; inactive voxels of an allocated leaf hold the background record
__global__ void k_leaves(const uint4* __restrict__ rec, const uint8_t* __restrict__ active, int nx, int ny, int nz,
                         const int32_t* __restrict__ leaf_table, int nlx, int nly, int nlz, uint4* __restrict__ leaves) {
  const size_t nleaf_total = (size_t)nlx * nly * nlz;
  for (size_t li = blockIdx.x; li < nleaf_total; li += gridDim.x) {
    const int32_t slot = leaf_table[li];
    if (slot < 0) continue;
    const int lx = (int)(li % nlx), ly = (int)((li / nlx) % nly), lz = (int)(li / ((size_t)nlx * nly));
    for (int v = threadIdx.x; v < kLeafVoxels; v += blockDim.x) {
      const int x = lx * 8 + (v & 7), y = ly * 8 + ((v >> 3) & 7), z = lz * 8 + (v >> 6);
      uint4 a = make_uint4(0u, 0u, 0u, 0u), b = make_uint4(0u, 0u, 0u, 0x3C00u);
      if (x < nx && y < ny && z < nz) {
        const size_t i = ((size_t)z * ny + y) * nx + x;
        if (active[i]) {
          a = rec[2 * i];
          b = rec[2 * i + 1];
        }
      }
      uint4* dst = leaves + ((size_t)slot * kLeafVoxels + v) * 2;
      dst[0] = a;
      dst[1] = b;
    }
  }
}

// packed form of every leaf-pool record (r2_packed.hpp)
__global__ void k_pack(const uint4* __restrict__ leaves, size_t n, uint4* __restrict__ hot, uint2* __restrict__ cold) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const Rec r = decode_record(leaves[2 * i], leaves[2 * i + 1]);
    PackIn in;
    in.density = r.density;
    for (int k = 0; k < 6; ++k) in.S[k] = r.S[k];
    in.diffuse[0] = r.diffuse.x; in.diffuse[1] = r.diffuse.y; in.diffuse[2] = r.diffuse.z;
    in.specular[0] = r.specular.x; in.specular[1] = r.specular.y; in.specular[2] = r.specular.z;
    in.normal[0] = r.normal.x; in.normal[1] = r.normal.y; in.normal[2] = r.normal.z;
    in.ior = r.ior;
    pack_record(in, &hot[i], &cold[i]);
  }
}

// majorant-DDA walk of the free-flight loop without its density decisions: the
// same position, mip and step arithmetic as traversal_step, recording (brick x, y,
// z, mip) of every majorant lookup; collide_every = k > 0 applies a tentative
// collision's mip drop after every k-th step
__global__ void k_dda_cells(const float* __restrict__ rays, int n, float sigma_t, int collide_every, int max_cells,
                            int32_t* __restrict__ cells, int32_t* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rays + (size_t)7 * i;
  const v3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
  const float tmax = r[6];
  const v3 ri = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
  float t = 0.f, mip = 3.f;
  int c = 0;
  int32_t* out = cells + (size_t)i * max_cells * 4;
  for (int it = 0; it < kLoopCap && t < tmax && c < max_cells; ++it) {
    const v3 cur = o + t * d;
    const int imip = (int)roundf(mip);
    out[4 * c + 0] = ((int)floorf(cur.x)) >> 3;
    out[4 * c + 1] = ((int)floorf(cur.y)) >> 3;
    out[4 * c + 2] = ((int)floorf(cur.z)) >> 3;
    out[4 * c + 3] = imip;
    ++c;
    const float dt = step_dda(cur, ri, imip);
    const float prev = t;
    t += dt / sigma_t;
    if (t == prev) break;
    mip = gmin(mip + 0.25f, 3.f);
    if (collide_every > 0 && c % collide_every == 0) mip = gmax(0.f, mip - 2.f);
  }
  counts[i] = c;
}

// filtered (trilinear, decoded) records at normalised positions: 17 floats each
// (density, S6, diffuse3, specular3, normal3, ior)
__global__ void k_sample(DGrid g, int n, const float* __restrict__ pos, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Cnt cnt;
  const Rec r = filtered_data<false>(g, mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), cnt);
  float* o = out + 17 * (size_t)i;
  o[0] = r.density;
  for (int k = 0; k < 6; ++k) o[1 + k] = r.S[k];
  o[7] = r.diffuse.x; o[8] = r.diffuse.y; o[9] = r.diffuse.z;
  o[10] = r.specular.x; o[11] = r.specular.y; o[12] = r.specular.z;
  o[13] = r.normal.x; o[14] = r.normal.y; o[15] = r.normal.z;
  o[16] = r.ior;
}

// density of a voxel as the brick grid sees it: the active record's density,
// 0 for inactive voxels (they hold the background record) and outside the grid
struct PoolDens {
  const int32_t* table;
  const uint4* leaves;
  int nx, ny, nz, nlx, nly;
  __device__ __forceinline__ float operator()(int x, int y, int z) const {
    if ((unsigned)x >= (unsigned)nx || (unsigned)y >= (unsigned)ny || (unsigned)z >= (unsigned)nz) return 0.f;
    const int32_t leaf = table[((size_t)(z >> 3) * nly + (y >> 3)) * nlx + (x >> 3)];
    if (leaf < 0) return 0.f;
    return __uint_as_float(leaves[((size_t)leaf * kLeafVoxels + (((z & 7) << 6) | ((y & 7) << 3) | (x & 7))) * 2].x);
  }
};

// one wavefront per brick: min/max over the dilated 10^3 neighbourhood
__global__ void k_brick_range(PoolDens dens, int nbx, int nby, int nbz, uint32_t* __restrict__ range,
                              uint32_t* __restrict__ nonempty) {
  const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const size_t nbt = (size_t)nbx * nby * nbz;
  if (wave >= nbt) return;
  const int bx = (int)(wave % nbx), by = (int)((wave / nbx) % nby), bz = (int)(wave / ((size_t)nbx * nby));
  // a brick with no leaf in its 3x3x3 leaf neighbourhood reads density 0 at every
  // voxel of its dilated range (the common case of a sparse grid)
  bool leaf = false;
  if (lane < 27) {
    const int lx = bx + lane % 3 - 1, ly = by + (lane / 3) % 3 - 1, lz = bz + lane / 9 - 1;
    const int nlx = (dens.nx + 7) >> 3, nly = (dens.ny + 7) >> 3, nlz = (dens.nz + 7) >> 3;
    if ((unsigned)lx < (unsigned)nlx && (unsigned)ly < (unsigned)nly && (unsigned)lz < (unsigned)nlz)
      leaf = dens.table[((size_t)lz * nly + ly) * nlx + lx] >= 0;
  }
  if (__ballot(leaf) == 0ull) {
    if (lane == 0) {
      range[wave] = encode_range(0.f, 0.f);
      nonempty[wave] = 0u;
    }
    return;
  }
  float lo = FLT_MAX, hi = -FLT_MAX;
  for (int k = lane; k < 1000; k += 64) {
    const int x = k % 10 - 1, y = (k / 10) % 10 - 1, z = k / 100 - 1;
    const float d = dens(bx * 8 + x, by * 8 + y, bz * 8 + z);
    lo = fminf(lo, d);
    hi = fmaxf(hi, d);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, off));
    hi = fmaxf(hi, __shfl_xor(hi, off));
  }
  if (lane == 0) {
    range[wave] = encode_range(lo, hi);
    nonempty[wave] = (hi == lo) ? 0u : 1u;
  }
}

// one wavefront per brick: quantise the 512 voxels of every non-empty brick
__global__ void k_brick_atlas(PoolDens dens, int nbx, int nby, int nbz, const uint32_t* __restrict__ range, const uint32_t* __restrict__ slot,
                              const uint32_t* __restrict__ nonempty, uint2* __restrict__ bricks,
                              uint8_t* __restrict__ atlas, float* __restrict__ maj0) {
  const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const size_t nbt = (size_t)nbx * nby * nbz;
  if (wave >= nbt) return;
  const int bx = (int)(wave % nbx), by = (int)((wave / nbx) % nby), bz = (int)(wave / ((size_t)nbx * nby));
  const uint32_t r = range[wave];
  const bool ne = nonempty[wave] != 0u;
  if (lane == 0) {
    bricks[wave] = make_uint2(r, ne ? slot[wave] : 0xffffffffu);
    maj0[wave] = range_hi(r);
  }
  if (!ne) return;
  const float lo = range_lo(r), hi = range_hi(r);
  uint8_t* dst = atlas + (size_t)slot[wave] * 512;
  for (int v = lane; v < 512; v += 64) {
    const int x = v & 7, y = (v >> 3) & 7, z = v >> 6;
    dst[v] = encode_voxel(dens(bx * 8 + x, by * 8 + y, bz * 8 + z), lo, hi);
  }
}

// min/max mip level from the level below (range encoded), plus its decoded max
__global__ void k_brick_mip(const uint32_t* __restrict__ src, int px, int py, uint32_t* __restrict__ dst, int sx,
                            int sy, int sz, float* __restrict__ maj) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sx * sy * sz) return;
  const int x = i % sx, y = (i / sx) % sy, z = i / (sx * sy);
  float lo = FLT_MAX, hi = -FLT_MAX;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const uint32_t c = src[((size_t)(2 * z + dz) * py + (2 * y + dy)) * px + (2 * x + dx)];
        lo = fminf(lo, range_lo(c));
        hi = fmaxf(hi, range_hi(c));
      }
  const uint32_t e = encode_range(lo, hi);
  dst[i] = e;
  maj[i] = range_hi(e);
}

// parity probe: the record the collapsed tree returns for integer voxel coordinates
__global__ void k_lookup(DGrid g, int n, const int32_t* __restrict__ xyz, uint4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  int32_t leaf = -1;
  if ((unsigned)x < (unsigned)g.nx && (unsigned)y < (unsigned)g.ny && (unsigned)z < (unsigned)g.nz)
    leaf = g.leaf_table[((size_t)(z >> 3) * g.nly + (y >> 3)) * g.nlx + (x >> 3)];
  uint4 a = make_uint4(0u, 0u, 0u, 0u), b = make_uint4(0u, 0u, 0u, 0x3C00u);
  if (leaf >= 0) {
    const uint4* p = g.leaves + ((size_t)leaf * kLeafVoxels + (((z & 7) << 6) | ((y & 7) << 3) | (x & 7))) * 2;
    a = p[0];
    b = p[1];
  }
  out[2 * i] = a;
  out[2 * i + 1] = b;
}

template <typename T>
int dmalloc(r2_ctx* ctx, r2_grid* g, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return r2host::hip_fail(ctx, e, "hipMalloc(grid)");
  g->allocs.push_back(q);
  *p = (T*)q;
  return R2_OK;
}

}  // namespace

namespace r2host {

int grid_build(r2_ctx* ctx, const r2_grid_desc* desc, r2_grid* g) {
  using namespace r2;
  const int nx = desc->nx, ny = desc->ny, nz = desc->nz;
  const bool sparse = desc->n_leaves > 0;
  hipStream_t st = ctx->stream;
  int rc;
  struct Tmp {
    std::vector<void*> p;
    ~Tmp() { for (void* q : p) (void)hipFree(q); }
  } tmp;
  auto tmalloc = [&](void** q, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(q, std::max<size_t>(bytes, 1));
    if (e == hipSuccess) tmp.p.push_back(*q);
    return e;
  };
  const int nlx = (nx + 7) / 8, nly = (ny + 7) / 8, nlz = (nz + 7) / 8;
  const size_t nleaf_total = (size_t)nlx * nly * nlz;
  // records staged in device memory: nx*ny*nz dense records, or n_leaves leaves of 512
  const size_t nrec = sparse ? (size_t)desc->n_leaves * kLeafVoxels : (size_t)nx * ny * nz;
  const uint4* d_rec = nullptr;
  if (desc->records_on_device) {
    d_rec = (const uint4*)desc->records;
  } else {
    void* staged = nullptr;
    R2_HIP(ctx, tmalloc(&staged, nrec * 32));
    R2_HIP(ctx, hipMemcpyAsync(staged, desc->records, nrec * 32, hipMemcpyHostToDevice, st));
    d_rec = (const uint4*)staged;
  }
  int* d_bbox = nullptr;
  uint32_t* d_max = nullptr;
  R2_HIP(ctx, tmalloc((void**)&d_bbox, 6 * 4));
  R2_HIP(ctx, tmalloc((void**)&d_max, 8));
  const int bbox_init[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
  R2_HIP(ctx, hipMemcpyAsync(d_bbox, bbox_init, sizeof(bbox_init), hipMemcpyHostToDevice, st));
  const uint32_t mm_init[2] = {0u, 0xFFFFFFFFu};
  R2_HIP(ctx, hipMemcpyAsync(d_max, mm_init, 8, hipMemcpyHostToDevice, st));

  std::vector<int32_t> table(nleaf_total, -1);
  int64_t nleaves = 0;
  const uint8_t* d_dense_active = nullptr;  // dense path: the active flags for the leaf scatter
  uint4* d_leaves = nullptr;
  bool covers_all = true;  // every voxel of the index box lies in a leaf (else some voxel reads 0)
  int bbox[6];
  uint32_t mm[2] = {0u, 0u};
  if (!sparse) {
    // dense: the active rule per voxel, leaves allocated in table order
    const size_t nv = nrec;
    uint8_t* d_active = nullptr;
    uint32_t* d_leaf_flag = nullptr;
    R2_HIP(ctx, tmalloc((void**)&d_active, nv));
    R2_HIP(ctx, tmalloc((void**)&d_leaf_flag, nleaf_total * 4));
    R2_HIP(ctx, hipMemsetAsync(d_leaf_flag, 0, nleaf_total * 4, st));
    // grid-stride over the voxels: 8 blocks per CU are enough to stream the records
    const int blocks = (int)std::min<size_t>((nv + 255) / 256, 2048);
    hipLaunchKernelGGL(k_active, dim3(blocks), dim3(256), 0, st, d_rec, nx, ny, nz, d_active, d_leaf_flag, nlx, nly,
                       d_bbox, d_max);
    R2_HIP(ctx, hipGetLastError());
    std::vector<uint32_t> flags(nleaf_total);
    R2_HIP(ctx, hipMemcpyAsync(flags.data(), d_leaf_flag, nleaf_total * 4, hipMemcpyDeviceToHost, st));
    R2_HIP(ctx, hipMemcpyAsync(bbox, d_bbox, sizeof(bbox), hipMemcpyDeviceToHost, st));
    R2_HIP(ctx, hipMemcpyAsync(mm, d_max, 8, hipMemcpyDeviceToHost, st));
    R2_HIP(ctx, hipStreamSynchronize(st));
    for (size_t i = 0; i < nleaf_total; ++i) table[i] = flags[i] ? (int32_t)(nleaves++) : -1;
    if (nleaves == 0) return fail(ctx, R2_ERR_INVALID, "grid has no active voxel");
    if ((rc = dmalloc(ctx, g, &d_leaves, (size_t)nleaves * kLeafVoxels * 2))) return rc;
    d_dense_active = d_active;
  } else {
    // sparse: the caller's leaves in their order
    nleaves = desc->n_leaves;
    if (!desc->leaf_origins) return fail(ctx, R2_ERR_INVALID, "grid: sparse records without leaf origins");
    uint64_t covered = 0;
    for (int64_t k = 0; k < nleaves; ++k) {
      const int32_t* o = desc->leaf_origins + 3 * k;
      if ((o[0] | o[1] | o[2]) & 7 || o[0] < 0 || o[1] < 0 || o[2] < 0 || o[0] >= nx || o[1] >= ny || o[2] >= nz)
        return fail(ctx, R2_ERR_INVALID, "grid: leaf origin " + std::to_string(k) +
                                             " is not a multiple of 8 inside the index extent");
      int32_t& t = table[((size_t)(o[2] >> 3) * nly + (o[1] >> 3)) * nlx + (o[0] >> 3)];
      if (t >= 0) return fail(ctx, R2_ERR_INVALID, "grid: two leaves at one origin");
      t = (int32_t)k;
      covered += (uint64_t)std::min(8, nx - o[0]) * std::min(8, ny - o[1]) * std::min(8, nz - o[2]);
    }
    covers_all = covered == (uint64_t)nx * ny * nz;
    int32_t* d_orig = nullptr;
    uint8_t* d_mask = nullptr;
    R2_HIP(ctx, tmalloc((void**)&d_orig, (size_t)nleaves * 12));
    R2_HIP(ctx, hipMemcpyAsync(d_orig, desc->leaf_origins, (size_t)nleaves * 12, hipMemcpyHostToDevice, st));
    if (desc->leaf_masks) {
      R2_HIP(ctx, tmalloc((void**)&d_mask, (size_t)nleaves * 64));
      R2_HIP(ctx, hipMemcpyAsync(d_mask, desc->leaf_masks, (size_t)nleaves * 64, hipMemcpyHostToDevice, st));
    }
    if ((rc = dmalloc(ctx, g, &d_leaves, (size_t)nleaves * kLeafVoxels * 2))) return rc;
    hipLaunchKernelGGL(k_sparse_leaves, dim3((unsigned)std::min<int64_t>(nleaves, 65536)), dim3(256), 0, st, d_rec,
                       d_mask, d_orig, nleaves, nx, ny, nz, d_leaves, d_bbox, d_max);
    R2_HIP(ctx, hipGetLastError());
    R2_HIP(ctx, hipMemcpyAsync(bbox, d_bbox, sizeof(bbox), hipMemcpyDeviceToHost, st));
    R2_HIP(ctx, hipMemcpyAsync(mm, d_max, 8, hipMemcpyDeviceToHost, st));
    R2_HIP(ctx, hipStreamSynchronize(st));
    if (bbox[0] > bbox[3]) return fail(ctx, R2_ERR_INVALID, "grid has no active voxel");
  }
  // read the copied extrema only after the stream has finished the copies
  const uint32_t maxbits = mm[0];
  uint32_t minbits = (mm[1] & 0x80000000u) ? (mm[1] & 0x7FFFFFFFu) : ~mm[1];
  if (!covers_all) {
    float mn;
    memcpy(&mn, &minbits, 4);
    if (!(mn <= 0.f)) minbits = 0u;  // the voxels outside every leaf read 0
  }

  DGrid& d = g->dev;
  d.nx = nx; d.ny = ny; d.nz = nz;
  d.nlx = nlx; d.nly = nly; d.nlz = nlz;
  d.precision = desc->precision;
  int32_t* d_table = nullptr;
  if ((rc = dmalloc(ctx, g, &d_table, nleaf_total))) return rc;
  R2_HIP(ctx, hipMemcpyAsync(d_table, table.data(), nleaf_total * 4, hipMemcpyHostToDevice, st));
  if (!sparse) {
    hipLaunchKernelGGL(k_leaves, dim3((unsigned)std::min<size_t>(nleaf_total, 65536)), dim3(256), 0, st, d_rec,
                       d_dense_active, nx, ny, nz, d_table, nlx, nly, nlz, d_leaves);
    R2_HIP(ctx, hipGetLastError());
  }
  d.leaf_table = d_table;
  d.leaves = d_leaves;
  d.leaves_packed = nullptr;
  d.leaves_cold = nullptr;
  if (desc->precision == R2_PREC_PACKED) {
    uint4* d_hot = nullptr;
    uint2* d_cold = nullptr;
    const size_t nv = (size_t)nleaves * kLeafVoxels;
    if ((rc = dmalloc(ctx, g, &d_hot, nv))) return rc;
    if ((rc = dmalloc(ctx, g, &d_cold, nv))) return rc;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)std::min<size_t>((nv + 255) / 256, 1u << 16)), dim3(256), 0, st, d_leaves,
                       nv, d_hot, d_cold);
    R2_HIP(ctx, hipGetLastError());
    d.leaves_packed = d_hot;
    d.leaves_cold = d_cold;
  }
  const PoolDens pool{d_table, d_leaves, nx, ny, nz, nlx, nly};

  // brick grid
  int nb[3];
  const int n3[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    const int b = (int)ceilf((float)n3[a] / 8.f);
    nb[a] = (int)ceilf((float)b / 8.f) * 8;
    if (nb[a] >= 1024) return fail(ctx, R2_ERR_UNSUPPORTED, "exceeded max brick count of 1024 per axis");
  }
  const size_t nbt = (size_t)nb[0] * nb[1] * nb[2];
  uint32_t *d_range = nullptr, *d_nonempty = nullptr, *d_slot = nullptr;
  R2_HIP(ctx, tmalloc((void**)&d_range, nbt * 4));
  R2_HIP(ctx, tmalloc((void**)&d_nonempty, nbt * 4));
  R2_HIP(ctx, tmalloc((void**)&d_slot, nbt * 4));
  {
    const size_t threads = nbt * 64;
    hipLaunchKernelGGL(k_brick_range, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, pool, nb[0], nb[1],
                       nb[2], d_range, d_nonempty);
    R2_HIP(ctx, hipGetLastError());
  }
  std::vector<uint32_t> ne(nbt), slot(nbt);
  R2_HIP(ctx, hipMemcpyAsync(ne.data(), d_nonempty, nbt * 4, hipMemcpyDeviceToHost, st));
  R2_HIP(ctx, hipStreamSynchronize(st));
  uint32_t nne = 0;
  for (size_t i = 0; i < nbt; ++i) slot[i] = ne[i] ? nne++ : 0u;
  R2_HIP(ctx, hipMemcpyAsync(d_slot, slot.data(), nbt * 4, hipMemcpyHostToDevice, st));
  uint2* d_bricks = nullptr;
  uint8_t* d_atlas = nullptr;
  float* d_maj[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t* d_mip[3] = {nullptr, nullptr, nullptr};
  if ((rc = dmalloc(ctx, g, &d_bricks, nbt))) return rc;
  if ((rc = dmalloc(ctx, g, &d_atlas, (size_t)std::max<uint32_t>(nne, 1) * 512))) return rc;
  // the four majorant levels in one buffer: level l (nbt / 8^l texels, nbt a
  // multiple of 512) starts at (nbt / 64) x {0, 64, 72, 73}[l], so that a lookup
  // finds its level's table by arithmetic
  if ((rc = dmalloc(ctx, g, &d_maj[0], nbt + nbt / 8 + nbt / 64 + nbt / 512))) return rc;
  d_maj[1] = d_maj[0] + nbt;
  d_maj[2] = d_maj[1] + nbt / 8;
  d_maj[3] = d_maj[2] + nbt / 64;
  {
    const size_t threads = nbt * 64;
    hipLaunchKernelGGL(k_brick_atlas, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, pool, nb[0], nb[1],
                       nb[2], d_range, d_slot, d_nonempty, d_bricks, d_atlas, d_maj[0]);
    R2_HIP(ctx, hipGetLastError());
  }
  for (int l = 0; l < 3; ++l) {
    const int sx = nb[0] >> (l + 1), sy = nb[1] >> (l + 1), sz = nb[2] >> (l + 1);
    const int px = nb[0] >> l, py = nb[1] >> l;
    const size_t cnt = (size_t)sx * sy * sz;
    if ((rc = dmalloc(ctx, g, &d_mip[l], cnt))) return rc;
    const uint32_t* src = l == 0 ? d_range : d_mip[l - 1];
    hipLaunchKernelGGL(k_brick_mip, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, src, px, py, d_mip[l], sx, sy,
                       sz, d_maj[l + 1]);
    R2_HIP(ctx, hipGetLastError());
  }
  // keep the level-0 ranges for the parity probe (d_range is temporary)
  uint32_t* d_range_keep = nullptr;
  if ((rc = dmalloc(ctx, g, &d_range_keep, nbt))) return rc;
  R2_HIP(ctx, hipMemcpyAsync(d_range_keep, d_range, nbt * 4, hipMemcpyDeviceToDevice, st));
  d.bricks = d_bricks;
  d.atlas = d_atlas;
  for (int l = 0; l < 4; ++l) d.maj[l] = d_maj[l];
  for (int a = 0; a < 3; ++a) d.nb[a] = nb[a];
  R2_HIP(ctx, hipStreamSynchronize(st));

  // transforms (CNVDBMedium::init, nvdb_medium.cu:152-170)
  r2_grid_info& info = g->info;
  for (int a = 0; a < 3; ++a) {
    info.ibb_min[a] = bbox[a];
    info.size[a] = bbox[3 + a] - bbox[a] + 1;
    info.n_bricks[a] = nb[a];
    d.ibb_min[a] = info.ibb_min[a];
    d.size[a] = info.size[a];
  }
  m44 nano = identity44();
  const float* mf = desc->map_mat;
  nano.c[0][0] = mf[0]; nano.c[0][1] = mf[3]; nano.c[0][2] = mf[6];
  nano.c[1][0] = mf[1]; nano.c[1][1] = mf[4]; nano.c[1][2] = mf[7];
  nano.c[2][0] = mf[2]; nano.c[2][1] = mf[5]; nano.c[2][2] = mf[8];
  nano.c[3][0] = desc->map_vec[0]; nano.c[3][1] = desc->map_vec[1]; nano.c[3][2] = desc->map_vec[2];
  m44 i2n = identity44();
  i2n.c[0][0] = (float)info.size[0]; i2n.c[1][1] = (float)info.size[1]; i2n.c[2][2] = (float)info.size[2];
  i2n.c[3][0] = (float)info.ibb_min[0]; i2n.c[3][1] = (float)info.ibb_min[1]; i2n.c[3][2] = (float)info.ibb_min[2];
  const m43 i2m = to43(mul44(nano, i2n));
  const m43 m2i = inverse43(i2m);
  m43 i2s;
  for (int k = 0; k < 12; ++k) i2s.m[k] = 0.f;
  i2s.m[0] = (float)info.size[0]; i2s.m[4] = (float)info.size[1]; i2s.m[8] = (float)info.size[2];
  i2s.m[9] = (float)info.ibb_min[0]; i2s.m[10] = (float)info.ibb_min[1]; i2s.m[11] = (float)info.ibb_min[2];
  const m43 s2i = inverse43(i2s);
  const m43 m2s = to43(mul44(to44(i2s), to44(m2i)));
  const m43 s2m = to43(mul44(to44(i2m), to44(s2i)));
  d.i2m = i2m; d.m2i = m2i; d.m2s = m2s; d.s2m = s2m;
  memcpy(info.index_to_model, i2m.m, 48);
  memcpy(info.model_to_index, m2i.m, 48);
  memcpy(info.model_to_scaled, m2s.m, 48);
  memcpy(info.scaled_to_model, s2m.m, 48);
  const v3 c0 = mul_point(i2m, mk(0.f, 0.f, 0.f)), c1 = mul_point(i2m, mk(1.f, 1.f, 1.f));
  info.bbox_min[0] = fminf(c0.x, c1.x); info.bbox_min[1] = fminf(c0.y, c1.y); info.bbox_min[2] = fminf(c0.z, c1.z);
  info.bbox_max[0] = fmaxf(c0.x, c1.x); info.bbox_max[1] = fmaxf(c0.y, c1.y); info.bbox_max[2] = fmaxf(c0.z, c1.z);
  float maxd;
  memcpy(&maxd, &maxbits, 4);
  info.max_density = maxd;
  memcpy(&info.min_density, &minbits, 4);
  info.inv_max_density = 1.f / maxd;
  info.n_leaves = nleaves;
  info.n_nonempty_bricks = nne;
  int64_t bytes = 0;
  bytes += (int64_t)nleaf_total * 4 + nleaves * kLeafVoxels * 32 + (int64_t)nbt * (8 + 4 + 4) + (int64_t)nne * 512;
  for (int l = 0; l < 3; ++l) bytes += (int64_t)(nb[0] >> (l + 1)) * (nb[1] >> (l + 1)) * (nb[2] >> (l + 1)) * 8;
  info.device_bytes = bytes;
  d.inv_max_density = info.inv_max_density;
  // glm::max(a, glm::max(b, c)) = (x < y) ? y : x
  {
    const float a = desc->sigma_a[0] + desc->sigma_s[0], b = desc->sigma_a[1] + desc->sigma_s[1],
                c = desc->sigma_a[2] + desc->sigma_s[2];
    const float bc = (b < c) ? c : b;
    d.sigma_t = (a < bc) ? bc : a;
  }
  {
    // a positive normal float is a power of two iff its mantissa bits are zero
    auto pow2 = [](float v) {
      uint32_t u;
      memcpy(&u, &v, 4);
      return v > 0.f && (u & 0x007fffffu) == 0u && ((u >> 23) & 0xffu) != 0u && ((u >> 23) & 0xffu) != 0xffu;
    };
    d.size_pow2 = 1;
    for (int a = 0; a < 3; ++a) {
      d.inv_size[a] = 1.f / (float)d.size[a];
      d.size_pow2 = d.size_pow2 && pow2((float)d.size[a]);
    }
    d.inv_sigma_t = 1.f / d.sigma_t;
    d.sigma_t_pow2 = pow2(d.sigma_t) ? 1 : 0;
  }
  for (int a = 0; a < 3; ++a) {
    d.sigma_s[a] = desc->sigma_s[a];
    g->sigma_a[a] = desc->sigma_a[a];
    g->sigma_s[a] = desc->sigma_s[a];
  }
  g->voxel_size_filtering = desc->voxel_size_filtering;
  // level-0 ranges and mip buffers stay reachable for r2_grid_download_bricks
  g->range0 = d_range_keep;
  for (int l = 0; l < 3; ++l) g->mip[l] = d_mip[l];
  return R2_OK;
}

int grid_lookup(r2_ctx* ctx, const r2_grid* g, int n, const int32_t* coords, uint8_t* out) {
  void *dc = nullptr, *dout = nullptr;
  struct F {
    void** q[2];
    ~F() { for (auto x : q) if (*x) (void)hipFree(*x); }
  } f{{&dc, &dout}};
  R2_HIP(ctx, hipMalloc(&dc, (size_t)n * 12));
  R2_HIP(ctx, hipMalloc(&dout, (size_t)n * 32));
  R2_HIP(ctx, hipMemcpy(dc, coords, (size_t)n * 12, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_lookup, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, g->dev, n, (const int32_t*)dc,
                     (uint4*)dout);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipStreamSynchronize(ctx->stream));
  R2_HIP(ctx, hipMemcpy(out, dout, (size_t)n * 32, hipMemcpyDeviceToHost));
  return R2_OK;
}

int grid_sample(r2_ctx* ctx, const r2_grid* g, int n, const float* pos, float* out) {
  void *dp = nullptr, *dout = nullptr;
  struct F {
    void** q[2];
    ~F() { for (auto x : q) if (*x) (void)hipFree(*x); }
  } f{{&dp, &dout}};
  R2_HIP(ctx, hipMalloc(&dp, (size_t)n * 12));
  R2_HIP(ctx, hipMalloc(&dout, (size_t)n * 17 * 4));
  R2_HIP(ctx, hipMemcpy(dp, pos, (size_t)n * 12, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sample, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, g->dev, n, (const float*)dp,
                     (float*)dout);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipStreamSynchronize(ctx->stream));
  R2_HIP(ctx, hipMemcpy(out, dout, (size_t)n * 17 * 4, hipMemcpyDeviceToHost));
  return R2_OK;
}

// scene-intersection probe: CDeviceScene::intersect of each ray against the
// scene's instance boxes (linear scan or hierarchy) and surfaces
__global__ void k_scene_intersect(const r2::DScene* __restrict__ sp, const float* __restrict__ rays, int n,
                                  float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const r2::DScene& s = r2::kconst(sp);
  const float* q = rays + 7 * (size_t)i;
  r2::Ray r = r2::mkray(r2::mk(q[0], q[1], q[2]), r2::mk(q[3], q[4], q[5]), q[6], -1);
  const r2::Hit h = r2::scene_intersect<true, false>(s, r);
  float* o = out + 4 * (size_t)i;
  o[0] = (float)h.hit;
  o[1] = h.t;
  o[2] = (float)h.medium;
  o[3] = (float)h.surf;
}

int scene_intersect_probe(r2_ctx* ctx, const HostScene& hs, int n, const float* rays7, float* out4) {
  hipStream_t st = ctx->stream;
  int rc = upload_scene(ctx, hs, true);
  if (rc) return rc;
  void *dr = nullptr, *dout = nullptr;
  R2_HIP(ctx, pool_get(ctx, 0, (size_t)n * 7 * 4, &dr));
  R2_HIP(ctx, pool_get(ctx, 1, (size_t)n * 4 * 4, &dout));
  R2_HIP(ctx, hipMemcpyAsync(dr, rays7, (size_t)n * 7 * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_scene_intersect, dim3((n + 255) / 256), dim3(256), 0, st, ctx->d_scene, (const float*)dr, n,
                     (float*)dout);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipMemcpyAsync(out4, dout, (size_t)n * 4 * 4, hipMemcpyDeviceToHost, st));
  R2_HIP(ctx, hipStreamSynchronize(st));
  return R2_OK;
}

int dda_cells(r2_ctx* ctx, int n, const float* rays, float sigma_t, int collide_every, int max_cells, int32_t* cells,
              int32_t* counts) {
  void *dr = nullptr, *dc = nullptr, *dn = nullptr;
  struct F {
    void** q[3];
    ~F() { for (auto x : q) if (*x) (void)hipFree(*x); }
  } f{{&dr, &dc, &dn}};
  const size_t cbytes = (size_t)n * max_cells * 16;
  R2_HIP(ctx, hipMalloc(&dr, (size_t)n * 28));
  R2_HIP(ctx, hipMalloc(&dc, cbytes));
  R2_HIP(ctx, hipMalloc(&dn, (size_t)n * 4));
  R2_HIP(ctx, hipMemcpy(dr, rays, (size_t)n * 28, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_dda_cells, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, (const float*)dr, n, sigma_t,
                     collide_every, max_cells, (int32_t*)dc, (int32_t*)dn);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipStreamSynchronize(ctx->stream));
  R2_HIP(ctx, hipMemcpy(cells, dc, cbytes, hipMemcpyDeviceToHost));
  R2_HIP(ctx, hipMemcpy(counts, dn, (size_t)n * 4, hipMemcpyDeviceToHost));
  return R2_OK;
}

}  // namespace r2host
