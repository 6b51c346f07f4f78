// NanoVDB grid files (.nvdb) of filtered voxel records.
//
// COpenvdbBackend::writeToFile (raytracer/src/filtering/openvdb_backend.cpp:134-157)
// turns the filter's OpenVDB Vec4d grid into a NanoVDB grid (nanovdb::openToNanoVDB)
// and writes it with nanovdb::io::writeGrid; CNVDBMedium reads it back
// (raytracer/src/medium/nvdb_medium.cu:130-173, 308-323) through the index bounding
// box, the map and voxel lookups.  The grid holds, per active voxel, one
// SFilteredDataCompact record in the bits of a Vec4d; a voxel is active when its
// density is > 0 or it is one of the 8 corners of the grid (setValues,
// openvdb_backend.cpp:108-132); the background is the all-zero record with ior = 1
// (openvdb_backend.cpp:30-42); the transform is the filter grid's index -> model
// matrix (setupGrid, openvdb_backend.cpp:84-95), which openToNanoVDB stores as the
// NanoVDB map (row-major 3x3 + translation, with its inverse, in float and double).
//
// The format is written and read here without the NanoVDB library, from the
// published layout of NanoVDB 32.3 (the version shipped with OpenVDB 9.0):
//   file:   Header {magic "NanoVDB0", version, grid count, codec} (16 B), per grid a
//           MetaData record (176 B) and the grid name, then the grid buffer;
//   buffer: GridData (672 B) | TreeData (64 B) | RootData + tile table | upper
//           internal nodes (32^3 children) | lower internal nodes (16^3) | leaves
//           (8^3), each level in tree order (root key, then child index: x major);
//           child references are signed byte offsets from the referring node; root
//           tiles are keyed by the upper node's origin, 21 bits per axis.
// The world bounding box is the index box (max + 1) mapped through the map in
// double, corner by corner; the voxel size is the length of each column of the
// map.  Node statistics (min / max / average / deviation) are written as absent:
// the grid flags carry HasBBox and IsBreadthFirst only, and the checksum field
// holds the "none" value (all ones).  Nodes are streamed to the file one at a time.
#include <algorithm>
#include <cmath>
#include <array>
#include <climits>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>
#include "r2_internal.hpp"

namespace {

constexpr uint64_t kMagic = 0x304244566f6e614eull;  // "NanoVDB0", little endian
constexpr uint32_t kMajor = 32;
constexpr uint32_t kVersion = (kMajor << 21) | (3u << 10) | 3u;
constexpr uint32_t kGridTypeVec4d = 18, kGridClassUnknown = 0;
constexpr uint32_t kFlagHasBBox = 1u << 1, kFlagBreadthFirst = 1u << 5;
constexpr size_t kFileHeader = 16, kMetaData = 176, kGridData = 672, kTreeData = 64;
// GridData fields
constexpr size_t kGdMap = 296, kGdWorldBBox = 560, kGdVoxelSize = 608, kGdClass = 632, kGdType = 636;
// RootData: bbox 24 | table size 4 | pad 4 | background, min, max (Vec4d) | average,
// deviation (double) = 144 B, 32-byte aligned; tile: key 8 | child 8 | state 4 | pad 4 |
// value 32 = 56 B, 32-byte aligned
constexpr size_t kRootData = 160, kRootTile = 64;
// InternalData: bbox 24 | flags 8 | value mask | child mask | min, max (Vec4d) | average,
// deviation (double), then the 32-byte aligned table of 32-byte entries (value or offset)
constexpr size_t internal_header(size_t mask_bytes) { return (24 + 8 + 2 * mask_bytes + 64 + 16 + 31) & ~size_t(31); }
constexpr size_t kUpperMask = 4096, kLowerMask = 512;
constexpr size_t kUpperHeader = internal_header(kUpperMask), kLowerHeader = internal_header(kLowerMask);
constexpr size_t kUpperSize = kUpperHeader + 32768 * 32, kLowerSize = kLowerHeader + 4096 * 32;
// LeafData: bbox min 12 | bbox extent 3 | flags 1 | value mask 64 | min, max (Vec4d) |
// average, deviation (double) = 160 B, then 512 values of 32 B
constexpr size_t kLeafHeader = 160, kLeafSize = kLeafHeader + 512 * 32;
constexpr int kMaxExtent = 1 << 16;

using Rec = std::array<uint8_t, 32>;

Rec background_record() {
  Rec r{};
  r[28] = 0x00;  // ior = half(1.0)
  r[29] = 0x3C;
  return r;
}

template <typename T>
void put(uint8_t* p, size_t off, const T& v) {
  memcpy(p + off, &v, sizeof(T));
}
template <typename T>
T get(const std::vector<uint8_t>& b, size_t off) {
  if (off + sizeof(T) > b.size()) throw std::runtime_error("nvdb: read past a node");
  T v;
  memcpy(&v, b.data() + off, sizeof(T));
  return v;
}

struct Box {
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  void add(int x, int y, int z) {
    lo[0] = std::min(lo[0], x), lo[1] = std::min(lo[1], y), lo[2] = std::min(lo[2], z);
    hi[0] = std::max(hi[0], x), hi[1] = std::max(hi[1], y), hi[2] = std::max(hi[2], z);
  }
  void add(const Box& b) {
    for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], b.lo[a]), hi[a] = std::max(hi[a], b.hi[a]);
  }
};

struct Leaf {
  int32_t o[3];
  uint8_t mask[64];
  Box box;
};
struct Inner {  // internal node: origin, first child and child count in the level below
  int32_t o[3];
  size_t first, count;
  Box box;
};

// position of a leaf in tree order: upper origin, child index in the upper node,
// child index in the lower node (coordinates are >= 0 here)
std::array<int32_t, 9> tree_key(const int32_t o[3]) {
  return {o[0] >> 12, o[1] >> 12, o[2] >> 12, (o[0] >> 7) & 31, (o[1] >> 7) & 31, (o[2] >> 7) & 31,
          (o[0] >> 3) & 15, (o[1] >> 3) & 15, (o[2] >> 3) & 15};
}

// internal nodes of one level over the (tree-ordered) nodes of the level below
template <typename Child>
std::vector<Inner> group(const std::vector<Child>& kids, int shift) {
  std::vector<Inner> out;
  const int32_t m = ~((1 << shift) - 1);
  for (size_t i = 0; i < kids.size(); ++i) {
    const int32_t o[3] = {kids[i].o[0] & m, kids[i].o[1] & m, kids[i].o[2] & m};
    if (out.empty() || memcmp(out.back().o, o, 12) != 0) out.push_back(Inner{{o[0], o[1], o[2]}, i, 0, Box{}});
    out.back().count++;
    out.back().box.add(kids[i].box);
  }
  return out;
}

void put_bbox(uint8_t* p, const Box& b) {
  for (int a = 0; a < 3; ++a) put<int32_t>(p, 4 * a, b.lo[a]), put<int32_t>(p, 12 + 4 * a, b.hi[a]);
}

void fill_records(uint8_t* p, size_t n, const Rec& r) {
  for (size_t i = 0; i < n; ++i) memcpy(p + 32 * i, r.data(), 32);
}

// one internal node: bbox, child mask, child offsets relative to the node
template <typename Child>
void write_internal(std::ofstream& f, std::vector<uint8_t>& nb, size_t header, size_t mask_bytes, int log2dim,
                    int child_shift, const Inner& n, size_t node_pos, const std::vector<Child>& kids,
                    size_t kid_base, size_t kid_size, const Rec& bg) {
  std::fill(nb.begin(), nb.end(), 0);
  uint8_t* p = nb.data();
  put_bbox(p, n.box);
  fill_records(p + 32 + 2 * mask_bytes, 2, bg);  // min, max (absent)
  fill_records(p + header, size_t(1) << (3 * log2dim), bg);
  for (size_t k = n.first; k < n.first + n.count; ++k) {
    const int32_t* c = kids[k].o;
    const uint32_t ix = (uint32_t)(c[0] - n.o[0]) >> child_shift, iy = (uint32_t)(c[1] - n.o[1]) >> child_shift,
                   iz = (uint32_t)(c[2] - n.o[2]) >> child_shift;
    const uint32_t idx = (ix << (2 * log2dim)) | (iy << log2dim) | iz;
    p[32 + mask_bytes + idx / 8] |= (uint8_t)(1u << (idx % 8));  // child mask
    put<int64_t>(p, header + 32 * idx, (int64_t)(kid_base + k * kid_size) - (int64_t)node_pos);
  }
  f.write((const char*)p, (std::streamsize)nb.size());
}

// NanoVDB Map of an affine index -> world transform: A (row-major 3x3), t, and the
// inverse of A (the inverse map subtracts t first, so it needs no translation).
struct MapD {
  double A[9], t[3], Ai[9];
  double voxel[3];
  void apply(const double x[3], double y[3]) const {
    for (int r = 0; r < 3; ++r) y[r] = A[3 * r] * x[0] + A[3 * r + 1] * x[1] + A[3 * r + 2] * x[2] + t[r];
  }
};
MapD make_map(const float* mat, const float* vec) {
  MapD m{};
  for (int i = 0; i < 9; ++i) m.A[i] = mat ? (double)mat[i] : (i % 4 == 0 ? 1.0 : 0.0);
  for (int i = 0; i < 3; ++i) m.t[i] = vec ? (double)vec[i] : 0.0;
  const double* a = m.A;
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  if (!(det != 0.0) || !std::isfinite(det)) throw std::invalid_argument("nvdb: the index -> world map is singular");
  const double id = 1.0 / det;
  m.Ai[0] = (a[4] * a[8] - a[5] * a[7]) * id;
  m.Ai[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  m.Ai[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  m.Ai[3] = (a[5] * a[6] - a[3] * a[8]) * id;
  m.Ai[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  m.Ai[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  m.Ai[6] = (a[3] * a[7] - a[4] * a[6]) * id;
  m.Ai[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  m.Ai[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  for (int c = 0; c < 3; ++c) m.voxel[c] = std::sqrt(a[c] * a[c] + a[3 + c] * a[3 + c] + a[6 + c] * a[6 + c]);
  return m;
}

void copy_err(char* err, int32_t err_len, const char* what) {
  if (err && err_len > 0) {
    strncpy(err, what, (size_t)err_len - 1);
    err[err_len - 1] = 0;
  }
}

}  // namespace

namespace r2host {

void nvdb_write(const std::string& path, int nx, int ny, int nz, const uint8_t* records, const float* map_mat,
                const float* map_vec) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nx > kMaxExtent || ny > kMaxExtent || nz > kMaxExtent)
    throw std::invalid_argument("nvdb: grid size outside 1..65536 per axis");
  const MapD map = make_map(map_mat, map_vec);
  const Rec bg = background_record();
  // leaves with at least one active voxel
  const int lx = (nx + 7) / 8, ly = (ny + 7) / 8, lz = (nz + 7) / 8;
  std::vector<int32_t> slot((size_t)lx * ly * lz, -1);
  std::vector<Leaf> leaves;
  uint64_t voxels = 0;
  Box all;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        float d;
        memcpy(&d, records + (((size_t)z * ny + y) * nx + x) * 32, 4);
        const bool corner = (x == 0 || x == nx - 1) && (y == 0 || y == ny - 1) && (z == 0 || z == nz - 1);
        if (!(d > 0.f || corner)) continue;
        int32_t& s = slot[((size_t)(z >> 3) * ly + (y >> 3)) * lx + (x >> 3)];
        if (s < 0) {
          s = (int32_t)leaves.size();
          leaves.push_back(Leaf{{x & ~7, y & ~7, z & ~7}, {}, Box{}});
        }
        Leaf& L = leaves[(size_t)s];
        const uint32_t n = (uint32_t)(((x & 7) << 6) | ((y & 7) << 3) | (z & 7));
        L.mask[n / 8] |= (uint8_t)(1u << (n % 8));
        L.box.add(x, y, z);
        all.add(x, y, z);
        ++voxels;
      }
  std::sort(leaves.begin(), leaves.end(), [](const Leaf& a, const Leaf& b) { return tree_key(a.o) < tree_key(b.o); });
  const std::vector<Inner> lowers = group(leaves, 7), uppers = group(lowers, 12);

  // world box: the 8 corners of the index box (max + 1) through the map
  double wlo[3] = {INFINITY, INFINITY, INFINITY}, whi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < 8; ++k) {
    const double x[3] = {(double)((k & 1) ? all.hi[0] + 1 : all.lo[0]), (double)((k & 2) ? all.hi[1] + 1 : all.lo[1]),
                         (double)((k & 4) ? all.hi[2] + 1 : all.lo[2])};
    double y[3];
    map.apply(x, y);
    for (int a = 0; a < 3; ++a) wlo[a] = std::min(wlo[a], y[a]), whi[a] = std::max(whi[a], y[a]);
  }
  const size_t n_leaf = leaves.size(), n_lower = lowers.size(), n_upper = uppers.size();
  const size_t off_tree = kGridData, off_root = off_tree + kTreeData;
  const size_t off_upper = off_root + kRootData + n_upper * kRootTile;
  const size_t off_lower = off_upper + n_upper * kUpperSize;
  const size_t off_leaf = off_lower + n_lower * kLowerSize;
  const size_t total = off_leaf + n_leaf * kLeafSize;

  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  // file header, meta data, grid name (empty)
  {
    uint8_t h[kFileHeader + kMetaData + 1] = {};
    put<uint64_t>(h, 0, kMagic);
    put<uint32_t>(h, 8, kVersion);
    put<uint16_t>(h, 12, 1);  // grid count
    put<uint16_t>(h, 14, 0);  // codec: none
    uint8_t* m = h + kFileHeader;
    put<uint64_t>(m, 0, (uint64_t)total);  // grid size
    put<uint64_t>(m, 8, (uint64_t)total);  // file size (uncompressed)
    put<uint64_t>(m, 16, 0ull);            // hash of the empty name
    put<uint64_t>(m, 24, voxels);
    put<uint32_t>(m, 32, kGridTypeVec4d);
    put<uint32_t>(m, 36, kGridClassUnknown);
    for (int a = 0; a < 3; ++a) {
      put<double>(m, 40 + 8 * a, wlo[a]);
      put<double>(m, 64 + 8 * a, whi[a]);
      put<int32_t>(m, 88 + 4 * a, all.lo[a]);
      put<int32_t>(m, 100 + 4 * a, all.hi[a]);
      put<double>(m, 112 + 8 * a, map.voxel[a]);
    }
    put<uint32_t>(m, 136, 1u);  // name size, terminating NUL included
    put<uint32_t>(m, 140, (uint32_t)n_leaf);
    put<uint32_t>(m, 144, (uint32_t)n_lower);
    put<uint32_t>(m, 148, (uint32_t)n_upper);
    put<uint32_t>(m, 152, 1u);  // roots
    put<uint32_t>(m, 172, kVersion);
    f.write((const char*)h, sizeof(h));
  }
  // GridData, TreeData, RootData and its tiles
  {
    std::vector<uint8_t> b(off_upper, 0);
    uint8_t* g = b.data();
    put<uint64_t>(g, 0, kMagic);
    put<uint64_t>(g, 8, ~0ull);  // checksum: none
    put<uint32_t>(g, 16, kVersion);
    put<uint32_t>(g, 20, kFlagHasBBox | kFlagBreadthFirst);
    put<uint32_t>(g, 24, 0u);  // grid index
    put<uint32_t>(g, 28, 1u);  // grid count
    put<uint64_t>(g, 32, (uint64_t)total);
    // map: mMatF, mInvMatF, mVecF, mTaperF, mMatD, mInvMatD, mVecD, mTaperD
    for (int i = 0; i < 9; ++i) {
      put<float>(g, kGdMap + 4 * i, (float)map.A[i]);
      put<float>(g, kGdMap + 36 + 4 * i, (float)map.Ai[i]);
      put<double>(g, kGdMap + 88 + 8 * i, map.A[i]);
      put<double>(g, kGdMap + 160 + 8 * i, map.Ai[i]);
    }
    for (int i = 0; i < 3; ++i) {
      put<float>(g, kGdMap + 72 + 4 * i, (float)map.t[i]);
      put<double>(g, kGdMap + 232 + 8 * i, map.t[i]);
    }
    put<float>(g, kGdMap + 84, 1.f);
    put<double>(g, kGdMap + 256, 1.0);
    for (int a = 0; a < 3; ++a) {
      put<double>(g, kGdWorldBBox + 8 * a, wlo[a]);
      put<double>(g, kGdWorldBBox + 24 + 8 * a, whi[a]);
      put<double>(g, kGdVoxelSize + 8 * a, map.voxel[a]);
    }
    put<uint32_t>(g, kGdClass, kGridClassUnknown);
    put<uint32_t>(g, kGdType, kGridTypeVec4d);
    uint8_t* t = g + off_tree;  // byte offsets from the tree to the first leaf, lower, upper node, root
    put<uint64_t>(t, 0, n_leaf ? (uint64_t)(off_leaf - off_tree) : 0);
    put<uint64_t>(t, 8, n_lower ? (uint64_t)(off_lower - off_tree) : 0);
    put<uint64_t>(t, 16, n_upper ? (uint64_t)(off_upper - off_tree) : 0);
    put<uint64_t>(t, 24, (uint64_t)(off_root - off_tree));
    put<uint32_t>(t, 32, (uint32_t)n_leaf);
    put<uint32_t>(t, 36, (uint32_t)n_lower);
    put<uint32_t>(t, 40, (uint32_t)n_upper);
    put<uint64_t>(t, 56, voxels);
    uint8_t* r = g + off_root;
    put_bbox(r, all);
    put<uint32_t>(r, 24, (uint32_t)n_upper);
    fill_records(r + 32, 3, bg);  // background, min, max
    for (size_t k = 0; k < n_upper; ++k) {
      uint8_t* tile = r + kRootData + k * kRootTile;
      const int32_t* o = uppers[k].o;
      const uint64_t key = (uint64_t)((uint32_t)o[2] >> 12) | ((uint64_t)((uint32_t)o[1] >> 12) << 21) |
                           ((uint64_t)((uint32_t)o[0] >> 12) << 42);
      put<uint64_t>(tile, 0, key);
      put<int64_t>(tile, 8, (int64_t)(off_upper + k * kUpperSize) - (int64_t)off_root);
      put<uint32_t>(tile, 16, 0u);
      memcpy(tile + 24, bg.data(), 32);
    }
    f.write((const char*)b.data(), (std::streamsize)b.size());
  }
  // upper nodes, lower nodes, leaves
  std::vector<uint8_t> nb(kUpperSize);
  for (size_t k = 0; k < n_upper; ++k)
    write_internal(f, nb, kUpperHeader, kUpperMask, 5, 7, uppers[k], off_upper + k * kUpperSize, lowers, off_lower,
                   kLowerSize, bg);
  nb.assign(kLowerSize, 0);
  for (size_t k = 0; k < n_lower; ++k)
    write_internal(f, nb, kLowerHeader, kLowerMask, 4, 3, lowers[k], off_lower + k * kLowerSize, leaves, off_leaf,
                   kLeafSize, bg);
  nb.assign(kLeafSize, 0);
  for (const Leaf& L : leaves) {
    std::fill(nb.begin(), nb.end(), 0);
    uint8_t* p = nb.data();
    for (int a = 0; a < 3; ++a) put<int32_t>(p, 4 * a, L.box.lo[a]);
    for (int a = 0; a < 3; ++a) p[12 + a] = (uint8_t)(L.box.hi[a] - L.box.lo[a]);
    p[15] = 0x2;  // the leaf has a bounding box
    memcpy(p + 16, L.mask, 64);
    fill_records(p + 80, 2, bg);  // min, max (absent)
    fill_records(p + kLeafHeader, 512, bg);
    for (uint32_t n = 0; n < 512; ++n) {
      if (!((L.mask[n / 8] >> (n % 8)) & 1u)) continue;
      const int x = L.o[0] + (int)(n >> 6), y = L.o[1] + (int)((n >> 3) & 7), z = L.o[2] + (int)(n & 7);
      memcpy(p + kLeafHeader + 32 * n, records + (((size_t)z * ny + y) * nx + x) * 32, 32);
    }
    f.write((const char*)p, (std::streamsize)nb.size());
  }
  if (!f) throw std::runtime_error("nvdb: write failed: " + path);
}

// Header, map and index box of the first grid of a NanoVDB file; with `records`,
// also the dense records over the index box (inactive voxels hold the background),
// x fastest.
// an opened NanoVDB file: its grid's header fields, node offsets and index box
struct NvdbFile {
  std::ifstream f;
  std::string path;
  uint64_t base = 0, grid_size = 0;
  uint64_t leaf_off = 0, lower_off = 0, upper_off = 0, root_off = 0;
  uint32_t n_leaf = 0, n_lower = 0, n_upper = 0;
  std::vector<uint8_t> root;
  int32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  std::vector<uint8_t> read_at(uint64_t off, size_t n) {
    if (off > grid_size || n > grid_size - off) throw std::runtime_error("nvdb: node outside the grid buffer");
    std::vector<uint8_t> b(n);
    f.seekg((std::streamoff)(base + off));
    f.read((char*)b.data(), (std::streamsize)n);
    if (!f) throw std::runtime_error("nvdb: read failed: " + path);
    return b;
  }
};

void nvdb_open(const std::string& path, r2_nvdb_info* info, NvdbFile& F) {
  F.path = path;
  F.f.open(path, std::ios::binary | std::ios::ate);
  std::ifstream& f = F.f;
  if (!f) throw std::runtime_error("cannot open " + path);
  const uint64_t file_bytes = (uint64_t)f.tellg();
  f.seekg(0);
  std::vector<uint8_t> h(kFileHeader + kMetaData);
  f.read((char*)h.data(), (std::streamsize)h.size());
  if (!f || get<uint64_t>(h, 0) != kMagic) throw std::runtime_error("not a NanoVDB file: " + path);
  if ((get<uint32_t>(h, 8) >> 21) != kMajor) throw std::runtime_error("nvdb: unsupported NanoVDB major version");
  if (get<uint16_t>(h, 12) < 1) throw std::runtime_error("nvdb: the file holds no grid");
  if (get<uint16_t>(h, 14) != 0) throw std::runtime_error("nvdb: compressed grids are not supported");
  F.grid_size = get<uint64_t>(h, kFileHeader + 0);
  const uint32_t name_size = get<uint32_t>(h, kFileHeader + 136);
  F.base = kFileHeader + kMetaData + (uint64_t)name_size;
  if (name_size > (1u << 20) || F.grid_size < kGridData + kTreeData + kRootData || F.base > file_bytes ||
      F.grid_size > file_bytes - F.base)
    throw std::runtime_error("nvdb: truncated file or bad grid size: " + path);
  const std::vector<uint8_t> g = F.read_at(0, kGridData + kTreeData);
  if (get<uint64_t>(g, 0) != kMagic || (get<uint32_t>(g, 16) >> 21) != kMajor)
    throw std::runtime_error("nvdb: bad grid header");
  if (get<uint32_t>(g, kGdType) != kGridTypeVec4d) throw std::runtime_error("nvdb: not a Vec4d grid");
  for (int i = 0; i < 9; ++i) info->map_mat[i] = get<float>(g, kGdMap + 4 * i);
  for (int i = 0; i < 3; ++i) info->map_vec[i] = get<float>(g, kGdMap + 72 + 4 * i);
  F.leaf_off = get<uint64_t>(g, kGridData + 0), F.lower_off = get<uint64_t>(g, kGridData + 8);
  F.upper_off = get<uint64_t>(g, kGridData + 16), F.root_off = get<uint64_t>(g, kGridData + 24);
  F.n_leaf = get<uint32_t>(g, kGridData + 32), F.n_lower = get<uint32_t>(g, kGridData + 36);
  F.n_upper = get<uint32_t>(g, kGridData + 40);
  F.root = F.read_at(kGridData + F.root_off, kRootData);
  for (int a = 0; a < 3; ++a) F.lo[a] = get<int32_t>(F.root, 4 * a), F.hi[a] = get<int32_t>(F.root, 12 + 4 * a);
  for (int a = 0; a < 3; ++a)
    if (F.hi[a] < F.lo[a] || (int64_t)F.hi[a] - F.lo[a] >= kMaxExtent)
      throw std::runtime_error("nvdb: empty or too large index bounding box");
  for (int a = 0; a < 3; ++a) info->ibb_min[a] = F.lo[a];
  info->nx = F.hi[0] - F.lo[0] + 1, info->ny = F.hi[1] - F.lo[1] + 1, info->nz = F.hi[2] - F.lo[2] + 1;
  info->active_voxels = (int64_t)get<uint64_t>(g, kGridData + 56);
  info->leaf_count = (int32_t)F.n_leaf, info->lower_count = (int32_t)F.n_lower, info->upper_count = (int32_t)F.n_upper;
  const uint32_t table = get<uint32_t>(F.root, 24);
  const std::vector<uint8_t> tiles = F.read_at(kGridData + F.root_off + kRootData, (size_t)table * kRootTile);
  for (uint32_t k = 0; k < table; ++k)
    if (get<int64_t>(tiles, k * kRootTile + 8) == 0 && get<uint32_t>(tiles, k * kRootTile + 16) != 0)
      throw std::runtime_error("nvdb: active root tiles are not supported");
}

// every active value of the tree: fn_tile(origin xyz, child dim, value) for the
// active tiles of the internal nodes (value mask set, no child), which cover their
// child's region, and fn_leaf(origin xyz, leaf node bytes) for the leaves
template <typename Tile, typename Leaf>
void nvdb_walk(NvdbFile& F, Tile fn_tile, Leaf fn_leaf) {
  auto tiles_of = [&](uint64_t first, uint32_t count, size_t size, size_t header, size_t mask_bytes, int log2dim,
                      int child_dim) {
    const uint32_t dim = 1u << log2dim;
    for (uint32_t k = 0; k < count; ++k) {
      const std::vector<uint8_t> nd = F.read_at(kGridData + first + (uint64_t)k * size, size);
      const int32_t m = ~(int32_t)(dim * child_dim - 1);
      const int ox = get<int32_t>(nd, 0) & m, oy = get<int32_t>(nd, 4) & m, oz = get<int32_t>(nd, 8) & m;
      for (uint32_t n = 0; n < dim * dim * dim; ++n) {
        const bool on = (nd[32 + n / 8] >> (n % 8)) & 1u, child = (nd[32 + mask_bytes + n / 8] >> (n % 8)) & 1u;
        if (!on || child) continue;
        fn_tile(ox + (int)(n >> (2 * log2dim)) * child_dim, oy + (int)((n >> log2dim) & (dim - 1)) * child_dim,
                oz + (int)(n & (dim - 1)) * child_dim, child_dim, nd.data() + header + 32 * n);
      }
    }
  };
  tiles_of(F.upper_off, F.n_upper, kUpperSize, kUpperHeader, kUpperMask, 5, 128);
  tiles_of(F.lower_off, F.n_lower, kLowerSize, kLowerHeader, kLowerMask, 4, 8);
  for (uint32_t k = 0; k < F.n_leaf; ++k) {
    const std::vector<uint8_t> nd = F.read_at(kGridData + F.leaf_off + (uint64_t)k * kLeafSize, kLeafSize);
    fn_leaf(get<int32_t>(nd, 0) & ~7, get<int32_t>(nd, 4) & ~7, get<int32_t>(nd, 8) & ~7, nd);
  }
}

void nvdb_read(const std::string& path, r2_nvdb_info* info, uint8_t* records, int64_t records_bytes) {
  NvdbFile F;
  nvdb_open(path, info, F);
  if (!records) return;
  const int32_t* lo = F.lo;
  const int32_t* hi = F.hi;
  const int nx = info->nx, ny = info->ny, nz = info->nz;
  const uint64_t need = (uint64_t)nx * ny * nz * 32;
  if (records_bytes < 0 || (uint64_t)records_bytes < need)
    throw std::invalid_argument("nvdb: record buffer smaller than the grid's index box");
  Rec bg;
  memcpy(bg.data(), F.root.data() + 32, 32);
  fill_records(records, (size_t)nx * ny * nz, bg);
  auto store = [&](int x, int y, int z, const uint8_t* v) {
    if (x < lo[0] || y < lo[1] || z < lo[2] || x > hi[0] || y > hi[1] || z > hi[2]) return;
    memcpy(records + ((((size_t)(z - lo[2]) * ny) + (y - lo[1])) * nx + (x - lo[0])) * 32, v, 32);
  };
  nvdb_walk(
      F,
      [&](int cx, int cy, int cz, int cd, const uint8_t* v) {
        for (int z = std::max(cz, lo[2]); z < std::min(cz + cd, hi[2] + 1); ++z)
          for (int y = std::max(cy, lo[1]); y < std::min(cy + cd, hi[1] + 1); ++y)
            for (int x = std::max(cx, lo[0]); x < std::min(cx + cd, hi[0] + 1); ++x) store(x, y, z, v);
      },
      [&](int ox, int oy, int oz, const std::vector<uint8_t>& nd) {
        for (uint32_t n = 0; n < 512; ++n)
          if ((nd[16 + n / 8] >> (n % 8)) & 1u)
            store(ox + (int)(n >> 6), oy + (int)((n >> 3) & 7), oz + (int)(n & 7), nd.data() + kLeafHeader + 32 * n);
      });
}

// The active values as 8^3 leaves in an index space shifted by `shift` (the
// index box minimum rounded down to a multiple of 8), so that leaf origins are
// non-negative multiples of 8: the file's leaves with their value masks, and
// the active tiles of the internal nodes as full leaves.  Voxels of a leaf are
// stored x fastest (NanoVDB's leaves store z fastest).  With records == NULL only
// the leaf count is returned.
int64_t nvdb_read_leaves(const std::string& path, r2_nvdb_info* info, int32_t shift[3], int64_t max_leaves,
                         int32_t* origins, uint8_t* records, uint8_t* masks) {
  NvdbFile F;
  nvdb_open(path, info, F);
  for (int a = 0; a < 3; ++a) shift[a] = (int32_t)(((int64_t)F.lo[a] >> 3) << 3);
  int64_t n = 0;
  Rec bg;
  memcpy(bg.data(), F.root.data() + 32, 32);
  auto leaf_at = [&](int ox, int oy, int oz) -> int64_t {
    if (!records) return n++;
    if (n >= max_leaves) throw std::invalid_argument("nvdb: leaf buffers smaller than the grid's leaves");
    origins[3 * n] = ox - shift[0], origins[3 * n + 1] = oy - shift[1], origins[3 * n + 2] = oz - shift[2];
    fill_records(records + (size_t)n * 512 * 32, 512, bg);
    memset(masks + (size_t)n * 64, 0, 64);
    return n++;
  };
  nvdb_walk(
      F,
      [&](int cx, int cy, int cz, int cd, const uint8_t* v) {
        for (int z = cz; z < cz + cd; z += 8)
          for (int y = cy; y < cy + cd; y += 8)
            for (int x = cx; x < cx + cd; x += 8) {
              const int64_t k = leaf_at(x, y, z);
              if (!records) continue;
              for (int i = 0; i < 512; ++i) memcpy(records + ((size_t)k * 512 + i) * 32, v, 32);
              memset(masks + (size_t)k * 64, 0xff, 64);
            }
      },
      [&](int ox, int oy, int oz, const std::vector<uint8_t>& nd) {
        const int64_t k = leaf_at(ox, oy, oz);
        if (!records) return;
        for (uint32_t m = 0; m < 512; ++m) {
          if (!((nd[16 + m / 8] >> (m % 8)) & 1u)) continue;
          const uint32_t v = (m >> 6) | (((m >> 3) & 7) << 3) | ((m & 7) << 6);
          memcpy(records + ((size_t)k * 512 + v) * 32, nd.data() + kLeafHeader + 32 * m, 32);
          masks[(size_t)k * 64 + v / 8] |= (uint8_t)(1u << (v % 8));
        }
      });
  return n;
}

}  // namespace r2host

extern "C" int r2_nvdb_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const void* records,
                             const float* map_mat, const float* map_vec, char* err, int32_t err_len) {
  if (!path || !records || nx <= 0 || ny <= 0 || nz <= 0) return R2_ERR_INVALID;
  try {
    r2host::nvdb_write(path, nx, ny, nz, (const uint8_t*)records, map_mat, map_vec);
  } catch (const std::invalid_argument& e) {
    copy_err(err, err_len, e.what());
    return R2_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    copy_err(err, err_len, "nvdb: out of host memory");
    return R2_ERR_OOM;
  } catch (const std::exception& e) {
    copy_err(err, err_len, e.what());
    return R2_ERR_IO;
  }
  return R2_OK;
}

extern "C" int r2_nvdb_read_leaves(const char* path, r2_nvdb_info* info, int32_t* origin_shift, int64_t* n_leaves,
                                   int32_t* leaf_origins, void* records, uint8_t* leaf_masks, char* err,
                                   int32_t err_len) {
  if (!path || !info || !origin_shift || !n_leaves) return R2_ERR_INVALID;
  if (records && (!leaf_origins || !leaf_masks)) return R2_ERR_INVALID;
  try {
    memset(info, 0, sizeof(*info));
    const int64_t cap = *n_leaves;
    *n_leaves = r2host::nvdb_read_leaves(path, info, origin_shift, cap, leaf_origins, (uint8_t*)records, leaf_masks);
  } catch (const std::invalid_argument& e) {
    copy_err(err, err_len, e.what());
    return R2_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    copy_err(err, err_len, "nvdb: out of host memory");
    return R2_ERR_OOM;
  } catch (const std::exception& e) {
    copy_err(err, err_len, e.what());
    return R2_ERR_IO;
  }
  return R2_OK;
}

extern "C" int r2_nvdb_read(const char* path, r2_nvdb_info* info, void* records, int64_t records_bytes, char* err,
                            int32_t err_len) {
  if (!path || !info) return R2_ERR_INVALID;
  try {
    memset(info, 0, sizeof(*info));
    r2host::nvdb_read(path, info, (uint8_t*)records, records_bytes);
  } catch (const std::invalid_argument& e) {
    copy_err(err, err_len, e.what());
    return R2_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    copy_err(err, err_len, "nvdb: out of host memory");
    return R2_ERR_OOM;
  } catch (const std::exception& e) {
    copy_err(err, err_len, e.what());
    return R2_ERR_IO;
  }
  return R2_OK;
}
