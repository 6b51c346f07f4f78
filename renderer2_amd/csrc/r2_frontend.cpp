// Front end over the C ABI: the reference's config.json / scenedescription.json
// driven pipeline (CPipeline::run, raytracer/pipeline.cu:19-93) for the
// volumetric path.
//   * CConfigLoader::loadConfig and its fill* / add* helpers
//     (raytracer/src/backend/config_loader.cu:12-368): same keys, same required
//     fields, comments allowed, "Mask" defaults to RENDER.
//   * CFilter::runFiltering (raytracer/filter.cu:74-123): LOD loop from
//     MinVoxelSize doubling per LOD over the FILTER-masked mesh, one grid file per
//     LOD in ./filtering/<mesh name>/filtered_mesh_<voxel size> and the volume
//     description JSON (raytracer/src/filtering/volume_description_manager.cpp:50-144).
//   * Raytracer::renderFrame + post-process (raytracer/raytracer.cu:161-206) and
//     CImageWriter's vertically flipped quality-100 JPEG and uncompressed PNG
//     (visualisation/src/image_writer.cpp:8-21, r2_image.cpp), plus a linear PFM.
// Assets: Wavefront OBJ + MTL for meshes (Kd, Ks, Ns -> roughness
// (2/(Ns+2))^0.5 and Ni -> ior with 1 mapped to 1.5, raytracer/src/material/material.cu:90-148;
// missing normals become flat face normals as with assimp's GenNormals),
// Radiance .hdr environment maps (RGBE, flat or run-length scanlines), and this
// build's grid files "<Path>.r2grid" (dense SFilteredDataCompact records with the
// index -> model transform) or, where none exists, the reference's NanoVDB file
// "<Path>.nvdb"; filtering writes both, and the .brickgrid beside them.  Circle and Sphere objects render as analytic
// surfaces with their material (r2_surface); meshes with the RENDER mask are parsed but not rendered (no triangle
// intersection in the render path).
#include "r2_internal.hpp"
#include "r2_json.hpp"
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace fs = std::filesystem;

namespace {

using r2json::Value;

struct SceneObject {
  std::string type;
  std::string path, directory, filename;
  float sigma_a[3] = {0, 0, 0}, sigma_s[3] = {0, 0, 0};
  bool has_g = false;
  float g = 0.f;
  float pos[3] = {0, 0, 0}, orientation[3] = {0, 1, 0}, scaling[3] = {1, 1, 1};
  unsigned mask = 1;  // RENDER
  r2_surface surf{};  // Circle / Sphere: shape and material
};

enum : unsigned { kMaskRender = 1u, kMaskFilter = 2u };

struct Config {
  int frame_width = 0, frame_height = 0, channels = 3, samples = 1;
  float gamma = 2.f;
  bool use_brick_grid = true;
  std::string debug_mode = "NONE";
  int max_depth = 50;  // the reference's bounce limit (path_integrator_impl.hpp:127)
  uint32_t seed = 0;
  bool filter = false;
  float min_voxel_size = 1.f;
  int lods = 1, samples_per_voxel = 1, debug_samples = 0, estimation_iterations = 0;
  bool filter_debug = false, clip_rays = true;
  float sigma_t = 1.f, alpha = 0.f;
  float fov = 40.f, cam_pos[3] = {0, 0, 0}, look_at[3] = {0, 0, -1}, up[3] = {0, 1, 0};
  std::string envmap, scenedescription;
  std::vector<SceneObject> objects;
  fs::path base;  // directory of the config file (fallback for relative paths)
};

std::string read_text(const fs::path& p) {
  std::ifstream f(p, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + p.string());
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// a relative path is taken from the working directory as in the reference,
// then from the config file's directory
fs::path resolve(const Config& c, const std::string& p) {
  fs::path q(p);
  if (q.is_absolute() || fs::exists(q)) return q;
  fs::path r = c.base / q;
  return fs::exists(r) ? r : q;
}

bool vec3(const Value* a, float* out) {
  if (!a || !a->is_array() || a->arr.size() != 3) return false;
  for (int i = 0; i < 3; ++i) {
    if (!a->arr[i].is_number()) return false;
    out[i] = (float)a->arr[i].num;
  }
  return true;
}
bool num(const Value* v, double* out) {
  if (!v || !v->is_number()) return false;
  *out = v->num;
  return true;
}
bool boolean(const Value* v, bool* out) {
  if (!v || !v->is_bool()) return false;
  *out = v->b;
  return true;
}
bool str(const Value* v, std::string* out) {
  if (!v || !v->is_string()) return false;
  *out = v->str;
  return true;
}

// getMask (raytracer/src/scene/sceneobject_mask.cpp:4-17), ORed over the array as
// config_loader.cu:34-41 does: "NONE" adds nothing, "FILTER" and "RENDER" their
// bit, and any other string (e.g. "ALL") both bits.  No "Mask" key: `dflt`.
unsigned mask_of(const Value* m, unsigned dflt) {
  if (!m || !m->is_array()) return dflt;
  unsigned r = 0;
  for (const Value& v : m->arr) {
    if (!v.is_string()) continue;
    if (v.str == "NONE") continue;
    if (v.str == "RENDER") r |= kMaskRender;
    else if (v.str == "FILTER") r |= kMaskFilter;
    else r |= kMaskRender | kMaskFilter;
  }
  return r;
}

bool fill_general(const Value* g, Config* c) {
  double w, h, ch, s, gamma;
  bool brick;
  std::string dbg;
  if (!g || !g->is_object()) return false;
  if (!num(g->find("FrameWidth"), &w) || !num(g->find("FrameHeight"), &h) || !num(g->find("ChannelsPerPixel"), &ch) ||
      !num(g->find("Samples"), &s) || !num(g->find("Gamma"), &gamma) || !boolean(g->find("UseBrickGrid"), &brick) ||
      !str(g->find("DebugMode"), &dbg))
    return false;
  c->frame_width = (int)w;
  c->frame_height = (int)h;
  c->channels = (int)ch;
  c->samples = (int)s;
  c->gamma = (float)gamma;
  c->use_brick_grid = brick;
  c->debug_mode = dbg;
  double v;
  if (num(g->find("MaxDepth"), &v)) c->max_depth = (int)v;  // extension of this build
  if (num(g->find("Seed"), &v)) c->seed = (uint32_t)v;       // extension of this build
  return true;
}

bool fill_filtering(const Value* f, Config* c) {
  if (!f || !f->is_object()) return false;
  double vs, lods, spv, ds, st, it, al;
  bool act, dbg, clip;
  if (!boolean(f->find("Active"), &act) || !num(f->find("MinVoxelSize"), &vs) || !num(f->find("LODs"), &lods) ||
      !num(f->find("SamplesPerVoxel"), &spv) || !boolean(f->find("Debug"), &dbg) ||
      !num(f->find("DebugSamples"), &ds) || !num(f->find("SigmaT"), &st) ||
      !num(f->find("EstimationIterations"), &it) || !num(f->find("Alpha"), &al) ||
      !boolean(f->find("ClipRays"), &clip))
    return false;
  c->filter = act;
  c->min_voxel_size = (float)vs;
  c->lods = (int)lods;
  c->samples_per_voxel = (int)spv;
  c->filter_debug = dbg;
  c->debug_samples = (int)ds;
  c->sigma_t = (float)st;
  c->estimation_iterations = (int)it;
  c->alpha = (float)al;
  c->clip_rays = clip;
  return true;
}

bool fill_camera(const Value* cam, Config* c) {
  double fov;
  if (!cam || !cam->is_object() || !num(cam->find("Fov"), &fov)) return false;
  if (!vec3(cam->find("Pos"), c->cam_pos) || !vec3(cam->find("LookAt"), c->look_at) || !vec3(cam->find("Up"), c->up))
    return false;
  c->fov = (float)fov;
  return true;
}

bool add_object(const Value& o, Config* c) {
  std::string type;
  if (!str(o.find("Type"), &type)) return false;
  SceneObject so;
  so.type = type;
  if (type == "Circle" || type == "Sphere") {
    double r, dr, ax, ay, ei, et;
    float n[3], d[3], sp[3];
    if (!vec3(o.find("Pos"), so.pos) || !num(o.find("Radius"), &r) || !vec3(o.find("Normal"), n) ||
        !vec3(o.find("DiffuseReflection"), d) || !num(o.find("DiffuseRoughness"), &dr) ||
        !vec3(o.find("SpecularReflection"), sp) || !num(o.find("AlphaX"), &ax) || !num(o.find("AlphaY"), &ay) ||
        !num(o.find("EtaI"), &ei) || !num(o.find("EtaT"), &et))
      return false;
    so.mask = mask_of(o.find("Mask"), kMaskRender);
    r2_surface& sf = so.surf;
    sf.shape = type == "Circle" ? R2_SHAPE_CIRCLE : R2_SHAPE_SPHERE;
    for (int a = 0; a < 3; ++a) {
      sf.pos[a] = so.pos[a];
      sf.normal[a] = n[a];
      sf.diffuse[a] = d[a];
      sf.specular[a] = sp[a];
    }
    sf.radius = (float)r;  // radius.get<float>(): 3.402823466e+38 (the forest ground) stays finite
    sf.diffuse_roughness = (float)dr;
    sf.alpha_x = (float)ax;
    sf.alpha_y = (float)ay;
    sf.eta_i = (float)ei;
    sf.eta_t = (float)et;
  } else if (type == "Medium") {
    if (!str(o.find("Path"), &so.path) || !vec3(o.find("SigmaA"), so.sigma_a) || !vec3(o.find("SigmaS"), so.sigma_s) ||
        !vec3(o.find("Pos"), so.pos) || !vec3(o.find("Orientation"), so.orientation) ||
        !vec3(o.find("Scaling"), so.scaling))
      return false;
    double g;
    if (num(o.find("G"), &g)) {
      so.has_g = true;
      so.g = (float)g;
    }
    so.mask = mask_of(o.find("Mask"), kMaskRender);
  } else if (type == "Mesh") {
    if (!str(o.find("Directory"), &so.directory) || !str(o.find("Filename"), &so.filename) ||
        !vec3(o.find("Pos"), so.pos) || !vec3(o.find("Orientation"), so.orientation) ||
        !vec3(o.find("Scaling"), so.scaling))
      return false;
    so.mask = mask_of(o.find("Mask"), kMaskRender);
  } else {
    return true;  // unknown types are ignored, as in the reference
  }
  c->objects.push_back(so);
  return true;
}

bool fill_scene(const Value* s, Config* c) {
  if (!s || !s->is_object() || !str(s->find("Envmap"), &c->envmap)) return false;
  if (!str(s->find("Scenedescription"), &c->scenedescription)) return false;
  const Value objs = r2json::parse(read_text(resolve(*c, c->scenedescription)));
  if (!objs.is_array()) return false;
  for (const Value& o : objs.arr)
    if (!add_object(o, c)) return false;
  return true;
}

Config load_config(const std::string& path) {
  Config c;
  c.base = fs::path(path).parent_path();
  const Value j = r2json::parse(read_text(path));
  bool ok = j.is_object();
  ok = ok && fill_general(j.find("General"), &c);
  ok = ok && fill_filtering(j.find("Filtering"), &c);
  ok = ok && fill_camera(j.find("Camera"), &c);
  ok = ok && fill_scene(j.find("Scene"), &c);
  if (!ok) throw std::runtime_error("Could not parse config");
  return c;
}

// ---------------------------------------------------------------- assets ---
struct HostMesh {
  std::vector<float> pos, nrm;
  std::vector<int32_t> idx, mat;
  std::vector<float> diffuse, specular, roughness, ior;  // per material
  int n_materials = 0;
  std::vector<std::string> warnings;
};

struct MtlDef {
  float kd[3] = {0.6f, 0.6f, 0.6f}, ks[3] = {0, 0, 0};
  float ns = 0.f, ni = 1.f;
};

void load_mtl(const fs::path& p, std::map<std::string, MtlDef>& out, std::vector<std::string>& warn) {
  std::ifstream f(p);
  if (!f) {
    warn.push_back("material library not found: " + p.string());
    return;
  }
  std::string line, cur;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string k;
    ls >> k;
    if (k == "newmtl") {
      ls >> cur;
      out[cur] = MtlDef();
    } else if (cur.empty()) {
      continue;
    } else if (k == "Kd") {
      ls >> out[cur].kd[0] >> out[cur].kd[1] >> out[cur].kd[2];
    } else if (k == "Ks") {
      ls >> out[cur].ks[0] >> out[cur].ks[1] >> out[cur].ks[2];
    } else if (k == "Ns") {
      ls >> out[cur].ns;
    } else if (k == "Ni") {
      ls >> out[cur].ni;
    } else if (k == "map_Kd" || k == "map_d" || k == "map_Bump" || k == "bump" || k == "norm") {
      warn.push_back("texture maps are not supported; using constant colours (" + k + ")");
    }
  }
}

int obj_index(const std::string& tok, int n) {
  const int v = std::atoi(tok.c_str());
  return v > 0 ? v - 1 : n + v;  // negative indices count from the end
}

HostMesh load_obj(const fs::path& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open mesh " + path.string());
  std::vector<float> vp, vn;
  std::map<std::string, MtlDef> mtl;
  std::map<std::string, int> mat_id;
  std::vector<MtlDef> mats;
  HostMesh m;
  int cur_mat = -1;
  std::map<std::pair<int, int>, int> vmap;  // (position, normal) -> output vertex
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string k;
    ls >> k;
    if (k == "v") {
      float x, y, z;
      ls >> x >> y >> z;
      vp.insert(vp.end(), {x, y, z});
    } else if (k == "vn") {
      float x, y, z;
      ls >> x >> y >> z;
      vn.insert(vn.end(), {x, y, z});
    } else if (k == "mtllib") {
      std::string name;
      ls >> name;
      load_mtl(path.parent_path() / name, mtl, m.warnings);
    } else if (k == "usemtl") {
      std::string name;
      ls >> name;
      auto it = mat_id.find(name);
      if (it == mat_id.end()) {
        const int id = (int)mats.size();
        mats.push_back(mtl.count(name) ? mtl[name] : MtlDef());
        mat_id[name] = id;
        cur_mat = id;
      } else {
        cur_mat = it->second;
      }
    } else if (k == "f") {
      if (cur_mat < 0) {
        cur_mat = (int)mats.size();
        mats.push_back(MtlDef());
        mat_id[""] = cur_mat;
      }
      std::vector<std::pair<int, int>> poly;
      std::string tok;
      while (ls >> tok) {
        const size_t s1 = tok.find('/');
        const int pi = obj_index(tok.substr(0, s1), (int)vp.size() / 3);
        int ni = -1;
        if (s1 != std::string::npos) {
          const size_t s2 = tok.find('/', s1 + 1);
          if (s2 != std::string::npos && s2 + 1 < tok.size()) ni = obj_index(tok.substr(s2 + 1), (int)vn.size() / 3);
        }
        if (pi < 0 || 3 * pi + 2 >= (int)vp.size() || (ni >= 0 && 3 * ni + 2 >= (int)vn.size()))
          throw std::runtime_error("mesh: face index out of range in " + path.string());
        poly.push_back({pi, ni});
      }
      if (poly.size() < 3) continue;
      const bool flat = std::any_of(poly.begin(), poly.end(), [](const std::pair<int, int>& q) { return q.second < 0; });
      float fnrm[3] = {0, 0, 0};
      if (flat) {  // face normal (Newell) for faces without vertex normals
        for (size_t a = 0; a < poly.size(); ++a) {
          const float* p0 = &vp[3 * poly[a].first];
          const float* p1 = &vp[3 * poly[(a + 1) % poly.size()].first];
          fnrm[0] += (p0[1] - p1[1]) * (p0[2] + p1[2]);
          fnrm[1] += (p0[2] - p1[2]) * (p0[0] + p1[0]);
          fnrm[2] += (p0[0] - p1[0]) * (p0[1] + p1[1]);
        }
        const float l = std::sqrt(fnrm[0] * fnrm[0] + fnrm[1] * fnrm[1] + fnrm[2] * fnrm[2]);
        for (float& q : fnrm) q = l > 0 ? q / l : 0.f;
      }
      std::vector<int> out;
      for (auto& q : poly) {
        if (flat) {  // flat-shaded faces get their own vertices
          out.push_back((int)m.pos.size() / 3);
          m.pos.insert(m.pos.end(), {vp[3 * q.first], vp[3 * q.first + 1], vp[3 * q.first + 2]});
          m.nrm.insert(m.nrm.end(), {fnrm[0], fnrm[1], fnrm[2]});
          continue;
        }
        auto it = vmap.find(q);
        if (it == vmap.end()) {
          const int id = (int)m.pos.size() / 3;
          m.pos.insert(m.pos.end(), {vp[3 * q.first], vp[3 * q.first + 1], vp[3 * q.first + 2]});
          m.nrm.insert(m.nrm.end(), {vn[3 * q.second], vn[3 * q.second + 1], vn[3 * q.second + 2]});
          it = vmap.emplace(q, id).first;
        }
        out.push_back(it->second);
      }
      for (size_t a = 1; a + 1 < out.size(); ++a) {  // fan triangulation
        m.idx.insert(m.idx.end(), {out[0], out[a], out[a + 1]});
        m.mat.push_back(cur_mat);
      }
    }
  }
  if (m.idx.empty()) throw std::runtime_error("mesh has no faces: " + path.string());
  m.n_materials = (int)mats.size();
  for (const MtlDef& d : mats) {
    m.diffuse.insert(m.diffuse.end(), d.kd, d.kd + 3);
    m.specular.insert(m.specular.end(), d.ks, d.ks + 3);
    m.roughness.push_back(std::pow(2.f / (d.ns + 2.f), 0.5f));
    m.ior.push_back(d.ni == 1.f ? 1.5f : d.ni);
  }
  return m;
}

// Radiance RGBE (.hdr), rows top to bottom, as stbi_loadf returns them
std::vector<float> load_hdr(const fs::path& p, int* w, int* h) {
  std::ifstream f(p, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open env map " + p.string());
  std::string line;
  std::getline(f, line);
  if (line.rfind("#?", 0) != 0) throw std::runtime_error("not a Radiance file: " + p.string());
  while (std::getline(f, line) && !line.empty()) {
    if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe")
      throw std::runtime_error("unsupported HDR format: " + line);
  }
  std::getline(f, line);
  char ys[3] = {0}, xs[3] = {0};
  if (sscanf(line.c_str(), "%2s %d %2s %d", ys, h, xs, w) != 4 || std::string(ys) != "-Y" || std::string(xs) != "+X")
    throw std::runtime_error("unsupported HDR orientation: " + line);
  const int W = *w, H = *h;
  if (W < 2 || H < 2 || (int64_t)W * H > (1ll << 28)) throw std::runtime_error("bad HDR size: " + line);
  std::vector<float> out((size_t)W * H * 3);
  std::vector<uint8_t> scan((size_t)W * 4);
  for (int y = 0; y < H; ++y) {
    uint8_t hdr4[4];
    f.read((char*)hdr4, 4);
    if (!f) throw std::runtime_error("truncated HDR file");
    if (W >= 8 && W < 32768 && hdr4[0] == 2 && hdr4[1] == 2 && ((hdr4[2] << 8) | hdr4[3]) == W) {
      for (int c = 0; c < 4; ++c) {  // run-length encoded channels
        int x = 0;
        while (x < W) {
          int n = f.get();
          if (n == EOF) throw std::runtime_error("truncated HDR file");
          if (n > 128) {
            n -= 128;
            const int v = f.get();
            if (x + n > W) throw std::runtime_error("bad HDR run");
            for (int k = 0; k < n; ++k) scan[4 * (x++) + c] = (uint8_t)v;
          } else {
            if (n == 0 || x + n > W) throw std::runtime_error("bad HDR run");
            for (int k = 0; k < n; ++k) scan[4 * (x++) + c] = (uint8_t)f.get();
          }
        }
      }
    } else {  // flat scanline
      memcpy(scan.data(), hdr4, 4);
      f.read((char*)scan.data() + 4, (std::streamsize)W * 4 - 4);
    }
    if (!f) throw std::runtime_error("truncated HDR file");
    for (int x = 0; x < W; ++x) {
      const uint8_t* q = &scan[4 * x];
      float* o = &out[3 * ((size_t)y * W + x)];
      if (q[3] == 0) {
        o[0] = o[1] = o[2] = 0.f;
      } else {
        const float s = std::ldexp(1.f, (int)q[3] - (128 + 8));
        o[0] = q[0] * s;
        o[1] = q[1] * s;
        o[2] = q[2] * s;
      }
    }
  }
  return out;
}

// grid file: "R2GRID01", int32 nx ny nz, float index_to_model[12] (glm mat4x3),
// float voxel size, then nx*ny*nz 32-byte records (x fastest)
void write_grid(const fs::path& p, const r2_filter_grid& g, float vs, const std::vector<uint8_t>& rec) {
  std::ofstream f(p, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + p.string());
  f.write("R2GRID01", 8);
  f.write((const char*)g.num_voxels, 12);
  f.write((const char*)g.index_to_model, 48);
  f.write((const char*)&vs, 4);
  f.write((const char*)rec.data(), (std::streamsize)rec.size());
}

struct GridFile {
  int32_t n[3];
  float i2m[12];
  float vs;
  std::vector<uint8_t> rec;
};

GridFile read_grid(const fs::path& p) {
  std::ifstream f(p, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open grid " + p.string());
  char magic[8];
  GridFile g;
  f.read(magic, 8);
  f.read((char*)g.n, 12);
  f.read((char*)g.i2m, 48);
  f.read((char*)&g.vs, 4);
  if (!f || memcmp(magic, "R2GRID01", 8) != 0 || g.n[0] <= 0 || g.n[1] <= 0 || g.n[2] <= 0)
    throw std::runtime_error("not a grid file: " + p.string());
  if (g.n[0] > 8192 || g.n[1] > 8192 || g.n[2] > 8192)
    throw std::runtime_error("grid file larger than 8192 voxels per axis: " + p.string());
  const size_t bytes = (size_t)g.n[0] * g.n[1] * g.n[2] * 32;
  g.rec.resize(bytes);
  f.read((char*)g.rec.data(), (std::streamsize)bytes);
  if (!f) throw std::runtime_error("truncated grid file: " + p.string());
  return g;
}

// ------------------------------------------------------------ brick grid ---
// CHostBrickGrid in cereal 1.3.2's portable binary archive (serialization.cpp:37-47,
// host_grid_brick.hpp:47-72): one byte 1 (little-endian data), then little-endian
// fields in declaration order; a std::vector is a uint64 length and its elements;
// a glm vector its components; std::pair first, second; std::atomic its value.
//   transform mat4 (identity, host_grid_brick.cpp:63), n_bricks uvec3, min_maj pair,
//   brick_counter size_t, indirection / range / atlas Buf3D (stride uvec3, data
//   vector), range_mipmaps vector<Buf3D<uint32>>.
// Atlas slots are handed out in the order of the reference's brick loop (z, y, x:
// one of the orders its parallel atomic allocation can produce); slot id -> atlas
// brick coordinates as indirection.to_coord (buf3d.hpp:34-36), pointers encoded as
// encode_ptr (common.hpp:31-38) and the atlas pruned in z (host_grid_brick.cpp:117).
struct LeWriter {
  std::vector<uint8_t> b;
  template <typename T>
  void put(const T& v) {
    const uint8_t* p = (const uint8_t*)&v;
    b.insert(b.end(), p, p + sizeof(T));
  }
  template <typename T>
  void vec(const std::vector<T>& v) {
    put<uint64_t>((uint64_t)v.size());
    const uint8_t* p = (const uint8_t*)v.data();
    b.insert(b.end(), p, p + v.size() * sizeof(T));
  }
  void uvec3(uint32_t x, uint32_t y, uint32_t z) { put(x), put(y), put(z); }
};

// reader of the same archive: little-endian fields in declaration order, a
// std::vector as a uint64 length and its elements
struct LeReader {
  std::vector<uint8_t> b;
  size_t pos = 0;
  template <typename T>
  T get() {
    if (b.size() - pos < sizeof(T)) throw std::runtime_error("brickgrid: truncated file");
    T v;
    memcpy(&v, b.data() + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  template <typename T>
  void vec(std::vector<T>& v, uint64_t want) {
    const uint64_t n = get<uint64_t>();
    if (n != want) throw std::runtime_error("brickgrid: a buffer's length does not match its dimensions");
    if ((b.size() - pos) / sizeof(T) < n) throw std::runtime_error("brickgrid: truncated file");
    v.resize(n);
    memcpy(v.data(), b.data() + pos, n * sizeof(T));
    pos += n * sizeof(T);
  }
  void uvec3(uint32_t* d) { d[0] = get<uint32_t>(), d[1] = get<uint32_t>(), d[2] = get<uint32_t>(); }
};

// rt::read_grid (serialization.cpp:63-70) of a CHostBrickGrid archive: brick
// counts, extrema, non-empty count; level-0 ranges, the atlas densities put
// back in brick layout through the indirection, and the mips (each optional)
void read_brickgrid(const std::string& path, r2_brickgrid_info* info, uint32_t* ranges, uint8_t* u8, uint32_t* mips) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + path);
  LeReader r;
  r.b.resize((size_t)f.tellg());
  f.seekg(0);
  f.read((char*)r.b.data(), (std::streamsize)r.b.size());
  if (!f) throw std::runtime_error("brickgrid: read failed: " + path);
  if (r.get<uint8_t>() != 1) throw std::runtime_error("brickgrid: not a little-endian portable archive");
  for (int k = 0; k < 16; ++k) (void)r.get<float>();  // the grid transform (identity when written here)
  uint32_t nb[3], d[3];
  r.uvec3(nb);
  for (int a = 0; a < 3; ++a)
    if (nb[a] == 0 || nb[a] > 1024 || nb[a] % 8) throw std::runtime_error("brickgrid: bad brick counts");
  info->n_bricks[0] = (int32_t)nb[0], info->n_bricks[1] = (int32_t)nb[1], info->n_bricks[2] = (int32_t)nb[2];
  info->min_density = r.get<float>();
  info->max_density = r.get<float>();
  info->n_nonempty_bricks = (int64_t)r.get<uint64_t>();
  const uint64_t n = (uint64_t)nb[0] * nb[1] * nb[2];
  std::vector<uint32_t> ind, rng;
  std::vector<uint8_t> atlas;
  r.uvec3(d);
  if (d[0] != nb[0] || d[1] != nb[1] || d[2] != nb[2]) throw std::runtime_error("brickgrid: bad indirection size");
  r.vec(ind, n);
  r.uvec3(d);
  if (d[0] != nb[0] || d[1] != nb[1] || d[2] != nb[2]) throw std::runtime_error("brickgrid: bad range size");
  r.vec(rng, n);
  r.uvec3(d);
  if (d[0] != 8 * nb[0] || d[1] != 8 * nb[1] || d[2] % 8) throw std::runtime_error("brickgrid: bad atlas size");
  r.vec(atlas, (uint64_t)d[0] * d[1] * d[2]);
  const uint32_t az = d[2];
  if (r.get<uint64_t>() != 3) throw std::runtime_error("brickgrid: expected 3 mip levels");
  std::vector<uint32_t> mip;
  size_t off = 0;
  for (int l = 0; l < 3; ++l) {
    r.uvec3(d);
    if (d[0] != nb[0] >> (l + 1) || d[1] != nb[1] >> (l + 1) || d[2] != nb[2] >> (l + 1))
      throw std::runtime_error("brickgrid: bad mip size");
    const uint64_t m = (uint64_t)d[0] * d[1] * d[2];
    r.vec(mip, m);
    if (mips) memcpy(mips + off, mip.data(), m * 4);
    off += m;
  }
  if (r.pos != r.b.size()) throw std::runtime_error("brickgrid: trailing bytes");
  if (ranges) memcpy(ranges, rng.data(), n * 4);
  if (u8) {
    const size_t ex = (size_t)nb[0] * 8, ey = (size_t)nb[1] * 8, ez = (size_t)nb[2] * 8;
    memset(u8, 0, ex * ey * ez);
    for (uint64_t bi = 0; bi < n; ++bi) {
      if ((rng[bi] & 0xFFFFu) == (rng[bi] >> 16)) continue;  // empty brick: no atlas slot
      const uint32_t p = ind[bi], px = (p >> 22) & 1023u, py = (p >> 12) & 1023u, pz = (p >> 2) & 1023u;
      if (px >= nb[0] || py >= nb[1] || 8 * pz + 8 > az) throw std::runtime_error("brickgrid: slot outside the atlas");
      const size_t bx = bi % nb[0], by = (bi / nb[0]) % nb[1], bz = bi / ((size_t)nb[0] * nb[1]);
      for (int v = 0; v < 512; ++v) {
        const size_t x = v & 7, y = (v >> 3) & 7, z = v >> 6;
        u8[((bz * 8 + z) * ey + by * 8 + y) * ex + bx * 8 + x] =
            atlas[(((size_t)pz * 8 + z) * ey + py * 8 + y) * ex + px * 8 + x];
      }
    }
  }
}

// a .brickgrid file against the brick grid the device built for g
bool brickgrid_matches(r2_ctx* ctx, const r2_grid* g, const fs::path& path) {
  r2_grid_info inf{};
  if (r2_grid_get_info(g, &inf) != R2_OK) return false;
  r2_brickgrid_info bi{};
  try {
    read_brickgrid(path.string(), &bi, nullptr, nullptr, nullptr);
  } catch (const std::exception&) {
    return false;
  }
  for (int a = 0; a < 3; ++a)
    if (bi.n_bricks[a] != inf.n_bricks[a]) return false;
  if (bi.n_nonempty_bricks != inf.n_nonempty_bricks || bi.max_density != inf.max_density ||
      bi.min_density != inf.min_density)
    return false;
  const size_t nb = (size_t)inf.n_bricks[0] * inf.n_bricks[1] * inf.n_bricks[2];
  size_t nmip = 0;
  for (int l = 0; l < 3; ++l)
    nmip += (size_t)(inf.n_bricks[0] >> (l + 1)) * (inf.n_bricks[1] >> (l + 1)) * (inf.n_bricks[2] >> (l + 1));
  std::vector<uint32_t> fr(nb), dr(nb), fm(nmip), dm(nmip);
  std::vector<uint8_t> fu(nb * 512), du(nb * 512);
  read_brickgrid(path.string(), &bi, fr.data(), fu.data(), fm.data());
  if (r2_grid_download_bricks(ctx, g, dr.data(), du.data(), dm.data()) != R2_OK) return false;
  return fr == dr && fm == dm && fu == du;
}

void write_brickgrid(r2_ctx* ctx, const r2_grid* g, const fs::path& path) {
  r2_grid_info inf{};
  if (r2_grid_get_info(g, &inf) != R2_OK) throw std::runtime_error("brickgrid: grid info");
  const uint32_t nbx = (uint32_t)inf.n_bricks[0], nby = (uint32_t)inf.n_bricks[1], nbz = (uint32_t)inf.n_bricks[2];
  const size_t nb = (size_t)nbx * nby * nbz;
  const size_t dx = (size_t)nbx * 8, dy = (size_t)nby * 8, dz = (size_t)nbz * 8;
  size_t nmip = 0;
  for (int l = 0; l < 3; ++l) nmip += (size_t)(nbx >> (l + 1)) * (nby >> (l + 1)) * (nbz >> (l + 1));
  std::vector<uint32_t> range(nb), mips(nmip);
  std::vector<uint8_t> dense(dx * dy * dz);
  if (r2_grid_download_bricks(ctx, g, range.data(), dense.data(), mips.data()) != R2_OK)
    throw std::runtime_error(std::string("brickgrid: ") + r2_last_error(ctx));
  auto empty = [](uint32_t r) { return (r & 0xFFFFu) == (r >> 16); };
  uint64_t count = 0;
  for (uint32_t r : range) count += empty(r) ? 0 : 1;
  const size_t zs = 8 * (size_t)std::lround(std::ceil((double)count / (double)(nbx * nby)));
  std::vector<uint32_t> ind(nb, 0u);
  std::vector<uint8_t> atlas(dx * dy * zs, 0);
  uint64_t id = 0;
  for (uint32_t bz = 0; bz < nbz; ++bz)
    for (uint32_t by = 0; by < nby; ++by)
      for (uint32_t bx = 0; bx < nbx; ++bx) {
        const size_t bi = ((size_t)bz * nby + by) * nbx + bx;
        if (empty(range[bi])) continue;
        const uint32_t px = (uint32_t)(id % nbx), py = (uint32_t)((id / nbx) % nby), pz = (uint32_t)(id / ((uint64_t)nbx * nby));
        ++id;
        ind[bi] = (std::min(px, 1023u) << 22) | (std::min(py, 1023u) << 12) | (std::min(pz, 1023u) << 2);
        for (uint32_t z = 0; z < 8; ++z)
          for (uint32_t y = 0; y < 8; ++y)
            for (uint32_t x = 0; x < 8; ++x)
              atlas[((size_t)(pz * 8 + z) * dy + (py * 8 + y)) * dx + (px * 8 + x)] =
                  dense[((size_t)(bz * 8 + z) * dy + (by * 8 + y)) * dx + (bx * 8 + x)];
      }
  LeWriter w;
  w.put<uint8_t>(1);
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) w.put<float>(c == r ? 1.f : 0.f);
  w.uvec3(nbx, nby, nbz);
  w.put<float>(inf.min_density);
  w.put<float>(inf.max_density);
  w.put<uint64_t>(count);
  w.uvec3(nbx, nby, nbz);
  w.vec(ind);
  w.uvec3(nbx, nby, nbz);
  w.vec(range);
  w.uvec3((uint32_t)dx, (uint32_t)dy, (uint32_t)zs);
  w.vec(atlas);
  w.put<uint64_t>(3);
  size_t off = 0;
  for (int l = 0; l < 3; ++l) {
    const uint32_t mx = nbx >> (l + 1), my = nby >> (l + 1), mz = nbz >> (l + 1);
    const size_t n = (size_t)mx * my * mz;
    w.uvec3(mx, my, mz);
    w.vec(std::vector<uint32_t>(mips.begin() + off, mips.begin() + off + n));
    off += n;
  }
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path.string() + " for writing");
  f.write((const char*)w.b.data(), (std::streamsize)w.b.size());
  if (!f) throw std::runtime_error("brickgrid: write failed: " + path.string());
}

// linear HDR frame as PFM (rows bottom to top, the frame's own order)
void write_pfm(const fs::path& p, int w, int h, const std::vector<float>& rgb) {
  std::ofstream f(p, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + p.string());
  f << "PF\n" << w << " " << h << "\n-1.0\n";
  f.write((const char*)rgb.data(), (std::streamsize)rgb.size() * 4);
}

std::string voxel_size_name(float vs) {
  char buf[64];
  snprintf(buf, sizeof(buf), "%f", vs);  // std::to_string(float)
  return buf;
}

// getVoxelSizeFiltering: the trailing decimal number of the path (nvdb_medium.cu:180-188)
float voxel_size_of_path(const std::string& p) {
  std::smatch m;
  static const std::regex r("(\\d+\\.\\d+)$");
  if (!std::regex_search(p, m, r)) throw std::runtime_error("No valid filepath: " + p);
  return std::stof(m.str(1));
}

struct Report {
  r2_pipeline_report r{};
};

void check(r2_ctx* ctx, int rc, const char* what) {
  if (rc != R2_OK) throw std::runtime_error(std::string(what) + ": " + r2_last_error(ctx));
}

// CFilter::runFiltering over the FILTER-masked mesh
void run_filter(r2_ctx* ctx, const Config& c, const fs::path& out_root, r2_pipeline_report* rep) {
  const SceneObject* mo = nullptr;
  for (const SceneObject& o : c.objects) {
    if (o.type == "Mesh" && (o.mask & kMaskFilter)) {
      if (mo && (mo->directory != o.directory || mo->filename != o.filename))
        throw std::runtime_error("Filtering only available for single mesh");
      mo = &o;
    }
  }
  if (!mo) {
    fprintf(stderr, "[WARNING]: No bounding boxes provided --> proceed without filtering.\n");
    return;
  }
  HostMesh hm = load_obj(resolve(c, (fs::path(mo->directory) / mo->filename).string()));
  for (const std::string& w : hm.warnings) fprintf(stderr, "[WARNING]: %s\n", w.c_str());
  r2_mesh_desc md{};
  md.n_vertices = (int32_t)(hm.pos.size() / 3);
  md.n_triangles = (int32_t)(hm.idx.size() / 3);
  md.positions = hm.pos.data();
  md.normals = hm.nrm.data();
  md.indices = hm.idx.data();
  md.material = hm.mat.data();
  md.n_materials = hm.n_materials;
  md.mat_diffuse = hm.diffuse.data();
  md.mat_specular = hm.specular.data();
  md.mat_roughness = hm.roughness.data();
  md.mat_ior = hm.ior.data();
  memcpy(md.pos, mo->pos, 12);
  memcpy(md.orientation, mo->orientation, 12);
  memcpy(md.scaling, mo->scaling, 12);
  r2_mesh* mesh = nullptr;
  check(ctx, r2_mesh_upload(ctx, &md, &mesh), "mesh upload");
  const std::string name = fs::path(mo->filename).stem().string();
  const fs::path dir = out_root / "filtering" / name;
  fs::create_directories(dir);
  const fs::path desc_path = out_root / "filtering" / "volume_description.json";
  Value desc;
  desc.kind = Value::Object;
  if (fs::exists(desc_path)) desc = r2json::parse(read_text(desc_path));
  const auto t0 = std::chrono::steady_clock::now();
  float vs = c.min_voxel_size;
  for (int lod = 0; lod < c.lods; ++lod, vs *= 2.f) {
    r2_filter_params p{vs, c.samples_per_voxel, c.estimation_iterations, c.sigma_t, c.alpha, c.clip_rays ? 1 : 0,
                       c.seed, lod};
    r2_filter_grid g{};
    check(ctx, r2_filter_setup(mesh, vs, &g), "filter set-up");
    std::vector<uint8_t> rec((size_t)g.num_voxels[0] * g.num_voxels[1] * g.num_voxels[2] * 32);
    check(ctx, r2_filter_lod(ctx, mesh, &p, rec.data(), 0, &g), "filter");
    write_grid(dir / ("filtered_mesh_" + voxel_size_name(vs) + ".r2grid"), g, vs, rec);
    {
      // the LOD as a NanoVDB grid (map = the grid's index -> model transform) and its
      // brick majorant grid beside it, as writeToFile does (openvdb_backend.cpp:134-157)
      r2_grid_desc d{};
      d.nx = g.num_voxels[0];
      d.ny = g.num_voxels[1];
      d.nz = g.num_voxels[2];
      d.records = rec.data();
      d.records_on_device = 0;
      for (int r = 0; r < 3; ++r) {
        for (int col = 0; col < 3; ++col) d.map_mat[r * 3 + col] = g.index_to_model[col * 3 + r];
        d.map_vec[r] = g.index_to_model[9 + r];
        d.sigma_s[r] = 1.f;
      }
      d.voxel_size_filtering = vs;
      d.precision = R2_PREC_FP32;
      r2host::nvdb_write((dir / ("filtered_mesh_" + voxel_size_name(vs) + ".nvdb")).string(), d.nx, d.ny, d.nz,
                         rec.data(), d.map_mat, d.map_vec);
      r2_grid* bg = nullptr;
      check(ctx, r2_grid_upload(ctx, &d, &bg), "brick grid build");
      try {
        write_brickgrid(ctx, bg, dir / ("filtered_mesh_" + voxel_size_name(vs) + ".brickgrid"));
      } catch (...) {
        r2_grid_free(ctx, bg);
        throw;
      }
      r2_grid_free(ctx, bg);
    }
    // model-space box of the LOD: getRotation(orientation) * worldToModel applied to the world box
    const r2::m44 R = r2host::get_rotation(r2::mk(mo->orientation[0], mo->orientation[1], mo->orientation[2]));
    r2::m43 w2m;
    memcpy(w2m.m, g.world_to_model, 48);
    const r2::m43 T = r2::to43(r2::mul44(R, r2::to44(w2m)));
    const r2::v3 a = r2::mul_point(T, r2::mk(g.world_bb_min[0], g.world_bb_min[1], g.world_bb_min[2]));
    const r2::v3 b = r2::mul_point(T, r2::mk(g.world_bb_max[0], g.world_bb_max[1], g.world_bb_max[2]));
    const float mn[3] = {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)};
    const float mx[3] = {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)};
    Value& e = desc.set(name).set(voxel_size_name(vs));
    auto arr3 = [](float x, float y, float z) {
      Value v = Value::array();
      v.arr = {Value::number(x), Value::number(y), Value::number(z)};
      return v;
    };
    e.set("NumVoxels") = arr3((float)g.num_voxels[0], (float)g.num_voxels[1], (float)g.num_voxels[2]);
    e.set("BBMin") = arr3(mn[0], mn[1], mn[2]);
    e.set("BBMax") = arr3(mx[0], mx[1], mx[2]);
    const float P[8][3] = {{mn[0], mn[1], mn[2]}, {mx[0], mn[1], mn[2]}, {mx[0], mn[1], mx[2]}, {mn[0], mn[1], mx[2]},
                           {mn[0], mx[1], mn[2]}, {mx[0], mx[1], mn[2]}, {mx[0], mx[1], mx[2]}, {mn[0], mx[1], mx[2]}};
    for (int k = 0; k < 8; ++k) e.set("P" + std::to_string(k)) = arr3(P[k][0], P[k][1], P[k][2]);
    ++rep->lods_filtered;
  }
  rep->filter_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::string text;
  r2json::write(desc, text);
  std::ofstream(desc_path) << text << "\n";
  r2_mesh_free(ctx, mesh);
}

std::vector<float> load_env(const Config& c, int* w, int* h) {
  if (c.envmap.rfind("synth:", 0) == 0) {  // this build's procedural maps: synth:white, synth:sky
    const bool sky = c.envmap == "synth:sky";
    *w = sky ? 512 : 32;
    *h = sky ? 256 : 16;
    std::vector<float> rgb((size_t)*w * *h * 3);
    if (r2_synth_env(sky ? 1 : 0, *w, *h, rgb.data()) != R2_OK) throw std::runtime_error("synth env failed");
    return rgb;
  }
  return load_hdr(resolve(c, c.envmap), w, h);
}

void run_render(r2_ctx* ctx, const Config& c, const fs::path& out_dir, r2_pipeline_report* rep) {
  std::vector<r2_grid*> grids;
  std::vector<r2_instance> inst;
  std::vector<r2_surface> surfs;
  std::map<std::string, r2_grid*> cache;  // CAssetManager::loadMedium reuses media by path
  for (const SceneObject& o : c.objects) {
    if (!(o.mask & kMaskRender)) continue;
    if (o.type == "Circle" || o.type == "Sphere") {
      if ((int)surfs.size() >= R2_MAX_SURFACES) throw std::runtime_error("too many surfaces in the scene");
      surfs.push_back(o.surf);
      continue;
    }
    if (o.type != "Medium") {
      fprintf(stderr, "[WARNING]: %s objects are not rendered by the volumetric path; skipped\n", o.type.c_str());
      ++rep->objects_skipped;
      continue;
    }
    r2_grid* g = nullptr;
    auto it = cache.find(o.path);
    if (it != cache.end()) {
      g = it->second;
    } else {
      // the library's own grid file, or else the reference's NanoVDB file of the
      // same path (CNVDBMedium loads "<Path>.nvdb", nvdb_medium.cu:130-173)
      const fs::path own = resolve(c, o.path + ".r2grid"), nano = resolve(c, o.path + ".nvdb");
      GridFile gf;
      std::vector<int32_t> leaf_org;
      std::vector<uint8_t> leaf_mask;
      int64_t n_leaves = 0;
      if (fs::exists(own) || !fs::exists(nano)) {
        gf = read_grid(own);
      } else {
        // the tree's leaves as they are (no densified copy): the sparse form of the grid
        r2_nvdb_info ni{};
        int32_t shift[3];
        n_leaves = r2host::nvdb_read_leaves(nano.string(), &ni, shift, 0, nullptr, nullptr, nullptr);
        for (int a = 0; a < 3; ++a)
          gf.n[a] = (a == 0 ? ni.nx : a == 1 ? ni.ny : ni.nz) + ni.ibb_min[a] - shift[a];
        if (gf.n[0] > 8192 || gf.n[1] > 8192 || gf.n[2] > 8192)
          throw std::runtime_error("grid file larger than 8192 voxels per axis: " + nano.string());
        leaf_org.resize((size_t)n_leaves * 3);
        leaf_mask.resize((size_t)n_leaves * 64);
        gf.rec.resize((size_t)n_leaves * 512 * 32);
        r2host::nvdb_read_leaves(nano.string(), &ni, shift, n_leaves, leaf_org.data(), gf.rec.data(), leaf_mask.data());
        // an index space shifted to a multiple of 8 below the index box: fold the shift into the translation
        const bool shifted = shift[0] != 0 || shift[1] != 0 || shift[2] != 0;
        for (int r = 0; r < 3; ++r) {
          for (int col = 0; col < 3; ++col) gf.i2m[col * 3 + r] = ni.map_mat[r * 3 + col];
          gf.i2m[9 + r] = !shifted ? ni.map_vec[r]
                                   : ni.map_vec[r] + ni.map_mat[r * 3] * (float)shift[0] +
                                         ni.map_mat[r * 3 + 1] * (float)shift[1] +
                                         ni.map_mat[r * 3 + 2] * (float)shift[2];
        }
      }
      r2_grid_desc d{};
      d.nx = gf.n[0];
      d.ny = gf.n[1];
      d.nz = gf.n[2];
      d.records = gf.rec.data();
      d.records_on_device = 0;
      d.n_leaves = n_leaves;
      d.leaf_origins = n_leaves ? leaf_org.data() : nullptr;
      d.leaf_masks = n_leaves ? leaf_mask.data() : nullptr;
      for (int r = 0; r < 3; ++r) {
        for (int col = 0; col < 3; ++col) d.map_mat[r * 3 + col] = gf.i2m[col * 3 + r];
        d.map_vec[r] = gf.i2m[9 + r];
      }
      memcpy(d.sigma_a, o.sigma_a, 12);
      memcpy(d.sigma_s, o.sigma_s, 12);
      d.voxel_size_filtering = voxel_size_of_path(o.path);
      d.precision = R2_PREC_FP32;
      check(ctx, r2_grid_upload(ctx, &d, &g), "grid upload");
      grids.push_back(g);
      cache[o.path] = g;
      // the reference loads the medium's .brickgrid instead of building it
      // (rt::read_grid, serialization.cpp:63-70): here the grid is built from the
      // records, and a file beside them is checked against that build
      const fs::path bgf = resolve(c, o.path + ".brickgrid");
      if (fs::exists(bgf)) {
        if (brickgrid_matches(ctx, g, bgf)) {
          ++rep->brickgrids_matched;
        } else {
          ++rep->brickgrids_differing;
          fprintf(stderr, "[WARNING]: %s differs from the brick grid of its medium's records; using the latter\n",
                  bgf.string().c_str());
        }
      }
    }
    if ((int)inst.size() >= r2::kMaxInstances) throw std::runtime_error("too many media in the scene");
    r2_instance in{};
    in.grid = g;
    memcpy(in.pos, o.pos, 12);
    memcpy(in.orientation, o.orientation, 12);
    memcpy(in.scaling, o.scaling, 12);
    inst.push_back(in);
  }
  int ew = 0, eh = 0;
  const std::vector<float> env_rgb = load_env(c, &ew, &eh);
  r2_env* env = nullptr;
  check(ctx, r2_env_upload(ctx, env_rgb.data(), ew, eh, &env), "env upload");
  r2_render_params p{};
  p.width = c.frame_width;
  p.height = c.frame_height;
  p.spp = c.samples;
  p.max_depth = c.max_depth;
  memcpy(p.cam_pos, c.cam_pos, 12);
  memcpy(p.cam_look_at, c.look_at, 12);
  memcpy(p.cam_up, c.up, 12);
  p.fov_deg = c.fov;
  p.seed = c.seed;
  p.n_instances = (int32_t)inst.size();
  p.instances = inst.data();
  p.n_surfaces = (int32_t)surfs.size();
  p.surfaces = surfs.data();
  p.env = env;
  p.flags = c.use_brick_grid ? R2_FLAG_BRICK_GRID : 0u;
  std::vector<float> hdr((size_t)p.width * p.height * 3);
  const auto t0 = std::chrono::steady_clock::now();
  check(ctx, r2_render(ctx, &p, hdr.data(), 0), "render");
  rep->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  rep->width = p.width;
  rep->height = p.height;
  rep->spp = p.spp;
  rep->media_rendered = (int32_t)inst.size();
  rep->surfaces_rendered = (int32_t)surfs.size();
  rep->mpaths_per_s = (double)p.width * p.height * p.spp / (rep->render_ms * 1e-3) / 1e6;
  std::vector<uint8_t> bytes(hdr.size());
  check(ctx, r2_postprocess(ctx, hdr.data(), p.width, p.height, c.gamma, nullptr, bytes.data(), 0), "postprocess");
  fs::create_directories(out_dir);
  r2host::write_jpg((out_dir / "output.jpg").string(), p.width, p.height, bytes.data(), 100);
  r2host::write_png((out_dir / "output.png").string(), p.width, p.height, bytes.data());
  write_pfm(out_dir / "output.pfm", p.width, p.height, hdr);
  r2_env_free(ctx, env);
  for (r2_grid* g : grids) r2_grid_free(ctx, g);
}

}  // namespace

extern "C" int r2_pipeline_run(const char* config_path, const char* out_dir, int32_t device, r2_pipeline_report* report,
                               char* err, int32_t err_len) {
  r2_pipeline_report rep{};
  r2_ctx* ctx = nullptr;
  int rc = R2_OK;
  auto set_err = [&](const std::string& m) {
    if (err && err_len > 0) {
      strncpy(err, m.c_str(), (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
  };
  if (!config_path) return R2_ERR_INVALID;
  try {
    const Config c = load_config(config_path);
    const fs::path out = out_dir ? fs::path(out_dir) : fs::path(".");
    rc = r2_create(&ctx, device);
    if (rc != R2_OK) throw std::runtime_error("no HIP device");
    if (c.filter) run_filter(ctx, c, out, &rep);
    run_render(ctx, c, out, &rep);
  } catch (const std::exception& e) {
    set_err(e.what());
    const std::string m = e.what();
    rc = (m.find("cannot open") != std::string::npos || m.find("not a") != std::string::npos ||
          m.find("truncated") != std::string::npos)
             ? R2_ERR_IO
             : (rc == R2_OK ? R2_ERR_INVALID : rc);
  }
  if (ctx) r2_destroy(ctx);
  if (report) *report = rep;
  return rc;
}

// Config parsing only (no device): fills a summary for front-end tests.
extern "C" int r2_grid_write_brickgrid(r2_ctx* ctx, const r2_grid* grid, const char* path, char* err,
                                       int32_t err_len) {
  if (!ctx || !grid || !path) return R2_ERR_INVALID;
  try {
    write_brickgrid(ctx, grid, fs::path(path));
  } catch (const std::exception& e) {
    if (err && err_len > 0) {
      strncpy(err, e.what(), (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
    return R2_ERR_IO;
  }
  return R2_OK;
}

extern "C" int r2_brickgrid_read(const char* path, r2_brickgrid_info* info, uint32_t* ranges, uint8_t* u8,
                                 uint32_t* mips, char* err, int32_t err_len) {
  if (!path || !info) return R2_ERR_INVALID;
  try {
    memset(info, 0, sizeof(*info));
    read_brickgrid(path, info, ranges, u8, mips);
  } catch (const std::bad_alloc&) {
    if (err && err_len > 0) {
      strncpy(err, "brickgrid: out of host memory", (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
    return R2_ERR_OOM;
  } catch (const std::exception& e) {
    if (err && err_len > 0) {
      strncpy(err, e.what(), (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
    return R2_ERR_IO;
  }
  return R2_OK;
}

extern "C" int r2_config_summary(const char* config_path, r2_config_summary_t* out, char* err, int32_t err_len) {
  if (!config_path || !out) return R2_ERR_INVALID;
  try {
    const Config c = load_config(config_path);
    memset(out, 0, sizeof(*out));
    out->frame_width = c.frame_width;
    out->frame_height = c.frame_height;
    out->samples = c.samples;
    out->gamma = c.gamma;
    out->use_brick_grid = c.use_brick_grid ? 1 : 0;
    out->filter_active = c.filter ? 1 : 0;
    out->min_voxel_size = c.min_voxel_size;
    out->lods = c.lods;
    out->samples_per_voxel = c.samples_per_voxel;
    out->sigma_t = c.sigma_t;
    out->estimation_iterations = c.estimation_iterations;
    out->alpha = c.alpha;
    out->clip_rays = c.clip_rays ? 1 : 0;
    out->fov = c.fov;
    memcpy(out->cam_pos, c.cam_pos, 12);
    memcpy(out->cam_look_at, c.look_at, 12);
    memcpy(out->cam_up, c.up, 12);
    for (const SceneObject& o : c.objects) {
      if (o.type == "Medium") ++out->n_media;
      else if (o.type == "Mesh") ++out->n_meshes;
      else ++out->n_surfaces;
      if (o.mask & kMaskFilter) ++out->n_filter_objects;
      if (o.mask & kMaskRender) ++out->n_render_objects;
    }
    out->max_depth = c.max_depth;
    return R2_OK;
  } catch (const std::exception& e) {
    if (err && err_len > 0) {
      strncpy(err, e.what(), (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
    return R2_ERR_INVALID;
  }
}
