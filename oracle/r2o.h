/*
 * Oracle (test infrastructure only): CPU restatement of the volumetric hot path of
 * FrederikBoehm/Renderer2.  Never linked by the product library.
 */
#ifndef R2O_H
#define R2O_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- precision-independent helpers (r2o_common.c) ---- */
typedef struct {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t buf[4];
  uint32_t n;
  uint32_t block;
} r2o_rng;

void r2o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void r2o_rng_init(r2o_rng* s, uint32_t key0, uint32_t key1, uint32_t c0, uint32_t c1, uint32_t c3);
uint32_t r2o_rng_next_u32(r2o_rng* s);

uint16_t r2o_half_from_float_glm(float f);
uint16_t r2o_half_from_float_rne(float f);
float r2o_half_to_float(uint16_t h);
uint32_t r2o_encode_range(float lo, float hi);
void r2o_decode_range(uint32_t d, float* lo, float* hi);
uint8_t r2o_encode_voxel(float value, float lo, float hi);
float r2o_decode_voxel(uint8_t v, float lo, float hi);
int r2o_voxel_active(const float* dens, int nx, int ny, int nz, int x, int y, int z);

typedef struct {
  int nx, ny, nz;   /* voxel counts of the source grid */
  int nb[3];        /* padded brick counts per axis */
  uint32_t* range;  /* nb^3 encoded half (min,max) */
  uint8_t* nonempty;
  uint8_t* u8;      /* dense (8*nb)^3 quantised densities */
  uint32_t* mip[3]; /* min/max mips, level l+1 has nb >> (l+1) per axis */
  size_t n_nonempty;
} r2o_bricks;

int r2o_brick_build(const float* dens, int nx, int ny, int nz, r2o_bricks* b);
void r2o_brick_free(r2o_bricks* b);
float r2o_brick_majorant(const r2o_bricks* b, int bx, int by, int bz, int mip);
float r2o_brick_density(const r2o_bricks* b, int ix, int iy, int iz);

/* ---- scene description shared by the FP32 / FP64 instantiations ---- */
/* distinct grids of a scene; instances (any number) refer to them by index */
#define R2O_MAX_GRIDS 16

typedef struct {
  int nx, ny, nz;
  const uint8_t* records;   /* nx*ny*nz 32-byte SFilteredDataCompact, x fastest */
  float map_mat[9];         /* NanoVDB Map mMatF (row-major 3x3): index -> model */
  float map_vec[3];         /* NanoVDB Map mVecF */
  float sigma_a[3], sigma_s[3];
  float voxel_size_filtering;
} r2o_grid_desc;

typedef struct {
  int grid;                 /* index into the grid list */
  float pos[3], orientation[3], scaling[3];
} r2o_instance_desc;

/* analytic surface with its material: CCircle / Sphere(worldPos, radius, normal)
 * and CMaterial(diffuse, specular, COrenNayarBRDF(roughness),
 * CMicrofacetBRDF(alphaX, alphaY, etaI, etaT)) (config_loader.cu:123-215,
 * sceneobject.cu:28-48) */
#define R2O_MAX_SURFACES 8
enum { R2O_SHAPE_CIRCLE = 0, R2O_SHAPE_SPHERE = 1 };
typedef struct {
  int shape;
  float pos[3], radius, normal[3];
  float diffuse[3], diffuse_roughness, specular[3], alpha_x, alpha_y, eta_i, eta_t;
} r2o_surface_desc;

typedef struct {
  int n_grids;
  r2o_grid_desc grids[R2O_MAX_GRIDS];
  int n_instances;
  const r2o_instance_desc* instances;  /* n_instances entries */
  const float* env_rgb;     /* env_h rows of env_w RGB float texels (row 0 = theta 0) */
  int env_w, env_h;
  int n_surfaces;
  r2o_surface_desc surfaces[R2O_MAX_SURFACES];
} r2o_scene_desc;

typedef struct {
  int width, height, spp, max_depth;
  float cam_pos[3], cam_look_at[3], cam_up[3], fov_deg;
  uint32_t seed, stream;
  int use_brick_grid;
  int sample_begin;         /* first sample index (for split renders) */
  int threads;              /* OpenMP threads (0 = default) */
} r2o_render_params;

/* per-path lookup counters (the reference's numLookups split by kind) */
enum { R2O_CNT_VOX = 0, R2O_CNT_MAJ, R2O_CNT_BRICK, R2O_CNT_ENV, R2O_CNT_DDA, R2O_CNT_FLIGHT, R2O_CNT_SCATTER, R2O_NCNT };

typedef struct r2o_scene_f32 r2o_scene_f32;
typedef struct r2o_scene_f64 r2o_scene_f64;

r2o_scene_f32* r2o_scene_new_f32(const r2o_scene_desc* d);
void r2o_scene_free_f32(r2o_scene_f32* s);
r2o_scene_f64* r2o_scene_new_f64(const r2o_scene_desc* d);
void r2o_scene_free_f64(r2o_scene_f64* s);

/* ---- mesh -> SGGX filter (r2o_filter.c) ---- */
/* RGBA float texture of a material (CTexture, texture.cu:29-99): height rows of
 * width texels, row 0 at v = 0 */
typedef struct {
  int width, height;
  const float* rgba;
} r2o_texture;

typedef struct {
  int n_vertices, n_triangles;
  const float* positions;   /* 3 per vertex, model space */
  const float* normals;     /* 3 per vertex, model space */
  const int32_t* indices;   /* 3 per triangle */
  const int32_t* material;  /* material id per triangle, or NULL (all 0) */
  int n_materials;
  const float* mat_diffuse;   /* 3 per material */
  const float* mat_specular;  /* 3 per material */
  const float* mat_roughness; /* 1 per material: SGGX sigma of buildS */
  const float* mat_ior;       /* 1 per material */
  float pos[3], orientation[3], scaling[3];
  const float* uvs;             /* 2 per vertex (texture coordinates), or NULL */
  int n_textures;
  const r2o_texture* textures;
  const int32_t* mat_textures;  /* 4 per material: diffuse, glossy, normal map, alpha mask (-1: none), or NULL */
} r2o_mesh_desc;

typedef struct {
  int nv[3];
  float index_to_model[12], model_to_index[12], model_to_world[12], world_to_model[12]; /* glm mat4x3, column major */
  float world_bb_min[3], world_bb_max[3];
} r2o_filter_grid;

typedef struct {
  int samples_per_voxel, estimation_iterations, clip_rays, lod, threads;
  float sigma_t, alpha;
  uint32_t seed;
} r2o_filter_params;

typedef struct r2o_filter_ctx r2o_filter_ctx;

int r2o_filter_setup(const r2o_mesh_desc* d, float voxel_size, r2o_filter_grid* g);
r2o_filter_ctx* r2o_filter_new(const r2o_mesh_desc* d);
void r2o_filter_free(r2o_filter_ctx* c);
int r2o_filter_run(const r2o_filter_ctx* c, const r2o_filter_grid* g, const r2o_filter_params* p, int64_t n_ids,
                   const int64_t* ids, uint8_t* out);
int r2o_filter_probe(const r2o_filter_ctx* c, const r2o_filter_grid* g, const r2o_filter_params* p, int64_t vid,
                     uint8_t* rec, float* o, float* d, float* tmax, int32_t* tri, float* t);
int r2o_filter_closest_hit(const r2o_filter_ctx* c, const float o[3], const float d[3], float tmax, float out[3]);
void r2o_tex_fetch(const r2o_texture* t, float x, float y, float out[4]);
float r2o_det_expf(float x);
float r2o_det_logf(float x);
float r2o_det_powf(float x, float y);
double r2o_det_exp_d(double x);
double r2o_det_log_d(double x);
void r2o_det_sincosf(float x, float* s, float* c);

/* ---- post-process (r2o_post.c) ---- */
void r2o_postprocess(const float* hdr, int w, int h, float gamma, float* ldr, uint8_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
