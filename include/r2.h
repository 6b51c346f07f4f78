/*
 * r2.h -- C ABI of the MI355X-native volumetric hot path of FrederikBoehm/Renderer2:
 * the per-voxel mesh -> SGGX filter and the volume path integrator, as HIP kernels
 * for gfx950 behind plain-pointer entry points.
 *
 * Conventions: every entry point returns R2_OK (0) or a negative R2_ERR_* code and
 * never throws; r2_last_error() holds the message of the last failure on a context.
 * A context is externally synchronised (one host thread at a time); the render
 * server (r2_server_*) is internally synchronised.  Host buffers belong to the
 * caller; device memory behind r2_grid / r2_env / r2_mesh handles belongs to the
 * library until the matching *_free call.
 *
 * Reference citations are relative to the reference repository root.
 */
#ifndef R2_H
#define R2_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define R2_OK 0
#define R2_ERR_INVALID (-1)     /* bad argument / shape */
#define R2_ERR_HIP (-2)         /* HIP runtime failure */
#define R2_ERR_OOM (-3)         /* device or host allocation failed / budget exceeded */
#define R2_ERR_UNSUPPORTED (-4) /* feature or size outside what the build supports */
#define R2_ERR_IO (-5)          /* file missing or malformed */
#define R2_ERR_CANCELLED (-6)   /* request cancelled / abort callback returned true */
#define R2_ERR_BUSY (-7)        /* server queue full */
#define R2_ERR_DEADLINE (-8)    /* server request not finished by its deadline (dropped) */

#define R2_VERSION 3

/* Grid record precision (r2_grid_desc.precision) */
#define R2_PREC_FP32 0   /* 32-byte SFilteredDataCompact voxel records (filtered_data.hpp:33-61) */
#define R2_PREC_PACKED 1 /* 24-byte records: FP16 density, the six SGGX coefficients of S in FP16 after an
                            FP64 eigen-decomposition (PSD projection), FP16 ior, u8 colours,
                            octahedral normal */

/* r2_render_params.flags */
#define R2_FLAG_BRICK_GRID 1u   /* majorant DDA + brick density (config "UseBrickGrid": true) */
#define R2_FLAG_FP64_ACCUM 2u   /* compensated FP64 framebuffer accumulation */

typedef struct r2_ctx r2_ctx;
typedef struct r2_grid r2_grid;
typedef struct r2_env r2_env;
typedef struct r2_mesh r2_mesh;
typedef struct r2_server r2_server;

/* ------------------------------------------------------------- context -- */
/* Replaces the device setup of rt::Raytracer / filter::CFilter
 * (raytracer/raytracer.cu:129-152, raytracer/filter.cu:20-57). */
int r2_create(r2_ctx** out, int device);
int r2_destroy(r2_ctx* ctx);
const char* r2_last_error(const r2_ctx* ctx);
int r2_version(void);
/* The HIP stream the context launches on (a hipStream_t). */
void* r2_stream(r2_ctx* ctx);

/* ---------------------------------------------------------------- grid -- */
/* A filtered volume: NanoVDB-layout record tree + brick majorant grid.
 * Replaces CNVDBMedium(path, sigmaA, sigmaS) (raytracer/src/medium/nvdb_medium.cu:45-61,
 * :130-173) with CAssetManager::loadMedium (raytracer/src/backend/asset_manager.cu:114-138)
 * and the brick grid build CHostBrickGrid (raytracer/src/grid_brick/host_grid_brick.cpp:61-152),
 * which here runs on the GPU. */
typedef struct {
  int32_t nx, ny, nz;          /* voxel counts; records are x fastest, then y, then z */
  const void* records;         /* nx*ny*nz SFilteredDataCompact records, 32 bytes each */
  int32_t records_on_device;   /* 1: `records` is a device pointer */
  float map_mat[9];            /* NanoVDB Map mMatF, row-major 3x3: index -> model */
  float map_vec[3];            /* NanoVDB Map mVecF */
  float sigma_a[3], sigma_s[3];
  float voxel_size_filtering;  /* filter voxel size of this LOD (nvdb_medium.cu:180-188) */
  int32_t precision;           /* R2_PREC_* */
  /* Sparse form (n_leaves > 0; since version 3), as a NanoVDB tree holds a medium: records
   * then holds n_leaves leaves of 8^3 records each (voxel x + 8 y + 64 z within the leaf, host
   * or device memory per records_on_device), leaf_origins their index-space origins (3 ints per
   * leaf, host memory: multiples of 8 inside [0, nx) x [0, ny) x [0, nz), no two alike) and
   * leaf_masks their active voxels (64 bytes per leaf, bit v % 8 of byte v / 8; host memory),
   * or NULL for the dense rule (density > 0 or a corner of the index box).  nx, ny, nz are
   * the index extent; voxels outside every leaf are background.  n_leaves = 0: dense records. */
  int64_t n_leaves;
  const int32_t* leaf_origins;
  const uint8_t* leaf_masks;
} r2_grid_desc;

typedef struct {
  float index_to_model[12], model_to_index[12];  /* glm::mat4x3, column major */
  float model_to_scaled[12], scaled_to_model[12];
  int32_t ibb_min[3], size[3];
  int32_t n_bricks[3];
  int64_t n_leaves, n_nonempty_bricks;
  float max_density, inv_max_density;
  float bbox_min[3], bbox_max[3];                /* model-space bbox of the grid */
  int64_t device_bytes;
  float min_density;                             /* over all voxels, inactive ones reading 0 */
} r2_grid_info;

int r2_grid_upload(r2_ctx* ctx, const r2_grid_desc* desc, r2_grid** out);
int r2_grid_free(r2_ctx* ctx, r2_grid* grid);
int r2_grid_get_info(const r2_grid* grid, r2_grid_info* out);
/* Parity probe: copy the brick grid back.  range: nb^3 u32 (half min | half max << 16),
 * u8: dense (8*nb)^3 quantised densities (meaningful in non-empty bricks),
 * mips: 3 levels of u32 ranges concatenated.  Any pointer may be NULL. */
int r2_grid_download_bricks(r2_ctx* ctx, const r2_grid* grid, uint32_t* range, uint8_t* u8, uint32_t* mips);
/* Parity probe: NanoVDB-layout voxel lookups. For each of n integer coordinates
 * (xyz triples) write the 32-byte record the tree returns (background when inactive);
 * packed grids keep these FP32 records beside their packed ones. */
int r2_grid_lookup(r2_ctx* ctx, const r2_grid* grid, int32_t n, const int32_t* coords, uint8_t* out_records);
/* Parity probe: the filtered record (trilinear over the 8 neighbours, decoded) at n
 * positions normalised to the index bbox, as CNVDBMedium::filteredData
 * (raytracer/include/medium/nvdb_medium_impl.hpp:47-64) computes it; 17 floats each:
 * density, S (xx yy zz xy xz yz), diffuse rgb, specular rgb, normal xyz, ior.
 * A packed grid answers from its packed records. */
int r2_grid_sample(r2_ctx* ctx, const r2_grid* grid, int32_t n, const float* pos, float* out);
/* Parity probe: the majorant-DDA walk of the free-flight loop of CNVDBMedium::sampleDDA
 * (raytracer/include/medium/nvdb_medium_impl.hpp:139-187) with stepDDA (:241-246) and the
 * mip schedule (start 3, +0.25 per step, -2 after a tentative collision, :13-15,158,184),
 * without density decisions.  rays7: n x (origin xyz, direction xyz, tmax) in voxel space;
 * collide_every = k > 0 applies a collision's mip drop after every k-th step (0: never).
 * cells: n x max_cells x (brick x, y, z, mip) of the majorant lookups; counts: cells per ray. */
int r2_dda_cells(r2_ctx* ctx, int32_t n, const float* rays7, float sigma_t, int32_t collide_every,
                 int32_t max_cells, int32_t* cells, int32_t* counts);

/* ----------------------------------------------------------------- env -- */
/* Lat-long HDR environment map with sin(theta)-weighted 2-D importance sampling.
 * Replaces CEnvironmentMap(path) + CDistribution2D (raytracer/src/scene/environmentmap.cu:11-29,
 * raytracer/src/sampling/distribution_2d.cu:7-14).  rgb: h rows of w RGB texels,
 * row 0 = theta 0 (+y). */
int r2_env_upload(r2_ctx* ctx, const float* rgb, int32_t w, int32_t h, r2_env** out);
int r2_env_free(r2_ctx* ctx, r2_env* env);

/* -------------------------------------------------------------- render -- */
/* A render takes up to R2_MAX_INSTANCES instances of at most R2_MAX_GRIDS
 * distinct grids (the reference's instance acceleration structure over its medium
 * objects, raytracer/src/scene/scene.cu:188-256).  Beyond 8 instances the boxes
 * are found through a hierarchy built per render, with the same closest hit as a
 * scan in instance order. */
#define R2_MAX_INSTANCES (1 << 20)
#define R2_MAX_GRIDS 16
typedef struct {
  const r2_grid* grid;
  float pos[3], orientation[3], scaling[3];   /* CHostSceneobject(medium, pos, orientation, scaling) */
} r2_instance;

/* Analytic surface with its material: CHostSceneobject(CCircle / Sphere(pos, radius,
 * normal), diffuseReflection, diffuseRoughness, specularReflection, alphaX, alphaY,
 * etaI, etaT) as the config loader's addCircle / addSphere build it
 * (raytracer/src/backend/config_loader.cu:122-215, raytracer/src/scene/sceneobject.cu:28-48):
 * an Oren-Nayar lobe of roughness diffuse_roughness (raytracer/include/material/oren_nayar_brdf.hpp:31-35)
 * and a GGX microfacet lobe (microfacet_brdf.hpp:42-45) mixed by the Fresnel weight
 * of (eta_i, eta_t) (material.hpp:67-100).  A circle is the disk of that radius in
 * the plane through pos with that normal (circle.hpp:88-113); a sphere is solved
 * in the frame translate(pos) * getRotation(normal) (shape.hpp:49-57, sphere.hpp:33-90). */
#define R2_SHAPE_CIRCLE 0
#define R2_SHAPE_SPHERE 1
#define R2_MAX_SURFACES 8
typedef struct {
  int32_t shape;                                          /* R2_SHAPE_* */
  float pos[3], radius, normal[3];
  float diffuse[3], diffuse_roughness, specular[3];
  float alpha_x, alpha_y, eta_i, eta_t;
} r2_surface;

typedef struct {
  int32_t width, height, spp, max_depth;
  float cam_pos[3], cam_look_at[3], cam_up[3], fov_deg;  /* CCamera (raytracer/src/camera/camera.cu:20-32) */
  uint32_t seed, stream;                                  /* RNG key */
  int32_t n_instances;
  const r2_instance* instances;
  const r2_env* env;
  uint32_t flags;                                         /* R2_FLAG_* */
  int32_t sample_begin;                                   /* first sample index */
  int32_t n_surfaces;                                     /* 0 .. R2_MAX_SURFACES */
  const r2_surface* surfaces;                             /* (since version 3) */
} r2_render_params;

/* Render one frame: linear HDR RGB, W*H*3 floats, row 0 = bottom (launch index y = 0).
 * Replaces rt::Raytracer::renderFrame's spp loop of optixLaunch(W, H, 1) with
 * __raygen__rg / CPathIntegrator::Li (raytracer/raytracer.cu:161-190,
 * raytracer/src/backend/shaders.optix.cu:240-256,
 * raytracer/include/integrators/path_integrator_impl.hpp:120-247).
 * out_on_device: 1 if out_rgb is a device pointer. */
int r2_render(r2_ctx* ctx, const r2_render_params* p, float* out_rgb, int32_t out_on_device);

/* Progress callback of r2_render_progressive: called after each completed sample
 * chunk with the samples per pixel done so far; a nonzero return stops the frame. */
typedef int (*r2_progress_fn)(void* user, int32_t samples_done, int32_t samples_total);

/* r2_render with a callback between sample chunks: rt::Raytracer::renderFrame's
 * keyCallback, polled after each spp launch, which returns the frame accumulated so
 * far when it asks to stop (raytracer/raytracer.cu:161-190).  The frame runs as one
 * launch per sample chunk (a chunk is ceil(spp W H / 2^24) samples of every pixel),
 * with two launches in flight.  Completed: R2_OK and the same frame as r2_render,
 * bit for bit.  Stopped: R2_ERR_CANCELLED, and out_rgb holds the sum over the
 * completed samples of L / spp, as the reference's framebuffer does. */
int r2_render_progressive(r2_ctx* ctx, const r2_render_params* p, float* out_rgb, int32_t out_on_device,
                          r2_progress_fn progress, void* user);

/* r2_render with no device allocation.  No counterpart in the reference: this
 * entry point is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * out_rgb must be device memory (W*H*3 floats).  The frame's chunk sums go to a
 * fixed 64 MiB device array that the library holds on each GPU it runs on until
 * the process ends, instead of the context's scratch.  Frames up to 5592405
 * pixels (2796202 with R2_FLAG_FP64_ACCUM) are accepted, larger ones return
 * R2_ERR_UNSUPPORTED; when the default chunking does not fit, items take more
 * samples each.  The result equals r2_render's bit for bit whenever r2_render's
 * chunk sums fit in 64 MiB.  All calls on one GPU share the array: calls that
 * overlap on one GPU (from several contexts or threads) race. */
int r2_render_noalloc(r2_ctx* ctx, const r2_render_params* p, float* out_rgb);
/* (r2_render_noalloc also takes the scene's instance table from the context's
 * preallocated scene buffer: scenes of more than 4096 instances return
 * R2_ERR_UNSUPPORTED there.) */

/* Parity probe: CDeviceScene::intersect (raytracer/include/scene/device_scene_impl.hpp:12-31, with
 * __intersection__volume / __intersection__surface, shaders.optix.cu:170-236) of n rays
 * (origin xyz, direction xyz, tmax) against the instance boxes and surfaces of p's scene
 * (the camera and frame fields are not used).  out: n x (hit, t, instance or -1, surface or -1). */
int r2_scene_intersect(r2_ctx* ctx, const r2_render_params* p, int32_t n, const float* rays7, float* out4);

/* Parity probe: one path per (pixel, sample) pair; out: n*3 (L / spp);
 * counters: n*8 u32 lookup counters (voxel records, majorants, brick densities,
 * env lookups, DDA steps, free-flight steps, scatter events, unused) or NULL. */
int r2_render_paths(r2_ctx* ctx, const r2_render_params* p, int32_t n, const int32_t* pixel,
                    const int32_t* sample, float* out, uint32_t* counters);

typedef struct {
  double last_kernel_ms;        /* integrator kernel time of the last render (HIP events) */
  double last_total_ms;         /* host wall time of the last r2_render call */
  int64_t last_paths;
  int32_t last_launches;
} r2_stats;
int r2_get_stats(const r2_ctx* ctx, r2_stats* out);

/* ------------------------------------------------------------- filter -- */
/* Triangle mesh for the mesh -> SGGX filter.  Replaces CMesh + OptiX GAS/IAS
 * (raytracer/include/mesh/mesh.hpp:18-59, raytracer/src/mesh/mesh.cu:140-168)
 * with per-LOD candidate lists: each voxel is tested against the triangles
 * within reach of its rays. */
/* RGBA float texture of a mesh material (CTexture: 8-bit images divided by 255,
 * raytracer/src/texture/texture.cu:29-99), sampled like tex2D<float4> with linear
 * filtering, wrap addressing and normalised coordinates (texture.cu:166-189). */
typedef struct {
  int32_t width, height;
  const float* rgba;           /* height rows of width RGBA texels, row 0 at v = 0 */
} r2_texture_desc;

typedef struct {
  int32_t n_vertices, n_triangles;
  const float* positions;      /* n_vertices * 3, model space */
  const float* normals;        /* n_vertices * 3 */
  const int32_t* indices;      /* n_triangles * 3 */
  const int32_t* material;     /* n_triangles material ids (NULL = all 0) */
  int32_t n_materials;
  const float* mat_diffuse;    /* n_materials * 3 */
  const float* mat_specular;   /* n_materials * 3 */
  const float* mat_roughness;  /* n_materials (specularRoughness, material.cu:146-148) */
  const float* mat_ior;        /* n_materials */
  float pos[3], orientation[3], scaling[3];  /* scene object transform */
  /* Textured materials (CMaterial's albedo, glossy, normal and alpha textures,
   * raytracer/include/material/material.hpp:54-57,111-156).  uvs: n_vertices * 2
   * texture coordinates, or NULL (every hit then has tc = (0, 0) and no alpha
   * test or normal map applies, as for a mesh without texture coordinates in
   * shaders.optix.cu:48-152).  mat_textures: n_materials * 4 indices into
   * textures (diffuse colour, glossy colour, normal map, alpha mask; -1 = the
   * constant colour / none), or NULL. */
  const float* uvs;
  int32_t n_textures;
  const r2_texture_desc* textures;
  const int32_t* mat_textures;
} r2_mesh_desc;
int r2_mesh_upload(r2_ctx* ctx, const r2_mesh_desc* desc, r2_mesh** out);
int r2_mesh_free(r2_ctx* ctx, r2_mesh* mesh);

typedef struct {
  float voxel_size;            /* this LOD's voxel size (world units) */
  int32_t samples_per_voxel;   /* "SamplesPerVoxel" (<= 1000, mesh_filter.hpp:66) */
  int32_t estimation_iterations;
  float sigma_t, alpha;
  int32_t clip_rays;
  uint32_t seed;
  int32_t lod;                 /* RNG stream tag */
} r2_filter_params;

typedef struct {
  int32_t num_voxels[3];
  float index_to_model[12], model_to_index[12], model_to_world[12], world_to_model[12];
  float world_bb_min[3], world_bb_max[3];
} r2_filter_grid;

/* Grid set-up of one LOD (COpenvdbBackend::init + setupGrid,
 * raytracer/src/filtering/openvdb_backend.cpp:26-106). */
int r2_filter_setup(const r2_mesh* mesh, float voxel_size, r2_filter_grid* out);
/* Filter one LOD: writes nx*ny*nz 32-byte SFilteredDataCompact records, voxel id
 * x + nx*(y + ny*z) (CFilter::runFiltering's optixLaunch(nx, ny, nz) of
 * __raygen__filtering / CMeshFilter::run, raytracer/filter.cu:74-123,
 * raytracer/src/backend/shaders.optix.cu:258-284,
 * raytracer/include/filtering/mesh_filter.hpp:51-122).
 * out_on_device: 1 if out_records is a device pointer. */
int r2_filter_lod(r2_ctx* ctx, const r2_mesh* mesh, const r2_filter_params* p, void* out_records,
                  int32_t out_on_device, r2_filter_grid* grid_out);
/* Filter only the listed voxels of the same LOD grid (host ids and host output,
 * n records in list order).  Probe entry point with the per-voxel results of
 * r2_filter_lod, for spot checks of large grids. */
int r2_filter_voxels(r2_ctx* ctx, const r2_mesh* mesh, const r2_filter_params* p, int32_t n,
                     const uint32_t* voxel_ids, void* out_records);

/* Counters of the last r2_filter_lod: grid voxels, voxels with candidate
 * triangles, (voxel, triangle) candidate pairs, rays cast, filter kernel time
 * and wall time of the call. */
typedef struct {
  int64_t voxels, active_voxels, pairs, rays;
  double filter_kernel_ms, total_ms;
} r2_filter_stats;
int r2_filter_get_stats(const r2_ctx* ctx, r2_filter_stats* out);

/* ------------------------------------------------------------- server -- */
/* Device-resident render server with continuous batching (no reference
 * counterpart beyond the interactive preview loop, raytracer/pipeline.cu:41-78). */
typedef struct {
  int64_t hbm_budget_bytes;    /* admission budget for per-request device memory: one buffer of this
                                  size is allocated by r2_server_start and sub-allocated per request
                                  (default 8 GiB; R2_ERR_OOM if it does not fit the device) */
  int32_t max_inflight;        /* max requests resident on the device at once */
  int32_t slice_paths;         /* paths per scheduling slice */
} r2_server_config;

typedef struct {
  r2_render_params params;     /* instances / env must stay valid until the result is taken */
  float* out_rgb;              /* host buffer W*H*3, filled when the request completes */
  double deadline_ms;          /* 0: none; else the time after submission by which the frame is
                                  wanted.  Queued requests are admitted earliest deadline first
                                  (those without one last, in arrival order); a request still
                                  unfinished at its deadline is dropped and reports
                                  R2_ERR_DEADLINE (checked at least once per slice). */
} r2_request;

typedef struct {
  int64_t submitted, completed, rejected, cancelled;
  double paths_per_sec;
  double p50_ms, p99_ms;
  int64_t peak_bytes;
  int64_t expired;             /* requests dropped at their deadline */
} r2_server_stats;

int r2_server_start(r2_ctx* ctx, const r2_server_config* cfg, r2_server** out);
/* Returns a request id >= 0, or R2_ERR_BUSY / R2_ERR_OOM (request larger than the budget). */
int64_t r2_server_submit(r2_server* s, const r2_request* req);
/* 1 = done (result written), 0 = pending, <0 = error / cancelled.  The outcome of a
 * finished request is reported once, by the first poll or wait that sees it; the id
 * is then forgotten (later calls return R2_ERR_INVALID). */
int r2_server_poll(r2_server* s, int64_t id, double* latency_ms);
int r2_server_wait(r2_server* s, int64_t id, double* latency_ms);
int r2_server_cancel(r2_server* s, int64_t id);
int r2_server_get_stats(r2_server* s, r2_server_stats* out);
int r2_server_stop(r2_server* s);

/* --------------------------------------------------------- post-process -- */
/* Tonemap, gamma and bytes of a rendered frame (SFrame), as rt::Raytracer::renderFrame
 * does after its spp loop (raytracer/raytracer.cu:53-127,192-206): k = exp(sum over
 * pixels of log(r+g+b+FLT_MIN) / (3 W H)) - W H FLT_MIN, c -> c / (c + k), then
 * c^(1/gamma) and round(255 c).  hdr / ldr / bytes hold W*H*3 values; ldr or bytes may
 * be NULL.  on_device: 1 if the three pointers are device pointers. */
int r2_postprocess(r2_ctx* ctx, const float* hdr, int32_t w, int32_t h, float gamma, float* ldr, uint8_t* bytes,
                   int32_t on_device);

/* ------------------------------------------------------------ front end -- */
/* The reference's program: CConfigLoader::loadConfig + CPipeline::run
 * (raytracer/src/backend/config_loader.cu:345-368, raytracer/pipeline.cu:19-93,
 * raytracer/main.cu:4-17).  Reads config.json and its scenedescription.json (same keys,
 * comments allowed), filters the FILTER-masked mesh to its LODs when
 * Filtering.Active (grid files <out_dir>/filtering/<mesh>/filtered_mesh_<vs>.r2grid and
 * filtering/volume_description.json), renders the RENDER-masked media under the env map
 * and writes <out_dir>/output.jpg (quality 100) and output.png (tonemapped, flipped
 * like the reference, raytracer/pipeline.cu:84-85) and output.pfm (linear).  err receives a message on failure. */
typedef struct {
  int32_t lods_filtered;
  double filter_ms;
  int32_t width, height, spp;
  double render_ms, mpaths_per_s;
  int32_t media_rendered, objects_skipped;
  int32_t surfaces_rendered;    /* Circle / Sphere objects rendered (since version 3) */
  /* media whose .brickgrid file (beside the grid file) equals / differs from the brick grid
   * built from their records (since version 3; a differing file is reported and not used) */
  int32_t brickgrids_matched, brickgrids_differing;
} r2_pipeline_report;
int r2_pipeline_run(const char* config_path, const char* out_dir, int32_t device, r2_pipeline_report* report,
                    char* err, int32_t err_len);

/* Image file of a post-processed frame: CImageWriter::writeToFile
 * (visualisation/src/image_writer.cpp:13-21).  rgb holds w*h 8-bit RGB pixels, rows
 * bottom to top as r2_postprocess writes them; the file is flipped to top-down, as
 * stbi_flip_vertically_on_write(true) makes the reference's.  R2_IMAGE_PNG: an
 * uncompressed PNG (stbi_write_png_compression_level = 0).  R2_IMAGE_JPG: a baseline
 * JFIF JPEG at quality 1..100 (the pipeline writes 100), chroma at full resolution above
 * 90 and 2x2-averaged at 90 and below.  R2_ERR_INVALID for a bad size, format or
 * quality; R2_ERR_IO when the file cannot be written.  Host only. */
#define R2_IMAGE_PNG 0
#define R2_IMAGE_JPG 1
int r2_image_write(const char* path, int32_t w, int32_t h, const uint8_t* rgb, int32_t format, int32_t quality,
                   char* err, int32_t err_len);

/* Write the brick majorant grid of a grid as the reference's .brickgrid file: a
 * CHostBrickGrid in cereal 1.3.2's portable binary archive (rt::write_grid,
 * raytracer/src/grid_brick/serialization.cpp:37-57; fields host_grid_brick.hpp:47-72).
 * r2_pipeline_run writes one beside each filtered LOD. */
int r2_grid_write_brickgrid(r2_ctx* ctx, const r2_grid* grid, const char* path, char* err, int32_t err_len);
/* Read a .brickgrid file (rt::read_grid, raytracer/src/grid_brick/serialization.cpp:63-70, which
 * CNVDBMedium uses in place of building the brick grid): brick counts, density extrema and the
 * non-empty brick count into info, and, where the pointers are not NULL, the level-0 ranges (one
 * encoded half pair per brick), the atlas densities put back in brick layout (8 nbx x 8 nby x
 * 8 nbz bytes, x fastest: the layout of r2_grid_download_bricks) and the 3 mip levels.  The
 * front end checks the file beside a medium against the brick grid built from its records.
 * R2_ERR_IO for a missing or malformed file.  Host only. */
typedef struct {
  int32_t n_bricks[3];
  float min_density, max_density;
  int64_t n_nonempty_bricks;
} r2_brickgrid_info;
int r2_brickgrid_read(const char* path, r2_brickgrid_info* info, uint32_t* ranges, uint8_t* u8, uint32_t* mips,
                      char* err, int32_t err_len);

/* NanoVDB grid files (.nvdb, NanoVDB 32.3 layout, one Vec4d grid, uncompressed), host
 * only.  r2_nvdb_write replaces COpenvdbBackend::setValues + writeToFile
 * (raytracer/src/filtering/openvdb_backend.cpp:108-157: openToNanoVDB +
 * nanovdb::io::writeGrid): records are nx*ny*nz 32-byte SFilteredDataCompact records,
 * x fastest; active voxels are those with density > 0 and the 8 grid corners;
 * map_mat (row-major 3x3) and map_vec are the index -> model transform that setupGrid
 * gives the grid (openvdb_backend.cpp:84-95), stored as the NanoVDB map (NULL: the
 * identity); node statistics are written as absent.
 * r2_nvdb_read replaces the grid load of CNVDBMedium (raytracer/src/medium/
 * nvdb_medium.cu:130-173): info gets the index box and map; with records != NULL it
 * also gets the dense records over the index box (x fastest, inactive voxels = the
 * grid's background); records_bytes must hold nx*ny*nz*32 (R2_ERR_INVALID if not).
 * R2_ERR_IO for a missing, truncated or unsupported file (other grid types, codecs,
 * active root tiles). */
typedef struct {
  int32_t nx, ny, nz;     /* extent of the index bounding box */
  int32_t ibb_min[3];     /* its lower corner */
  int64_t active_voxels;
  float map_mat[9];       /* index -> model, row major (mMatF) */
  float map_vec[3];       /* translation (mVecF) */
  int32_t leaf_count, lower_count, upper_count;
} r2_nvdb_info;
int r2_nvdb_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const void* records, const float* map_mat,
                  const float* map_vec, char* err, int32_t err_len);
int r2_nvdb_read(const char* path, r2_nvdb_info* info, void* records, int64_t records_bytes, char* err,
                 int32_t err_len);
/* The same file as sparse leaves, for the sparse form of r2_grid_desc (the grid load of
 * CNVDBMedium without densifying the tree, raytracer/src/medium/nvdb_medium.cu:130-173):
 * the file's leaves with their value masks, and the active tiles of its internal nodes as
 * full leaves, in an index space shifted by origin_shift (the index box minimum rounded down
 * to a multiple of 8, so origins are non-negative multiples of 8; fold map_mat * shift into
 * map_vec).  In: *n_leaves = capacity of the buffers (leaf_origins 3 ints, records 512 x 32 B,
 * leaf_masks 64 B per leaf, voxels x fastest).  Out: *n_leaves = the leaves of the file; with
 * records == NULL only that count (and info, origin_shift) is returned.  Host only. */
int r2_nvdb_read_leaves(const char* path, r2_nvdb_info* info, int32_t* origin_shift, int64_t* n_leaves,
                        int32_t* leaf_origins, void* records, uint8_t* leaf_masks, char* err, int32_t err_len);

/* Parse a config (and its scene description) without a device: the fields of SConfig
 * (raytracer/include/backend/config_loader.hpp:9-33) that drive this library. */
typedef struct {
  int32_t frame_width, frame_height, samples;
  float gamma;
  int32_t use_brick_grid;
  int32_t filter_active;
  float min_voxel_size;
  int32_t lods, samples_per_voxel;
  float sigma_t;
  int32_t estimation_iterations;
  float alpha;
  int32_t clip_rays;
  float fov, cam_pos[3], cam_look_at[3], cam_up[3];
  int32_t n_media, n_meshes, n_surfaces, n_filter_objects;
  int32_t max_depth;
  int32_t n_render_objects;    /* objects whose mask has the RENDER bit */
} r2_config_summary_t;
int r2_config_summary(const char* config_path, r2_config_summary_t* out, char* err, int32_t err_len);

/* ----------------------------------------------------- host utilities -- */
/* Synthetic inputs of the benchmark configurations (seeded, deterministic):
 * kind 1 = isotropic value-noise volume (S = I, normal +y), kind 2 = anisotropic
 * fibre volume (S = R diag(1,1,0.05) R^T, von Mises-Fisher fibre field).
 * Writes n^3 32-byte records.  density_scale: max density of the noise field. */
int r2_synth_volume(int32_t kind, int32_t n, uint32_t seed, float density_scale, int32_t octaves,
                    void* out_records);
/* kind 0 = constant white (w x h of 1.0), kind 1 = procedural sky gradient + 0.5 degree sun. */
int r2_synth_env(int32_t kind, int32_t w, int32_t h, float* out_rgb);
/* Procedural mesh: noise-displaced cube-sphere (12*res^2 triangles, material 0) plus
 * n_quads random thin square quads (2 triangles each, material 1) in a shell around it.
 * With positions == NULL only the counts are written. */
int r2_synth_mesh(int32_t res, int32_t n_quads, float quad_size, uint32_t seed, int32_t* n_vertices,
                  int32_t* n_triangles, float* positions, float* normals, int32_t* indices, int32_t* material);

#ifdef __cplusplus
}
#endif
#endif
