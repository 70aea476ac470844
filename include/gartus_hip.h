/*
 * gartus_hip.h — C ABI of the MI355X path tracer.
 *
 * This is the drop-in boundary for gartus's path-tracing hot path:
 *   PathTracer::render / render_hdr_image / render_denoising_aovs /
 *   render_with_lights* / render_with_environment* over a compiled RayScene
 *   (src/graphics/raytracing/renderer.rs:101-224,374-609).
 * Every entry point below names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns int status: GT_OK (0) or a GT_E* code. The
 *     message of the last failure on the calling thread is gt_last_error().
 *     Conditions that panic (assert!) in the reference return GT_EINVAL.
 *   - Inputs are copied; outputs are caller-owned buffers.
 *   - Host-pointer outputs ("out_*") are written after the device work ends.
 *     The *_device variants take device pointers and leave the result on the
 *     context's stream; call gt_context_synchronize before reading it.
 *   - A context owns one HIP stream and is not thread-safe; separate
 *     contexts are independent.
 *   - Scenes, light lists and environments belong to the context that
 *     created them.
 */
#ifndef GARTUS_HIP_H
#define GARTUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GT_OK = 0,
  GT_EINVAL = 1,   /* invalid argument (reference: assert! panic) */
  GT_ENOMEM = 2,   /* host or device allocation failed */
  GT_EDEVICE = 3,  /* HIP runtime error */
  GT_ESTATE = 4,   /* object in the wrong state (e.g. scene not built) */
  GT_EUNSUPPORTED = 5
};

typedef struct gt_context gt_context;
typedef struct gt_scene gt_scene;
typedef struct gt_lights gt_lights;
typedef struct gt_env gt_env;

/* ---- camera ------------------------------------------------------------
 * Plain-data mirror of RayCamera's builder state
 * (src/graphics/camera.rs:612-643, defaults :1034-1064).
 */
enum { GT_PIXEL_RANDOM = 0, GT_PIXEL_STRATIFIED = 1, GT_PIXEL_STRATIFIED_GRID = 2 };
enum { GT_LIGHTING_CONTINUATION = 0, GT_LIGHTING_NEE = 1 };
enum { GT_BACKGROUND_CONSTANT = 0, GT_BACKGROUND_VERTICAL_GRADIENT = 1 };
/* RNG streams: GT_RNG_COUNTER is the per-(pixel, sample, bounce) stream the
 * GPU uses; GT_RNG_PIXEL_STREAM is the reference's one sequential stream per
 * pixel (camera.rs:2669), available in the CPU oracle only. */
enum { GT_RNG_COUNTER = 0, GT_RNG_PIXEL_STREAM = 1 };
/* gt_camera.flags. 0 renders exactly the reference's estimator.
 * GT_FLAG_GGX_REJECT_BELOW_HORIZON ends a path whose GGX continuation
 * sample reflects below the surface. The reference's
 * GgxReflectionPdf::generate (sampling.rs:313-318) returns the outgoing
 * direction there, and ray_color then weights that retro-reflection by
 * scattering_pdf / pdf as if it were a density sample. Between facing
 * GGX surfaces this compounds (x88 per bounce measured), giving a biased
 * estimator with unbounded variance. Rejecting is the unbiased choice,
 * since the BSDF is zero below the horizon; the RNG draws are unchanged. */
enum { GT_FLAG_GGX_REJECT_BELOW_HORIZON = 1 };

typedef struct gt_camera {
  uint32_t image_width;             /* camera.rs:1035, > 0 */
  uint32_t samples_per_pixel;       /* clamped to >= 1 */
  double aspect_ratio;              /* finite, > 0 */
  uint32_t pixel_sample_mode;       /* GT_PIXEL_* (camera.rs:40-49) */
  uint32_t stratified_grid_width;   /* for GT_PIXEL_STRATIFIED_GRID */
  uint32_t max_depth;               /* default 10 */
  int32_t russian_roulette_min_depth; /* default 5; -1 = off */
  uint64_t rng_seed;                /* default 1 */
  double vertical_fov;              /* degrees, (0, 180) */
  double lookfrom[3];
  double lookat[3];
  double view_up[3];
  double defocus_angle;             /* degrees, [0, 180) */
  double focus_distance;            /* > 0 */
  double shutter_start, shutter_end;
  uint32_t background_kind;         /* GT_BACKGROUND_* (camera.rs:212-322) */
  uint32_t direct_lighting_mode;    /* GT_LIGHTING_* (camera.rs:92-196) */
  double background[9];             /* constant: [0..2]; gradient: nadir, horizon, zenith */
  double light_pdf_weight;          /* mixture weight for continuation mode */
  uint32_t rng_mode;                /* GT_RNG_* */
  uint32_t flags;                   /* GT_FLAG_* (0: the reference's behaviour) */
  /* RayCamera::with_adaptive_sampling (camera.rs:1244-1264; AdaptiveSampling
   * :362-394). Off while adaptive_error_threshold is 0; otherwise it must be
   * finite and > 0. In GT_PIXEL_RANDOM mode each pixel then runs
   * render_world_pixel_adaptive (:2937-2982): up to adaptive_max_samples
   * samples, stopping at the first accepted sample count >= adaptive_min_samples
   * whose largest per-channel standard error of the mean (Welford, FP64) is
   * below the threshold; the pixel is that running mean. min is raised to 1
   * and max to min, as AdaptiveSampling::new does. Stratified modes ignore it. */
  uint32_t adaptive_min_samples;
  uint32_t adaptive_max_samples;
  double adaptive_error_threshold;
} gt_camera;

/* RayCamera::new(100, 1.0) defaults (camera.rs:1034-1064). */
void gt_camera_default(gt_camera* camera);
/* floor(width / aspect) clamped to >= 1 (camera.rs:1071). */
uint32_t gt_camera_image_height(const gt_camera* camera);
/* samples actually taken per pixel (camera.rs:1547-1557). */
uint32_t gt_camera_effective_spp(const gt_camera* camera);

/* ---- context ----------------------------------------------------------- */
int gt_context_create(int device, gt_context** out);
int gt_context_destroy(gt_context* ctx);
int gt_context_synchronize(gt_context* ctx);
/* The context's hipStream_t, as an opaque pointer. */
void* gt_context_stream(gt_context* ctx);
/* Device memory on the context's device, for the *_device entry points. */
int gt_device_malloc(gt_context* ctx, size_t bytes, void** out);
int gt_device_free(gt_context* ctx, void* ptr);
int gt_memcpy_to_host(gt_context* ctx, void* dst, const void* src, size_t bytes);
int gt_memcpy_to_device(gt_context* ctx, void* dst, const void* src, size_t bytes);
/* Thread-local message of the last failure ("" when none). */
const char* gt_last_error(void);
const char* gt_version(void);

/* Tunables picked by the autotuner (persistent grid, queue size, BVH width). */
typedef struct gt_tuning {
  uint32_t block_size;      /* threads per block of the wavefront kernels: 64..1024 */
  uint32_t blocks_per_cu;   /* persistent blocks per CU; 0 = occupancy maximum */
  uint64_t max_paths;       /* wavefront queue capacity in paths; 0 = automatic */
  uint32_t bvh_width;       /* 2 or 4 */
  uint32_t use_graph;       /* replay the render's launch sequence as a hipGraph (default 1): a
                               sequence is captured the second time a render repeats it (same
                               scene, camera, outputs and tuning) and replayed while it stays
                               the same; renders that ask for statistics run eagerly */
  uint32_t count_traversal; /* count node fetches and primitive tests (gt_render_stats) */
  uint32_t stack_lds;       /* traversal-stack entries per lane kept in LDS (4..32; deeper levels
                               spill to a per-lane global slice); 0 = 8 */
} gt_tuning;
int gt_context_get_tuning(gt_context* ctx, gt_tuning* out);
int gt_context_set_tuning(gt_context* ctx, const gt_tuning* tuning);

/* ---- scene (RayScene, src/graphics/raytracing/scene.rs:1102-1578) ------- */
int gt_scene_create(gt_context* ctx, gt_scene** out);
int gt_scene_destroy(gt_scene* scene);
/* RayScene::add_material with RayMaterial::{Lambertian (material.rs:353),
 * DiffuseLight (:940), GgxMicrofacet (:1383), Metal (:1962),
 * Dielectric (:2180)}; colours must be finite. */
int gt_scene_add_lambertian(gt_scene* scene, double r, double g, double b, uint32_t* id);
int gt_scene_add_diffuse_light(gt_scene* scene, double r, double g, double b, uint32_t* id);
int gt_scene_add_ggx(gt_scene* scene, double r, double g, double b, double roughness, uint32_t* id);
int gt_scene_add_metal(gt_scene* scene, double r, double g, double b, double fuzz, uint32_t* id);
int gt_scene_add_dielectric(gt_scene* scene, double refraction_index, uint32_t* id);
/* RayMaterial::Isotropic (material.rs:1048-1141): scatters uniformly over the
 * sphere (SpherePdf), scattering_pdf 1/(4 pi); albedo must be finite. */
int gt_scene_add_isotropic(gt_scene* scene, double r, double g, double b, uint32_t* id);
/* RayMaterial::HenyeyGreenstein (material.rs:1182-1333): the Henyey-Greenstein
 * phase function around the incident direction (HenyeyGreensteinPdf,
 * sampling.rs:91-166); albedo finite, anisotropy finite and in (-1, 1). */
int gt_scene_add_henyey_greenstein(gt_scene* scene, double r, double g, double b, double anisotropy,
                                   uint32_t* id);
/* RayScene::add_{sphere,moving_sphere,triangle,quad}[s] (scene.rs:1237-1286).
 * Layouts per primitive: sphere (cx,cy,cz,r); moving sphere
 * (c0x,c0y,c0z,c1x,c1y,c1z,r); triangle (p0,p1,p2); quad (corner,u,v).
 * material_ids[i] must name an existing material (scene.rs:1158-1161).
 * Primitive ids are assigned in insertion order over all kinds. */
int gt_scene_add_spheres(gt_scene* scene, const double* data, size_t n, const uint32_t* material_ids);
int gt_scene_add_moving_spheres(gt_scene* scene, const double* data, size_t n, const uint32_t* material_ids);
int gt_scene_add_triangles(gt_scene* scene, const double* data, size_t n, const uint32_t* material_ids);
int gt_scene_add_quads(gt_scene* scene, const double* data, size_t n, const uint32_t* material_ids);
/* Triangles with per-vertex shading normals: the triangles of a TriangleMesh
 * (MeshTriangle::with_vertex_normals, mesh.rs:24-151), added to the RayScene
 * (the reference traverses a mesh with its own BVH; closest hits do not
 * depend on the tree). positions: (p0, p1, p2) per triangle; normals: the
 * three vertex normals (n0, n1, n2), 9 values per triangle. The geometric
 * normal orients the face; the shading normal is n0 w + n1 u + n2 v at the
 * hit's barycentrics, normalized and flipped with the face, or the geometric
 * normal when that sum is (near) zero (mesh.rs:105-119, object.rs:470-479).
 * Materials use it where the reference reads HitRecord::shading_normal. */
int gt_scene_add_mesh_triangles(gt_scene* scene, const double* positions, const double* normals, size_t n,
                                const uint32_t* material_ids);
/* RayScene::build_bvh_with_options (scene.rs:1562, bvh.rs:233-288): SAH
 * build with the reference's exact splits, then upload to the GPU.
 * leaf_size 0 is clamped to 1 (bvh.rs:73-75). Any later add invalidates it. */
int gt_scene_build(gt_scene* scene, uint32_t leaf_size);
/* No reference counterpart: the scene compile on the device, for large and
 * changing scenes. Replaces the host SAH build of SurfaceScene::to_ray_scene
 * and FlatBvh::build (scene.rs:1517-1547, bvh.rs:233-288) with a linear BVH
 * (Karras 2012) over 63-bit Morton codes, leaves of at most leaf_size
 * primitives. Primitive records are bit-identical to the host compile's and
 * closest hits are the same (they do not depend on the tree); traversal
 * counts differ because the tree does. A later mutation invalidates it like
 * any build; the next render then builds on the host unless this is called
 * again. */
int gt_scene_build_device(gt_scene* scene, uint32_t leaf_size);

typedef struct gt_scene_info {
  uint64_t primitive_count;
  uint64_t material_count;
  uint64_t bvh_node_count;      /* reference FlatBvh node count (scene.rs:1438) */
  uint64_t bvh_depth;           /* deepest root-to-leaf path, root = 1 */
  uint64_t device_bytes;        /* nodes + primitives + materials on the GPU */
  double bounds_min[3], bounds_max[3];
  double build_ms;              /* device time of gt_scene_build_device; 0 after a host build */
  uint32_t built_on_device;     /* 1 after gt_scene_build_device */
  uint32_t reserved;
} gt_scene_info;
int gt_scene_get_info(const gt_scene* scene, gt_scene_info* out);

/* ---- closest-hit queries (the intersection oracle) ---------------------
 * RayScene::hit (scene.rs:1639-1657) with the hit primitive exposed
 * (scene.rs:849-874). One record per ray; prim_id -1 means no hit.
 */
typedef struct gt_ray {
  float origin[3];
  float time;
  float direction[3];
  float t_max;   /* ray interval is (t_min, t_max) */
} gt_ray;
typedef struct gt_hit {
  float t, u, v;
  int32_t prim_id;
  float normal[3];     /* oriented against the ray (object.rs:127-149) */
  uint32_t front_face;
} gt_hit;
/* BvhTraversalStats (bvh.rs:9-17), summed over the rays. */
typedef struct gt_traversal_stats {
  uint64_t rays;
  uint64_t node_bounds_tests;
  uint64_t node_bounds_hits;
  uint64_t leaf_visits;
  uint64_t primitive_candidates;
  uint64_t leaf_hit_results;
  uint64_t max_stack_depth;
} gt_traversal_stats;
int gt_intersect(gt_context* ctx, const gt_scene* scene, const gt_ray* rays, size_t n,
                 float t_min, gt_hit* out_hits, gt_traversal_stats* out_stats /* nullable */);
/* No counterpart in the reference. gt_intersect for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc. It allocates nothing (no hipMalloc,
 * hipFree or allocating library call) and does not synchronise: rays, hits
 * and stats are device pointers, and the results land on the context's
 * stream (gt_context_synchronize before reading them). Same kernels and
 * results as gt_intersect, bit for bit.
 *   - The scene must already be built (GT_ESTATE otherwise).
 *   - d_stats (nullable) is a device gt_traversal_stats; it is cleared by a
 *     kernel, then summed into.
 *   - Its only scratch, the traversal-stack overflow, is a fixed device array
 *     of 2^21 8-byte entries (16 MiB) per GPU, held until the process ends.
 *     Deeper trees run on fewer blocks; a tree needing more than
 *     2^21 / (CUs * block_size) levels beyond the LDS stack is rejected with
 *     GT_EINVAL (at 256 CUs and 128 threads: 64 levels).
 *   - Every call on a GPU shares that GPU's array: calls that overlap on one
 *     GPU (two contexts or streams at once) race. */
int gt_intersect_device_noalloc(gt_context* ctx, const gt_scene* scene, const gt_ray* d_rays,
                                size_t n, float t_min, gt_hit* d_hits,
                                gt_traversal_stats* d_stats /* nullable, device */);

/* ---- light targets (SamplingTargetList / WeightedSamplingTargetList,
 *      scene.rs:363-609) ------------------------------------------------- */
int gt_lights_create(gt_context* ctx, gt_lights** out);
int gt_lights_destroy(gt_lights* lights);
/* weight must be finite and > 0 (scene.rs:529-537); add_* without weight = 1. */
int gt_lights_add_sphere(gt_lights* lights, const double center[3], double radius, double weight);
int gt_lights_add_moving_sphere(gt_lights* lights, const double c0[3], const double c1[3],
                                double radius, double weight);
int gt_lights_add_quad(gt_lights* lights, const double corner[3], const double u[3],
                       const double v[3], double weight);
int gt_lights_add_triangle(gt_lights* lights, const double p0[3], const double p1[3],
                           const double p2[3], double weight);
size_t gt_lights_len(const gt_lights* lights);
/* RayScene::emissive_targets (scene.rs:1421-1423): every primitive whose
 * material is emissive, weight 1, in primitive order. */
int gt_scene_emissive_targets(const gt_scene* scene, gt_lights** out);

/* ---- environment light (EnvironmentLight, environment.rs:22-195) -------
 * gt_env_create_constant mirrors EnvironmentLight::constant (:64-77).
 * gt_env_create_rgb32f is the FP32 HDR lat-long form: linear RGB texels,
 * row 0 at theta = 0 (+y), same lookup, CDF and PDF as the 8-bit form. */
int gt_env_create_constant(gt_context* ctx, double r, double g, double b, gt_env** out);
int gt_env_create_rgb32f(gt_context* ctx, const float* rgb, uint32_t width, uint32_t height, gt_env** out);
int gt_env_destroy(gt_env* env);

/* ---- rendering --------------------------------------------------------
 * gt_render_hdr: RayCamera::render_world*_hdr_image (camera.rs:4246-4336)
 * with the camera's strategy as given (the PathTracer layer forces NEE for
 * environment and light-connection renders, camera.rs:4167-4173).
 * lights and env may be NULL. out_rgb: W*H*3 floats, row-major, top row
 * first (HdrImage::from_pixels, display.rs:295).
 */
typedef struct gt_render_stats {
  uint64_t camera_paths;      /* paths started (= W*H*spp) */
  uint64_t extension_rays;    /* closest-hit rays traced, primary included */
  uint64_t shadow_rays;       /* NEE rays traced */
  uint64_t accepted_samples;  /* finite samples (camera.rs:2984-2991) */
  double device_ms;           /* device time of the render, HIP events */
  double extend_ms;           /* summed duration of the extend kernel launches */
  uint64_t extend_launches;
  double shade_ms, connect_ms;
  /* with gt_tuning.count_traversal: BVH nodes fetched and primitives tested */
  uint64_t ext_node_fetches, ext_prim_tests;
  uint64_t shadow_node_fetches, shadow_prim_tests;
} gt_render_stats;
int gt_render_hdr(gt_context* ctx, const gt_scene* scene, const gt_camera* camera,
                  const gt_lights* lights, const gt_env* env, float* out_rgb,
                  gt_render_stats* out_stats /* nullable */);
/* Same render; result stays on the device (float W*H*3, or double when
 * out_is_double != 0, the FP64-accumulated form used for references). */
int gt_render_hdr_device(gt_context* ctx, const gt_scene* scene, const gt_camera* camera,
                         const gt_lights* lights, const gt_env* env, void* d_out,
                         int out_is_double, gt_render_stats* out_stats);
/* RayCamera::render_world*_denoising_aovs (camera.rs:4175-4244): beauty as
 * gt_render_hdr, albedo = first-hit denoise_albedo, normal = 0.5*(n+1),
 * both 0 on a miss, averaged over the same samples (camera.rs:3064-3116). */
int gt_render_aovs(gt_context* ctx, const gt_scene* scene, const gt_camera* camera,
                   const gt_lights* lights, const gt_env* env,
                   float* out_beauty, float* out_albedo, float* out_normal,
                   gt_render_stats* out_stats);
int gt_render_aovs_device(gt_context* ctx, const gt_scene* scene, const gt_camera* camera,
                          const gt_lights* lights, const gt_env* env,
                          float* d_beauty, float* d_albedo, float* d_normal,
                          gt_render_stats* out_stats);

/* ---- rendering a subset of the pixels ----------------------------------
 * No direct reference counterpart: the reference renders whole images or
 * tiles (render_tiles, camera.rs:3272-3294). Renders only the n listed
 * pixels (indices y * W + x, each < W*H, none listed twice: GT_EINVAL
 * otherwise) with the same per-(pixel, sample, bounce) streams, so a listed
 * pixel's value equals the same pixel of gt_render_hdr_device bit for bit.
 * Used for converged references on crops. pixels is a host array; d_out is
 * the device image (W*H*3 float, or double when out_is_double) of which only
 * the listed pixels are written; the result is on the context's stream.
 * stats->camera_paths counts n * spp. An adaptive camera samples the listed
 * pixels adaptively. */
int gt_render_pixels_device(gt_context* ctx, const gt_scene* scene, const gt_camera* camera,
                            const gt_lights* lights, const gt_env* env, const uint32_t* pixels,
                            size_t n, void* d_out, int out_is_double,
                            gt_render_stats* out_stats /* nullable */);

/* ---- progressive rendering ----------------------------------------------
 * RayCamera::render_world*_tiled_progressive (camera.rs:3172-3294,
 * 3468-3500, 4084-4165, 4414-4437). The image is split into tiles of
 * tile_size pixels in rows of tiles, top to bottom, left to right
 * (render_tiles :3272-3294). After each tile is final, progress(user,
 * &update) runs on the calling thread with the display canvas so far
 * (Rgb::from_linear_color per pixel; pixels of tiles not yet reported are
 * 0). The GPU renders bands of whole tile rows at a time, so updates come in
 * bursts, in tile order. A nonzero return from progress stops the render:
 * the call returns GT_ECANCELLED and out_canvas holds the tiles reported so
 * far. The finished canvas equals the display image of gt_render_hdr. */
enum { GT_ECANCELLED = 6 };
typedef struct gt_render_tile {  /* RenderTile (camera.rs:398-409) */
  uint32_t x, y, width, height;
} gt_render_tile;
typedef struct gt_render_progress {  /* ProgressiveRenderUpdate (camera.rs:430-490) */
  gt_render_tile tile;       /* the tile that just completed */
  uint64_t completed_tiles;  /* including tile */
  uint64_t total_tiles;
  uint32_t image_width, image_height;
  const uint8_t* pixels;     /* image_width * image_height * 3 bytes, row-major */
} gt_render_progress;
typedef int (*gt_progress_fn)(void* user, const gt_render_progress* update);
/* tile_size 0 = 16 (DEFAULT_RENDER_TILE_SIZE, camera.rs:34); out_canvas is a
 * host buffer of W*H*3 bytes. */
int gt_render_progressive(gt_context* ctx, const gt_scene* scene, const gt_camera* camera,
                          const gt_lights* lights, const gt_env* env, uint32_t tile_size,
                          gt_progress_fn progress, void* user, uint8_t* out_canvas);

/* ---- batched rendering: many scenes and cameras in one wavefront ------
 * No reference counterpart: the reference renders one scene per call, and a
 * caller loops over PathTracer::render_* (renderer.rs:101-224,374-609). Here
 * the samples of all n frames share each launch: every batch of paths mixes
 * frames, and each kernel looks up its path's frame. Each result equals
 * gt_render_hdr of that frame alone (bit for bit: same streams, same
 * accumulation order). lights and envs are arrays of n entries (entries and
 * the arrays themselves may be NULL). out_rgb[i] is a host buffer of
 * W_i*H_i*3 floats; the _device form takes device pointers and leaves the
 * results on the context's stream. */
int gt_render_batch(gt_context* ctx, size_t n, const gt_scene* const* scenes,
                    const gt_camera* cameras, const gt_lights* const* lights,
                    const gt_env* const* envs, float* const* out_rgb,
                    gt_render_stats* out_stats);
int gt_render_batch_device(gt_context* ctx, size_t n, const gt_scene* const* scenes,
                           const gt_camera* cameras, const gt_lights* const* lights,
                           const gt_env* const* envs, void* const* d_out, int out_is_double,
                           gt_render_stats* out_stats);

/* ---- autotuner ---------------------------------------------------------
 * No reference counterpart (the reference has no device tunables). Times the
 * given frames, without output, at their own spp (or at sample_spp random
 * samples when nonzero), under candidate tunings. Stage 1 tries threads per
 * block 64/128/256 x BVH width 2/4 x LDS stack 8/16 at a 2^25-path queue;
 * stage 2 tries queue capacities 2^23/2^24/2^26 for the winner (capacities
 * above twice the workload are skipped). The fastest becomes the context's
 * tuning. Results do not depend on the tuning, only speed does.
 * out_best and out_best_ms may be NULL. */
int gt_autotune(gt_context* ctx, size_t n, const gt_scene* const* scenes,
                const gt_camera* cameras, const gt_lights* const* lights,
                const gt_env* const* envs, uint32_t sample_spp,
                gt_tuning* out_best, double* out_best_ms);

/* ---- output boundary: HdrImage / ToneMap encodings (display.rs:164-468) --
 * Per-pixel conversions of a linear RGB image (width*height*3 values,
 * row-major, top row first) into the formats the reference writes, computed
 * on the device. File headers are text and come from gt_encoded_header. */
enum {
  GT_TONE_NONE = 0,     /* ToneMappingOperator::None: clip */
  GT_TONE_REINHARD = 1, /* x / (1 + x) */
  GT_TONE_ACES = 2      /* ACES filmic fit, clamped to [0, 1] */
};
/* ToneMap (display.rs:191-208; builders :224-258 panic where these fail
 * with GT_EINVAL): exposure finite and >= 0, gamma finite and > 0. */
typedef struct gt_tone_map {
  double exposure; /* default 1 */
  double gamma;    /* default 2 (sqrt, Rgb::linear_to_gamma_component) */
  uint32_t op;     /* GT_TONE_*, default GT_TONE_NONE */
  uint32_t reserved;
} gt_tone_map;
enum {
  /* HdrImage::to_canvas_tone_mapped (display.rs:330-344): ToneMap::map_color
   * (:264-277) per channel; w*h*3 bytes, top row first. */
  GT_ENCODE_CANVAS = 0,
  /* Rgb::from_raw_linear_color (colors.rs:412-422), the albedo and normal
   * canvases of DenoisingAovs (camera.rs:4212-4243); w*h*3 bytes. */
  GT_ENCODE_RAW = 1,
  /* encode_rgbe (display.rs:444-468) per pixel: write_radiance_hdr's body
   * (:397-409); w*h*4 bytes, top row first. */
  GT_ENCODE_RGBE = 2,
  /* write_pfm's body (display.rs:365-387): w*h*3 little-endian floats,
   * bottom row first, non-finite values as 0. */
  GT_ENCODE_PFM = 3
};
/* Size in bytes of the encoded body; 0 for an unknown format. */
uint64_t gt_encoded_size(int format, uint32_t width, uint32_t height);
/* The file header the reference writes before the body: "PF\n<w> <h>\n-1.0\n"
 * (display.rs:366-369) for GT_ENCODE_PFM, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe
 * \n\n-Y <h> +X <w>\n" (:398-401) for GT_ENCODE_RGBE, "" for the canvases.
 * Writes at most cap bytes (no terminator) and the full length to *out_len. */
int gt_encoded_header(int format, uint32_t width, uint32_t height, char* buf, size_t cap,
                      size_t* out_len);
/* Encodes src (float, or double when src_is_double; host memory, or device
 * memory when src_on_device) into dst (gt_encoded_size bytes; host, or
 * device when dst_on_device). tone_map is read by GT_ENCODE_CANVAS only
 * (NULL = defaults). Returns when dst holds the result. */
int gt_encode_image(gt_context* ctx, const void* src, int src_is_double, int src_on_device,
                    uint32_t width, uint32_t height, int format, const gt_tone_map* tone_map,
                    void* dst, int dst_on_device);

#ifdef __cplusplus
}
#endif
#endif /* GARTUS_HIP_H */
