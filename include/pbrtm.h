/* pbrtm — MI355X (gfx950) HLBVH construction and traversal behind a C ABI.
 *
 * Every function returns 0 (PBRTM_OK) on success or a negative PBRTM_ERR_*
 * code; the message is then available from pbrtm_last_error(ctx). No C++
 * exception crosses this boundary and nothing calls exit(): the reference's
 * CHECK_CUDA_ERROR / REPORT_FATAL_ERROR abort paths (gpu/macro.h:19-51)
 * become error codes here.
 *
 * Pointers are host pointers unless PBRTM_DEVICE_PTRS is set in `flags`, in
 * which case they are device pointers on the context's device; every call is
 * complete (stream synchronised) when it returns. A context is used by one
 * host thread at a time and owns one HIP stream.
 *
 * Data formats
 *   triangles  [n][3][3] float: three vertices per triangle (an unindexed soup),
 *              or vertices [nv][3] float + indices [n][3] uint32 (a TriangleMesh,
 *              shapes/triangle_mesh.cu:39-106, vertices already in render space).
 *   rays       [n][8] float: o.x o.y o.z t_max d.x d.y d.z unused. t_max is the
 *              reference's intersect(ray, t_max) argument; a ray with
 *              !(t_max > 0) is inactive and reports a miss.
 *   hits       prim id uint32 (0xFFFFFFFF = miss), t float, barycentrics
 *              (b0, b1, b2) float — the reference's ShapeIntersection::t_hit and
 *              TriangleIntersection (shapes/triangle.h:15-22).
 */
#ifndef PBRTM_H
#define PBRTM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBRTM_OK 0
#define PBRTM_ERR_INVALID (-1)   /* bad argument */
#define PBRTM_ERR_HIP (-2)       /* HIP runtime error */
#define PBRTM_ERR_DEVICE (-3)    /* a kernel reported a condition (see message) */
#define PBRTM_ERR_NOMEM (-4)
#define PBRTM_ERR_UNSUPPORTED (-5)

#define PBRTM_DEVICE_PTRS 1u
/* pbrtm_bvh_intersect: traverse the binary build tree in the reference's node
 * order instead of the 8-wide layout (same hits; visit counts comparable with
 * the reference's). */
#define PBRTM_TRACE_BINARY 2u

#define PBRTM_NO_HIT 0xFFFFFFFFu

typedef struct pbrtm_ctx pbrtm_ctx;
typedef struct pbrtm_bvh pbrtm_bvh;

/* Per-phase build timings (HIP events on the context stream) and shape. */
typedef struct pbrtm_build_stats {
    double ms_total;      /* host wall clock of the whole call */
    double ms_sort;       /* centroid bounds + Morton codes + radix sort */
    double ms_treelets;   /* sorted records, treelet emission */
    double ms_top;        /* SAH over treelets */
    double ms_bottom;     /* midpoint splits inside treelets */
    double ms_layout;     /* traversal layout (leaf-ordered triangles, 8-wide collapse) */
    uint64_t num_prims;
    uint64_t num_treelets;
    uint64_t top_levels;      /* level-synchronous steps for nodes > 64 treelets */
    uint64_t bottom_levels;   /* level-synchronous steps for nodes > 64 primitives */
    uint64_t small_top;       /* wave-finished top subtrees */
    uint64_t small_bottom;    /* wave-finished bottom subtrees */
    uint64_t wide_nodes;      /* 8-wide traversal nodes */
    uint64_t wide_depth;      /* wide levels on the longest path */
} pbrtm_build_stats;

/* No counterpart in the reference (its objects live in one process-wide CUDA
 * context, main.cu:124-146). device >= 0 selects a HIP device. */
int pbrtm_create(pbrtm_ctx **out, int device);
void pbrtm_destroy(pbrtm_ctx *ctx);
const char *pbrtm_last_error(const pbrtm_ctx *ctx);
int pbrtm_device_count(int *count);
/* The HIP stream the context launches on (a hipStream_t), for callers that
 * time or order work around it. */
void *pbrtm_stream(pbrtm_ctx *ctx);

/* Morton codes and the canonical order. Replaces hlbvh.cu:101-141 (codes) and
 * the treelet counting sort hlbvh.cu:457-502: codes_out[i] is the 30-bit code
 * of the i-th triangle in sorted order and order_out[i] its input index; the
 * order sorts by full code with ties by input index, which refines the
 * reference's treelet-only order. */
int pbrtm_morton_sort(pbrtm_ctx *ctx, const float *tri, uint64_t n, uint32_t flags, uint32_t *codes_out,
                      uint32_t *order_out);

/* Builds the aggregate. Replaces `static const HLBVH *HLBVH::create(const
 * std::vector<const Primitive *> &, GPUMemoryAllocator &)` (accelerator/hlbvh.h:76-77,
 * hlbvh.cu:171-177, build_bvh :384-624). idx == NULL means `verts` is a soup of
 * 3*nt vertices. With PBRTM_DEVICE_PTRS and idx == NULL the handle keeps
 * `verts` by pointer (as HLBVH keeps its primitives, hlbvh.h:118-121): it must
 * outlive the handle. n == 0 builds an empty aggregate (hlbvh.cu:392-395). */
int pbrtm_bvh_build(pbrtm_ctx *ctx, const float *verts, uint64_t nv, const uint32_t *idx, uint64_t nt,
                    uint32_t flags, pbrtm_bvh **out, pbrtm_build_stats *stats);
/* The context keeps the freed aggregate's device buffers (one set, released by
 * pbrtm_destroy) and hands them to its next build: building, using and freeing
 * in a loop allocates no device memory after the first pass. */
void pbrtm_bvh_free(pbrtm_bvh *bvh);

/* Bounds of the whole aggregate: lo.xyz, hi.xyz. Replaces HLBVH::bounds()
 * (hlbvh.h:79-86); an empty aggregate gives the empty box. */
int pbrtm_bvh_bounds(const pbrtm_bvh *bvh, float out6[6]);

/* The tree as a canonical byte string, independent of node numbering and of
 * primitive order inside leaves: preorder, left (= "< split" side) first; per
 * node u32 tag (1 = leaf, else axis << 8), six u32 bound bit patterns (-0.0
 * written as +0.0), u32 primitive count, and for a leaf its primitive ids
 * ascending. buf == NULL returns the size in *len. No reference counterpart:
 * it makes the reference's atomic-ordered tree (hlbvh.cu:65, :167, :796)
 * comparable. */
int pbrtm_bvh_export_canonical(const pbrtm_bvh *bvh, uint8_t *buf, uint64_t *len);

/* Batched ray queries. any_hit = 0 replaces HLBVH::intersect (closest hit,
 * hlbvh.cu:231-289); any_hit = 1 replaces HLBVH::fast_intersect (hlbvh.cu:179-229),
 * whose bool is prim_out != PBRTM_NO_HIT. t_out and bary_out ([n][3]) may be NULL.
 * The 8-wide traversal tests triangles in exactly the reference's order, so hits
 * (ties included) equal the binary traversal's. stats_out (host, may be NULL)
 * receives [nodes visited, box tests, triangle tests] summed over the rays; it
 * selects a counting instance of the same traversal. */
int pbrtm_bvh_intersect(pbrtm_bvh *bvh, const float *rays, uint64_t n, int any_hit, uint32_t flags,
                        uint32_t *prim_out, float *t_out, float *bary_out, uint64_t *stats_out);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pbrtm_bvh_intersect with PBRTM_DEVICE_PTRS (8-wide
 * traversal, no statistics), but the ray queue, the device error word and the
 * traversal's stack spill live in a fixed library scratch array (32 MiB + 256 B
 * per device, held until the process ends) instead of buffers it allocates.
 * rays, prim_out, t_out and bary_out are device pointers (t_out, bary_out
 * nullable); results equal pbrtm_bvh_intersect's bit for bit. At most 2^22 rays
 * per call (PBRTM_ERR_INVALID beyond), and the tree's stack spill for n rays must
 * fit 16 MiB (trees deeper than 12 wide levels; PBRTM_ERR_INVALID otherwise).
 * Calls that overlap on one device share the scratch and race: serialise them. */
int pbrtm_bvh_intersect_noalloc(pbrtm_bvh *bvh, const float *rays, uint64_t n, int any_hit, uint32_t *prim_out,
                                float *t_out, float *bary_out);

/* From closest hits (prim, barycentrics), one cosine-weighted diffuse bounce
 * per ray, as integrators/ambient_occlusion.cu:36-52 spawns it: u = the first
 * two PCG32 draws of sequence = ray index, frame around the face-forwarded
 * geometric normal, origin offset past the hit's error bounds (base/ray.h:22-43).
 * Misses and zero-pdf samples give inactive rays. */
int pbrtm_ao_bounce(pbrtm_bvh *bvh, const float *rays, const uint32_t *prim, const float *bary, uint64_t n,
                    uint32_t flags, float *rays_out);

/* Workload generators (host, deterministic, thread-count independent).
 * Triangle soup: centroid c ~ U[0,1)^3, each vertex c + l*u with
 * l ~ LogNormal(mu, sigma) (Box-Muller) and u uniform on the sphere; triangle i
 * uses draws 16i.. of the PCG32 sequence `seed`. */
int pbrtm_gen_soup(uint64_t seed, uint64_t n, float mu, float sigma, int threads, float *tri_out);
/* One pinhole ray per pixel centre, row-major from the top-left pixel,
 * vertical field of view fov_deg, t_max = +inf. */
int pbrtm_gen_camera_rays(uint32_t width, uint32_t height, const float eye[3], const float look[3],
                          const float up[3], float fov_deg, float *rays_out);

/* ------------------------------------------------------------------------
 * Wavefront path tracer. A scene is the flat result of the reference's scene
 * front end (scene/scene_builder.cu): object-space triangles with their
 * object-to-world transform, a material and an optional diffuse area light per triangle, a LookAt + perspective
 * camera and the film resolution. Render space is camera-world
 * (base/camera.h:26-45); the Mitchell filter (radius 2, B = C = 1/3), the
 * stratified sampler and the sRGB RGB film are the reference's defaults
 * (scene_builder.cu:210-218, :275-290; films/rgb_film.cu:19-60).
 * ------------------------------------------------------------------------ */
typedef struct pbrtm_scene pbrtm_scene;

#define PBRTM_MATERIAL_DIFFUSE 0u    /* materials/diffuse_material.cu */
#define PBRTM_MATERIAL_CONDUCTOR 1u  /* materials/conductor_material.cu */
#define PBRTM_MATERIAL_DIELECTRIC 2u /* materials/dielectric_material.cu */
#define PBRTM_METAL_AG 0u           /* named spectra "metal-Ag-eta"/"metal-Ag-k" (scene_builder.cu:137-168) */
#define PBRTM_METAL_AL 1u
#define PBRTM_METAL_AU 2u
#define PBRTM_METAL_CU 3u

typedef struct pbrtm_material {
    uint32_t type;            /* PBRTM_MATERIAL_* */
    float rgb[3];             /* diffuse "rgb reflectance" (each <= 1) */
    uint32_t metal;           /* conductor eta/k: PBRTM_METAL_* */
    float uroughness, vroughness; /* conductor and dielectric */
    uint32_t remap_roughness; /* "bool remaproughness" (default true) */
    float eta;                /* dielectric "float eta" (default 1.5) */
    uint32_t eta_named;       /* dielectric: 0 = the constant eta, 1 = "spectrum eta" "glass-BK7" */
} pbrtm_material;

typedef struct pbrtm_area_light { /* AreaLightSource "diffuse" (lights/diffuse_area_light.cu:9-35) */
    float rgb[3];                 /* "rgb L" */
    float scale;                  /* "float scale" */
    uint32_t two_sided;           /* "bool twosided" */
} pbrtm_area_light;

typedef struct pbrtm_infinite_light { /* LightSource "infinite" without "filename"
                                         (lights/uniform_infinite_light.cu:4-43) */
    float rgb[3];                 /* "rgb L" (an RGBIlluminantSpectrum) */
    uint32_t has_rgb;             /* 0: no "L", the colour space illuminant itself */
    float scale;                  /* "float scale" (default 1) */
    float illuminance;            /* "float illuminance" (<= 0: not given) */
    uint32_t light_index;         /* its place in the scene's light list: the number of
                                     emissive triangles created before it (the reference keeps
                                     LightSource and area lights in file order) */
} pbrtm_infinite_light;

#define PBRTM_SHAPE_SPHERE 1u   /* Shape "sphere" (shapes/sphere.cu) */
#define PBRTM_SHAPE_DISK 2u     /* Shape "disk" (shapes/disk.cu) */
#define PBRTM_SHAPE_INSTANCE 3u /* a triangle of an ObjectInstance: TransformedPrimitive
                                   (primitives/transformed_primitive.cu, scene_builder.cu:775-810) */

typedef struct pbrtm_shape {
    uint32_t type;        /* PBRTM_SHAPE_* */
    uint32_t xform;       /* sphere, disk: the CTM at the Shape; instance: the CTM at the ObjectInstance.
                             An index into xforms / xforms_inv, UINT32_MAX = identity */
    float radius;         /* "float radius" (default 1) */
    float zmin, zmax;     /* sphere "zmin" / "zmax" (defaults -radius / radius) */
    float phimax;         /* "phimax" in degrees (default 360) */
    float height;         /* disk "height" (default 0) */
    float innerradius;    /* disk "innerradius" (default 0) */
} pbrtm_shape;

typedef struct pbrtm_scene_desc {
    float eye[3], look[3], up[3]; /* LookAt */
    float fov;                    /* Camera "perspective" "float fov" (degrees) */
    uint32_t width, height;       /* Film "integer xresolution" / "yresolution" */
    uint64_t num_tris;
    const float *tri;             /* [n][3][3] world space (host) */
    const uint32_t *tri_material; /* [n] index into materials */
    const int32_t *tri_light;     /* [n] index into lights, -1 = not emissive; may be NULL */
    const uint8_t *tri_reverse;   /* [n] ReverseOrientation; may be NULL */
    uint32_t num_materials;
    const pbrtm_material *materials;
    uint32_t num_lights;
    const pbrtm_area_light *lights;
    uint32_t num_xforms;          /* object-to-world matrices (the CTM at each Shape) */
    const float *xforms;          /* [num_xforms][4][4] row-major; may be NULL */
    const uint32_t *tri_xform;    /* [n] index into xforms, UINT32_MAX = identity; NULL = all identity */
    uint32_t num_infinite_lights; /* at most 4 */
    const pbrtm_infinite_light *infinite_lights;
    /* Analytic shapes and instancing. The primitive list stays the triangle list above,
     * in the front end's order (which orders the aggregate and the lights); entry i is
     * shapes[tri_shape[i]] when tri_shape[i] != UINT32_MAX: a sphere or a disk (its
     * vertices are ignored; its material, light and ReverseOrientation are entry i's),
     * or a triangle of an object instance (its vertices, placed by tri_xform[i] as the
     * definition's CTM placed them, are the instanced triangle; instances cannot emit). */
    const float *xforms_inv;      /* [num_xforms][4][4]: each CTM's inverse as the front end composed it
                                     (Transform keeps both); NULL = inverted here */
    uint32_t num_shapes;
    const pbrtm_shape *shapes;
    const uint32_t *tri_shape;    /* [n]; NULL = every entry is a triangle */
} pbrtm_scene_desc;

typedef struct pbrtm_render_params {
    uint32_t spp;       /* samples per pixel, a square (stratified) */
    uint32_t max_depth; /* Integrator "path" "integer maxdepth" (default 5) */
    uint32_t pool_size; /* path slots in flight; 0 = 2^21 as the reference (wavefront_path.cu:18) for
                           the wavefront integrator, whose film order depends on it; for the megakernel
                           integrators (film in sample order) 2^22 (megakernelpath) or 2^23 slots,
                           at most the render's path count */
    uint32_t flags;     /* PBRTM_DEVICE_PTRS: film_rgb / film_aux are device pointers;
                           one PBRTM_INTEGRATOR_* value (default: the wavefront path integrator) */
} pbrtm_render_params;

/* pbrtm_render_params.flags: the integrator (SceneBuilder::build_integrator,
 * scene_builder.cu:289-313; MegakernelIntegrator::create, megakernel_integrator.cu:82-111).
 * The megakernel integrators add each pixel's samples to the film in sample order
 * (megakernel_integrator.cu:14-42); here they run on the same path pool. Only
 * pbrtm_render takes them; pbrtm_render_grad and pbrtm_render_batch render with
 * the wavefront path integrator. */
/* pbrtm_render_params.flags: trace with the counting instance of the traversal and
 * report its totals in pbrtm_render_stats (same films and ray counts; for byte
 * accounting, not for timing). */
#define PBRTM_RENDER_COUNT 0x100u
#define PBRTM_INTEGRATOR_MASK 0xF0u
#define PBRTM_INTEGRATOR_PATH 0x00u              /* "path", "volpath": WavefrontPathIntegrator */
#define PBRTM_INTEGRATOR_MEGAKERNEL_PATH 0x10u   /* "megakernelpath": MegakernelPathIntegrator
                                                    (integrators/megakernel_path.cu:35-213) */
#define PBRTM_INTEGRATOR_AMBIENT_OCCLUSION 0x20u /* "ambientocclusion": AmbientOcclusionIntegrator
                                                    (integrators/ambient_occlusion.cu:9-58) */
#define PBRTM_INTEGRATOR_SURFACE_NORMAL 0x30u    /* "surfacenormal": SurfaceNormalIntegrator
                                                    (integrators/surface_normal.cu:8-32) */

/* The PBRTM_INTEGRATOR_* value of an integrator name as the reference's
 * --integrator option and Integrator directive take it; PBRTM_ERR_UNSUPPORTED
 * for names the library does not render (bdpt, mlt, ...). */
int pbrtm_integrator_from_name(const char *name, uint32_t *flags_out);

typedef struct pbrtm_render_stats {
    double ms_total, ms_trace, ms_shadow, ms_shade, ms_film;
    uint64_t rounds, rays, shadow_rays, paths;
    uint32_t ray_order;  /* the ray-order autotuner's choice: bit 0 closest-hit rays sorted, bit 1 shadow rays */
    uint32_t grid_quarters;  /* the occupancy autotuner's choice: traversal grid in quarters of the maximum */
    /* with PBRTM_RENDER_COUNT: [wide nodes visited, child box tests, primitive tests]
       summed over the closest-hit and the shadow traversals (else zero) */
    uint64_t closest_counts[3], shadow_counts[3];
} pbrtm_render_stats;

/* Builds the scene tables and its aggregate on the context's device. Replaces
 * SceneBuilder::preprocess + build_integrator's set-up (scene_builder.cu:267-314,
 * :829-870): render-space geometry, spectra, camera, filter, lights and the
 * power light sampler (light_samplers/power_light_sampler.cu:7-45). */
int pbrtm_scene_create(pbrtm_ctx *ctx, const pbrtm_scene_desc *desc, pbrtm_scene **out);
void pbrtm_scene_free(pbrtm_scene *scene);

/* Renders with the wavefront path integrator. Replaces
 * WavefrontPathIntegrator::create + render (integrators/wavefront_path.h:87-92,
 * wavefront_path.cu:541-572, :626-707) and RGBFilm::get_pixel_rgb
 * (films/rgb_film.cu:129-144): film_rgb [h][w][3] is the output-space RGB of
 * every pixel; film_aux [h][w][4] (may be NULL) the sensor-RGB sum and the
 * filter-weight sum. Paths, pixels and samples are the reference's: the same
 * (pixel, sample) gives the same sampler stream. */
int pbrtm_render(pbrtm_ctx *ctx, pbrtm_scene *scene, const pbrtm_render_params *params, float *film_rgb,
                 float *film_aux, pbrtm_render_stats *stats);

/* Differentiable rendering (no counterpart in the reference): the loss
 * sum over pixels and channels of (film - target)^2 of pbrtm_render's film and
 * its gradient by path replay. The forward render keeps every path's final
 * spectral radiance; a second render of the same paths (same sampler streams)
 * accumulates dLoss/d(diffuse albedo RGB) per material (through the
 * RGB-to-spectrum table) and dLoss/d(scale) per emitter (pbrtm_area_light).
 * Each sample's estimator (fixed sampler streams) is differentiated path by
 * path, the light pmf attached wherever it appears (MIS weights, the light
 * choice's 1/pmf) and every discrete choice held (which light, Russian
 * roulette), so the result is the derivative of the fixed-seed loss between
 * the parameter values where some sample's light choice flips. d_albedo [num_materials][3] (zero for non-diffuse materials) and
 * d_scale [num_lights] may be NULL; film_rgb (may be NULL) receives the film.
 * Host pointers; at most 255 emitters. */
typedef struct pbrtm_grad_stats {
    double ms_forward, ms_adjoint, ms_replay, ms_total;
    uint64_t rounds, rays, shadow_rays, paths;
} pbrtm_grad_stats;

int pbrtm_render_grad(pbrtm_ctx *ctx, pbrtm_scene *scene, const pbrtm_render_params *params, const float *target_rgb,
                      float *film_rgb, double *loss, float *d_albedo, float *d_scale, pbrtm_grad_stats *stats);

/* Batched rendering (no counterpart in the reference, which renders one scene
 * per process): every job's paths share one pool of path slots, so many small
 * renders fill the device together. Each job's film is bit-identical to
 * pbrtm_render of the same scene and parameters with pool_size = the job's
 * pool_size (0 = 2^21): finished paths keep their length in rounds, and a
 * per-job pass recovers the round each path would have finished in that
 * one-by-one schedule, which orders the film's per-pixel sums. */
typedef struct pbrtm_job {
    pbrtm_scene *scene;            /* scenes may repeat across jobs */
    pbrtm_render_params params;    /* spp, max_depth, pool_size as pbrtm_render; flags unused */
    float *film_rgb;               /* [h][w][3], required */
    float *film_aux;               /* [h][w][4] or NULL */
    uint32_t rounds;               /* out: rounds of the job's one-by-one render */
    uint32_t reserved;
} pbrtm_job;

/* pbrtm_batch_params.flags */
#define PBRTM_BATCH_NO_GRAPH 2u    /* enqueue each round's kernels instead of replaying a captured graph */
/* Autotune the shared pool size (queue capacity: 2^23, 2^24, 2^25 slots) and the
 * persistent traversal grid (4, 3, 2 quarters of the resident workgroups) from
 * measured time, across calls: the context remembers each batch (the same
 * scenes, spp, depths and pool sizes, in order); its first calls each try one
 * candidate (the pool sizes at the full grid, then the best pool at smaller
 * grids) and later calls use the fastest. pool_size is then ignored. Films do not
 * depend on either choice. */
#define PBRTM_BATCH_AUTOTUNE 4u

typedef struct pbrtm_batch_params {
    uint32_t pool_size;           /* slots shared by the batch; 0 = 2^24 (capped at the batch's path count) */
    uint32_t rounds_per_launch;   /* rounds per graph launch between done checks; 0 = 8 */
    uint32_t flags;               /* PBRTM_DEVICE_PTRS (films are device pointers), PBRTM_BATCH_NO_GRAPH */
    uint32_t solo_paths;          /* jobs with at least this many paths render alone, before the shared
                                     pool (they fill the GPU by themselves); 0 or 0xFFFFFFFF = never (the
                                     default: one shared pool measured fastest on mixed jobs) */
} pbrtm_batch_params;

typedef struct pbrtm_batch_stats {
    double ms_total, ms_rounds, ms_schedule, ms_film;  /* ms_rounds .. ms_film: the shared pool's */
    uint64_t rounds, rays, shadow_rays, paths;
    uint32_t pool_size, launches;
    uint32_t solo_jobs;
    uint32_t grid_quarters;  /* the shared pool's traversal grid, in quarters of the resident maximum */
    uint32_t tune_state;     /* PBRTM_BATCH_AUTOTUNE: 1 = this call tried a candidate, 2 = the tuned choice */
    uint32_t reserved;
} pbrtm_batch_stats;

int pbrtm_render_batch(pbrtm_ctx *ctx, pbrtm_job *jobs, uint32_t njobs, const pbrtm_batch_params *params,
                       pbrtm_batch_stats *stats);

/* Spectral set-up exposed for tests: the sigmoid-polynomial coefficients the
 * reference's RGB-to-spectrum table gives an RGB (spectrum_util/rgb_to_spectrum_data.cu:361-400),
 * and the sRGB colour space matrices and white point (rgb_color_space.h:12-44). */
int pbrtm_rgb_to_sigmoid(const float rgb[3], float out3[3]);
int pbrtm_srgb_matrices(float xyz_from_rgb9[9], float rgb_from_xyz9[9], float white_xy2[2]);

/* The .pbrt front end (host only). Replaces SceneBuilder's parse
 * (scene/scene_builder.cu:316-720) for the directives the path tracer renders;
 * anything else that would change the image is an error (message in err).
 * The handle owns the description's arrays; pbrtm_pbrt_render_params gives the
 * file's samples per pixel (rounded to a square, scene_builder.cu:283-287) and
 * maxdepth, pbrtm_pbrt_output the Film filename. */
typedef struct pbrtm_pbrt_file pbrtm_pbrt_file;
int pbrtm_pbrt_parse(const char *path, pbrtm_pbrt_file **out, char *err, uint64_t errlen);
const pbrtm_scene_desc *pbrtm_pbrt_desc(const pbrtm_pbrt_file *file);
int pbrtm_pbrt_render_params(const pbrtm_pbrt_file *file, pbrtm_render_params *out);
const char *pbrtm_pbrt_output(const pbrtm_pbrt_file *file);
void pbrtm_pbrt_free(pbrtm_pbrt_file *file);

/* Image output. Replaces Film::write_to_png (base/film.cu:142-181): each linear
 * output-space RGB value goes through SRGBColorEncoding::from_linear
 * (spectrum_util/color_encoding.h:24-60) to 8 bits, alpha 255, and the image is
 * written as an RGBA8 PNG. pbrtm_srgb8 returns the 8-bit pixels ([n][4]). */
int pbrtm_srgb8(const float *rgb, uint64_t n, uint8_t *rgba_out);
int pbrtm_write_png(const float *rgb, uint32_t width, uint32_t height, const char *path);
/* The same linear output-space film, unencoded, as a little-endian PFM (no
 * counterpart in the reference, whose films leave only as 8-bit PNG): for
 * comparing images at full precision. */
int pbrtm_write_pfm(const float *rgb, uint32_t width, uint32_t height, const char *path);

#ifdef __cplusplus
}
#endif

#endif /* PBRTM_H */
