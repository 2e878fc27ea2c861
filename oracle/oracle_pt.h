/* TEST INFRASTRUCTURE ONLY — the CPU oracle's wavefront path tracer.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * liboracle.so, and only as the checker or the timed CPU baseline.
 *
 * A host-only restatement of integrators/wavefront_path.cu and everything it
 * calls for diffuse, conductor and dielectric surfaces lit by diffuse area lights: the
 * stratified sampler, the tabulated Mitchell filter, the perspective camera,
 * spectral sampling and the RGB film. Paths are scheduled in rounds exactly as
 * the reference's path pool schedules them, so the film receives its samples
 * in the reference's order (round, then pixel, then sample).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_material {
    uint32_t type;          /* 0 = diffuse, 1 = conductor, 2 = dielectric */
    float rgb[3];           /* diffuse: RGB reflectance (albedo, each <= 1) */
    uint32_t metal;         /* conductor: eta/k of 0 = Ag, 1 = Al, 2 = Au, 3 = Cu */
    float uroughness, vroughness;
    uint32_t remap_roughness;
    float eta;              /* dielectric: constant index of refraction */
    uint32_t eta_named;     /* dielectric: 0 = constant eta, 1 = "glass-BK7" */
} orc_material;

typedef struct orc_light {  /* AreaLightSource "diffuse" */
    float rgb[3];           /* "rgb L" */
    float scale;
    uint32_t two_sided;
} orc_light;

typedef struct orc_infinite_light { /* LightSource "infinite" (uniform_infinite_light.cu:4-43) */
    float rgb[3];           /* "rgb L" */
    uint32_t has_rgb;       /* 0: no "L", the colour space illuminant */
    float scale;
    float illuminance;      /* <= 0: not given */
    uint32_t light_index;   /* place in the light list: emissive triangles created before it */
} orc_infinite_light;

typedef struct orc_shape {  /* Shape "sphere" (1), "disk" (2), or an ObjectInstance's triangle (3) */
    uint32_t type;
    uint32_t xform;         /* the CTM (instance: at the ObjectInstance): index into xforms, UINT32_MAX = identity */
    float radius, zmin, zmax, phimax /* degrees */, height, innerradius;
} orc_shape;

typedef struct orc_scene {
    float eye[3], look[3], up[3]; /* LookAt */
    float fov;                    /* Camera "perspective" "float fov" */
    uint32_t width, height;       /* Film "rgb" resolution */
    uint32_t spp;                 /* stratified samples per pixel (a square) */
    uint32_t max_depth;           /* Integrator "path" "integer maxdepth" */
    uint32_t pool_size;           /* path slots in flight */
    uint64_t num_tris;
    const float *tri;             /* [n][3][3] object space (world space without xforms) */
    const uint32_t *tri_material; /* [n] */
    const int32_t *tri_light;     /* [n] index into lights or -1; nullable */
    const uint8_t *tri_reverse;   /* [n] ReverseOrientation; nullable */
    uint32_t num_materials;
    const orc_material *materials;
    uint32_t num_lights;
    const orc_light *lights;
    uint32_t num_xforms;          /* object-to-world 4x4 row-major matrices */
    const float *xforms;
    const uint32_t *tri_xform;    /* [n] index or UINT32_MAX = identity; nullable */
    const float *select_scale;    /* [num_lights] nullable: emitter scales the light choice's cdf
                                     is built from (test knob: freezes the choice under a change of
                                     the scales; pmf values in the estimator stay the scene's) */
    uint32_t num_infinite_lights;
    const orc_infinite_light *infinite_lights;
    const float *xforms_inv;      /* [num_xforms] inverses as composed by the front end; nullable */
    uint32_t num_shapes;
    const orc_shape *shapes;
    const uint32_t *tri_shape;    /* [n] UINT32_MAX = triangle, else an index into shapes; nullable */
} orc_scene;

/* film_rgb: [height][width][3] = RGBFilm::get_pixel_rgb; aux (nullable):
 * [height][width][4] = rgb_sum, weight_sum; stats (nullable, 5 entries): rounds,
 * rays, shadow rays, finished paths, microseconds spent rendering (scene set-up
 * and the BVH build excluded). */
int orc_render(const orc_scene *scene, int threads, float *film_rgb, float *aux, uint64_t *stats);
/* The same render, reporting per path id (sample * W * H + pixel) the round its
 * first ray is traced in and the round it is recorded in the film. */
int orc_render_paths(const orc_scene *scene, int threads, uint32_t *path_start, uint32_t *path_finish,
                     uint64_t *stats);

/* The wavefront render's per-path spectral radiance and wavelength pdfs
 * ([W*H*spp][4] by path id; path_pdf nullable). */
int orc_render_path_L(const orc_scene *scene, int threads, float *path_L, float *path_pdf);

/* The megakernel integrators (base/megakernel_integrator.cu:14-147): integrator
 * 1 = "megakernelpath" (integrators/megakernel_path.cu:35-213), 2 =
 * "ambientocclusion" (ambient_occlusion.cu:25-58), 3 = "surfacenormal"
 * (surface_normal.cu:17-32). Every pixel's samples in sample order, each added
 * to the film as it finishes; film_rgb / aux as orc_render; path_L (nullable)
 * as orc_render_path_L; pool_size unused. */
int orc_render_megakernel(const orc_scene *scene, int integrator, int threads, float *film_rgb, float *aux,
                          float *path_L, uint64_t *stats);

/* Loss and gradient of the path-traced film (FP64 path replay): loss = sum over
 * pixels and channels of (film - target)^2; d_rgb [num_materials][3] is
 * dLoss/d(diffuse albedo RGB) (0 for other materials), d_scale [num_lights]
 * dLoss/d(emitter scale). The fixed-seed estimator is differentiated path by
 * path with the light pmf attached everywhere it appears (MIS weights and the
 * light choice's 1/pmf); the discrete choices themselves are held, so this is
 * the derivative central differences measure while no sample's light choice
 * flips. hold_selection != 0 holds the choice's 1/pmf instead. film_rgb
 * receives the film. */
int orc_render_grad(const orc_scene *scene, int threads, const float *target, float *film_rgb, double *loss,
                    double *d_rgb, double *d_scale, uint64_t *stats, int hold_selection);

/* Intermediate tables, for tests: the RGB-to-spectrum sigmoid coefficients of an
 * albedo RGB, the sRGB rgb_from_xyz matrix, the filter sample of a 2D u. */
int orc_rgb_to_sigmoid(const float rgb[3], float out3[3]);
int orc_rgb_to_sigmoid_jacobian(const float rgb[3], double J9[9]); /* dc_q/drgb_k at [3q + k] */
int orc_srgb_matrices(float xyz_from_rgb9[9], float rgb_from_xyz9[9], float white_xy2[2]);
int orc_filter_sample(const float u[2], float out3[3]);
float orc_spectrum_eval(int which, float lambda); /* 0-2 CIE XYZ, 3 D65 (normalised), 4-11 eta/k of Ag, Al,
                                                     Au, Cu, 12 glass-BK7 eta */
int orc_sample_visible(float u, float lam4[4], float pdf4[4]);
/* A stratified sampler started at (pixel, sample, dimension 0); kinds[i] = 1
 * draws get_1d, 2 draws get_2d; returns the number of floats written. */
int orc_sampler_sequence(uint32_t spd, uint32_t pixel, uint32_t sample, const int *kinds, int n, float *out);
/* The camera ray through raster point (px, py): world-space origin, render-space direction. */
int orc_camera_ray(const orc_scene *scene, float px, float py, float out6[6]);
int orc_spherical_triangle(const float v9[9], const float p[3], const float u[2], float b3[3], float *pdf);
int orc_invert_spherical_triangle(const float v9[9], const float p[3], const float w3[3], float u2[2]);
float orc_fr_complex(float cos_i, float eta, float k);
/* Shading functions in the local frame (shading normal = +z), one direction
 * pair per entry, for the tests that pin them to the mathematics. A BxDF:
 * type 0 diffuse (r = the sampled reflectance at 4 wavelengths), 1 conductor
 * (eta4, k4), 2 dielectric (eta; 0 means 1). ax, ay are the Trowbridge-Reitz
 * alphas after any roughness remapping; a rough lobe clamps them to >= 1e-4
 * as the material does. flags: BxDFFlags (reflection 1, transmission 2,
 * diffuse 4, glossy 8, specular 16). */
typedef struct orc_bxdf {
    uint32_t type;
    float r[4];
    float eta4[4], k4[4];
    float eta;
    float ax, ay;
} orc_bxdf;
uint32_t orc_bxdf_flags(const orc_bxdf *b);
/* BSDF::f and BSDF::pdf of (wo, wi): f4 [n][4], pdf [n]. */
int orc_bxdf_eval(const orc_bxdf *b, uint64_t n, const float *wo, const float *wi, float *f4, float *pdf);
/* BSDF::sample_f of wo with (uc, u0, u1) = u3[i]: f4, wi, pdf, eta (the relative
 * IOR of a transmission, else 1); flags[i] bit 0 = a sample was returned,
 * bit 1 specular, bit 2 transmission. */
int orc_bxdf_sample(const orc_bxdf *b, uint64_t n, const float *wo, const float *u3, float *f4, float *wi,
                    float *pdf, float *eta, uint32_t *flags);
float orc_fr_dielectric(float cos_i, float eta);
/* refract(wi, n, eta): returns 1 and the transmitted direction and relative
 * IOR, or 0 at total internal reflection. */
int orc_refract(const float wi[3], const float n[3], float eta, float *etap, float wt[3]);
/* Trowbridge-Reitz: out4[i] = D(wm), Lambda(w), G1(w), D_w(wm) (the visible
 * normal density); orc_tr_sample: Sample_wm(w, u). */
int orc_tr_eval(float ax, float ay, uint64_t n, const float *w, const float *wm, float *out4);
int orc_tr_sample(float ax, float ay, uint64_t n, const float *w, const float *u2, float *wm);
/* The power light sampler's pmf over the scene's lights in light order, and
 * each light's emissive triangle (-1 = an infinite light). */
int orc_light_pmf(const orc_scene *scene, float *pmf, int32_t *tri, uint32_t cap, uint32_t *count);

/* Analytic shapes one at a time (tests), given render_from_object (m16) and its
 * inverse (inv16), row-major: closest hits of rays [n][8] (t, render-space point
 * and normal), Shape::sample(ctx, u) and Shape::pdf(ctx, wi) with ctx [n][9] =
 * (p, n, ns); and the directed-rounding primitives (op 0 mul, 1 div, 2 sqrt, 3 add). */
int orc_shape_intersect(const orc_shape *sh, const float *m16, const float *inv16, int reverse, const float *rays,
                        uint64_t n, float *t_out, float *p_out, float *n_out, uint8_t *hit_out);
int orc_shape_sample(const orc_shape *sh, const float *m16, const float *inv16, int reverse, const float *ctx,
                     const float *u, uint64_t n, float *p_out, float *n_out, float *pdf_out, uint8_t *ok_out);
int orc_shape_pdf(const orc_shape *sh, const float *m16, const float *inv16, int reverse, const float *ctx,
                  const float *wi, uint64_t n, float *pdf_out);
int orc_directed(int op, float a, float b, float out[2]);

/* Film::write_to_png's per-channel encoding: linear RGB [n][3] -> sRGB RGBA8 [n][4]. */
int orc_srgb8(const float *rgb, uint64_t n, uint8_t *rgba);

#ifdef __cplusplus
}
#endif
