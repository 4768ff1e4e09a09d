/*
 * spectral_oracle.h -- CPU restatement of spectral-core's light-transport path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.  The product
 * (spectral_amd, libspectral_hip.so) never links, loads or calls it.
 *
 * Every function follows a file:line of the reference crate spectral-core
 * (paths relative to the reference checkout).  The float expressions are
 * written in the reference's evaluation order and the library is compiled
 * with -ffp-contract=off, so each +,-,*,/,sqrt rounds exactly as the
 * reference's Rust does (Rust never contracts to FMA).  Transcendentals
 * (sinf, cosf, tanf, expf, powf) are glibc's, which is what Rust's f32
 * methods call on Linux.
 *
 * The POD structs are byte-compatible with the product's C ABI
 * (include/spectral_hip.h), which follows spectral-gpu/src/upload.rs:9-32.
 */
#ifndef SPECTRAL_ORACLE_H
#define SPECTRAL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scene PODs (upload.rs:9-32) ---------------------------------------- */
typedef struct {
    float center[3];
    float radius;
    uint32_t plane_start;
    uint32_t plane_count;
    uint32_t kind;      /* 0 sphere, 1 convex solid */
    uint32_t material;  /* index into materials */
} or_prim;

typedef struct {
    float normal[3];
    float d;
} or_plane;

typedef struct {
    float reflectance;  /* Lambertian albedo */
    uint32_t glass;     /* 0 BK7, 1 SF11, 2 Water */
    uint32_t kind;      /* 0 Lambertian, 1 Dielectric */
    uint32_t pad;
} or_material;

/* Scene::{background, horizon} (scene.rs:19-26) plus a sun-disk term
 * (extension, zero = absent) used by the rainbow configuration. */
typedef struct {
    float background;
    float horizon;
    uint32_t has_horizon;
    float sun_radiance;
    float sun_dir[3];
    float sun_cos;
} or_scene_env;

typedef struct {
    const or_prim *prims;
    uint32_t n_prims;
    const or_plane *planes;
    uint32_t n_planes;
    const or_material *mats;
    uint32_t n_mats;
    or_scene_env env;
} or_scene;

/* Camera::view_basis (camera.rs:36-38). */
typedef struct {
    float origin[3];
    float lower_left[3];
    float horizontal[3];
    float vertical[3];
} or_camera;

/* lighttrace::Beam (lighttrace.rs:116-121). */
typedef struct {
    float corner[3];
    float u[3];
    float v[3];
    float dir[3];
} or_beam;

/* lighttrace::Screen geometry (lighttrace.rs:15-34); the sum lives apart. */
typedef struct {
    float corner[3];
    float u[3];
    float v[3];
    uint32_t width;
    uint32_t height;
} or_screen;

/* Render options shared by all three entry points.  Zero-initialised
 * options plus max_depth=8 reproduce the reference exactly. */
typedef struct {
    uint32_t illuminant;  /* 0 D65, 1 A, 2 equal energy (photons only) */
    uint32_t dispersion;  /* 1 = n(lambda); 0 = n(550 nm) for every lambda */
    uint32_t max_depth;   /* bounce cap; the reference fixes 8 */
    uint32_t rr_depth;    /* camera only: 0 = off, k = roulette from bounce k */
    uint32_t hero_lanes;  /* photons: 0/1 = one wavelength; 4 = the hero comb (extension) */
} or_opts;

/* Volume estimator parameters (volume.rs:99-112). */
typedef struct {
    float sigma_s;
    float sigma_t;
    float g;
    float max_dist;
} or_vol_params;

/* Work counters (segments = bounce-loop iterations, splats = film deposits). */
typedef struct {
    uint64_t segments;
    uint64_t splats;
    uint64_t paths;
} or_stats;

/* Per-photon state record (vol_photons.rs:87-93, 932-991). */
#define OR_MAX_PAIRS 9
typedef struct {
    uint32_t num_states;
    float lambda;
    uint32_t tir_count;
    uint32_t pad;
    float states[OR_MAX_PAIRS][2][4]; /* [origin.xyz,_][dir.xyz,_] */
} or_photon_record;

/* ---- rng.rs -------------------------------------------------------------- */
uint32_t or_pcg32(uint32_t input);
uint32_t or_mix2(uint32_t a, uint32_t b);
uint32_t or_rng_new(uint32_t key, uint32_t seed);
void or_rng_draws_u32(uint32_t key, uint32_t seed, uint32_t n, uint32_t *out);
void or_rng_draws_f32(uint32_t key, uint32_t seed, uint32_t n, float *out);

/* ---- spectrum.rs / sellmeier.rs / cie.rs ----------------------------------- */
void or_hero_lambdas(float u, float out[4]);
float or_glass_n(uint32_t glass, float lambda_nm);
int or_load_tables(const char *data_dir);
void or_cmf(float nm, float out[3]);
float or_illuminant(uint32_t ill, float nm);
void or_integrate_illuminant(uint32_t ill, float out[3]);
float or_illuminant_scale(uint32_t ill);
void or_xyz_to_linear_srgb(const float xyz[3], float out[3]);
void or_tonemap_to_u8(const float linear[3], float exposure, uint8_t out[3]);

/* ---- geom.rs / optics.rs / material.rs / camera.rs / scene.rs ------------- */
/* hit_out: t, point.xyz, normal.xyz, front_face(0/1) */
int or_sphere_intersect(const float center[3], float radius, const float o[3], const float d[3],
                        float t_min, float t_max, float hit_out[8]);
int or_solid_intersect(const or_plane *planes, uint32_t n, const float o[3], const float d[3],
                       float t_min, float t_max, float hit_out[8]);
int or_scene_intersect(const or_scene *s, const float o[3], const float d[3], float hit_out[8],
                       uint32_t *mat_out);
void or_reflect(const float d[3], const float n[3], float out[3]);
int or_refract(const float d[3], const float n[3], float n1, float n2, float out[3]);
float or_fresnel_reflectance(float cos_i, float n1, float n2);
int or_material_scatter(const or_material *m, const float wo[3], const float hit[8], float n_hero,
                        uint32_t *rng_state, float dir_out[3], float *weight_out);
void or_camera_look_at(const float origin[3], const float target[3], const float up[3],
                       float vfov_deg, float aspect, or_camera *out);
void or_camera_primary_ray(const or_camera *c, float s, float t, float dir_out[3]);
int or_camera_project(const or_camera *c, const float p[3], float out[3]);
float or_background_radiance(const or_scene_env *env, const float dir[3]);

/* ---- tracer.rs: Renderer::accumulate ------------------------------------- */
/* Adds spp samples per pixel (sample indices sample_base..sample_base+spp-1)
 * into sum (3*w*h, row 0 = top), as CpuTracer::accumulate does. */
void or_camera_accumulate(const or_scene *s, const or_camera *cam, uint32_t w, uint32_t h,
                          uint32_t spp, uint32_t sample_base, uint32_t seed, const or_opts *o,
                          float *sum, uint32_t threads, or_stats *stats);
void or_camera_sample(const or_scene *s, const or_camera *cam, uint32_t w, uint32_t h,
                      uint32_t px, uint32_t py, uint32_t sample_idx, uint32_t seed,
                      const or_opts *o, float out[3]);

/* ---- lighttrace.rs -------------------------------------------------------- */
void or_lighttrace(const or_scene *s, const or_screen *scr, const or_beam *beam, uint32_t n,
                   uint32_t photon_base, uint32_t seed, const or_opts *o, float *sum,
                   uint32_t threads, or_stats *stats);

int or_lighttrace_photon(const or_scene *sc, const or_screen *scr, const float o[3], const float d[3],
                         float lambda, uint32_t *rng, const or_opts *op, float out[3]);

/* ---- volume.rs ------------------------------------------------------------ */
/* estimator: 0 point (render_volumetric_scene_point), 1 beam
 * (render_volumetric_scene), 2 legacy (render_volumetric). */
void or_volume(const or_scene *s, const or_camera *cam, const or_beam *beam, uint32_t w,
               uint32_t h, uint32_t n, uint32_t photon_base, const or_vol_params *vp,
               const float *zbuf, uint32_t seed, uint32_t estimator, const or_opts *o,
               float *img, uint32_t threads, or_stats *stats);
float or_phase_hg(float g, float cos_theta);
/* Test hook: beam estimator on one thread into img; returns the largest
 * bounding-box area of `group` consecutive steps of one segment. */
uint64_t or_volume_beam_max_box(const or_scene *s, const or_camera *cam, const or_beam *beam, uint32_t w,
                                uint32_t h, uint32_t n, uint32_t photon_base, const or_vol_params *vp,
                                const float *zbuf, uint32_t seed, const or_opts *o, uint32_t group, float *img);

/* Per-photon ray states (vol_photons.rs:932-991).  mode: 0 volume point /
 * legacy (4 dist draws per segment), 1 volume beam (no dist draws),
 * 2 lighttrace (screen tested first; scr must be non-NULL). */
void or_photon_record_states(const or_scene *s, const or_beam *beam, const or_screen *scr,
                             uint32_t photon_idx, uint32_t seed, float max_dist, uint32_t mode,
                             const or_opts *o, or_photon_record *rec);

/* Surface rim + depth pass (prism_dsotm.rs:59-74, vol_photons.rs:1517-1538). */
void or_surface_rim(const or_scene *s, const or_camera *cam, uint32_t w, uint32_t h,
                    float *rim_xyz, float *depth);

/* The viewer's display pass and exposure quantile (spectral-viewer
 * main.rs:131-174, 614-644) on an FP32 film (SCALE = 1). */
void or_display_resolve(const float *vol, const float *surf, uint32_t w, uint32_t h, uint32_t total_photons,
                        float target_n, float exposure, float kradius, uint8_t *rgba);
float or_display_quantile(const float *vol, const float *surf, uint32_t w, uint32_t h, uint32_t total_photons,
                          float target_n, float quantile);
/* The floor-caustic product image (prism_on_plane.rs:16-185). */
void or_box_blur(const float *in, float *out, uint32_t w, uint32_t h, int radius);
void or_floor_surface(const or_scene *s, const or_camera *cam, const or_screen *floor, const float *floor_sum,
                      float gain, const float ambient[3], const float background[3], const float tint[3],
                      uint32_t w, uint32_t h, float *surface, float *depth);
void or_composite_srgb8(const float *surface, const float *vol, uint32_t w, uint32_t h, float exposure,
                        uint8_t *rgb);
uint32_t or_struct_sizes(uint32_t which);

#ifdef __cplusplus
}
#endif
#endif
