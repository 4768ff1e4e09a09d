/*
 * spectral_hip.h -- C ABI of the MI355X (gfx950) light-transport core.
 *
 * Drop-in for spectral-core's three light-transport entry points:
 *   camera render  CpuTracer::new + Renderer::accumulate   (spectral-core/src/tracer.rs:24-41,115-141;
 *                                                           spectral-core/src/lib.rs:30-34)
 *   lighttrace     lighttrace::trace                       (spectral-core/src/lighttrace.rs:157-166)
 *   volume         volume::render_volumetric_scene_point,  (spectral-core/src/volume.rs:99-168,
 *                  render_volumetric_scene, render_volumetric   233-329, 22-85)
 * and for the GPU mirror's constructors/methods (spectral-gpu/src/lib.rs:179-388,
 * spectral-gpu/src/vol_photons.rs:275-779).
 *
 * Conventions
 *  - Every entry point returns an int status: SPX_OK (0) or a negative code.
 *    The message of the last failure is spx_last_error(ctx).  Nothing
 *    unwinds across the boundary.
 *  - Films are caller-owned DEVICE buffers of 3*w*h floats, interleaved XYZ,
 *    row 0 = top (the reference's Vec<[f32;3]> layout, lib.rs:10).  Render
 *    calls ADD into them (AccumBuffer::sum semantics): the caller clears a
 *    film with spx_fill_f32 and divides by its sample count.
 *  - `stream` is a hipStream_t (NULL = the null stream).  Calls are
 *    asynchronous on that stream unless documented otherwise.
 *  - One context per thread; concurrency comes from streams.
 *  - Struct layouts are byte-compatible with spectral-gpu/src/upload.rs:9-32.
 */
#ifndef SPECTRAL_HIP_H
#define SPECTRAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SPX_OK 0
#define SPX_EINVAL (-1)   /* invalid argument (null pointer, zero size, bad enum) */
#define SPX_ENOMEM (-2)   /* device allocation failed */
#define SPX_EDEVICE (-3)  /* HIP runtime error */
#define SPX_ECAP (-4)     /* a fixed capacity was exceeded (scene size, debug count) */
#define SPX_EBUSY (-5)    /* serving: request not admitted under the memory cap */

/* ---- enums ------------------------------------------------------------- */
#define SPX_PRIM_SPHERE 0u
#define SPX_PRIM_SOLID 1u
#define SPX_MAT_LAMBERTIAN 0u
#define SPX_MAT_DIELECTRIC 1u
#define SPX_GLASS_BK7 0u
#define SPX_GLASS_SF11 1u
#define SPX_GLASS_WATER 2u
#define SPX_ILLUM_D65 0u
#define SPX_ILLUM_A 1u
#define SPX_ILLUM_E 2u /* equal energy: the reference's photon power of 1 */
#define SPX_EST_POINT 0u
#define SPX_EST_BEAM 1u
#define SPX_EST_LEGACY 2u
#define SPX_ENGINE_PLAIN 0u     /* one thread per path/photon, one launch */
#define SPX_ENGINE_WAVEFRONT 1u /* per-bounce SoA queues + compaction */

/* Scene capacity (the scene is staged in LDS by every kernel). */
#define SPX_MAX_PRIMS 16u
#define SPX_MAX_PLANES 64u
#define SPX_MAX_MATERIALS 16u

/* ---- PODs ---------------------------------------------------------------- */
/* GpuPrimitive, upload.rs:9-16 */
typedef struct {
    float center[3];
    float radius;
    uint32_t plane_start;
    uint32_t plane_count;
    uint32_t kind;     /* SPX_PRIM_* */
    uint32_t material; /* index into the material array */
} spx_prim;

/* GpuPlane, upload.rs:29-32: half-space {x : normal.x + d <= 0}, normal out. */
typedef struct {
    float normal[3];
    float d;
} spx_plane;

/* GpuMaterial, upload.rs:20-25 */
typedef struct {
    float reflectance; /* Lambertian albedo */
    uint32_t glass;    /* SPX_GLASS_* */
    uint32_t kind;     /* SPX_MAT_* */
    uint32_t pad;
} spx_material;

/* Scene::{background, horizon} (scene.rs:19-26, 44-50).  sun_* is an
 * extension (a directional sun disk, zero radiance = absent): escaped rays
 * with dot(dir, sun_dir) >= sun_cos add sun_radiance. */
typedef struct {
    float background;
    float horizon;
    uint32_t has_horizon;
    float sun_radiance;
    float sun_dir[3];
    float sun_cos;
} spx_env;

/* Camera::view_basis (camera.rs:36-38). */
typedef struct {
    float origin[3];
    float lower_left[3];
    float horizontal[3];
    float vertical[3];
} spx_camera;

/* lighttrace::Beam (lighttrace.rs:116-121). */
typedef struct {
    float corner[3];
    float u[3];
    float v[3];
    float dir[3];
} spx_beam;

/* lighttrace::Screen geometry (lighttrace.rs:15-34). */
typedef struct {
    float corner[3];
    float u[3];
    float v[3];
    uint32_t width;
    uint32_t height;
} spx_screen;

/* Render options.  {illuminant, dispersion=1, max_depth=8, rr_depth=0}
 * reproduces the reference. */
typedef struct {
    uint32_t illuminant; /* SPX_ILLUM_*; camera: background SPD; photons: beam SPD */
    uint32_t dispersion; /* 1 = n(lambda); 0 = n(550 nm) (vol_photons.rs:77 `spectral`) */
    uint32_t max_depth;  /* bounces (tracer.rs:39, lighttrace.rs:133, volume.rs:126): 1..8 */
    uint32_t rr_depth;   /* camera: 0 = no roulette; k = Russian roulette from bounce k */
    uint32_t hero_lanes; /* photons: 0 or 1 = one wavelength per photon (the reference);
                            4 = the stratified hero comb (spectrum.rs:20-32) per photon, its
                            companion lanes collapsing at the first dielectric as a camera
                            path's do (tracer.rs:86-91); not with SPX_EST_BEAM.  Camera
                            paths always carry the comb. */
} spx_opts;

/* Volume weights (volume.rs:99-112; VolWeights, vol_photons.rs:256-267). */
typedef struct {
    float sigma_s;
    float sigma_t;
    float g;
    float max_dist;
} spx_vol_params;

/* Per-photon ray states (DebugPhotonGpu, vol_photons.rs:87-93). */
typedef struct {
    uint32_t num_states; /* 1 + successful scatters */
    float lambda;
    uint32_t tir_count;  /* bounces whose refract() had no solution */
    uint32_t pad;
    float states[9][2][4]; /* [origin.xyz,_][dir.xyz,_] */
} spx_photon_record;

typedef struct spx_ctx spx_ctx;
typedef struct spx_scene spx_scene;

/* Per-context tuning of the wavefront and batched engines (the plain engine
 * ignores it).  spx_autotune picks these values by timing a job table. */
typedef struct {
    uint32_t max_replicas;   /* photon film copies per job (1 = deposit straight into the film) */
    uint32_t replica_mib;    /* per-job budget for the copies, MiB */
    uint32_t splat_mode;     /* photon deposits: 0 = direct fp32 atomics, 1 = LDS hash first,
                                2 = binned records sorted by film tile and summed in LDS as
                                fixed point (no fp32 atomics; films bit-reproducible), 3 = the
                                LDS hash draining to binned records (fewer records; fp32 sums
                                in LDS, so not bit-reproducible); default 2 */
    uint32_t chunk_log2;     /* wavefront queue capacity: 2^chunk_log2 paths / photons per bounce */
    uint32_t bin_log2;       /* binned deposits: at most 2^bin_log2 records per round (32 B of
                                device memory each; buffers sized to the work); default 30 */
    uint32_t tile_wlog2;     /* binned deposits: film tiles of 1024 pixels, 2^tile_wlog2 wide and
                                2^(10 - tile_wlog2) tall (2..8); default 5 = 32x32 */
    uint32_t batch_blocks;   /* batched engine: persistent grid size in blocks of 256 threads
                                (0 = 4 per compute unit) */
    uint32_t chunk_paths;    /* batched engine: camera work item, in paths (default 16384) */
    uint32_t chunk_photons;  /* batched engine: photon work item, in photons (default 8192) */
    uint32_t overlap_bins;   /* batched engine, binned deposits: N = 1 or 2 overlapped rounds, the
                                bin passes of round r on a second stream while rounds r+1 .. r+N
                                trace (N + 1 record buffers, round r+N+1 reuses round r's);
                                default 0 */
    uint32_t camera_blocks;  /* batched engine: 0 = the camera round runs first, on the whole grid;
                                N > 0 = it runs on its own stream with N blocks, concurrently with
                                the photon rounds; default 0 */
    uint32_t photon_threads; /* batched engine, deposit modes 1 and 3: threads per photon-kernel block,
                                256 (LDS hash of 2048 pixels per block) or 1024 (one block per
                                compute unit sharing a hash of 8192 pixels: more deposits meet
                                in LDS); the grid keeps batch_blocks x 256 threads; default 256 */
} spx_tuning;

/* ---- context --------------------------------------------------------------- */
/* GpuContext::new (spectral-gpu/src/lib.rs:21-62).  Uploads the CIE 1931 CMF
 * and D65/A tables (cie.rs:6-8, data.rs:56-66). */
int spx_ctx_create(int device, spx_ctx **out);
void spx_ctx_destroy(spx_ctx *ctx);
const char *spx_last_error(const spx_ctx *ctx);
/* Number of compute units of the context's device. */
int spx_ctx_num_cus(const spx_ctx *ctx);
int spx_ctx_set_tuning(spx_ctx *ctx, const spx_tuning *t);
int spx_ctx_get_tuning(const spx_ctx *ctx, spx_tuning *t);
/* Measurement hook (no counterpart in the reference): while enabled, every
 * spx_batch_run brackets its kernel groups with HIP events on the caller's
 * stream: the camera trace kernel (k_batch_camera), each photon trace
 * launch (k_batch_photons) and each round's bin passes (count, scan,
 * scatter, accumulate).  spx_ctx_kernel_time waits for the launches recorded
 * so far, returns the summed duration (ms) and count of the two trace
 * kernels, and clears the record; spx_ctx_kernel_times returns the same per
 * group (index SPX_KT_*) and clears it.  Enabling clears it too.  Graph
 * replays (spx_batch_plan_launch) are not bracketed. */
#define SPX_KT_CAMERA 0u
#define SPX_KT_PHOTON 1u
#define SPX_KT_BINS 2u
int spx_ctx_kernel_timing(spx_ctx *ctx, int enable);
int spx_ctx_kernel_time(spx_ctx *ctx, double *total_ms, uint32_t *launches);
int spx_ctx_kernel_times(spx_ctx *ctx, double ms[3], uint32_t launches[3]);

/* Camera::look_at (camera.rs:15-27), evaluated on the host.  Needs no device. */
int spx_camera_look_at(const float origin[3], const float target[3], const float up[3],
                       float vfov_deg, float aspect, spx_camera *out);

/* ---- scene ------------------------------------------------------------------ */
/* flatten_scene + upload (vol_photons.rs:118-183, lib.rs:91-158). */
int spx_scene_create(spx_ctx *ctx, const spx_env *env, const spx_prim *prims, uint32_t n_prims,
                     const spx_plane *planes, uint32_t n_planes, const spx_material *mats,
                     uint32_t n_mats, spx_scene **out);
void spx_scene_destroy(spx_scene *scene);

/* ---- camera render: Renderer::accumulate (tracer.rs:115-141) -------------- */
/* Adds spp hero-wavelength samples per pixel, sample indices
 * sample_base .. sample_base+spp-1 (tracer.rs:119,127), into film_sum. */
int spx_camera_accumulate(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam,
                          uint32_t width, uint32_t height, uint32_t spp, uint32_t sample_base,
                          uint32_t seed, const spx_opts *opts, uint32_t engine, float *film_sum,
                          void *stream);

/* ---- lighttrace::trace (lighttrace.rs:157-166) ------------------------------ */
/* Photons photon_base .. photon_base+n-1; screen_sum is 3*width*height. */
int spx_lighttrace(spx_ctx *ctx, const spx_scene *scene, const spx_screen *screen,
                   const spx_beam *beam, uint32_t n, uint32_t photon_base, uint32_t seed,
                   const spx_opts *opts, uint32_t engine, float *screen_sum, void *stream);

/* ---- volume (volume.rs:22-329) ---------------------------------------------- */
/* estimator: SPX_EST_POINT (render_volumetric_scene_point), SPX_EST_BEAM
 * (render_volumetric_scene), SPX_EST_LEGACY (render_volumetric; ignores
 * sigma_t, g and zbuf).  zbuf: device, w*h floats, NULL = all +INF.  With
 * SPX_EST_BEAM the wavefront engine queues the photons' segments and
 * spreads each segment's midpoint steps over a wave. */
int spx_volume(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam, const spx_beam *beam,
               uint32_t width, uint32_t height, uint32_t n, uint32_t photon_base,
               const spx_vol_params *vp, const float *zbuf, uint32_t seed, uint32_t estimator,
               const spx_opts *opts, uint32_t engine, float *film_sum, void *stream);

/* Per-photon parity records (GpuPhotonFilm::trace_debug, vol_photons.rs:727-779)
 * for photons first .. first+count-1 (count <= 65536), written to host
 * memory.  mode 0: volume point/legacy RNG order; 1: volume beam; 2:
 * lighttrace (screen must be non-NULL).  Synchronous. */
int spx_photon_records(spx_ctx *ctx, const spx_scene *scene, const spx_beam *beam,
                       const spx_screen *screen, uint32_t first, uint32_t count, uint32_t seed,
                       float max_dist, uint32_t mode, const spx_opts *opts,
                       spx_photon_record *out_host);

/* Surface rim pass + z-buffer (GpuPhotonFilm::recompute_surface,
 * vol_photons.rs:685-714; prism_dsotm.rs:59-74).  rim_xyz (3*w*h) and
 * depth (w*h, +INF on a miss) are device buffers, either may be NULL. */
int spx_surface_rim(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam, uint32_t width,
                    uint32_t height, float *rim_xyz, float *depth, void *stream);

/* XYZ film -> sRGB8 (xyz_to_linear_srgb + tonemap_to_u8, cie.rs:121-144)
 * of film_sum * scale.  rgb is a device buffer of 3*w*h bytes. */
int spx_resolve_srgb8(spx_ctx *ctx, const float *film_sum, uint32_t width, uint32_t height,
                      float scale, float exposure, uint8_t *rgb, void *stream);

/* The viewer's progressive display pass (spectral-viewer main.rs:131-174,
 * the blit shader; its fixed-point accum is an FP32 film here, so its SCALE
 * is 1).  Per pixel: the Gaussian-weighted mean of vol_sum over the
 * (2r+1)^2 neighbourhood inside the film, r = (int)kradius, sigma =
 * max(kradius / 2, 0.5), times target_n / max(total_photons, 1); plus the
 * surface layer (3*w*h XYZ, NULL = none); times exposure; XYZ -> linear
 * sRGB, Reinhard c/(1+c), clamp below at 0, gamma 1/2.2.  rgba (4*w*h
 * bytes, 4-byte aligned) receives round(255 * clamp(v, 0, 1)) per channel
 * (the Rgba8Unorm store) and alpha 255.  All buffers are device memory;
 * row 0 = top. */
int spx_display_resolve(spx_ctx *ctx, const float *vol_sum, const float *surface, uint32_t width,
                        uint32_t height, uint32_t total_photons, float target_n, float exposure,
                        float kradius, uint8_t *rgba, void *stream);

/* The exposure rules' luminance quantile (viewer compute_exposure_p997,
 * main.rs:614-644 with q = 0.997; prism_on_plane.rs:172-176 with q = 0.995
 * and total_photons = target_n = 1): with L_i = vol_sum_Y[i] * (target_n /
 * max(total_photons, 1)) + surface_Y[i] (surface may be NULL), *lum_q =
 * max(L_(k), 1e-9) where L_(k) is the k-th smallest, k = min(floor(fp32(w*h)
 * * q), w*h - 1).  The caller's exposure is c / *lum_q (1 in the viewer, 1.1
 * in prism_on_plane).  Exact selection on the device (radix select over the
 * float bits); device buffers in, host float out.  Its kernels and copies
 * run on `stream`, ordered after earlier work there, and the call waits for
 * that stream before it returns. */
int spx_display_quantile(spx_ctx *ctx, const float *vol_sum, const float *surface, uint32_t width,
                         uint32_t height, uint32_t total_photons, float target_n, float quantile,
                         float *lum_q, void *stream);

/* No counterpart in the reference: spx_display_quantile for callers that
 * cannot allocate device memory (tight per-frame display loops, harnesses
 * that forbid hipMalloc).  Same arguments and result; it calls no hipMalloc
 * or hipFree: the histograms and the select's state live in a fixed 4 KiB
 * device array of the library (one per GPU, held until the process ends),
 * cleared by the call itself, and the four radix passes decide their digits
 * on the device, so the only host round trip is the final 4-byte result.
 * Films up to 2^28 pixels (SPX_ECAP above: the counts are 32-bit).  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (two streams, two threads) race and must be serialised by the caller.
 * Waits for `stream` before it returns. */
int spx_display_quantile_noalloc(spx_ctx *ctx, const float *vol_sum, const float *surface, uint32_t width,
                                 uint32_t height, uint32_t total_photons, float target_n, float quantile,
                                 float *lum_q, void *stream);

/* ---- the floor-caustic product image (prism_on_plane.rs:16-185) ---------------- */
/* box_blur (prism_on_plane.rs:40-74): separable box filter of radius r over
 * a 3*w*h XYZ buffer, clamp-to-edge; the horizontal pass into tmp, the
 * vertical pass into out, each tap sum in tap order then times 1/(2r+1).
 * Device buffers of 3*w*h floats; out may be in, tmp must differ from both. */
int spx_box_blur(spx_ctx *ctx, const float *in, float *tmp, float *out, uint32_t width, uint32_t height,
                 uint32_t radius, void *stream);

/* prism_on_plane's surface pass (prism_on_plane.rs:16-38 shade_floor,
 * 121-154): the camera ray at each pixel centre against the scene's nearest
 * solid and the floor screen's rectangle.  The floor shades as ambient +
 * gain * Screen::sample(floor_sum) (bilinear, lighttrace.rs:99-113), a miss
 * as background; a solid nearer than the floor as 0.7 * the floor behind +
 * glass_tint + (1 - |n.d|)^3 * (0.6, 0.6, 0.7).  surface (3*w*h) and depth
 * (w*h, distance to the shaded surface, +INF on a miss: the volume pass's
 * z-buffer) are device buffers; floor_sum is the screen's device film. */
int spx_floor_surface(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam, const spx_screen *floor,
                      const float *floor_sum, float floor_gain, const float ambient[3], const float background[3],
                      const float glass_tint[3], uint32_t width, uint32_t height, float *surface, float *depth,
                      void *stream);

/* tonemap_to_u8(xyz_to_linear_srgb(surface + vol), exposure) per pixel
 * (prism_on_plane.rs:166-182, cie.rs:121-144); rgb: 3*w*h bytes, device. */
int spx_composite_srgb8(spx_ctx *ctx, const float *surface, const float *vol, uint32_t width, uint32_t height,
                        float exposure, uint8_t *rgb, void *stream);

/* ---- known-answer kernels ------------------------------------------------------ */
/* PathRng::new(key, seed) then n next_u32 draws on the device (rng_parity.rs:4-16). */
int spx_rng_draws(spx_ctx *ctx, uint32_t key, uint32_t seed, uint32_t n, uint32_t *out_host);
/* Glass::n on the device (sellmeier.rs:28-44), for n wavelengths. */
int spx_glass_n(spx_ctx *ctx, uint32_t glass, const float *lambda_host, uint32_t n,
                float *out_host);

/* ---- batched jobs (persistent launch over a device job table) ------------------ */
#define SPX_JOB_CAMERA 0u
#define SPX_JOB_VOLUME 1u
#define SPX_JOB_LIGHTTRACE 2u

typedef struct {
    uint32_t kind;        /* SPX_JOB_* */
    uint32_t width;       /* film size (camera, volume); a lighttrace job's */
    uint32_t height;      /* film is its screen's width x height */
    uint32_t spp;         /* camera */
    uint32_t sample_base; /* camera RNG counter base */
    uint32_t n_photons;   /* volume, lighttrace */
    uint32_t photon_base; /* photon RNG counter base */
    uint32_t seed;
    uint32_t estimator;   /* volume: SPX_EST_POINT only in batches */
    uint32_t pad;
    spx_opts opts;
    spx_camera camera;    /* camera, volume */
    spx_beam beam;        /* volume, lighttrace */
    spx_screen screen;    /* lighttrace */
    spx_vol_params vol;   /* volume */
    const spx_scene *scene;
    float *film;          /* device, 3*width*height, accumulated into */
    const float *zbuf;    /* volume, device, may be NULL */
} spx_job;

typedef struct {
    uint32_t splat_mode;   /* photon deposits: as spx_tuning.splat_mode */
    uint32_t blocks;       /* persistent grid size (0 = tuned default) */
    uint32_t chunk_paths;  /* camera work item size in paths (0 = tuned default) */
    uint32_t chunk_photons;/* photon work item size (0 = tuned default) */
} spx_batch_cfg;

/* Runs every job of the table in ONE persistent launch on `stream`.  Work
 * items are (job, range) chunks taken from a device queue; results equal
 * the per-job calls at the same seeds and bases within float summation
 * order.  cfg NULL = the context's tuning. */
int spx_batch_run(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, const spx_batch_cfg *cfg,
                  void *stream);

/* A batch plan: the launch schedule of one job table -- device job table,
 * work items per XCD queue, the camera round and the photon deposit rounds
 * with their record, replica and queue buffers, all owned by the plan --
 * captured once as a hipGraph (queue reset, trace kernel and bin passes of
 * every round) and replayed by spx_batch_plan_launch.  No counterpart in the
 * reference (its GPU mirror records one command buffer per dispatch).
 * Creation uploads on `stream` and waits for it. */
typedef struct spx_batch_plan spx_batch_plan;
int spx_batch_plan_create(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, const spx_batch_cfg *cfg,
                          void *stream, spx_batch_plan **out);
/* Rewrites the per-job values that leave the schedule unchanged -- RNG
 * bases, seeds, cameras, beams, screen geometry, volume weights, options,
 * scenes, film and z-buffer pointers -- with a stream-ordered upload on
 * `stream`.  Kinds, film sizes, spp, photon counts, estimators and which jobs
 * share a film must equal the plan's (SPX_EINVAL otherwise: create a new
 * plan). */
int spx_batch_plan_update(spx_batch_plan *plan, const spx_job *jobs, uint32_t n_jobs, void *stream);
/* Replays the captured graph on `stream` (ordered after earlier work on it). */
int spx_batch_plan_launch(spx_batch_plan *plan, void *stream);
/* Rounds in the plan's schedule (camera round included) and graph nodes. */
int spx_batch_plan_info(const spx_batch_plan *plan, uint32_t *rounds, uint32_t *graph_nodes);
void spx_batch_plan_destroy(spx_batch_plan *plan);

/* Autotuner of the batched engine (no counterpart in the reference).  Times
 * the job table under candidate tunings -- persistent grid size, camera and
 * photon work-item sizes, deposit mode, record capacity per round and film
 * tile shape -- one coordinate at a time for `sweeps` passes (1..4) starting
 * from the context's tuning.  Each candidate runs the table into scratch
 * films (the jobs' own films are not touched) once untimed and twice timed
 * with HIP events on `stream`; the faster timed run counts.  The best tuning
 * is installed in the context and returned in *out; *best_ms (optional) is
 * its time per table.  Synchronous. */
int spx_autotune(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, uint32_t sweeps, void *stream,
                 spx_tuning *out, double *best_ms);

/* ---- serving: continuous batching over resident films ------------------------- */
/* A request is K frames of one job.  Frame f traces the job with its RNG
 * counter advanced by f frames (sample_base + f*spp, photon_base + f*n, as
 * tracer.rs:119 continues the sample index), its camera turned by
 * f*orbit_deg about the vertical axis through orbit_center.  With
 * orbit_deg == 0 the frames accumulate into one progressive film; otherwise
 * the film is reset before each frame (a camera move resets accumulation,
 * as the viewer does).  The server owns each request's film from admission
 * until spx_server_release. */
typedef struct {
    spx_job job;        /* frame-0 job; job.film is ignored */
    uint32_t frames;    /* K >= 1 */
    float orbit_deg;
    float orbit_center[3];
} spx_request;

#define SPX_REQ_PENDING 0u
#define SPX_REQ_ACTIVE 1u
#define SPX_REQ_DONE 2u

typedef struct {
    uint32_t state;       /* SPX_REQ_* */
    uint32_t frames_done;
    uint64_t footprint;   /* film + deposit-queue bytes reserved at admission */
    double t_submit;      /* seconds on the server clock */
    double t_admit;
    double t_done;
} spx_request_info;

typedef struct spx_server spx_server;

int spx_server_create(spx_ctx *ctx, uint64_t mem_cap_bytes, spx_server **out);
void spx_server_destroy(spx_server *srv);
/* Queues a request (validated now; an invalid request fails alone with
 * SPX_EINVAL).  SPX_EBUSY if its footprint exceeds the whole cap. */
int spx_server_submit(spx_server *srv, const spx_request *req, uint64_t *id);
/* One continuous-batching iteration: admits pending requests (oldest first)
 * whose footprint fits under the cap, runs the next frame of every active
 * request in ONE batched launch on `stream`, waits for it and retires the
 * requests that finished.  *n_active (optional) = requests in that launch. */
int spx_server_step(spx_server *srv, void *stream, uint32_t *n_active);
int spx_server_info(const spx_server *srv, uint64_t id, spx_request_info *out);
/* The camera basis frame `frame` of a request uses. */
int spx_server_frame_camera(const spx_server *srv, uint64_t id, uint32_t frame, spx_camera *out);
/* Copies a request's film (3*w*h floats) to host memory.  Synchronous. */
int spx_server_film(spx_server *srv, uint64_t id, float *host_out, size_t floats);
int spx_server_release(spx_server *srv, uint64_t id);
uint64_t spx_server_bytes_in_use(const spx_server *srv);
double spx_server_clock(const spx_server *srv);
/* Steps replay a captured batch plan while the active set is unchanged
 * (default 1); 0 = every step is one spx_batch_run. */
int spx_server_set_graphs(spx_server *srv, int enable);
/* Steps that captured a new plan and steps that replayed one. */
int spx_server_stats(const spx_server *srv, uint64_t *captures, uint64_t *replays);

/* ---- device memory and streams (so callers need no HIP bindings) --------------- */
int spx_malloc(spx_ctx *ctx, size_t bytes, void **out);
int spx_free(spx_ctx *ctx, void *ptr);
int spx_memcpy_h2d(spx_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream);
int spx_memcpy_d2h(spx_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream);
/* fp32 atomic-add throughput probe: threads x ops hardware atomics spread by
 * a hash over n_slots floats of buf (device).  Asynchronous on stream. */
int spx_atomic_probe(spx_ctx *ctx, float *buf, size_t n_slots, uint32_t ops, uint32_t threads,
                     void *stream);
/* Copies n floats device-to-device with a plain streaming kernel. */
int spx_copy_f32(spx_ctx *ctx, float *dst, const float *src, size_t n, void *stream);
/* Fills n floats with value using a kernel (safe inside graph capture). */
int spx_fill_f32(spx_ctx *ctx, float *dst, float value, size_t n, void *stream);
int spx_stream_create(spx_ctx *ctx, void **out);
int spx_stream_destroy(spx_ctx *ctx, void *stream);
int spx_stream_sync(spx_ctx *ctx, void *stream);
int spx_event_create(spx_ctx *ctx, void **out);
int spx_event_destroy(spx_ctx *ctx, void *ev);
int spx_event_record(spx_ctx *ctx, void *ev, void *stream);
int spx_event_elapsed_ms(spx_ctx *ctx, void *ev_start, void *ev_end, float *ms);

#ifdef __cplusplus
}
#endif
#endif /* SPECTRAL_HIP_H */
