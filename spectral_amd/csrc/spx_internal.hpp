// spx_internal.hpp -- host-side internals shared by the kernel and ABI units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "spectral_hip.h"
#include "spx_bins.hpp"
#include "spx_device.hpp"

namespace spx {

// Record buffers of overlapped deposit rounds at most (spx_tuning.overlap_bins + 1).
constexpr int kRoundBufs = 3;

// Per-launch parameter blocks (passed by value as kernel arguments).
struct CamParams {
    spx_camera cam;
    uint32_t w, h, spp, sample_base;
    uint32_t seed, illuminant, dispersion, max_depth;
    uint32_t rr_depth, pad0, pad1, pad2;
    const SceneData *scene;
    const float4 *tbl;  // 71 rows {xbar,ybar,zbar,S} for the job's illuminant
    float *film;
};

struct PhotonParams {
    spx_beam beam;
    Proj proj;             // volume: Camera::project terms
    float scr_corner[3], scr_uu;
    float scr_u[3], scr_vv;
    float scr_v[3], pad0;
    float scr_n[3], pad1;  // lighttrace: Screen::normal
    uint32_t scr_w, scr_h;
    uint32_t replicas;     // film copies receiving deposits (>= 1); copy r = film + r * rep_stride
    uint32_t rep_stride;   // floats between copies
    spx_vol_params vp;
    uint32_t w, h, n, photon_base;
    uint32_t seed, estimator, illuminant, dispersion;
    uint32_t max_depth;
    uint32_t key_base;     // binned deposits: first record key of this job's film
    float ill_scale;       // photon power weight S(lambda)/mean S (1 for equal energy)
    uint32_t hero;         // 1: photons carry the 4-wavelength hero comb (spx_opts.hero_lanes = 4)
    const SceneData *scene;
    const float4 *tbl;
    const float *zbuf;     // may be null (= +INF)
    float *film;
};

}  // namespace spx

struct spx_ctx {
    int device = 0;
    int num_cus = 0;
    float4 *tbl_dev = nullptr;  // [2][71]: D65 rows then A rows (w = S)
    float tbl_host[2][spx::kTblN][4];
    uint32_t *quantile_scratch = nullptr;  // the device's copy of the quantile's file-scope scratch array
    double ill_mean[2];
    std::string last_error;
    void *scratch = nullptr;   // grow-only work buffer for the wavefront queues
    size_t scratch_bytes = 0;
    hipStream_t scratch_stream = nullptr;
    float *replica = nullptr;     // grow-only film replica buffer
    size_t replica_floats = 0;
    spx_tuning tuning;
    void *bins[spx::kRoundBufs] = {};  // grow-only binned-deposit buffers (more than one: overlapped rounds)
    size_t bins_bytes[spx::kRoundBufs] = {};
    hipStream_t aux = nullptr;           // second stream of overlapped bin passes
    hipStream_t aux_cam = nullptr;       // stream of a concurrent camera round
    std::vector<hipEvent_t> sync_ev;     // cross-stream ordering events (no timing)
    void *stage = nullptr;        // pinned staging for small stream-ordered uploads
    size_t stage_bytes = 0;
    hipEvent_t stage_done = nullptr;
    bool stage_pending = false;
    void *batch_dev = nullptr;    // grow-only device job table + items + queues
    size_t batch_dev_bytes = 0;
    void *batch_host = nullptr;   // pinned staging for the job table
    size_t batch_host_bytes = 0;
    hipEvent_t batch_done = nullptr;  // recorded after the staging copy
    bool batch_pending = false;
    uint64_t buf_gen = 0;              // bumped whenever bins / replica / batch_dev move
    bool ktime_on = false;             // bracket batched kernel groups with events
    std::vector<hipEvent_t> ktime_ev;  // pool; pairs (start, end)
    std::vector<uint32_t> ktime_kind;  // SPX_KT_* of each recorded pair
    size_t ktime_used = 0;             // events recorded since the last read
};

struct spx_scene {
    spx_ctx *ctx = nullptr;
    spx::SceneData host;
    spx::SceneData *dev = nullptr;
};

namespace spx {

int set_error(spx_ctx *ctx, int code, const std::string &msg);
int check_hip(spx_ctx *ctx, hipError_t e, const char *what);
// Ensures ctx->scratch holds at least `bytes`; synchronises `stream` first
// when it has to grow.
int ensure_scratch(spx_ctx *ctx, size_t bytes, hipStream_t stream);

void make_proj(const spx_camera &c, Proj &P);
float illuminant_scale(const spx_ctx *ctx, uint32_t ill);
const float4 *table_for(const spx_ctx *ctx, uint32_t ill);

// Kernel launchers (return hipGetLastError()).
hipError_t launch_camera_plain(const CamParams &p, hipStream_t s);
hipError_t launch_photon_plain(const PhotonParams &p, int lighttrace, hipStream_t s);
hipError_t launch_photon_records(const PhotonParams &p, uint32_t first, uint32_t count,
                                 uint32_t mode, spx_photon_record *out, hipStream_t s);
hipError_t launch_surface_rim(const SceneData *scene, const spx_camera &cam, uint32_t w, uint32_t h,
                              float *rim, float *depth, hipStream_t s);
hipError_t launch_resolve_srgb8(const float *film, uint32_t w, uint32_t h, float scale,
                                float exposure, uint8_t *rgb, hipStream_t s);
// prism_on_plane's floor shading constants.
struct FloorShade {
    float gain;
    float ambient[3];
    float background[3];
    float tint[3];
};
hipError_t launch_box_blur(const float *in, float *tmp, float *out, uint32_t w, uint32_t h, uint32_t radius,
                           hipStream_t s);
hipError_t launch_floor_surface(const SceneData *scene, const spx_camera &cam, const PhotonParams &F,
                                const float *floor_sum, const FloorShade &sh, uint32_t w, uint32_t h,
                                float *surface, float *depth, hipStream_t s);
hipError_t launch_composite_srgb8(const float *surf, const float *vol, uint32_t w, uint32_t h, float exposure,
                                  uint8_t *rgb, hipStream_t s);
hipError_t launch_display_resolve(const float *vol, const float *surf, uint32_t w, uint32_t h, float k,
                                  int radius, float inv2s2, float exposure, uint8_t *rgba, hipStream_t s);
// The exposure quantile over the file-scope scratch array of k_plain.hip:
// its device address for the current device (0 or a hipError_t), the four
// radix passes with their decisions kept on the device, and the word of the
// array that ends holding the selected key.
int quantile_scratch_address(uint32_t **out);
hipError_t launch_quantile_select_dev(const float *vol, const float *surf, uint32_t n, float k,
                                      unsigned long long rank, uint32_t *q, hipStream_t s);
uint32_t quantile_scratch_key_offset();
hipError_t launch_lum_digit_hist(const float *vol, const float *surf, uint32_t n, float k, uint32_t prefix,
                                 uint32_t mask, uint32_t shift, uint32_t *hist, hipStream_t s);
hipError_t launch_rng_draws(uint32_t key, uint32_t seed, uint32_t n, uint32_t *out, hipStream_t s);
hipError_t launch_glass_n(uint32_t glass, const float *lam, uint32_t n, float *out, hipStream_t s);
hipError_t launch_fill(float *dst, float v, size_t n, hipStream_t s);
hipError_t launch_copy(float *dst, const float *src, size_t n, hipStream_t s);

// Film replicas: photon deposits of one job are spread over R copies (block b
// writes copy b % R) to divide same-pixel atomic contention by R; the copies
// are then summed into the caller's film.
struct ReplicaJob {
    float *dst;         // caller's film
    const float *src;   // R copies of n floats each
    uint32_t n, reps;
    uint32_t first_block, pad;
};
hipError_t launch_reduce_replicas(const ReplicaJob *jobs_dev, uint32_t n_jobs, uint32_t total_blocks,
                                  hipStream_t s);
int ensure_replicas(spx_ctx *ctx, size_t floats, hipStream_t stream);
hipError_t launch_reduce_one(const ReplicaJob &J, hipStream_t s);
// Replica count for a film of n floats under the context's tuning.
uint32_t replica_count(const spx_ctx *ctx, size_t film_floats);
// Binned deposits: per-call job map (key ranges -> films) and the passes.
struct BinJobMap {
    std::vector<uint32_t> key, npix, width, shared, strip_job;
    std::vector<float *> film;
    uint32_t n_strips = 0;
    uint32_t tile_wlog2 = kTileLog2;  // tiles 2^tile_wlog2 x 2^(10 - tile_wlog2)
    explicit BinJobMap(uint32_t tw = kTileLog2) : tile_wlog2(tw) {}
    // Adds a film of w x h pixels; returns its key base.
    uint32_t add(float *f, uint32_t w, uint32_t h) {
        uint32_t k = n_strips << kStripLog2;
        key.push_back(k);
        npix.push_back(w * h);
        width.push_back(w);
        shared.push_back(0);
        film.push_back(f);
        uint32_t strips = film_strips(w, h, tile_wlog2);
        for (uint32_t i = 0; i < strips; ++i) strip_job.push_back((uint32_t)film.size() - 1);
        n_strips += strips;
        return k;
    }
    static uint32_t film_strips(uint32_t w, uint32_t h, uint32_t tw) {
        const uint32_t th = kStripLog2 - tw;
        return ((w + (1u << tw) - 1) >> tw) * ((h + (1u << th) - 1) >> th);
    }
};
// Byte layout of one deposit round's buffers: records, sorted records,
// counters, group lists, split-strip accumulators, then the job map.
struct BinLayout {
    size_t cap = 0, n = 0, nj = 0, zero_words = 0, max_multi = 0;
    size_t o_recs = 0, o_vals = 0, o_sorted = 0, o_pool = 0, o_partn = 0, o_fill = 0, o_zero = 0, o_tot = 0;
    size_t o_off = 0, o_poff = 0, o_glist = 0, o_gcount = 0, o_macc = 0, o_smacc = 0, o_maccs = 0, o_maccn = 0;
    size_t o_map = 0, m_key = 0, m_npix = 0, m_w = 0, m_sh = 0, m_film = 0, m_end = 0;
    size_t total = 0;
};
size_t a256(size_t x);
// Layout for `map` with room for `want` records (capacity: a power of two
// within [2^15, 2^bin_log2]).
BinLayout bins_layout(const BinJobMap &map, size_t want, uint32_t bin_log2);
// Points `st` / `job_shared` into the buffer at `base`.
void bins_bind(char *base, const BinLayout &L, const BinJobMap &map, BinState &st, const uint32_t **job_shared);
// Stream-ordered upload of the job map into the buffer at `base`.
int bins_upload_map(spx_ctx *ctx, char *base, const BinLayout &L, const BinJobMap &map, hipStream_t s);
// The above on the context's grow-only buffer.
int prepare_bins(spx_ctx *ctx, const BinJobMap &map, size_t want, hipStream_t s, BinState &st,
                 const uint32_t **job_shared, int slot = 0);
// Share of the record capacity a round is planned to fill.
constexpr double kBinFill = 0.85;
// Expected film deposits per photon: the prism-and-floor volume renders
// average 13.9 (point estimator), screen photons under 4.
inline double records_per_photon(bool lighttrace) { return lighttrace ? 4.0 : 14.0; }
hipError_t launch_bins_reset(const BinState &st, hipStream_t s);
hipError_t launch_bins_resolve(const BinState &st, const uint32_t *job_shared, uint32_t acc_blocks, hipStream_t s);
// Stream-ordered upload of a small host block through pinned staging.
int upload_async(spx_ctx *ctx, void *dev, const void *host, size_t bytes, hipStream_t s);

// Validates a batch job exactly as spx_batch_run would (no launch).
int validate_job(spx_ctx *ctx, const spx_job &job);
// A batch plan on the context's own grow-only buffers (no allocation when
// they are large enough): for callers that update the plan before every
// launch (the server).  Launching it after the context reallocated those
// buffers, or after another batch ran on the context without an update in
// between, is invalid; spx_batch_plan_update refreshes the uploads and
// fails with SPX_EINVAL once the buffers have moved.
int batch_plan_create_shared(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, hipStream_t s,
                             spx_batch_plan **out);
// Records per frame a photon job is expected to deposit (its queue share).
double job_queue_bytes(const spx_ctx *ctx, const spx_job &job);

// Argument checks and parameter blocks shared by the entry points.
bool valid_opts(const spx_opts *o, bool camera);
spx_opts default_opts(bool camera);
bool finite3(const float *v);
int check_film(spx_ctx *ctx, const float *film, uint32_t w, uint32_t h, const char *who);
int check_screen(spx_ctx *ctx, const spx_screen &scr, const char *who);
int photon_params(spx_ctx *ctx, const spx_scene *scene, const spx_beam *beam, const spx_opts *opts, uint32_t n,
                  uint32_t photon_base, uint32_t seed, PhotonParams &p, const char *who);
void set_screen(PhotonParams &p, const spx_screen &scr);
int atomic_probe(spx_ctx *ctx, float *buf, size_t n_slots, uint32_t ops, uint32_t threads, hipStream_t s);

// Wavefront engine (per-bounce SoA queues).
size_t camera_wavefront_scratch(uint32_t chunk);
hipError_t run_camera_wavefront(spx_ctx *ctx, const CamParams &p, void *scratch, uint32_t chunk,
                                hipStream_t s);
size_t photon_wavefront_scratch(uint32_t chunk);
size_t beam_wavefront_scratch(uint32_t chunk, uint32_t max_depth);
hipError_t run_beam_wavefront(spx_ctx *ctx, const PhotonParams &p, void *scratch, uint32_t chunk, hipStream_t s);
// bins: binned deposits (hash: an LDS hash in front of the records), else
// the LDS hash draining to film atomics.
hipError_t run_photon_wavefront(spx_ctx *ctx, const PhotonParams &p, int lighttrace, void *scratch,
                                uint32_t chunk, hipStream_t s, const BinState *bins,
                                const uint32_t *job_shared, bool hash = false);

}  // namespace spx
