// spx_api.hip -- the C ABI (include/spectral_hip.h): contexts, scenes,
// validation and dispatch to the plain / wavefront / batched kernels.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include <algorithm>
#include <map>
#include <vector>

#include "spx_batch.hpp"
#include "spx_internal.hpp"
#include "_gen_cie.inc"

using namespace spx;

namespace spx {

int set_error(spx_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int check_hip(spx_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return SPX_OK;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, e == hipErrorOutOfMemory ? SPX_ENOMEM : SPX_EDEVICE, m);
}

int ensure_scratch(spx_ctx *ctx, size_t bytes, hipStream_t stream) {
    if (ctx->scratch_bytes >= bytes) return SPX_OK;
    if (ctx->scratch) {
        // Work queued on any stream may still read the old buffer.
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return check_hip(ctx, e, "hipDeviceSynchronize");
        hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    (void)stream;
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&ctx->scratch, want);
    if (e != hipSuccess) return check_hip(ctx, e, "hipMalloc(scratch)");
    ctx->scratch_bytes = want;
    return SPX_OK;
}

// Camera::project's per-camera terms, camera.rs:51-56 (camera_proj_params,
// vol_photons.rs:100-115).
void make_proj(const spx_camera &c, Proj &P) {
    f3 hz = ld3(c.horizontal), vt = ld3(c.vertical), o = ld3(c.origin);
    f3 u = normalize(hz), v = normalize(vt);
    f3 w = o - hz * 0.5f - vt * 0.5f - ld3(c.lower_left);
    P.o[0] = o.x; P.o[1] = o.y; P.o[2] = o.z;
    P.u[0] = u.x; P.u[1] = u.y; P.u[2] = u.z;
    P.v[0] = v.x; P.v[1] = v.y; P.v[2] = v.z;
    P.w[0] = w.x; P.w[1] = w.y; P.w[2] = w.z;
    P.vw = length(hz);
    P.vh = length(vt);
    P.pad[0] = P.pad[1] = 0.0f;
}

int ensure_replicas(spx_ctx *ctx, size_t floats, hipStream_t stream) {
    if (ctx->replica_floats >= floats) return SPX_OK;
    if (ctx->replica) {
        hipError_t e = hipDeviceSynchronize();  // earlier launches may still use the copies
        if (e != hipSuccess) return check_hip(ctx, e, "hipDeviceSynchronize");
        hipFree(ctx->replica);
        ctx->replica = nullptr;
        ctx->replica_floats = 0;
    }
    (void)stream;
    size_t want = floats + floats / 4;
    hipError_t e = hipMalloc(&ctx->replica, want * sizeof(float));
    ctx->buf_gen++;
    if (e != hipSuccess) return check_hip(ctx, e, "hipMalloc(replicas)");
    ctx->replica_floats = want;
    return SPX_OK;
}

uint32_t replica_count(const spx_ctx *ctx, size_t film_floats) {
    const spx_tuning &t = ctx->tuning;
    if (t.max_replicas <= 1 || film_floats == 0) return 1;
    size_t budget = (size_t)t.replica_mib << 20;
    size_t r = budget / (film_floats * sizeof(float));
    if (r > t.max_replicas) r = t.max_replicas;
    return r < 2 ? 1u : (uint32_t)r;
}

int upload_async(spx_ctx *ctx, void *dev, const void *host, size_t bytes, hipStream_t s) {
    if (bytes == 0) return SPX_OK;
    if (ctx->stage_pending) {
        hipError_t e = hipEventSynchronize(ctx->stage_done);
        if (e != hipSuccess) return check_hip(ctx, e, "upload wait");
        ctx->stage_pending = false;
    }
    if (!ctx->stage_done) {
        hipError_t e = hipEventCreateWithFlags(&ctx->stage_done, hipEventDisableTiming);
        if (e != hipSuccess) return check_hip(ctx, e, "upload event");
    }
    if (ctx->stage_bytes < bytes) {
        if (ctx->stage) hipHostFree(ctx->stage);
        ctx->stage = nullptr;
        ctx->stage_bytes = 0;
        hipError_t e = hipHostMalloc(&ctx->stage, bytes * 2, hipHostMallocDefault);
        if (e != hipSuccess) return check_hip(ctx, e, "upload staging");
        ctx->stage_bytes = bytes * 2;
    }
    memcpy(ctx->stage, host, bytes);
    hipError_t e = hipMemcpyAsync(dev, ctx->stage, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(ctx->stage_done, s);
    if (e != hipSuccess) return check_hip(ctx, e, "upload copy");
    ctx->stage_pending = true;
    return SPX_OK;
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

BinLayout bins_layout(const BinJobMap &map, size_t want, uint32_t bin_log2) {
    BinLayout L;
    // Capacity: the smallest power of two >= want, within [2^15, 2^bin_log2].
    size_t cap = (size_t)1 << 15;
    while (cap < want && cap < ((size_t)1 << bin_log2)) cap <<= 1;
    L.cap = cap;
    const size_t chunks = cap / kBinChunk;
    const size_t n = map.n_strips, nj = map.film.size();
    L.n = n;
    L.nj = nj;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += a256(bytes); return o; };
    L.o_recs = take(cap * 4);       // keys
    L.o_vals = take(cap * 12);      // values
    L.o_sorted = take(sorted_capacity(cap, n) * sizeof(BinRec));
    const size_t groups = cap / kGroupRecs;
    L.zero_words = n * (kSub + 3);  // sub_count | ovf_count | ovf_cur | strip_max
    L.o_pool = take(kBinPools * 4);
    L.o_partn = take(4);
    L.o_fill = take(chunks * 4);
    L.o_zero = take(L.zero_words * 4);
    L.o_tot = take(n * 4);
    L.o_off = take((n + 1) * 4);
    L.o_poff = take((n + 1) * 4);
    L.o_glist = take(groups * kGroupSlots * sizeof(uint2));
    L.o_gcount = take(groups * 4);
    // A split strip holds more than kPartRecs records, so at most cap / kPartRecs exist.
    L.max_multi = cap / kPartRecs + 1;
    L.o_macc = take(L.max_multi * 3 * kStripPix * 8);
    L.o_smacc = take(n * 4);
    L.o_maccs = take(L.max_multi * 4);
    L.o_maccn = take(4);
    // the uploaded map: strip_job | job_key | job_npix | job_width | job_shared | job_film
    L.o_map = off;
    L.m_key = a256(n * 4);
    L.m_npix = L.m_key + a256(nj * 4);
    L.m_w = L.m_npix + a256(nj * 4);
    L.m_sh = L.m_w + a256(nj * 4);
    L.m_film = L.m_sh + a256(nj * 4);
    L.m_end = L.m_film + a256(nj * sizeof(float *));
    L.total = off + L.m_end;
    return L;
}

void bins_bind(char *b, const BinLayout &L, const BinJobMap &map, BinState &st, const uint32_t **job_shared) {
    const size_t n = L.n;
    st.rkey = (uint32_t *)(b + L.o_recs);
    st.rval = (float *)(b + L.o_vals);
    st.sorted = (BinRec *)(b + L.o_sorted);
    st.pool_next = (uint32_t *)(b + L.o_pool);
    st.chunk_fill = (uint32_t *)(b + L.o_fill);
    uint32_t *z = (uint32_t *)(b + L.o_zero);
    st.sub_count = z;
    st.ovf_count = z + n * kSub;
    st.ovf_cur = st.ovf_count + n;
    st.strip_max = st.ovf_cur + n;
    st.zero_words = L.zero_words;
    st.strip_total = (uint32_t *)(b + L.o_tot);
    st.strip_off = (uint32_t *)(b + L.o_off);
    st.part_off = (uint32_t *)(b + L.o_poff);
    st.glist = (uint2 *)(b + L.o_glist);
    st.gcount = (uint32_t *)(b + L.o_gcount);
    st.macc = (unsigned long long *)(b + L.o_macc);
    st.strip_macc = (uint32_t *)(b + L.o_smacc);
    st.macc_strip = (uint32_t *)(b + L.o_maccs);
    st.macc_count = (uint32_t *)(b + L.o_maccn);
    st.max_multi = (uint32_t)L.max_multi;
    st.part_next = (uint32_t *)(b + L.o_partn);
    st.strip_job = (const uint32_t *)(b + L.o_map);
    st.job_key = (const uint32_t *)(b + L.o_map + L.m_key);
    st.job_npix = (const uint32_t *)(b + L.o_map + L.m_npix);
    st.job_width = (const uint32_t *)(b + L.o_map + L.m_w);
    st.job_film = (float *const *)(b + L.o_map + L.m_film);
    st.pool_cap = (uint32_t)(L.cap / kBinPools);
    st.n_strips = (uint32_t)n;
    st.tile_wlog2 = map.tile_wlog2;
    st.atomic_film = 0;
    *job_shared = (const uint32_t *)(b + L.o_map + L.m_sh);
}

int bins_upload_map(spx_ctx *ctx, char *b, const BinLayout &L, const BinJobMap &map, hipStream_t s) {
    const size_t n = L.n, nj = L.nj;
    std::vector<char> host(L.m_end, 0);
    memcpy(host.data(), map.strip_job.data(), n * 4);
    memcpy(host.data() + L.m_key, map.key.data(), nj * 4);
    memcpy(host.data() + L.m_npix, map.npix.data(), nj * 4);
    memcpy(host.data() + L.m_w, map.width.data(), nj * 4);
    memcpy(host.data() + L.m_sh, map.shared.data(), nj * 4);
    memcpy(host.data() + L.m_film, map.film.data(), nj * sizeof(float *));
    return upload_async(ctx, b + L.o_map, host.data(), L.m_end, s);
}

int prepare_bins(spx_ctx *ctx, const BinJobMap &map, size_t want, hipStream_t s, BinState &st,
                 const uint32_t **job_shared, int slot) {
    const BinLayout L = bins_layout(map, want, ctx->tuning.bin_log2);
    void *&buf = ctx->bins[slot];
    size_t &bytes = ctx->bins_bytes[slot];
    if (bytes < L.total) {
        hipError_t e = hipDeviceSynchronize();  // earlier rounds may still use the old buffers
        if (e != hipSuccess) return check_hip(ctx, e, "prepare_bins sync");
        if (buf) hipFree(buf);
        buf = nullptr;
        bytes = 0;
        size_t want_bytes = L.total + L.total / 8;
        e = hipMalloc(&buf, want_bytes);
        ctx->buf_gen++;
        if (e != hipSuccess) return check_hip(ctx, e, "hipMalloc(bins)");
        bytes = want_bytes;
    }
    char *b = (char *)buf;
    int rc = bins_upload_map(ctx, b, L, map, s);
    if (rc) return rc;
    bins_bind(b, L, map, st, job_shared);
    return SPX_OK;
}

// Single-job photon call through binned deposits; `photons` is the most
// one deposit round holds (the wavefront resolves once per chunk).
template <class F>
static hipError_t with_bins(spx_ctx *ctx, PhotonParams &p, uint32_t w, uint32_t h, uint32_t photons, int lighttrace,
                            hipStream_t s, F launch, int *rc) {
    BinJobMap map(ctx->tuning.tile_wlog2);
    p.key_base = map.add(p.film, w, h);
    p.replicas = 1;
    p.rep_stride = 0;
    BinState st;
    const uint32_t *shared = nullptr;
    const size_t want = (size_t)((double)photons * records_per_photon(lighttrace) / kBinFill);
    if ((*rc = prepare_bins(ctx, map, want, s, st, &shared))) return hipSuccess;
    return launch(p, &st, shared);
}

// Runs `launch` with deposits spread over film replicas, then sums them
// into p.film.  Falls back to the film itself when one copy is chosen.
template <class F>
static hipError_t with_replicas(spx_ctx *ctx, PhotonParams &p, size_t film_floats, hipStream_t s, F launch,
                                int *rc) {
    *rc = SPX_OK;
    uint32_t R = replica_count(ctx, film_floats);
    p.replicas = 1;
    p.rep_stride = 0;
    if (R <= 1) return launch(p);
    if ((*rc = ensure_replicas(ctx, R * film_floats, s))) return hipSuccess;
    float *film = p.film;
    hipError_t e = launch_fill(ctx->replica, 0.0f, R * film_floats, s);
    if (e != hipSuccess) return e;
    p.film = ctx->replica;
    p.replicas = R;
    p.rep_stride = (uint32_t)film_floats;
    e = launch(p);
    if (e != hipSuccess) return e;
    ReplicaJob J;
    J.dst = film;
    J.src = ctx->replica;
    J.n = (uint32_t)film_floats;
    J.reps = R;
    J.first_block = 0;
    J.pad = 0;
    p.film = film;
    return launch_reduce_one(J, s);
}

float illuminant_scale(const spx_ctx *ctx, uint32_t ill) {
    if (ill > SPX_ILLUM_A) return 1.0f;
    return (float)(1.0 / ctx->ill_mean[ill]);
}

const float4 *table_for(const spx_ctx *ctx, uint32_t ill) {
    return ctx->tbl_dev + (ill == SPX_ILLUM_A ? kTblN : 0);
}

}  // namespace spx

// ---------------------------------------------------------------------------
// Table::parse (cie.rs:24-45) over the embedded CSV text.
static int parse_table(const char *text, int cols, float *dst /* [71][cols] */) {
    const char *p = text;
    int n = 0;
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
        if (!*p) break;
        if (n >= kTblN) return -1;
        char *e;
        (void)strtof(p, &e);
        p = e;
        for (int c = 0; c < cols; ++c) {
            while (*p == ',' || *p == ' ') ++p;
            dst[n * cols + c] = strtof(p, &e);
            p = e;
        }
        while (*p && *p != '\n') ++p;
        ++n;
    }
    return n == kTblN ? 0 : -2;
}

bool spx::valid_opts(const spx_opts *o, bool camera) {
    if (o->max_depth < 1 || o->max_depth > 32) return false;
    if (o->illuminant > SPX_ILLUM_E) return false;
    if (o->dispersion > 1) return false;
    if (o->hero_lanes != 0 && o->hero_lanes != 1 && o->hero_lanes != 4) return false;
    (void)camera;
    return true;
}

spx_opts spx::default_opts(bool camera) {
    spx_opts o;
    o.illuminant = camera ? SPX_ILLUM_D65 : SPX_ILLUM_E;
    o.dispersion = 1;
    o.max_depth = 8;
    o.rr_depth = 0;
    o.hero_lanes = 0;
    return o;
}

bool spx::finite3(const float *v) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }

int spx::check_film(spx_ctx *ctx, const float *film, uint32_t w, uint32_t h, const char *who) {
    if (!film) return set_error(ctx, SPX_EINVAL, std::string(who) + ": null film");
    if (w == 0 || h == 0) return set_error(ctx, SPX_EINVAL, std::string(who) + ": zero-size film");
    if ((uint64_t)w * h > (1ull << 28))
        return set_error(ctx, SPX_ECAP, std::string(who) + ": film larger than 2^28 pixels");
    return SPX_OK;
}

int spx::photon_params(spx_ctx *ctx, const spx_scene *scene, const spx_beam *beam, const spx_opts *opts,
                       uint32_t n, uint32_t photon_base, uint32_t seed, PhotonParams &p, const char *who) {
    if (!beam) return set_error(ctx, SPX_EINVAL, std::string(who) + ": null beam");
    if (!finite3(beam->corner) || !finite3(beam->u) || !finite3(beam->v) || !finite3(beam->dir))
        return set_error(ctx, SPX_EINVAL, std::string(who) + ": non-finite beam");
    spx_opts o = opts ? *opts : default_opts(false);
    if (!valid_opts(&o, false)) return set_error(ctx, SPX_EINVAL, std::string(who) + ": bad options");
    memset(&p, 0, sizeof p);
    p.beam = *beam;
    p.n = n;
    p.photon_base = photon_base;
    p.seed = seed;
    p.illuminant = o.illuminant;
    p.dispersion = o.dispersion;
    p.max_depth = o.max_depth;
    p.ill_scale = illuminant_scale(ctx, o.illuminant);
    p.hero = o.hero_lanes == 4 ? 1u : 0u;
    p.replicas = 1;
    p.rep_stride = 0;
    p.scene = scene->dev;
    p.tbl = table_for(ctx, o.illuminant);
    return SPX_OK;
}

// A screen whose rectangle is finite and spans an area (Screen::new,
// lighttrace.rs:27-34, divides by |u|^2 and |v|^2 and normalises u x v).
int spx::check_screen(spx_ctx *ctx, const spx_screen &scr, const char *who) {
    if (!finite3(scr.corner) || !finite3(scr.u) || !finite3(scr.v))
        return set_error(ctx, SPX_EINVAL, std::string(who) + ": non-finite screen");
    PhotonParams p{};
    set_screen(p, scr);
    if (!(p.scr_uu > 0.0f) || !(p.scr_vv > 0.0f) || !finite3(p.scr_n))
        return set_error(ctx, SPX_EINVAL, std::string(who) + ": degenerate screen rectangle");
    return SPX_OK;
}

void spx::set_screen(PhotonParams &p, const spx_screen &scr) {
    // Screen::new, lighttrace.rs:27-34
    f3 u = ld3(scr.u), v = ld3(scr.v), n = normalize(cross(u, v));
    memcpy(p.scr_corner, scr.corner, sizeof p.scr_corner);
    memcpy(p.scr_u, scr.u, sizeof p.scr_u);
    memcpy(p.scr_v, scr.v, sizeof p.scr_v);
    p.scr_n[0] = n.x; p.scr_n[1] = n.y; p.scr_n[2] = n.z;
    p.scr_uu = dot(u, u);
    p.scr_vv = dot(v, v);
    p.scr_w = scr.width;
    p.scr_h = scr.height;
}

extern "C" {

int spx_ctx_create(int device, spx_ctx **out) {
    if (!out) return SPX_EINVAL;
    *out = nullptr;
    spx_ctx *ctx = new spx_ctx();
    ctx->device = device;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || device < 0 || device >= ndev) {
        delete ctx;
        return e != hipSuccess ? SPX_EDEVICE : SPX_EINVAL;
    }
    if ((e = hipSetDevice(device)) != hipSuccess) {
        delete ctx;
        return SPX_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    // this device's address of the exposure quantile's scratch array (the
    // entry point that allocates nothing uses it)
    if (quantile_scratch_address(&ctx->quantile_scratch) != 0 || !ctx->quantile_scratch) {
        delete ctx;
        return SPX_EDEVICE;
    }
    float cmf[kTblN][3], d65[kTblN], a[kTblN];
    if (parse_table(kCmfCsv, 3, &cmf[0][0]) || parse_table(kD65Csv, 1, d65) || parse_table(kACsv, 1, a)) {
        delete ctx;
        return SPX_EINVAL;
    }
    double s65 = 0.0, sa = 0.0;
    for (int i = 0; i < kTblN; ++i) {
        for (int t = 0; t < 2; ++t) {
            ctx->tbl_host[t][i][0] = cmf[i][0];
            ctx->tbl_host[t][i][1] = cmf[i][1];
            ctx->tbl_host[t][i][2] = cmf[i][2];
        }
        ctx->tbl_host[0][i][3] = d65[i];
        ctx->tbl_host[1][i][3] = a[i];
        s65 += (double)d65[i];
        sa += (double)a[i];
    }
    ctx->ill_mean[0] = s65 / (double)kTblN;
    ctx->ill_mean[1] = sa / (double)kTblN;
    if ((e = hipMalloc(&ctx->tbl_dev, sizeof(ctx->tbl_host))) != hipSuccess ||
        (e = hipMemcpy(ctx->tbl_dev, ctx->tbl_host, sizeof(ctx->tbl_host), hipMemcpyHostToDevice)) !=
            hipSuccess) {
        if (ctx->tbl_dev) hipFree(ctx->tbl_dev);
        delete ctx;
        return e == hipErrorOutOfMemory ? SPX_ENOMEM : SPX_EDEVICE;
    }
    ctx->tuning.max_replicas = 16;
    ctx->tuning.replica_mib = 64;
    ctx->tuning.splat_mode = 2;
    ctx->tuning.chunk_log2 = 22;
    ctx->tuning.bin_log2 = 30;
    ctx->tuning.tile_wlog2 = kTileLog2;
    ctx->tuning.batch_blocks = 0;
    ctx->tuning.chunk_paths = 16384;
    ctx->tuning.chunk_photons = 8192;
    ctx->tuning.overlap_bins = 0;
    ctx->tuning.camera_blocks = 0;
    ctx->tuning.photon_threads = 256;
    *out = ctx;
    return SPX_OK;
}

int spx_ctx_set_tuning(spx_ctx *ctx, const spx_tuning *t) {
    if (!ctx || !t) return SPX_EINVAL;
    if (t->max_replicas < 1 || t->max_replicas > 256 || t->splat_mode > 3 || t->chunk_log2 < 12 ||
        t->chunk_log2 > 26 || t->replica_mib > 65536 || t->bin_log2 < 15 || t->bin_log2 > 30 ||
        t->tile_wlog2 < kTileLog2Min || t->tile_wlog2 > kTileLog2Max || t->batch_blocks > 65536 ||
        t->chunk_paths < 64 || t->chunk_paths > (1u << 22) || t->chunk_photons < 64 ||
        t->chunk_photons > (1u << 24) || t->overlap_bins >= (uint32_t)kRoundBufs || t->camera_blocks > 65536 ||
        (t->photon_threads != 256 && t->photon_threads != 1024))
        return set_error(ctx, SPX_EINVAL, "spx_ctx_set_tuning: value out of range");
    ctx->tuning = *t;
    return SPX_OK;
}

int spx_ctx_get_tuning(const spx_ctx *ctx, spx_tuning *t) {
    if (!ctx || !t) return SPX_EINVAL;
    *t = ctx->tuning;
    return SPX_OK;
}

int spx_atomic_probe(spx_ctx *ctx, float *buf, size_t n_slots, uint32_t ops, uint32_t threads, void *stream) {
    if (!ctx || !buf || n_slots == 0 || n_slots > 0xFFFFFFFFu || threads == 0)
        return set_error(ctx, SPX_EINVAL, "spx_atomic_probe: bad arguments");
    return atomic_probe(ctx, buf, n_slots, ops, threads, (hipStream_t)stream);
}

void spx_ctx_destroy(spx_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    if (ctx->scratch) hipFree(ctx->scratch);
    if (ctx->replica) hipFree(ctx->replica);
    for (int k = 0; k < kRoundBufs; ++k)
        if (ctx->bins[k]) hipFree(ctx->bins[k]);
    if (ctx->aux) hipStreamDestroy(ctx->aux);
    if (ctx->aux_cam) hipStreamDestroy(ctx->aux_cam);
    for (hipEvent_t ev : ctx->sync_ev) hipEventDestroy(ev);
    if (ctx->stage) hipHostFree(ctx->stage);
    if (ctx->stage_done) hipEventDestroy(ctx->stage_done);
    if (ctx->batch_dev) hipFree(ctx->batch_dev);
    if (ctx->batch_host) hipHostFree(ctx->batch_host);
    if (ctx->batch_done) hipEventDestroy(ctx->batch_done);
    for (hipEvent_t ev : ctx->ktime_ev) hipEventDestroy(ev);
    if (ctx->tbl_dev) hipFree(ctx->tbl_dev);
    delete ctx;
}

int spx_ctx_kernel_timing(spx_ctx *ctx, int enable) {
    if (!ctx) return SPX_EINVAL;
    ctx->ktime_on = enable != 0;
    ctx->ktime_used = 0;
    return SPX_OK;
}

int spx_ctx_kernel_times(spx_ctx *ctx, double ms[3], uint32_t launches[3]) {
    if (!ctx || !ms || !launches) return set_error(ctx, SPX_EINVAL, "spx_ctx_kernel_times: null argument");
    for (int g = 0; g < 3; ++g) {
        ms[g] = 0.0;
        launches[g] = 0;
    }
    for (size_t k = 0; k + 1 < ctx->ktime_used; k += 2) {
        hipError_t e = hipEventSynchronize(ctx->ktime_ev[k + 1]);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ktime_ev[k], ctx->ktime_ev[k + 1]);
        if (e != hipSuccess) {
            ctx->ktime_used = 0;
            return check_hip(ctx, e, "spx_ctx_kernel_times");
        }
        uint32_t g = ctx->ktime_kind[k / 2];
        ms[g] += t;
        launches[g] += 1;
    }
    ctx->ktime_used = 0;
    return SPX_OK;
}

int spx_ctx_kernel_time(spx_ctx *ctx, double *total_ms, uint32_t *launches) {
    if (!ctx || !total_ms || !launches) return set_error(ctx, SPX_EINVAL, "spx_ctx_kernel_time: null argument");
    double ms[3];
    uint32_t n[3];
    int rc = spx_ctx_kernel_times(ctx, ms, n);
    if (rc) return rc;
    *total_ms = ms[SPX_KT_CAMERA] + ms[SPX_KT_PHOTON];
    *launches = n[SPX_KT_CAMERA] + n[SPX_KT_PHOTON];
    return SPX_OK;
}

const char *spx_last_error(const spx_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int spx_camera_look_at(const float origin[3], const float target[3], const float up[3], float vfov_deg,
                       float aspect, spx_camera *out) {
    if (!origin || !target || !up || !out) return SPX_EINVAL;
    const float rads_per_deg = 3.14159265358979323846f / 180.0f;  // f32::to_radians
    f3 o = ld3(origin), tg = ld3(target), up3 = ld3(up);
    float theta = vfov_deg * rads_per_deg;
    float hh = tanf(theta / 2.0f);
    float viewport_h = 2.0f * hh;
    float viewport_w = aspect * viewport_h;
    f3 w = normalize(o - tg);
    f3 u = normalize(cross(up3, w));
    f3 v = cross(w, u);
    f3 horizontal = u * viewport_w;
    f3 vertical = v * viewport_h;
    f3 ll = o - horizontal / 2.0f - vertical / 2.0f - w;
    const f3 *src[4] = {&o, &ll, &horizontal, &vertical};
    float *dst[4] = {out->origin, out->lower_left, out->horizontal, out->vertical};
    for (int k = 0; k < 4; ++k) {
        dst[k][0] = src[k]->x;
        dst[k][1] = src[k]->y;
        dst[k][2] = src[k]->z;
    }
    if (!finite3(out->origin) || !finite3(out->lower_left) || !finite3(out->horizontal) ||
        !finite3(out->vertical))
        return SPX_EINVAL;
    return SPX_OK;
}

int spx_ctx_num_cus(const spx_ctx *ctx) { return ctx ? ctx->num_cus : 0; }

int spx_scene_create(spx_ctx *ctx, const spx_env *env, const spx_prim *prims, uint32_t n_prims,
                     const spx_plane *planes, uint32_t n_planes, const spx_material *mats, uint32_t n_mats,
                     spx_scene **out) {
    if (!ctx || !out || !env) return set_error(ctx, SPX_EINVAL, "spx_scene_create: null argument");
    *out = nullptr;
    if (n_prims > SPX_MAX_PRIMS || n_planes > SPX_MAX_PLANES || n_mats > SPX_MAX_MATERIALS)
        return set_error(ctx, SPX_ECAP, "spx_scene_create: scene exceeds the LDS capacity");
    if ((n_prims && !prims) || (n_planes && !planes) || (n_mats && !mats))
        return set_error(ctx, SPX_EINVAL, "spx_scene_create: null array with nonzero count");
    for (uint32_t i = 0; i < n_prims; ++i) {
        const spx_prim &p = prims[i];
        if (p.kind > SPX_PRIM_SOLID || p.material >= n_mats)
            return set_error(ctx, SPX_EINVAL, "spx_scene_create: bad primitive kind or material");
        if (p.kind == SPX_PRIM_SOLID && ((uint64_t)p.plane_start + p.plane_count > n_planes))
            return set_error(ctx, SPX_EINVAL, "spx_scene_create: plane range out of bounds");
    }
    for (uint32_t i = 0; i < n_mats; ++i)
        if (mats[i].kind > SPX_MAT_DIELECTRIC || mats[i].glass > SPX_GLASS_WATER)
            return set_error(ctx, SPX_EINVAL, "spx_scene_create: bad material");
    spx_scene *sc = new spx_scene();
    sc->ctx = ctx;
    memset(&sc->host, 0, sizeof sc->host);
    sc->host.n_prims = n_prims;
    sc->host.n_planes = n_planes;
    sc->host.n_mats = n_mats;
    sc->host.env = *env;
    if (n_prims) memcpy(sc->host.prims, prims, n_prims * sizeof(spx_prim));
    if (n_planes) memcpy(sc->host.planes, planes, n_planes * sizeof(spx_plane));
    if (n_mats) memcpy(sc->host.mats, mats, n_mats * sizeof(spx_material));
    hipError_t e = hipMalloc(&sc->dev, sizeof(SceneData));
    if (e == hipSuccess) e = hipMemcpy(sc->dev, &sc->host, sizeof(SceneData), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (sc->dev) hipFree(sc->dev);
        delete sc;
        return check_hip(ctx, e, "spx_scene_create upload");
    }
    *out = sc;
    return SPX_OK;
}

void spx_scene_destroy(spx_scene *scene) {
    if (!scene) return;
    hipDeviceSynchronize();
    if (scene->dev) hipFree(scene->dev);
    delete scene;
}

static int check_common(spx_ctx *ctx, const spx_scene *scene, const char *who) {
    if (!ctx) return SPX_EINVAL;
    if (!scene || scene->ctx != ctx)
        return set_error(ctx, SPX_EINVAL, std::string(who) + ": scene is null or from another context");
    return SPX_OK;
}


int spx_camera_accumulate(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam, uint32_t width,
                          uint32_t height, uint32_t spp, uint32_t sample_base, uint32_t seed,
                          const spx_opts *opts, uint32_t engine, float *film_sum, void *stream) {
    int rc = check_common(ctx, scene, "spx_camera_accumulate");
    if (rc) return rc;
    if (!cam) return set_error(ctx, SPX_EINVAL, "spx_camera_accumulate: null camera");
    if ((rc = check_film(ctx, film_sum, width, height, "spx_camera_accumulate"))) return rc;
    spx_opts o = opts ? *opts : default_opts(true);
    if (!valid_opts(&o, true)) return set_error(ctx, SPX_EINVAL, "spx_camera_accumulate: bad options");
    if (!finite3(cam->origin) || !finite3(cam->lower_left) || !finite3(cam->horizontal) ||
        !finite3(cam->vertical))
        return set_error(ctx, SPX_EINVAL, "spx_camera_accumulate: non-finite camera");
    if (spp == 0) return SPX_OK;
    CamParams p;
    memset(&p, 0, sizeof p);
    p.cam = *cam;
    p.w = width; p.h = height; p.spp = spp; p.sample_base = sample_base; p.seed = seed;
    p.illuminant = o.illuminant; p.dispersion = o.dispersion; p.max_depth = o.max_depth;
    p.rr_depth = o.rr_depth;
    p.scene = scene->dev;
    p.tbl = table_for(ctx, o.illuminant);
    p.film = film_sum;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (engine == SPX_ENGINE_PLAIN) {
        e = launch_camera_plain(p, s);
    } else if (engine == SPX_ENGINE_WAVEFRONT) {
        // up to 2^24 paths per chunk (two queues of 10 words per path and
        // 12 B of results: 1.5 GB), at least one pixel's samples
        const uint64_t paths = (uint64_t)width * height * spp;
        uint32_t chunk = (uint32_t)std::min<uint64_t>(paths, 1u << 24);
        chunk = chunk > spp ? chunk : spp;
        if ((rc = ensure_scratch(ctx, camera_wavefront_scratch(chunk), s))) return rc;
        e = run_camera_wavefront(ctx, p, ctx->scratch, chunk, s);
    } else {
        return set_error(ctx, SPX_EINVAL, "spx_camera_accumulate: unknown engine");
    }
    return check_hip(ctx, e, "spx_camera_accumulate launch");
}



int spx_lighttrace(spx_ctx *ctx, const spx_scene *scene, const spx_screen *screen, const spx_beam *beam,
                   uint32_t n, uint32_t photon_base, uint32_t seed, const spx_opts *opts, uint32_t engine,
                   float *screen_sum, void *stream) {
    int rc = check_common(ctx, scene, "spx_lighttrace");
    if (rc) return rc;
    if (!screen) return set_error(ctx, SPX_EINVAL, "spx_lighttrace: null screen");
    if ((rc = check_film(ctx, screen_sum, screen->width, screen->height, "spx_lighttrace"))) return rc;
    if ((rc = check_screen(ctx, *screen, "spx_lighttrace"))) return rc;
    PhotonParams p;
    if ((rc = photon_params(ctx, scene, beam, opts, n, photon_base, seed, p, "spx_lighttrace"))) return rc;
    set_screen(p, *screen);
    p.film = screen_sum;
    if (n == 0) return SPX_OK;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (engine == SPX_ENGINE_PLAIN) {
        e = launch_photon_plain(p, 1, s);
    } else if (engine == SPX_ENGINE_WAVEFRONT) {
        uint32_t chunk = 1u << ctx->tuning.chunk_log2;
        if ((rc = ensure_scratch(ctx, photon_wavefront_scratch(chunk), s))) return rc;
        if (ctx->tuning.splat_mode >= 2)
            e = with_bins(ctx, p, screen->width, screen->height, std::min(n, chunk), 1, s,
                          [&](const PhotonParams &q, const BinState *b, const uint32_t *sh) {
                              return run_photon_wavefront(ctx, q, 1, ctx->scratch, chunk, s, b, sh,
                                                          ctx->tuning.splat_mode == 3);
                          }, &rc);
        else
            e = with_replicas(ctx, p, (size_t)3 * screen->width * screen->height, s,
                              [&](const PhotonParams &q) {
                                  return run_photon_wavefront(ctx, q, 1, ctx->scratch, chunk, s, nullptr, nullptr);
                              }, &rc);
        if (rc) return rc;
    } else {
        return set_error(ctx, SPX_EINVAL, "spx_lighttrace: unknown engine");
    }
    return check_hip(ctx, e, "spx_lighttrace launch");
}

int spx_volume(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam, const spx_beam *beam,
               uint32_t width, uint32_t height, uint32_t n, uint32_t photon_base, const spx_vol_params *vp,
               const float *zbuf, uint32_t seed, uint32_t estimator, const spx_opts *opts, uint32_t engine,
               float *film_sum, void *stream) {
    int rc = check_common(ctx, scene, "spx_volume");
    if (rc) return rc;
    if (!cam || !vp) return set_error(ctx, SPX_EINVAL, "spx_volume: null camera or parameters");
    if ((rc = check_film(ctx, film_sum, width, height, "spx_volume"))) return rc;
    if (estimator > SPX_EST_LEGACY) return set_error(ctx, SPX_EINVAL, "spx_volume: unknown estimator");
    if (!isfinite(vp->sigma_s) || !isfinite(vp->sigma_t) || !isfinite(vp->g) || !(vp->max_dist > 0.0f) ||
        !isfinite(vp->max_dist))
        return set_error(ctx, SPX_EINVAL, "spx_volume: bad volume parameters");
    PhotonParams p;
    if ((rc = photon_params(ctx, scene, beam, opts, n, photon_base, seed, p, "spx_volume"))) return rc;
    if (p.hero && estimator == SPX_EST_BEAM)
        return set_error(ctx, SPX_EINVAL, "spx_volume: the hero comb is not offered with the beam estimator");
    make_proj(*cam, p.proj);
    p.vp = *vp;
    p.w = width;
    p.h = height;
    p.estimator = estimator;
    p.zbuf = zbuf;
    p.film = film_sum;
    if (n == 0) return SPX_OK;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (engine == SPX_ENGINE_PLAIN) {
        e = launch_photon_plain(p, 0, s);
    } else if (engine == SPX_ENGINE_WAVEFRONT && estimator == SPX_EST_BEAM) {
        // segments queued in HBM (up to 2^chunk_log2 per chunk of photons),
        // their steps spread over waves
        const uint32_t depth = p.max_depth > 0 ? p.max_depth : 1;
        const uint32_t chunk = std::min(n, std::max(1u, (1u << ctx->tuning.chunk_log2) / depth));
        if ((rc = ensure_scratch(ctx, beam_wavefront_scratch(chunk, depth), s))) return rc;
        e = with_replicas(ctx, p, (size_t)3 * width * height, s,
                          [&](const PhotonParams &q) { return run_beam_wavefront(ctx, q, ctx->scratch, chunk, s); },
                          &rc);
        if (rc) return rc;
    } else if (engine == SPX_ENGINE_WAVEFRONT) {
        uint32_t chunk = 1u << ctx->tuning.chunk_log2;
        if ((rc = ensure_scratch(ctx, photon_wavefront_scratch(chunk), s))) return rc;
        if (ctx->tuning.splat_mode >= 2)
            e = with_bins(ctx, p, width, height, std::min(n, chunk), 0, s,
                          [&](const PhotonParams &q, const BinState *b, const uint32_t *sh) {
                              return run_photon_wavefront(ctx, q, 0, ctx->scratch, chunk, s, b, sh,
                                                          ctx->tuning.splat_mode == 3);
                          }, &rc);
        else
            e = with_replicas(ctx, p, (size_t)3 * width * height, s,
                              [&](const PhotonParams &q) {
                                  return run_photon_wavefront(ctx, q, 0, ctx->scratch, chunk, s, nullptr, nullptr);
                              }, &rc);
        if (rc) return rc;
    } else {
        return set_error(ctx, SPX_EINVAL, "spx_volume: unknown engine");
    }
    return check_hip(ctx, e, "spx_volume launch");
}

int spx_photon_records(spx_ctx *ctx, const spx_scene *scene, const spx_beam *beam, const spx_screen *screen,
                       uint32_t first, uint32_t count, uint32_t seed, float max_dist, uint32_t mode,
                       const spx_opts *opts, spx_photon_record *out_host) {
    int rc = check_common(ctx, scene, "spx_photon_records");
    if (rc) return rc;
    if (!out_host) return set_error(ctx, SPX_EINVAL, "spx_photon_records: null output");
    if (count > 65536) return set_error(ctx, SPX_ECAP, "spx_photon_records: count > 65536");
    if (mode > 2 || (mode == 2 && !screen)) return set_error(ctx, SPX_EINVAL, "spx_photon_records: bad mode");
    PhotonParams p;
    if ((rc = photon_params(ctx, scene, beam, opts, count, first, seed, p, "spx_photon_records"))) return rc;
    if (p.max_depth > 8) return set_error(ctx, SPX_ECAP, "spx_photon_records: max_depth > 8");
    if (screen) set_screen(p, *screen);
    p.vp.max_dist = max_dist;
    if (count == 0) return SPX_OK;
    spx_photon_record *dev = nullptr;
    hipError_t e = hipMalloc(&dev, count * sizeof(spx_photon_record));
    if (e != hipSuccess) return check_hip(ctx, e, "spx_photon_records alloc");
    e = launch_photon_records(p, first, count, mode, dev, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out_host, dev, count * sizeof(spx_photon_record), hipMemcpyDeviceToHost);
    hipFree(dev);
    return check_hip(ctx, e, "spx_photon_records");
}

int spx_surface_rim(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam, uint32_t width,
                    uint32_t height, float *rim_xyz, float *depth, void *stream) {
    int rc = check_common(ctx, scene, "spx_surface_rim");
    if (rc) return rc;
    if (!cam || width == 0 || height == 0) return set_error(ctx, SPX_EINVAL, "spx_surface_rim: bad arguments");
    if (!rim_xyz && !depth) return SPX_OK;
    return check_hip(ctx, launch_surface_rim(scene->dev, *cam, width, height, rim_xyz, depth, (hipStream_t)stream),
                     "spx_surface_rim");
}

int spx_resolve_srgb8(spx_ctx *ctx, const float *film_sum, uint32_t width, uint32_t height, float scale,
                      float exposure, uint8_t *rgb, void *stream) {
    if (!ctx) return SPX_EINVAL;
    if (!film_sum || !rgb || width == 0 || height == 0)
        return set_error(ctx, SPX_EINVAL, "spx_resolve_srgb8: bad arguments");
    return check_hip(ctx, launch_resolve_srgb8(film_sum, width, height, scale, exposure, rgb, (hipStream_t)stream),
                     "spx_resolve_srgb8");
}

int spx_box_blur(spx_ctx *ctx, const float *in, float *tmp, float *out, uint32_t width, uint32_t height,
                 uint32_t radius, void *stream) {
    if (!ctx) return SPX_EINVAL;
    int rc;
    if ((rc = check_film(ctx, in, width, height, "spx_box_blur"))) return rc;
    if (!tmp || !out || tmp == in || tmp == out || radius > 1024)
        return set_error(ctx, SPX_EINVAL, "spx_box_blur: null or aliased buffers, or radius above 1024");
    return check_hip(ctx, launch_box_blur(in, tmp, out, width, height, radius, (hipStream_t)stream), "spx_box_blur");
}

int spx_floor_surface(spx_ctx *ctx, const spx_scene *scene, const spx_camera *cam, const spx_screen *floor,
                      const float *floor_sum, float floor_gain, const float ambient[3], const float background[3],
                      const float glass_tint[3], uint32_t width, uint32_t height, float *surface, float *depth,
                      void *stream) {
    int rc = check_common(ctx, scene, "spx_floor_surface");
    if (rc) return rc;
    if (!cam || !floor || !ambient || !background || !glass_tint || !surface)
        return set_error(ctx, SPX_EINVAL, "spx_floor_surface: null argument");
    if ((rc = check_film(ctx, floor_sum, floor->width, floor->height, "spx_floor_surface (floor)"))) return rc;
    if ((rc = check_film(ctx, surface, width, height, "spx_floor_surface"))) return rc;
    if (!finite3(cam->origin) || !finite3(cam->lower_left) || !finite3(cam->horizontal) || !finite3(cam->vertical) ||
        !finite3(floor->corner) || !finite3(floor->u) || !finite3(floor->v) || !std::isfinite(floor_gain) ||
        !finite3(ambient) || !finite3(background) || !finite3(glass_tint))
        return set_error(ctx, SPX_EINVAL, "spx_floor_surface: non-finite camera, floor or shading value");
    PhotonParams F{};
    set_screen(F, *floor);
    if (!(F.scr_uu > 0.0f) || !(F.scr_vv > 0.0f) || !finite3(F.scr_n))
        return set_error(ctx, SPX_EINVAL, "spx_floor_surface: degenerate floor rectangle");
    FloorShade sh;
    sh.gain = floor_gain;
    memcpy(sh.ambient, ambient, sizeof sh.ambient);
    memcpy(sh.background, background, sizeof sh.background);
    memcpy(sh.tint, glass_tint, sizeof sh.tint);
    return check_hip(ctx,
                     launch_floor_surface(scene->dev, *cam, F, floor_sum, sh, width, height, surface, depth,
                                          (hipStream_t)stream),
                     "spx_floor_surface");
}

int spx_composite_srgb8(spx_ctx *ctx, const float *surface, const float *vol, uint32_t width, uint32_t height,
                        float exposure, uint8_t *rgb, void *stream) {
    if (!ctx) return SPX_EINVAL;
    int rc;
    if ((rc = check_film(ctx, vol, width, height, "spx_composite_srgb8"))) return rc;
    if (!surface || !rgb || !std::isfinite(exposure))
        return set_error(ctx, SPX_EINVAL, "spx_composite_srgb8: null buffer or non-finite exposure");
    return check_hip(ctx, launch_composite_srgb8(surface, vol, width, height, exposure, rgb, (hipStream_t)stream),
                     "spx_composite_srgb8");
}

// k = (1 / SCALE) * inv_n * TARGET_N of the viewer with SCALE = 1 (FP32 film).
static float display_k(uint32_t total_photons, float target_n) {
    const float inv_n = 1.0f / (float)(total_photons > 1u ? total_photons : 1u);
    return inv_n * target_n;
}

int spx_display_resolve(spx_ctx *ctx, const float *vol_sum, const float *surface, uint32_t width, uint32_t height,
                        uint32_t total_photons, float target_n, float exposure, float kradius, uint8_t *rgba,
                        void *stream) {
    if (!ctx) return SPX_EINVAL;
    int rc;
    if ((rc = check_film(ctx, vol_sum, width, height, "spx_display_resolve"))) return rc;
    if (!rgba || ((uintptr_t)rgba & 3u))
        return set_error(ctx, SPX_EINVAL, "spx_display_resolve: null or unaligned output");
    if (!std::isfinite(target_n) || !std::isfinite(exposure) || !(kradius >= 0.0f) || !(kradius <= 32.0f))
        return set_error(ctx, SPX_EINVAL, "spx_display_resolve: bad target_n, exposure or kradius (0..32)");
    const float sigma = fmaxf(kradius * 0.5f, 0.5f);
    const float inv2s2 = 1.0f / (2.0f * sigma * sigma);
    return check_hip(ctx,
                     launch_display_resolve(vol_sum, surface, width, height, display_k(total_photons, target_n),
                                            (int)kradius, inv2s2, exposure, rgba, (hipStream_t)stream),
                     "spx_display_resolve");
}

// Radix select of the rank-th smallest composited luminance over `hist`
// (256 words of device memory), stream-ordered on s: each pass clears the
// histogram with a kernel, counts one key digit, copies the counts back and
// waits for s.  Returns the selected float's key.
static hipError_t quantile_select(const float *vol_sum, const float *surface, uint32_t n, float k, uint64_t rank,
                                  uint32_t *hist, hipStream_t s, uint32_t *key_out) {
    uint32_t prefix = 0, mask = 0, h_hist[256];
    hipError_t e = hipSuccess;
    for (int pass = 0; pass < 4; ++pass) {
        const uint32_t shift = 24u - 8u * (uint32_t)pass;
        e = launch_fill(reinterpret_cast<float *>(hist), 0.0f, 256, s);  // 0.0f is all-zero bits
        if (e == hipSuccess) e = launch_lum_digit_hist(vol_sum, surface, n, k, prefix, mask, shift, hist, s);
        if (e == hipSuccess) e = hipMemcpyAsync(h_hist, hist, sizeof(h_hist), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        uint32_t b = 0;
        uint64_t below = 0;
        while (b < 255u && rank >= below + h_hist[b]) below += h_hist[b++];
        rank -= below;
        prefix |= b << shift;
        mask |= 0xFFu << shift;
    }
    *key_out = prefix;
    return hipSuccess;
}

static float key_to_lum(uint32_t key) {
    const uint32_t u = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
    float L;
    memcpy(&L, &u, sizeof(L));
    return fmaxf(L, 1e-9f);
}

static int quantile_args(spx_ctx *ctx, const float *vol_sum, uint32_t width, uint32_t height, float target_n,
                         float quantile, const float *lum_q, const char *who) {
    int rc;
    if ((rc = check_film(ctx, vol_sum, width, height, who))) return rc;
    if (!lum_q || !std::isfinite(target_n) || !(quantile >= 0.0f) || !(quantile <= 1.0f))
        return set_error(ctx, SPX_EINVAL, std::string(who) + ": bad output, target_n or quantile (0..1)");
    return SPX_OK;
}

static uint64_t quantile_rank(uint32_t n, float quantile) {
    uint64_t rank = (uint64_t)((float)n * quantile);
    return rank > n - 1u ? n - 1u : rank;
}

int spx_display_quantile(spx_ctx *ctx, const float *vol_sum, const float *surface, uint32_t width, uint32_t height,
                         uint32_t total_photons, float target_n, float quantile, float *lum_q, void *stream) {
    if (!ctx) return SPX_EINVAL;
    int rc = quantile_args(ctx, vol_sum, width, height, target_n, quantile, lum_q, "spx_display_quantile");
    if (rc) return rc;
    const uint32_t n = width * height;
    uint32_t *hist = nullptr;
    hipError_t e = hipMalloc(&hist, 256 * sizeof(uint32_t));
    if (e != hipSuccess) return check_hip(ctx, e, "spx_display_quantile alloc");
    uint32_t key = 0;
    e = quantile_select(vol_sum, surface, n, display_k(total_photons, target_n), quantile_rank(n, quantile), hist,
                        (hipStream_t)stream, &key);
    hipFree(hist);
    if (e != hipSuccess) return check_hip(ctx, e, "spx_display_quantile");
    *lum_q = key_to_lum(key);
    return SPX_OK;
}

int spx_display_quantile_noalloc(spx_ctx *ctx, const float *vol_sum, const float *surface, uint32_t width,
                                 uint32_t height, uint32_t total_photons, float target_n, float quantile, float *lum_q,
                                 void *stream) {
    if (!ctx) return SPX_EINVAL;
    // films above 2^28 pixels are refused here (SPX_ECAP): the scratch
    // histograms count in 32 bits
    int rc = quantile_args(ctx, vol_sum, width, height, target_n, quantile, lum_q, "spx_display_quantile_noalloc");
    if (rc) return rc;
    const uint32_t n = width * height;
    hipStream_t s = (hipStream_t)stream;
    uint32_t *q = ctx->quantile_scratch;
    hipError_t e = launch_quantile_select_dev(vol_sum, surface, n, display_k(total_photons, target_n),
                                              quantile_rank(n, quantile), q, s);
    uint32_t key = 0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&key, q + quantile_scratch_key_offset(), sizeof(key), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return check_hip(ctx, e, "spx_display_quantile_noalloc");
    *lum_q = key_to_lum(key);
    return SPX_OK;
}

int spx_rng_draws(spx_ctx *ctx, uint32_t key, uint32_t seed, uint32_t n, uint32_t *out_host) {
    if (!ctx) return SPX_EINVAL;
    if (!out_host || n == 0 || n > 1u << 20) return set_error(ctx, SPX_EINVAL, "spx_rng_draws: bad arguments");
    uint32_t *dev = nullptr;
    hipError_t e = hipMalloc(&dev, n * sizeof(uint32_t));
    if (e != hipSuccess) return check_hip(ctx, e, "spx_rng_draws alloc");
    e = launch_rng_draws(key, seed, n, dev, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out_host, dev, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    hipFree(dev);
    return check_hip(ctx, e, "spx_rng_draws");
}

int spx_glass_n(spx_ctx *ctx, uint32_t glass, const float *lambda_host, uint32_t n, float *out_host) {
    if (!ctx) return SPX_EINVAL;
    if (!lambda_host || !out_host || n == 0 || glass > SPX_GLASS_WATER)
        return set_error(ctx, SPX_EINVAL, "spx_glass_n: bad arguments");
    float *dev = nullptr;
    hipError_t e = hipMalloc(&dev, 2 * (size_t)n * sizeof(float));
    if (e != hipSuccess) return check_hip(ctx, e, "spx_glass_n alloc");
    e = hipMemcpy(dev, lambda_host, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_glass_n(glass, dev, n, dev + n, nullptr);
    if (e == hipSuccess) e = hipMemcpy(out_host, dev + n, n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(dev);
    return check_hip(ctx, e, "spx_glass_n");
}

// ---- memory / streams / events ----------------------------------------------
int spx_malloc(spx_ctx *ctx, size_t bytes, void **out) {
    if (!ctx || !out || bytes == 0) return set_error(ctx, SPX_EINVAL, "spx_malloc: bad arguments");
    *out = nullptr;
    return check_hip(ctx, hipMalloc(out, bytes), "spx_malloc");
}
int spx_free(spx_ctx *ctx, void *ptr) {
    if (!ctx) return SPX_EINVAL;
    if (!ptr) return SPX_OK;
    return check_hip(ctx, hipFree(ptr), "spx_free");
}
int spx_memcpy_h2d(spx_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return set_error(ctx, SPX_EINVAL, "spx_memcpy_h2d: null");
    if (!bytes) return SPX_OK;
    hipError_t e = stream ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream)
                          : hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return check_hip(ctx, e, "spx_memcpy_h2d");
}
int spx_memcpy_d2h(spx_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return set_error(ctx, SPX_EINVAL, "spx_memcpy_d2h: null");
    if (!bytes) return SPX_OK;
    hipError_t e = stream ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream)
                          : hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return check_hip(ctx, e, "spx_memcpy_d2h");
}
int spx_fill_f32(spx_ctx *ctx, float *dst, float value, size_t n, void *stream) {
    if (!ctx || (!dst && n)) return set_error(ctx, SPX_EINVAL, "spx_fill_f32: null");
    return check_hip(ctx, launch_fill(dst, value, n, (hipStream_t)stream), "spx_fill_f32");
}
int spx_copy_f32(spx_ctx *ctx, float *dst, const float *src, size_t n, void *stream) {
    if (!ctx || (n && (!dst || !src))) return set_error(ctx, SPX_EINVAL, "spx_copy_f32: null");
    if (n & 3) return set_error(ctx, SPX_EINVAL, "spx_copy_f32: n must be a multiple of 4");
    return check_hip(ctx, launch_copy(dst, src, n, (hipStream_t)stream), "spx_copy_f32");
}
int spx_stream_create(spx_ctx *ctx, void **out) {
    if (!ctx || !out) return SPX_EINVAL;
    hipStream_t s = nullptr;
    int rc = check_hip(ctx, hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "spx_stream_create");
    *out = s;
    return rc;
}
int spx_stream_destroy(spx_ctx *ctx, void *stream) {
    if (!ctx) return SPX_EINVAL;
    if (!stream) return SPX_OK;
    return check_hip(ctx, hipStreamDestroy((hipStream_t)stream), "spx_stream_destroy");
}
int spx_stream_sync(spx_ctx *ctx, void *stream) {
    if (!ctx) return SPX_EINVAL;
    return check_hip(ctx, hipStreamSynchronize((hipStream_t)stream), "spx_stream_sync");
}
int spx_event_create(spx_ctx *ctx, void **out) {
    if (!ctx || !out) return SPX_EINVAL;
    hipEvent_t ev = nullptr;
    int rc = check_hip(ctx, hipEventCreate(&ev), "spx_event_create");
    *out = ev;
    return rc;
}
int spx_event_destroy(spx_ctx *ctx, void *ev) {
    if (!ctx) return SPX_EINVAL;
    if (!ev) return SPX_OK;
    return check_hip(ctx, hipEventDestroy((hipEvent_t)ev), "spx_event_destroy");
}
int spx_event_record(spx_ctx *ctx, void *ev, void *stream) {
    if (!ctx || !ev) return SPX_EINVAL;
    return check_hip(ctx, hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "spx_event_record");
}
int spx_event_elapsed_ms(spx_ctx *ctx, void *a, void *b, float *ms) {
    if (!ctx || !a || !b || !ms) return SPX_EINVAL;
    hipError_t e = hipEventSynchronize((hipEvent_t)b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b);
    return check_hip(ctx, e, "spx_event_elapsed_ms");
}

}  // extern "C"
