// k_wave.hip -- wavefront engine: per-bounce SoA queues in HBM, wave-ballot +
// prefix-sum compaction of live paths, scene + CIE table in LDS, photon
// film deposits aggregated in an LDS hash before fp32 global atomics.
//
// Bounce schedule per chunk: [reset counters] -> first (emit + bounce 0)
// -> bounce 1 .. max_depth-1 -> (camera) resolve.  A camera chunk holds
// whole pixels (all spp samples of each), and the resolve kernel sums each
// pixel's samples in sample order, so the camera film equals the plain
// kernel's bit for bit.  Camera paths are numbered sample-major (slot =
// sample * pixels + pixel): neighbouring lanes trace neighbouring pixels,
// and the resolve reads each sample's results coalesced.  A bounce kernel's
// grid is sized before its queue length is known, so blocks past the queue
// return before they stage the scene.
#include "spx_paths.hpp"
#include "spx_splat.hpp"

#include <string.h>

#include <algorithm>

namespace spx {

// ---------------------------------------------------------------------------
// Camera queues: SoA {ox,oy,oz,dx,dy,dz,rng,u_hero,tp0,slot|collapsed<<31}.
// Live paths have zero radiance (radiance only changes at the escape that
// ends a path) and lanes 1..3 equal lane 0 until the collapse, so ten words
// carry the whole state.
// ---------------------------------------------------------------------------
struct CamQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    uint32_t *rng;
    float *u, *tp;
    uint32_t *slot;
};

__device__ inline void cam_store(const CamQueue &q, uint32_t i, const CamState &st, uint32_t slot) {
    q.ox[i] = st.ro.x; q.oy[i] = st.ro.y; q.oz[i] = st.ro.z;
    q.dx[i] = st.rd.x; q.dy[i] = st.rd.y; q.dz[i] = st.rd.z;
    q.rng[i] = st.rng;
    q.u[i] = st.u_hero;
    q.tp[i] = st.tp[0];
    q.slot[i] = slot | (st.valid == 1 ? 0x80000000u : 0u);
}

__device__ inline uint32_t cam_load(const CamQueue &q, uint32_t i, CamState &st) {
    st.ro = mk(q.ox[i], q.oy[i], q.oz[i]);
    st.rd = mk(q.dx[i], q.dy[i], q.dz[i]);
    st.rng = q.rng[i];
    st.u_hero = q.u[i];
    st.lam0 = hero_lambda(st.u_hero, 0);
    uint32_t s = q.slot[i];
    bool collapsed = (s & 0x80000000u) != 0;
    float t0 = q.tp[i];
    st.tp[0] = t0;
    st.tp[1] = st.tp[2] = st.tp[3] = collapsed ? 0.0f : t0;
    st.esc = 0.0f;
    st.valid = collapsed ? 1u : (uint32_t)kHeroN;
    return s & 0x7FFFFFFFu;
}

// Block-level compaction into a block-private queue region: one LDS atomic
// per wave reserves the wave's slots, at = the slot within the region.
__device__ inline uint32_t region_enqueue(bool alive, uint32_t *lds_count) {
    uint64_t mask = __ballot(alive);
    uint32_t lane = __lane_id();
    uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t total = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (total) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane == leader) base = atomicAdd(lds_count, total);
        base = __builtin_amdgcn_readlane(base, leader);
    }
    return base + prefix;
}

// Camera chunks run on a fixed grid of G blocks; block g owns queue region
// [g * seg, (g + 1) * seg) of both queues and keeps its paths there from
// bounce to bounce (the survivors of a region never outnumber its paths),
// so compaction needs no global atomic -- a counter every wave of the chip
// would hit -- and a block's paths stay in its XCD's L2.  counts[b * G + g]
// = paths of region g entering bounce b.
struct CamWave {
    CamQueue q[2];
    float *rx, *ry, *rz;   // per-path XYZ results, chunk-local slot order
    uint32_t *counts;      // [33 * G]
    uint32_t G, seg;       // regions and their capacity
};

__device__ inline void cam_finish(const CamWave &W, const float4 *tbl, const CamState &st, uint32_t slot,
                                  bool equal_energy) {
    f3 c = camera_xyz(tbl, st, equal_energy);
    W.rx[slot] = c.x;
    W.ry[slot] = c.y;
    W.rz[slot] = c.z;
}

// Emission + bounce 0 of paths [g * seg, (g + 1) * seg) of the chunk
// (sample-major: path j is sample j / npix of chunk pixel j % npix).
__global__ __launch_bounds__(256) void k_cam_first(CamParams P, CamWave W, uint32_t pix0, uint32_t npix) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    __shared__ uint32_t s_cnt;
    stage_scene(S, P.scene, tbl, P.tbl);
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t total = npix * P.spp;
    const uint32_t j0 = blockIdx.x * W.seg;
    const uint32_t j1 = min(j0 + W.seg, total);
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        uint32_t pixel = pix0 + j % npix;
        uint32_t k = j / npix;
        CamState st;
        camera_begin(P, pixel % P.w, pixel / P.w, P.sample_base + k, st);
        bool alive = camera_bounce(S, tbl, P, 0, st) && P.max_depth > 1;
        uint32_t at = region_enqueue(alive, &s_cnt);
        if (alive) cam_store(W.q[1], j0 + at, st, j);
        else cam_finish(W, tbl, st, j, P.illuminant == SPX_ILLUM_E);
    }
    __syncthreads();
    if (threadIdx.x == 0) W.counts[1u * W.G + blockIdx.x] = s_cnt;
}

// Bounce b of region g's surviving paths.
__global__ __launch_bounds__(256) void k_cam_bounce(CamParams P, CamWave W, uint32_t b) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    __shared__ uint32_t s_cnt;
    const uint32_t n_in = W.counts[b * W.G + blockIdx.x];
    if (n_in == 0) return;  // count of the next bounce stays 0 (reset per chunk)
    stage_scene(S, P.scene, tbl, P.tbl);
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const CamQueue &qin = W.q[b & 1u];
    const CamQueue &qout = W.q[(b + 1u) & 1u];
    const uint32_t r0 = blockIdx.x * W.seg;
    for (uint32_t i = threadIdx.x; i < n_in; i += blockDim.x) {
        CamState st;
        uint32_t slot = cam_load(qin, r0 + i, st);
        bool alive = camera_bounce(S, tbl, P, b, st) && (b + 1u < P.max_depth);
        uint32_t at = region_enqueue(alive, &s_cnt);
        if (alive) cam_store(qout, r0 + at, st, slot);
        else cam_finish(W, tbl, st, slot, P.illuminant == SPX_ILLUM_E);
    }
    __syncthreads();
    if (threadIdx.x == 0) W.counts[(b + 1u) * W.G + blockIdx.x] = s_cnt;
}

// The tail: region g's paths entering bounce b, each traced to its end in
// one thread.  Past the first bounces few paths are left in a chunk, and a
// launch per remaining bounce would cost more in launches and empty blocks
// than compaction saves; the bounce sequence of every path is unchanged.
__global__ __launch_bounds__(256) void k_cam_tail(CamParams P, CamWave W, uint32_t b0) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    const uint32_t n_in = W.counts[b0 * W.G + blockIdx.x];
    if (n_in == 0) return;
    stage_scene(S, P.scene, tbl, P.tbl);
    __syncthreads();
    const CamQueue &qin = W.q[b0 & 1u];
    const uint32_t r0 = blockIdx.x * W.seg;
    for (uint32_t i = threadIdx.x; i < n_in; i += blockDim.x) {
        CamState st;
        uint32_t slot = cam_load(qin, r0 + i, st);
        for (uint32_t b = b0; b < P.max_depth; ++b)
            if (!camera_bounce(S, tbl, P, b, st)) break;
        cam_finish(W, tbl, st, slot, P.illuminant == SPX_ILLUM_E);
    }
}

// Per-pixel in-order sum of the chunk's samples (tracer.rs:125-139).
__global__ __launch_bounds__(256) void k_cam_resolve(CamParams P, CamWave W, uint32_t pix0, uint32_t npix) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (uint32_t k = 0; k < P.spp; ++k) {  // sample order; sample-major slots
        const size_t at = (size_t)k * npix + j;
        ax += W.rx[at];
        ay += W.ry[at];
        az += W.rz[at];
    }
    uint32_t i = pix0 + j;
    P.film[3 * i + 0] += ax;
    P.film[3 * i + 1] += ay;
    P.film[3 * i + 2] += az;
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Regions: at most kCamRegions blocks, each region a multiple of 256 paths
// (queues sized chunk + kCamRegions * 256 to hold the rounding).
constexpr uint32_t kCamRegions = 2048;
constexpr uint32_t kMaxBounces = 33;
// Bounces from this one on run in one tail launch (k_cam_tail).
constexpr uint32_t kCamTailFrom = 3;

size_t camera_wavefront_scratch(uint32_t chunk) {
    const size_t cap = (size_t)chunk + (size_t)kCamRegions * 256;
    return align_up((size_t)kMaxBounces * kCamRegions * 4) + 2 * 10 * align_up(cap * 4) + 3 * align_up(cap * 4);
}

static CamWave carve_cam(void *scratch, uint32_t chunk) {
    char *p = (char *)scratch;
    const size_t a = align_up(((size_t)chunk + (size_t)kCamRegions * 256) * 4);
    CamWave W;
    W.counts = (uint32_t *)p;
    p += align_up((size_t)kMaxBounces * kCamRegions * 4);
    for (int k = 0; k < 2; ++k) {
        CamQueue &q = W.q[k];
        q.ox = (float *)p; p += a;
        q.oy = (float *)p; p += a;
        q.oz = (float *)p; p += a;
        q.dx = (float *)p; p += a;
        q.dy = (float *)p; p += a;
        q.dz = (float *)p; p += a;
        q.rng = (uint32_t *)p; p += a;
        q.u = (float *)p; p += a;
        q.tp = (float *)p; p += a;
        q.slot = (uint32_t *)p; p += a;
    }
    W.rx = (float *)p; p += a;
    W.ry = (float *)p; p += a;
    W.rz = (float *)p;
    W.G = 0;
    W.seg = 0;
    return W;
}

static uint32_t persistent_blocks(const spx_ctx *ctx, uint32_t work, uint32_t per_cu) {
    uint32_t cap = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * per_cu;
    uint32_t need = (work + 255) / 256;
    return need < cap ? (need ? need : 1) : cap;
}

hipError_t run_camera_wavefront(spx_ctx *ctx, const CamParams &p, void *scratch, uint32_t chunk,
                                hipStream_t s) {
    CamWave W = carve_cam(scratch, chunk);
    uint32_t npix_total = p.w * p.h;
    uint32_t per = chunk / p.spp;
    if (per == 0) per = 1;  // spp > chunk: a chunk holds one pixel (needs chunk >= spp)
    if ((uint64_t)per * p.spp > chunk || p.max_depth + 1u > kMaxBounces) return hipErrorInvalidValue;
    for (uint32_t pix0 = 0; pix0 < npix_total; pix0 += per) {
        uint32_t npix = npix_total - pix0 < per ? npix_total - pix0 : per;
        uint32_t total = npix * p.spp;
        W.G = std::min(persistent_blocks(ctx, total, 8), kCamRegions);
        W.seg = (((total + W.G - 1) / W.G) + 255u) & ~255u;  // G * seg <= chunk + G * 256
        hipError_t e = launch_fill(reinterpret_cast<float *>(W.counts), 0.0f, (size_t)(p.max_depth + 1u) * W.G, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_cam_first, dim3(W.G), dim3(256), 0, s, p, W, pix0, npix);
        const uint32_t tail = std::min(p.max_depth, kCamTailFrom);
        for (uint32_t b = 1; b < tail; ++b)
            hipLaunchKernelGGL(k_cam_bounce, dim3(W.G), dim3(256), 0, s, p, W, b);
        if (tail < p.max_depth) hipLaunchKernelGGL(k_cam_tail, dim3(W.G), dim3(256), 0, s, p, W, tail);
        hipLaunchKernelGGL(k_cam_resolve, dim3((npix + 255) / 256), dim3(256), 0, s, p, W, pix0, npix);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Photon queues: SoA {ox,oy,oz,dx,dy,dz,rng,lambda,power}.
// ---------------------------------------------------------------------------
struct PhQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    uint32_t *rng;
    float *lam, *pw;
};

// The wavelength travels as its draw u (lambda = 380 + 350 u, recomputed
// bit for bit) and the power's sign bit marks a live hero comb (powers are
// >= 0), so the spectrum of either option is rebuilt on load.
__device__ inline void ph_store(const PhQueue &q, uint32_t i, const PhotonState &st) {
    q.ox[i] = st.ro.x; q.oy[i] = st.ro.y; q.oz[i] = st.ro.z;
    q.dx[i] = st.rd.x; q.dy[i] = st.rd.y; q.dz[i] = st.rd.z;
    q.rng[i] = st.rng;
    q.lam[i] = st.u;
    q.pw[i] = __uint_as_float(__float_as_uint(st.power) | (st.comb ? 0x80000000u : 0u));
}

__device__ inline void ph_load(const PhQueue &q, uint32_t i, const PhotonParams &P, const float4 *tbl,
                               PhotonState &st) {
    st.ro = mk(q.ox[i], q.oy[i], q.oz[i]);
    st.rd = mk(q.dx[i], q.dy[i], q.dz[i]);
    st.rng = q.rng[i];
    st.u = q.lam[i];
    st.lambda = kLambdaMin + kLambdaRange * st.u;
    const uint32_t pw = __float_as_uint(q.pw[i]);
    st.power = __uint_as_float(pw & 0x7FFFFFFFu);
    st.comb = pw >> 31;
    st.n_glass = kNoGlass;
    st.cmf = st.comb ? photon_comb(P, tbl, st.u) : (P.hero ? photon_hero(P, tbl, st.lambda)
                                                            : cmf_lookup(tbl, st.lambda));
}

// Photon chunks use block-owned queue regions as camera chunks do:
// counts[b * G + g] = photons of region g entering bounce b.
struct PhWave {
    PhQueue q[2];
    uint32_t *counts;  // [33 * G]
    uint32_t G, seg;
};

constexpr int kSplatSlots = 2048;

// MODE: 0 volume point, 2 volume legacy, 3 lighttrace.
template <int MODE, class Sink>
__device__ inline bool ph_bounce(const SceneData &S, const float4 *tbl, const PhotonParams &P,
                                 PhotonState &st, Sink &sink) {
    if (MODE == 3) return lighttrace_bounce(S, tbl, P, st, sink) == 0;
    if (MODE == 2) return volume_bounce<SPX_EST_LEGACY>(S, tbl, P, st, sink);
    return volume_bounce<SPX_EST_POINT>(S, tbl, P, st, sink);
}

// Per-kernel deposit sink: SINK 1 = LDS hash + atomics, 2 = binned records,
// 3 = LDS hash draining to binned records.
template <int SINK>
struct PhSink;

template <>
struct PhSink<1> {
    uint32_t sk[kSplatSlots];
    float sv[3 * kSplatSlots];
    uint32_t sfill;
};

template <>
struct PhSink<2> {
    uint32_t wstate[2 * 4];
};

template <>
struct PhSink<3> {
    uint32_t sk[kSplatSlots];
    float sv[3 * kSplatSlots];
    uint32_t sfill;
    uint32_t wstate[2 * 4];
};

// Bounce b of region g = blockIdx.x: the region's n_in photons (bounce 0:
// chunk photons [g * seg, (g + 1) * seg)), survivors compacted into the
// same region of the other queue.
template <int MODE, int SINK>
__device__ inline void ph_body(const SceneData &S, const float4 *tbl, const PhotonParams &P, PhWave &W,
                               uint32_t b, uint32_t start, uint32_t n_in, const BinState &bins,
                               PhSink<SINK> &L, uint32_t *s_cnt) {
    const bool first = b == 0;
    const PhQueue &qin = W.q[b & 1u];
    const PhQueue &qout = W.q[(b + 1u) & 1u];
    const uint32_t r0 = blockIdx.x * W.seg;
    auto step = [&](auto &sink) {
        // Block-uniform trip count so that the flush barriers line up.
        for (uint32_t base = 0; base < n_in; base += blockDim.x) {
            uint32_t i = base + threadIdx.x;
            if (i < n_in) {
                PhotonState st;
                if (first) photon_begin(P, tbl, P.photon_base + start + r0 + i, st);
                else ph_load(qin, r0 + i, P, tbl, st);
                bool alive = ph_bounce<MODE>(S, tbl, P, st, sink) && (b + 1u < P.max_depth);
                uint32_t at = region_enqueue(alive, s_cnt);
                if (alive) ph_store(qout, r0 + at, st);
            }
            if constexpr (SINK == 1 || SINK == 3) sink.maybe_flush();
        }
    };
    if constexpr (SINK == 1) {
        LdsSplat<kSplatSlots> sink{L.sk, L.sv, &L.sfill, deposit_film(P)};
        sink.init();
        __syncthreads();
        step(sink);
        sink.flush();
    } else if constexpr (SINK == 3) {
        LdsSplat<kSplatSlots, kHashBinProbes> hash{L.sk, L.sv, &L.sfill, nullptr};
        hash.init();
        if (threadIdx.x < 8) L.wstate[threadIdx.x] = (threadIdx.x & 1) ? 0u : kChunkEmpty;
        __syncthreads();
        const uint32_t fw = MODE == 3 ? P.scr_w : P.w;
        HashBinSink<kSplatSlots> sink{hash, BinSink{bins.rkey, bins.rval, bins.pool_next, bins.chunk_fill, bins.pool_cap,
                                                    xcc_id() % kBinPools, P.key_base, P.film, L.wstate, fw,
                                                    tiles_across(fw, bins.tile_wlog2), bins.tile_wlog2}};
        step(sink);
        sink.flush();
        bin_wave_finish(L.wstate, bins.chunk_fill);
    } else {
        if (threadIdx.x < 8) L.wstate[threadIdx.x] = (threadIdx.x & 1) ? 0u : kChunkEmpty;
        __syncthreads();
        const uint32_t fw = MODE == 3 ? P.scr_w : P.w;
        BinSink sink{bins.rkey, bins.rval, bins.pool_next, bins.chunk_fill, bins.pool_cap, xcc_id() % kBinPools, P.key_base,
                     P.film, L.wstate, fw, tiles_across(fw, bins.tile_wlog2), bins.tile_wlog2};
        step(sink);
        bin_wave_finish(L.wstate, bins.chunk_fill);
    }
}

template <int MODE, int SINK>
__global__ __launch_bounds__(256) void k_ph(PhotonParams P, PhWave W, uint32_t b, uint32_t start, uint32_t count,
                                            BinState bins) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    __shared__ PhSink<SINK> L;
    __shared__ uint32_t s_cnt;
    const uint32_t r0 = blockIdx.x * W.seg;
    const uint32_t n_in = b == 0 ? (r0 < count ? min(W.seg, count - r0) : 0u) : W.counts[b * W.G + blockIdx.x];
    // an empty region returns at once (no barrier, flush or record chunk of
    // its block is pending; its next count stays 0)
    if (n_in == 0) return;
    stage_scene(S, P.scene, tbl, P.tbl);
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    ph_body<MODE, SINK>(S, tbl, P, W, b, start, n_in, bins, L, &s_cnt);
    __syncthreads();
    if (threadIdx.x == 0) W.counts[(b + 1u) * W.G + blockIdx.x] = s_cnt;
}

constexpr uint32_t kPhRegions = 1024;

size_t photon_wavefront_scratch(uint32_t chunk) {
    const size_t cap = (size_t)chunk + (size_t)kPhRegions * 256;
    return align_up((size_t)kMaxBounces * kPhRegions * 4) + 2 * 9 * align_up(cap * 4);
}

static PhWave carve_ph(void *scratch, uint32_t chunk) {
    char *p = (char *)scratch;
    const size_t a = align_up(((size_t)chunk + (size_t)kPhRegions * 256) * 4);
    PhWave W;
    W.counts = (uint32_t *)p;
    W.G = 0;
    W.seg = 0;
    p += align_up((size_t)kMaxBounces * kPhRegions * 4);
    for (int k = 0; k < 2; ++k) {
        PhQueue &q = W.q[k];
        q.ox = (float *)p; p += a;
        q.oy = (float *)p; p += a;
        q.oz = (float *)p; p += a;
        q.dx = (float *)p; p += a;
        q.dy = (float *)p; p += a;
        q.dz = (float *)p; p += a;
        q.rng = (uint32_t *)p; p += a;
        q.lam = (float *)p; p += a;
        q.pw = (float *)p; p += a;
    }
    return W;
}

template <int MODE, int SINK>
static void run_ph(spx_ctx *ctx, const PhotonParams &p, PhWave W, uint32_t chunk, hipStream_t s,
                   const BinState *bins, const uint32_t *job_shared) {
    BinState none;
    memset(&none, 0, sizeof none);
    const BinState &B = bins ? *bins : none;
    for (uint32_t start = 0; start < p.n; start += chunk) {
        uint32_t count = p.n - start < chunk ? p.n - start : chunk;
        W.G = std::min(persistent_blocks(ctx, count, 4), kPhRegions);
        W.seg = (((count + W.G - 1) / W.G) + 255u) & ~255u;  // G * seg <= chunk + G * 256
        launch_fill(reinterpret_cast<float *>(W.counts), 0.0f, (size_t)(p.max_depth + 1u) * W.G, s);
        if (SINK >= 2) launch_bins_reset(B, s);
        for (uint32_t b = 0; b < p.max_depth; ++b)
            hipLaunchKernelGGL((k_ph<MODE, SINK>), dim3(W.G), dim3(256), 0, s, p, W, b, start, count, B);
        if (SINK >= 2) launch_bins_resolve(B, job_shared, (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * 4, s);
    }
}

hipError_t run_photon_wavefront(spx_ctx *ctx, const PhotonParams &p, int lighttrace, void *scratch,
                                uint32_t chunk, hipStream_t s, const BinState *bins,
                                const uint32_t *job_shared, bool hash) {
    PhWave W = carve_ph(scratch, chunk);
    if (bins && hash) {
        if (lighttrace) run_ph<3, 3>(ctx, p, W, chunk, s, bins, job_shared);
        else if (p.estimator == SPX_EST_LEGACY) run_ph<2, 3>(ctx, p, W, chunk, s, bins, job_shared);
        else run_ph<0, 3>(ctx, p, W, chunk, s, bins, job_shared);
    } else if (bins) {
        if (lighttrace) run_ph<3, 2>(ctx, p, W, chunk, s, bins, job_shared);
        else if (p.estimator == SPX_EST_LEGACY) run_ph<2, 2>(ctx, p, W, chunk, s, bins, job_shared);
        else run_ph<0, 2>(ctx, p, W, chunk, s, bins, job_shared);
    } else {
        if (lighttrace) run_ph<3, 1>(ctx, p, W, chunk, s, bins, job_shared);
        else if (p.estimator == SPX_EST_LEGACY) run_ph<2, 1>(ctx, p, W, chunk, s, bins, job_shared);
        else run_ph<0, 1>(ctx, p, W, chunk, s, bins, job_shared);
    }
    return hipGetLastError();
}

}  // namespace spx
