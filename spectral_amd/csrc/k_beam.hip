// k_beam.hip -- the wavefront form of the beam estimator
// (volume::render_volumetric_scene, volume.rs:233-329).
//
// The thread-per-photon form walks each segment's up to 4096 midpoint steps
// and their tent splats in one lane, so a wave waits for its longest
// segment.  Here the photons are traced first, each segment (origin,
// direction, length, power, wavelength, step count) appended to a queue in
// HBM; then persistent waves take one segment at a time and spread its
// steps over their 64 lanes, summing overlapping splats in LDS first.  The
// RNG is drawn only by the trace (the beam estimator draws none per step),
// and each step's splat is the same function as in the per-photon form, so
// films differ only in the order of their fp32 sums and in one product per
// texel, reassociated (the tile sums weight * tent weight and applies the
// segment's colour once per pixel).
#include "spx_internal.hpp"
#include "spx_paths.hpp"

namespace spx {

// One queued segment, three float4: {o, len}, {d, power}, {lambda, m, -, -}.
constexpr uint32_t kBeamSegFloat4 = 3;

__global__ __launch_bounds__(256) void k_beam_trace(PhotonParams P, uint32_t start, uint32_t count, float4 *segs,
                                                    uint32_t *n_segs) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    stage_scene(S, P.scene, tbl, P.tbl);
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    PhotonState st;
    photon_begin(P, tbl, P.photon_base + start + i, st);
    for (uint32_t b = 0; b < P.max_depth; ++b) {  // volume_bounce<SPX_EST_BEAM>
        Hit h;
        const int mi = scene_hit(S, st.ro, st.rd, h);
        const float seg_len = mi >= 0 ? h.t : P.vp.max_dist;
        const uint32_t m = beam_step_count(P, st.ro, st.rd, seg_len);
        // append: one atomic per wave
        const uint64_t act = __ballot(1);
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((unsigned long long)act) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(n_segs, (uint32_t)__popcll(act));
        base = __shfl(base, leader) + (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
        float4 *q = segs + (size_t)base * kBeamSegFloat4;
        q[0] = make_float4(st.ro.x, st.ro.y, st.ro.z, seg_len);
        q[1] = make_float4(st.rd.x, st.rd.y, st.rd.z, st.power);
        q[2] = make_float4(st.lambda, __uint_as_float(m), 0.0f, 0.0f);
        if (mi < 0) break;
        photon_scatter(S, tbl, P, mi, h, st);
    }
}

// A wave's LDS tile over the bounding box of 64 consecutive steps'
// footprints: consecutive steps are half a pixel apart and their tents
// overlap, so summing them in LDS first turns several film atomics per
// pixel into one.  The 64 steps lie on one segment, so their colours share
// its wavelength: each is cmf * weight, and the tile sums the scalar
// weight * tent weight, multiplied by cmf once per pixel when it is
// flushed (the same terms, reassociated: one LDS add per texel instead of
// three).  A deposit outside the box goes to the film directly.
// 5 KiB: 26 one-wave blocks (tile + CIE table) per compute unit, six or
// seven waves per SIMD; groups whose box is larger are split
// (k_beam_deposit).  Measured on the product scene's beam pass (3.5e7
// photons, no z-buffer): 512, 768,
// 1024, 1280, 1536, 2048, 2560, 3072 pixels gave 3.76, 3.65, 3.62, 3.59,
// 3.64, 3.83, 3.95, 4.32 s: more resident waves outweigh more flushes.
constexpr uint32_t kBeamTilePix = 1280;
constexpr uint32_t kBeamBlocksPerCu = (160u * 1024u) / (kBeamTilePix * 4u + kTblN * 16u);
// The values that scale a splat -- the tent's texel weights (tent_weight:
// a quarter of the IEEE form's instructions per texel), the step's weight
// and its radius (beam_step_prep) -- from hardware square roots,
// reciprocals and exp2, each within a few ulp; where a splat lands stays
// IEEE-rounded.
constexpr bool kFastTent = true;

struct TileSink {
    float *tile;
    float *film;
    f3 cmf;
    int bx0, by0, bw, bh;
};

// v = (weight * tent weight) in every channel (splat_transverse's colour
// times the texel's weight, with colour = the step's scalar weight).
// Every deposit of a sub-group's splats lies in its box: the box is the
// union of the steps' [x0, x1] x [y0, y1] (the same expressions as
// splat_transverse's), the tent rows stay inside those, and the
// rounded-centre fallback round(c) lies in [floor(c - r), ceil(c + r)] for
// r >= 1/2, clamped to the film like the box.
__device__ inline void deposit_xy(TileSink &s, uint32_t px, uint32_t py, uint32_t idx, f3 v) {
    const int x = (int)px - s.bx0, y = (int)py - s.by0;
    __hip_atomic_fetch_add(s.tile + y * s.bw + x, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The steps of a segment that can land on the film: [k_lo, k_hi], empty
// when k_lo > k_hi.  project() rejects a point past any of five planes
// through the camera -- behind it (dir.w >= -1e-6), or projecting left of,
// right of, below or above the film (s < 0 is dir.u < (dir.w) vw / 2, and so
// on) -- and each is a linear function f of the point, positive on the
// rejected side, so linear in the step index along the segment.  Where f
// exceeds 2M at both end steps no step lands; where it exceeds 2M at one
// end, the steps up to one before where it falls to 2M cannot land.  M,
// 1e-4 of a bound on every term (the points' l1 distance from the camera
// times 1 + vw + vh), is far above the rounding of f here, of the steps'
// points and of project()'s own arithmetic (a few 2^-24 of that bound
// each), so every step cut is one that project() rejects: the film is the
// one of walking every step.
__device__ inline void beam_plane_values(const Proj &J, f3 q, float f[5]) {
    const f3 dir = q - ld3(J.o);
    const float a = dot(dir, ld3(J.u)), b = dot(dir, ld3(J.v)), c = dot(dir, ld3(J.w));
    f[0] = c + 1e-6f;
    f[1] = 0.5f * c * J.vw - a;
    f[2] = a + 0.5f * c * J.vw;
    f[3] = 0.5f * c * J.vh - b;
    f[4] = b + 0.5f * c * J.vh;
}

__device__ inline void beam_visible_steps(const Proj &J, f3 ro, f3 rd, float ds, uint32_t m, uint32_t &k_lo,
                                          uint32_t &k_hi) {
    const f3 o = ld3(J.o);
    const float da = 0.5f * ds, db = ((float)(m - 1) + 0.5f) * ds;
    const float l1 = fabsf(ro.x) + fabsf(ro.y) + fabsf(ro.z) + fabsf(o.x) + fabsf(o.y) + fabsf(o.z) +
                     fmaxf(fabsf(da), fabsf(db)) * (fabsf(rd.x) + fabsf(rd.y) + fabsf(rd.z));
    const float m2 = 2e-4f * l1 * (1.0f + fabsf(J.vw) + fabsf(J.vh));
    float fa[5], fb[5];
    beam_plane_values(J, ro + rd * da, fa);
    beam_plane_values(J, ro + rd * db, fb);
    const float span = (float)(m - 1);
    int lo = 0, hi = (int)m - 1;
    for (int i = 0; i < 5; ++i) {
        const bool ha = fa[i] > m2, hb = fb[i] > m2;
        if (ha && hb) {
            lo = (int)m;
            hi = -1;
        } else if (ha) {  // the first steps lie past plane i
            const int kc = (int)floorf((fa[i] - m2) / (fa[i] - fb[i]) * span) - 1;
            lo = max(lo, kc + 1);
        } else if (hb) {  // the last steps do
            const int kc = (int)floorf((fb[i] - m2) / (fb[i] - fa[i]) * span) - 1;
            hi = min(hi, (int)m - 2 - kc);
        }
    }
    k_lo = (uint32_t)lo;
    k_hi = hi < lo ? 0u : (uint32_t)hi;
    if (hi < lo) k_lo = 1u;
}

// The deposit kernel's block is one wave, and one wave's LDS operations
// execute in the order it issues them, so its tile needs no barrier between
// zeroing, splatting and flushing -- only the compiler kept from moving
// memory accesses across these points.  (A block barrier would also make
// the wave wait for its flush's film atomics.)
__device__ inline void wave_lds_order() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Wave-wide min / max for a block of one full wave (every lane active):
// DPP within each 16-lane row (pairs, quads, rotations by 4 and 8), then
// the four rows' results read into scalar registers -- no LDS round trip.
template <int CTRL>
__device__ inline int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <bool MAX>
__device__ inline int wave_extreme(int v) {
    auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
    v = op(v, dpp_i<0xB1>(v));   // quad_perm [1, 0, 3, 2]
    v = op(v, dpp_i<0x4E>(v));   // quad_perm [2, 3, 0, 1]
    v = op(v, dpp_i<0x124>(v));  // row_ror:4
    v = op(v, dpp_i<0x128>(v));  // row_ror:8
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}

// Persistent one-wave blocks: a wave claims the next queued segment and
// takes its steps 64 at a time, lane k0 + lane; each group of 64 splats
// into the LDS tile over their footprints' bounding box (as in
// splat_transverse, volume.rs:178-218), which is then added to the block's
// film copy.
__global__ __launch_bounds__(64) void k_beam_deposit(PhotonParams P, const float4 *segs, const uint32_t *n_segs,
                                                     uint32_t *cursor) {
    __shared__ float4 tbl[kTblN];
    __shared__ float tile[kBeamTilePix];
    for (uint32_t k = threadIdx.x; k < (uint32_t)kTblN; k += blockDim.x) tbl[k] = P.tbl[k];
    __syncthreads();
    const uint32_t total = *n_segs;
    float *film = deposit_film(P);
    GlobalSink direct{film};
    const uint32_t lane = threadIdx.x;
    const int W = (int)P.w, H = (int)P.h;
    while (true) {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(cursor, 1u);
        j = __builtin_amdgcn_readlane(j, 0);
        if (j >= total) break;
        const float4 *q = segs + (size_t)j * kBeamSegFloat4;
        const float4 a = q[0], b = q[1], c = q[2];
        const f3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
        const uint32_t m = __float_as_uint(c.y);
        const float ds = a.w / (float)m;
        const f3 cmf = cmf_lookup(tbl, c.x);
        uint32_t k_lo, k_hi;
        beam_visible_steps(P.proj, ro, rd, ds, m, k_lo, k_hi);
        for (uint32_t k0 = k_lo; k0 <= k_hi; k0 += 64) {
            const uint32_t k = k0 + lane;
            BeamSplat sp;
            const bool ok = k <= k_hi && beam_step_prep<kFastTent>(P, ro, rd, b.w, cmf, ds, k, sp);
            // the footprint splat_transverse walks (plus its rounded-centre fallback)
            int x0 = 0x7FFFFFFF, x1 = -1, y0 = 0x7FFFFFFF, y1 = -1;
            if (ok) {
                const float cx = sp.s * (float)W - 0.5f, cy = (1.0f - sp.t) * (float)H - 0.5f;
                const float r = fmaxf(sp.radius, 0.5f);
                x0 = (int)fmaxf(floorf(cx - r), 0.0f);
                x1 = (int)ceilf(cx + r);
                x1 = x1 < 0 ? 0 : (x1 > W - 1 ? W - 1 : x1);
                y0 = (int)fmaxf(floorf(cy - r), 0.0f);
                y1 = (int)ceilf(cy + r);
                y1 = y1 < 0 ? 0 : (y1 > H - 1 ? H - 1 : y1);
            }
            if (__ballot(ok) == 0) continue;  // no step of the group deposits
            // The group's steps in sub-groups: all 64 when their footprint box
            // fits the tile, else halves, quarters, ... (a single step's box,
            // at most 50 x 50 texels for the 24-pixel radius cap, fits any
            // tile of 2500 pixels or more; below that it splats to the film).
            uint32_t lo = 0, S = 64;
            while (lo < 64) {
                const bool in = lane >= lo && lane < lo + S;
                const int bx0 = wave_extreme<false>(in ? x0 : 0x7FFFFFFF);
                const int bx1 = wave_extreme<true>(in ? x1 : -1);
                const int by0 = wave_extreme<false>(in ? y0 : 0x7FFFFFFF);
                const int by1 = wave_extreme<true>(in ? y1 : -1);
                if (bx1 < 0) {  // no step of the sub-group is on the film
                    lo += S;
                    continue;
                }
                const int bw = bx1 - bx0 + 1, bh = by1 - by0 + 1;
                if (bw * bh > (int)kBeamTilePix) {
                    if (S > 1) {
                        S >>= 1;
                        continue;
                    }
                    if (in && ok) splat_transverse<GlobalSink, kFastTent>(P, sp.s, sp.t, sp.color, sp.radius, direct);
                    lo += 1;
                    continue;
                }
                const int area = bw * bh;
                for (int e = (int)lane; e < area; e += 64) tile[e] = 0.0f;
                wave_lds_order();
                if (in && ok) {
                    TileSink ts{tile, film, cmf, bx0, by0, bw, bh};
                    splat_transverse<TileSink, kFastTent>(P, sp.s, sp.t, mk(sp.weight, sp.weight, sp.weight),
                                                          sp.radius, ts);
                }
                wave_lds_order();
                // Flush: lane l takes float f = f0 + l of the box's 3 * area
                // film floats (pixel f / 3, channel f % 3), so one atomic
                // instruction covers 64 consecutive floats of a film row (two
                // or three cache lines) rather than one channel of 64 pixels.
                // The quotients from float reciprocals are exact: (f + 1/2) / 3
                // lies at least 1/6 from an integer and below 1280 + 22, and
                // (e + 1/2) / bw at least 1/(2 bw) from one and below
                // area / bw <= 1280 / bw; the reciprocal and the product round
                // by less than 2^-22 relative, 1302 * 2^-22 < 1/6 and
                // 1280 / bw * 2^-22 < 1/(2 bw), so the floors are the integer
                // quotients.
                const int n3 = 3 * area;
                const float rbw = __builtin_amdgcn_rcpf((float)bw);
                for (int f0 = 0; f0 < n3; f0 += 64) {
                    const int f = f0 + (int)lane;
                    const int e = (int)(((float)f + 0.5f) * (1.0f / 3.0f));
                    const int ch = f - 3 * e;
                    const float v = f < n3 ? tile[e] : 0.0f;
                    if (v != 0.0f) {
                        const int row = (int)(((float)e + 0.5f) * rbw);
                        const int col = e - (int)__umul24((uint32_t)row, (uint32_t)bw);
                        const uint32_t pix = __umul24((uint32_t)(by0 + row), (uint32_t)W) + (uint32_t)(bx0 + col);
                        const float cc = ch == 0 ? cmf.x : (ch == 1 ? cmf.y : cmf.z);
                        unsafeAtomicAdd(film + (3u * pix + (uint32_t)ch), cc * v);
                    }
                }
                lo += S;
                wave_lds_order();
            }
        }
    }
}

size_t beam_wavefront_scratch(uint32_t chunk, uint32_t max_depth) {
    return 256 + (size_t)chunk * max_depth * kBeamSegFloat4 * sizeof(float4);
}

// Photons in chunks: per chunk, zero the two counters, trace, deposit.
hipError_t run_beam_wavefront(spx_ctx *ctx, const PhotonParams &p, void *scratch, uint32_t chunk, hipStream_t s) {
    uint32_t *counters = (uint32_t *)scratch;  // [0] segments queued, [1] deposit cursor
    float4 *segs = (float4 *)((char *)scratch + 256);
    const uint32_t cus = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256);
    for (uint32_t start = 0; start < p.n; start += chunk) {
        const uint32_t count = p.n - start < chunk ? p.n - start : chunk;
        hipError_t e = launch_fill((float *)counters, 0.0f, 2, s);  // +0.0f is all zero bits
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_beam_trace, dim3((count + 255) / 256), dim3(256), 0, s, p, start, count, segs,
                           counters);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // one-wave blocks, as many as the compute units' LDS holds
        hipLaunchKernelGGL(k_beam_deposit, dim3(cus * kBeamBlocksPerCu), dim3(64), 0, s, p, segs, counters, counters + 1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace spx
