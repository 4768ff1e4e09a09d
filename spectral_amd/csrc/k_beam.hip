// k_beam.hip -- the wavefront form of the beam estimator
// (volume::render_volumetric_scene, volume.rs:233-329).
//
// The thread-per-photon form walks each segment's up to 4096 midpoint steps
// and their tent splats in one lane, so a wave waits for its longest
// segment.  Here the photons are traced first, each segment (origin,
// direction, length, power, wavelength, step count) appended to a queue in
// HBM; then persistent waves take one segment at a time and spread its
// steps over their 64 lanes.  The RNG is drawn only by the trace (the beam
// estimator draws none per step), and each step's splat is the same
// function as in the per-photon form, so films differ only in the order of
// their fp32 sums.
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
        photon_scatter(S, P, mi, h, st);
    }
}

// Persistent waves: a wave claims the next queued segment and its lanes
// take steps lane, lane + 64, ... of it; deposits are fp32 atomics into the
// block's film copy.
__global__ __launch_bounds__(256) void k_beam_deposit(PhotonParams P, const float4 *segs, const uint32_t *n_segs,
                                                      uint32_t *cursor) {
    __shared__ float4 tbl[kTblN];
    for (uint32_t k = threadIdx.x; k < (uint32_t)kTblN; k += blockDim.x) tbl[k] = P.tbl[k];
    __syncthreads();
    const uint32_t total = *n_segs;
    GlobalSink sink{deposit_film(P)};
    const uint32_t lane = __lane_id();
    while (true) {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(cursor, 1u);
        j = __builtin_amdgcn_readfirstlane(__shfl(j, 0));
        if (j >= total) break;
        const float4 *q = segs + (size_t)j * kBeamSegFloat4;
        const float4 a = q[0], b = q[1], c = q[2];
        const f3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
        const uint32_t m = __float_as_uint(c.y);
        const float ds = a.w / (float)m;
        const f3 cmf = cmf_lookup(tbl, c.x);
        for (uint32_t k = lane; k < m; k += 64) beam_step(P, ro, rd, b.w, cmf, ds, k, sink);
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
        hipLaunchKernelGGL(k_beam_deposit, dim3(cus * 8), dim3(256), 0, s, p, segs, counters, counters + 1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace spx
