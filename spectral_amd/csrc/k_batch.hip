// k_batch.hip -- the batched engine: many uneven camera / volume / lighttrace
// jobs in ONE persistent launch fed from a device job table.
//
//  * Work items are (job, range) chunks: camera items are pixel ranges with
//    all their spp samples, photon items are photon-index ranges.  The host
//    sorts items by cost (longest first) and deals them to 8 queues, one per
//    XCD; block b starts on queue b % 8 (blocks are dispatched round-robin
//    over the XCDs), so a job's film stays warm in one XCD's L2, and steals
//    from the other queues when its own runs dry.
//  * Inside an item every lane runs its own path / photon bounce by bounce
//    and regenerates a new one the moment its current one ends, so the wave
//    stays full even though path lengths differ (no per-bounce barrier).
//  * Camera items and photon items run in separate kernels (each sized for
//    its own path's registers): the camera round first, then the photon
//    rounds.
//  * Camera samples are summed per pixel in an LDS accumulator and written
//    once per item (one owner per pixel: no global atomics).  Photon
//    deposits go to the sink the deposit mode names: binned records, the
//    LDS splat hash drained with fp32 atomics, or direct fp32 atomics.
#include "spx_batch.hpp"
#include "spx_paths.hpp"
#include "spx_splat.hpp"

namespace spx {

constexpr int kBatchSlots = 2048;
constexpr uint32_t kMaxCamPix = 2048;  // LDS pixel accumulator capacity per item

__device__ inline uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Camera item: pixels [start, start+count) of job J, all spp samples each.
// Unit f is sample f / count of item pixel f % count: the lanes of a wave
// work on neighbouring pixels, so their LDS sums never contend for one
// address, and when count divides the block size each lane keeps one pixel
// and sums its samples in sample order (as the reference does).
__device__ inline void batch_camera_item(const SceneData &S, const float4 *tbl, const CamParams &P,
                                         uint32_t start, uint32_t count, float *acc) {
    for (uint32_t k = threadIdx.x; k < 3 * count; k += blockDim.x) acc[k] = 0.0f;
    __syncthreads();
    const uint32_t total = count * P.spp;
    uint32_t f = threadIdx.x;
    if (f < total) {
        CamState st;
        uint32_t lp = f % count;
        camera_begin(P, (start + lp) % P.w, (start + lp) / P.w, P.sample_base + f / count, st);
        uint32_t b = 0;
        while (true) {
            bool alive = camera_bounce(S, tbl, P, b, st);
            ++b;
            if (alive && b < P.max_depth) continue;
            f3 c = camera_xyz(tbl, st, P.illuminant == SPX_ILLUM_E);
            __hip_atomic_fetch_add(acc + 3 * lp + 0, c.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc + 3 * lp + 1, c.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc + 3 * lp + 2, c.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            f += blockDim.x;
            if (f >= total) break;
            lp = f % count;
            camera_begin(P, (start + lp) % P.w, (start + lp) / P.w, P.sample_base + f / count, st);
            b = 0;
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) {
        uint32_t i = start + k;
        P.film[3 * i + 0] += acc[3 * k + 0];
        P.film[3 * i + 1] += acc[3 * k + 1];
        P.film[3 * i + 2] += acc[3 * k + 2];
    }
}

template <class Sink>
__device__ inline void batch_photon_item(const SceneData &S, const float4 *tbl, const PhotonParams &P,
                                         uint32_t kind, uint32_t start, uint32_t count, Sink &sink) {
    uint32_t f = threadIdx.x;
    if (f >= count) return;
    PhotonState st;
    photon_begin(P, tbl, P.photon_base + start + f, st);
    sink_photon_begin(sink);
    uint32_t b = 0;
    while (true) {
        bool alive;
        if (kind == SPX_JOB_LIGHTTRACE) alive = lighttrace_bounce(S, tbl, P, st, sink) == 0;
        else if (P.estimator == SPX_EST_LEGACY) alive = volume_bounce<SPX_EST_LEGACY>(S, tbl, P, st, sink);
        else alive = volume_bounce<SPX_EST_POINT>(S, tbl, P, st, sink);
        ++b;
        if (alive && b < P.max_depth) continue;
        f += blockDim.x;
        if (f >= count) break;
        photon_begin(P, tbl, P.photon_base + start + f, st);
        sink_photon_begin(sink);
        b = 0;
    }
}

// The same photon item with every lane of the wave in the loop until the
// wave's last photon ends, so that each of a bounce's deposit candidates
// reaches the sink with the wave converged: the transport and the deposit
// arithmetic stay per lane (the same operations, and random draws in the
// same order, as the bounce functions above), the sink's bookkeeping is
// wave-uniform.  Each wave takes an equal share of the item's photons and
// hands a lane whose photon has ended the share's next one from a
// wave-uniform counter, so the wave's lanes run out of photons within one
// photon's life of each other (a fixed stride of photons per lane leaves
// the wave waiting on its longest lane).  A photon's result depends only on
// its index, not on the lane that traces it.
//
// POOL: the few lanes that finish in one bounce (about 13 of 64) would each
// compute their next photon's start -- its draws, table lookups and the fp64
// Sellmeier index of the scene's glass -- with the rest of the wave idle.
// Instead the wave keeps the starts of its next photons in a pool (`pool`:
// kPhotonStartWords x 64 words of LDS, photon p's start at slot p % 64):
// when the finishing lanes need photons the pool does not hold, the wave
// computes the starts of all its next photons up to 64 ahead at once, one
// per lane, and a finishing lane then reads its photon's start back.
constexpr uint32_t kPoolSlots = 64;
constexpr uint32_t kPoolWords = kPhotonStartWords * kPoolSlots;
constexpr uint32_t kPhotonPoolBytes = 4 * kPoolWords * 4;  // four waves of 256-thread blocks

__device__ inline void pool_store(float *pool, uint32_t slot, const PhotonStart &s) {
    pool[0 * kPoolSlots + slot] = __uint_as_float(s.rng);
    pool[1 * kPoolSlots + slot] = s.a;
    pool[2 * kPoolSlots + slot] = s.b;
    pool[3 * kPoolSlots + slot] = s.u;
    pool[4 * kPoolSlots + slot] = s.power;
    pool[5 * kPoolSlots + slot] = s.n;
    pool[6 * kPoolSlots + slot] = s.cmf.x;
    pool[7 * kPoolSlots + slot] = s.cmf.y;
    pool[8 * kPoolSlots + slot] = s.cmf.z;
}

__device__ inline PhotonStart pool_load(const float *pool, uint32_t slot) {
    PhotonStart s;
    s.rng = __float_as_uint(pool[0 * kPoolSlots + slot]);
    s.a = pool[1 * kPoolSlots + slot];
    s.b = pool[2 * kPoolSlots + slot];
    s.u = pool[3 * kPoolSlots + slot];
    s.power = pool[4 * kPoolSlots + slot];
    s.n = pool[5 * kPoolSlots + slot];
    s.cmf = mk(pool[6 * kPoolSlots + slot], pool[7 * kPoolSlots + slot], pool[8 * kPoolSlots + slot]);
    return s;
}

// Orders one wave's LDS stores before its later LDS loads of other lanes'
// words (the LDS executes a wave's instructions in order; this keeps the
// compiler from moving the loads up).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool POOL, class WS>
__device__ inline void conv_photon_item(const SceneData &S, const float4 *tbl, const PhotonParams &P,
                                        uint32_t kind, uint32_t start, uint32_t count, WS &sink, float *pool) {
    const uint32_t waves = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = __lane_id();
    const uint32_t per = (count + waves - 1) / waves;
    const uint32_t w0 = min(wave * per, count), w1 = min(w0 + per, count);
    uint32_t next = w0 + 64;  // the share's next photon not yet handed out (wave-uniform)
    uint32_t filled = next;   // POOL: the pool holds the starts of photons [next, filled)
    const uint32_t g0 = POOL ? uniform_u32(scene_first_glass(S)) : kNoGlass;
    uint32_t f = w0 + lane;
    bool live = f < w1;
    bool spread = false;  // the photon has scattered diffusely (its deposits skip the LDS table)
    uint32_t b = 0;
    PhotonState st;
    if (live) photon_resume(P, photon_start(P, tbl, P.photon_base + start + f, g0), g0, st);
    const bool screen = kind == SPX_JOB_LIGHTTRACE;
    const bool legacy = uniform_u32(P.estimator) == SPX_EST_LEGACY;
    const bool iso = uniform_f32(P.vp.g) == 0.0f;
    const f3 cam_o = ld3(P.proj.o);
    const float *zbuf = uniform_ptr(P.zbuf);
    while (__ballot(live)) {
        Hit h;
        int mi = -1;
        bool ends = false;  // the photon ends after this bounce's deposits
        if (screen) {  // lighttrace.rs:133-151
            bool landed = false;
            ScreenTexels T;
            if (live) {
                mi = scene_hit(S, st.ro, st.rd, h);
                float ss, rr, stt;
                const bool on = screen_coords(P, st.ro, st.rd, ss, rr, stt);
                const float scene_t = mi >= 0 ? h.t : INFINITY;
                const float screen_t = on ? stt : INFINITY;
                landed = screen_t < scene_t;
                if (landed) T = screen_texels(P, ss, rr, st.cmf, st.power);
                ends = landed || mi < 0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                Dep d;
                d.ok = false;
                if (landed) d = screen_texel(P, T, k);
                sink.deposit(d, spread);
            }
        } else {  // volume.rs:130-150 (point) / 51-64 (legacy)
            float seg_len = 0.0f;
            if (live) {
                mi = scene_hit(S, st.ro, st.rd, h);
                seg_len = mi >= 0 ? h.t : P.vp.max_dist;
                ends = mi < 0;
            }
            const f3 cmf = st.cmf;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                Dep d;
                d.ok = false;
                if (live) {
                    if (legacy) d = point_sample<true, true>(P, st, cmf, seg_len, st.rng, cam_o, zbuf);
                    else if (iso) d = point_sample<false, true>(P, st, cmf, seg_len, st.rng, cam_o, zbuf);
                    else d = point_sample<false, false>(P, st, cmf, seg_len, st.rng, cam_o, zbuf);
                }
                sink.deposit(d, spread);
            }
        }
        bool done = false;
        if (live) {
            if (!ends) {
                photon_scatter(S, tbl, P, mi, h, st);
                if (S.mats[mi].kind == SPX_MAT_LAMBERTIAN) spread = true;
            }
            ++b;
            done = ends || b >= P.max_depth;
        }
        const uint64_t m = __ballot(done);
        if (m) {
            const uint32_t mine = next + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (POOL && min(next + (uint32_t)__popcll(m), w1) > filled) {
                // (wave-uniform) the starts of photons [filled, next + 64), the pool's free slots
                const uint32_t nf = min(next + kPoolSlots, w1) - filled;
                if (lane < nf) {
                    const uint32_t p = filled + lane;
                    pool_store(pool, p % kPoolSlots, photon_start(P, tbl, P.photon_base + start + p, kNoGlass));
                }
                if (g0 != kNoGlass && lane < nf) {  // the fp64 index apart: fewer registers live at once
                    const uint32_t slot = (filled + lane) % kPoolSlots;
                    const float lambda = kLambdaMin + kLambdaRange * pool[3 * kPoolSlots + slot];
                    pool[5 * kPoolSlots + slot] = glass_n(g0, P.dispersion ? lambda : 550.0f);
                }
                filled += nf;
                wave_lds_sync();
            }
            next += (uint32_t)__popcll(m);
            if (done) {
                f = mine;
                live = f < w1;
                if (live) {
                    if (POOL) photon_resume(P, pool_load(pool, f % kPoolSlots), g0, st);
                    else photon_begin(P, tbl, P.photon_base + start + f, st);
                    spread = false;
                    b = 0;
                }
            }
        }
    }
}

// Shared by both kernels: block-level item claim (home XCD queue first, then
// the others) and scene staging when the item's job changes scene.
struct BatchBlock {
    SceneData S;
    float4 tbl[kTblN];
    uint32_t item;
    const SceneData *scene;
    const float4 *tbl_src;
};

__device__ inline uint32_t claim_item(BatchQueues *q, BatchBlock &B, uint32_t home, uint32_t &qi) {
    if (threadIdx.x == 0) {
        uint32_t it = 0xFFFFFFFFu;
        while (qi < kXcdQueues) {
            uint32_t qq = (home + qi) % kXcdQueues;
            uint32_t k = atomicAdd(&q->next[qq], 1u);
            if (k < q->size[qq]) {
                it = q->begin[qq] + k;
                break;
            }
            ++qi;
        }
        B.item = it;
    }
    __syncthreads();
    return uniform_u32(B.item);
}

__device__ inline void stage_for(BatchBlock &B, const SceneData *scene, const float4 *jt) {
    if (scene != B.scene || jt != B.tbl_src) {  // restage only when the job's scene changes
        __syncthreads();
        stage_scene(B.S, scene, B.tbl, jt);
        __syncthreads();
        if (threadIdx.x == 0) {
            B.scene = scene;
            B.tbl_src = jt;
        }
    }
    __syncthreads();
}

// Camera items only (sized for the camera path's registers: capped at 128
// VGPRs for four waves per SIMD, which measured 10 % faster than three).
__global__ __launch_bounds__(256, 4) void k_batch_camera(const BatchJob *__restrict__ jobs,
                                                      const BatchItem *__restrict__ items, BatchQueues *q) {
    __shared__ BatchBlock B;
    __shared__ float acc[3 * kMaxCamPix];
    if (threadIdx.x == 0) {
        B.scene = nullptr;
        B.tbl_src = nullptr;
    }
    const uint32_t home = blockIdx.x % kXcdQueues;
    uint32_t qi = 0;
    while (true) {
        const uint32_t it = claim_item(q, B, home, qi);
        if (it == 0xFFFFFFFFu) break;
        const BatchItem item = items[it];
        const uint32_t j = uniform_u32(item.job);
        const CamParams P = jobs[j].cam;
        stage_for(B, P.scene, P.tbl);
        batch_camera_item(B.S, B.tbl, P, item.start, item.count, acc);
        __syncthreads();
    }
}

// Photon items only (volume and lighttrace), deposits through sink SPLAT:
// 0 = direct film atomics, 1 = LDS hash, 2 = binned records, 3 = LDS hash
// draining to binned records.
// THREADS: block size; the LDS hash holds 8 slots per thread (2048 at 256
// threads; 8192 at 1024 threads, one block filling a compute unit, so all of
// the unit's photons share one 128 KiB table).  Capped at 128 VGPRs (four
// waves per SIMD), so a concurrent camera round's waves still fit beside
// three photon blocks per compute unit.  The binned forms of 256 threads
// take each wave's next photons from a pool of precomputed starts
// (conv_photon_item) in 9 KiB of dynamic LDS (kPhotonPoolBytes at launch),
// which still leaves a camera block room beside three photon blocks; being
// dynamic, it does not lower the four waves per SIMD the register budget is
// planned for.
template <int SPLAT, int THREADS = 256>
__global__ __launch_bounds__(THREADS, 4) void k_batch_photons(const BatchJob *__restrict__ jobs,
                                                           const BatchItem *__restrict__ items, BatchQueues *q,
                                                           BinState bins) {
    constexpr int kSlots = THREADS / 256 * kBatchSlots;
    constexpr int kWaves = THREADS / 64;
    constexpr bool kPool = SPLAT >= 2 && THREADS == 256;
    __shared__ BatchBlock B;
    __shared__ uint32_t sk[SPLAT == 1 || SPLAT == 3 ? kSlots : 1];
    __shared__ float sv[SPLAT == 1 || SPLAT == 3 ? 3 * kSlots : 1];
    extern __shared__ float spool[];  // kPool: kWaves * kPoolWords (the launch's dynamic LDS)
    __shared__ uint32_t sfill;
    LdsSplat<kSlots, SPLAT == 3 ? kHashBinProbes : 8, false> hash{sk, sv, &sfill, nullptr};
    if (threadIdx.x == 0) {
        B.scene = nullptr;
        B.tbl_src = nullptr;
    }
    if (SPLAT == 1 || SPLAT == 3) hash.init();
    // binned deposits: the wave-uniform record sink (its chunk state in registers)
    WaveSink<kSlots, SPLAT == 3, THREADS> ws{{sk, sv, &sfill, nullptr}, bins.rkey, bins.rval, bins.pool_next,
                                              bins.chunk_fill, bins.pool_cap, xcc_id() % kBinPools,
                                              bins.tile_wlog2};
    const uint32_t home = blockIdx.x % kXcdQueues;
    uint32_t qi = 0;
    while (true) {
        const uint32_t it = claim_item(q, B, home, qi);
        if (it == 0xFFFFFFFFu) break;
        const BatchItem item = items[it];
        const uint32_t j = uniform_u32(item.job);
        const uint32_t kind = uniform_u32(jobs[j].kind);
        const PhotonParams &P = jobs[j].ph;
        stage_for(B, P.scene, P.tbl);
        if constexpr (SPLAT == 0) {
            GlobalSink g{deposit_film(P)};
            batch_photon_item(B.S, B.tbl, P, kind, item.start, item.count, g);
        } else if constexpr (SPLAT >= 2) {
            ws.job(P.key_base, kind == SPX_JOB_LIGHTTRACE ? P.scr_w : P.w, P.film);
            conv_photon_item<kPool>(B.S, B.tbl, P, kind, item.start, item.count, ws,
                                    kPool ? spool + (threadIdx.x >> 6) * kPoolWords : nullptr);
            ws.flush();
        } else {
            hash.film = deposit_film(P);
            batch_photon_item(B.S, B.tbl, P, kind, item.start, item.count, hash);
            hash.flush();
        }
        __syncthreads();
    }
    if (SPLAT >= 2) ws.finish();
}

// Zeroes the claim counters of every round's queues (the graph replays the
// same image, so the counters are reset by a kernel, not by the upload).
__global__ __launch_bounds__(256) void k_queues_reset(char *base, uint32_t n_rounds, uint32_t stride) {
    for (uint32_t i = threadIdx.x; i < n_rounds * kXcdQueues; i += blockDim.x) {
        BatchQueues *q = reinterpret_cast<BatchQueues *>(base + (size_t)(i / kXcdQueues) * stride);
        q->next[i % kXcdQueues] = 0u;
    }
}

hipError_t launch_queues_reset(BatchQueues *q, uint32_t n_rounds, size_t stride, hipStream_t s) {
    hipLaunchKernelGGL(k_queues_reset, dim3(1), dim3(256), 0, s, reinterpret_cast<char *>(q), n_rounds,
                       (uint32_t)stride);
    return hipGetLastError();
}

hipError_t launch_batch(bool camera, const BatchJob *jobs, const BatchItem *items, BatchQueues *q,
                        uint32_t blocks, uint32_t splat_mode, uint32_t photon_threads, const BinState &bins,
                        hipStream_t s) {
    const bool wide = photon_threads == 1024 && (splat_mode == 1 || splat_mode == 3) && blocks >= 4;
    if (camera)
        hipLaunchKernelGGL(k_batch_camera, dim3(blocks), dim3(256), 0, s, jobs, items, q);
    else if (splat_mode == 0)
        hipLaunchKernelGGL(k_batch_photons<0>, dim3(blocks), dim3(256), 0, s, jobs, items, q, bins);
    else if (wide && splat_mode == 1)
        hipLaunchKernelGGL((k_batch_photons<1, 1024>), dim3(blocks / 4), dim3(1024), 0, s, jobs, items, q, bins);
    else if (wide)
        hipLaunchKernelGGL((k_batch_photons<3, 1024>), dim3(blocks / 4), dim3(1024), 0, s, jobs, items, q, bins);
    else if (splat_mode == 1)
        hipLaunchKernelGGL(k_batch_photons<1>, dim3(blocks), dim3(256), 0, s, jobs, items, q, bins);
    else if (splat_mode == 2)
        hipLaunchKernelGGL(k_batch_photons<2>, dim3(blocks), dim3(256), kPhotonPoolBytes, s, jobs, items, q, bins);
    else
        hipLaunchKernelGGL(k_batch_photons<3>, dim3(blocks), dim3(256), kPhotonPoolBytes, s, jobs, items, q, bins);
    return hipGetLastError();
}

// Sums R film copies into the caller's film: one block per 1024 floats.
__global__ __launch_bounds__(256) void k_reduce_replicas(const ReplicaJob *jobs, uint32_t n_jobs) {
    // binary search the job owning this block
    uint32_t lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (jobs[mid].first_block <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const ReplicaJob J = jobs[lo];
    uint32_t base = (blockIdx.x - J.first_block) * 1024u;
    for (uint32_t k = threadIdx.x; k < 1024u; k += blockDim.x) {
        uint32_t i = base + k;
        if (i >= J.n) break;
        float s = J.dst[i];
        for (uint32_t r = 0; r < J.reps; ++r) s += J.src[(size_t)r * J.n + i];
        J.dst[i] = s;
    }
}

__global__ __launch_bounds__(256) void k_reduce_one(ReplicaJob J) {
    uint32_t base = blockIdx.x * 1024u;
    for (uint32_t k = threadIdx.x; k < 1024u; k += blockDim.x) {
        uint32_t i = base + k;
        if (i >= J.n) break;
        float s = J.dst[i];
        for (uint32_t r = 0; r < J.reps; ++r) s += J.src[(size_t)r * J.n + i];
        J.dst[i] = s;
    }
}

hipError_t launch_reduce_one(const ReplicaJob &J, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_one, dim3((J.n + 1023) / 1024), dim3(256), 0, s, J);
    return hipGetLastError();
}

hipError_t launch_reduce_replicas(const ReplicaJob *jobs_dev, uint32_t n_jobs, uint32_t total_blocks,
                                  hipStream_t s) {
    if (!n_jobs || !total_blocks) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_replicas, dim3(total_blocks), dim3(256), 0, s, jobs_dev, n_jobs);
    return hipGetLastError();
}

// fp32 atomic throughput probe: each of `threads` threads adds 1.0 to `ops`
// addresses drawn by a hash from [0, n_slots).
__global__ __launch_bounds__(256) void k_atomic_probe(float *buf, uint32_t n_slots, uint32_t ops, uint32_t threads) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= threads) return;
    for (uint32_t k = 0; k < ops; ++k) {
        uint32_t h = pcg32(t * ops + k);
        unsafeAtomicAdd(buf + (h % n_slots), 1.0f);
    }
}

int atomic_probe(spx_ctx *ctx, float *buf, size_t n_slots, uint32_t ops, uint32_t threads, hipStream_t s) {
    (void)ctx;
    hipLaunchKernelGGL(k_atomic_probe, dim3((threads + 255) / 256), dim3(256), 0, s, buf, (uint32_t)n_slots,
                       ops, threads);
    return hipGetLastError() == hipSuccess ? SPX_OK : SPX_EDEVICE;
}

}  // namespace spx
