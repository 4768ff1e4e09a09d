// k_bins.hip -- the count / scan / scatter / accumulate passes of binned
// photon deposits (spx_bins.hpp).
#include "spx_bins.hpp"
#include "spx_internal.hpp"

namespace spx {

constexpr uint32_t kHashSlots = 1024;  // distinct strips per chunk tracked in LDS
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;

__global__ __launch_bounds__(256) void k_bins_reset(BinState st) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kBinPools) st.pool_next[t] = 0;
    if (t == 0) *st.part_next = 0;
    for (uint32_t i = t; i < st.n_strips; i += gridDim.x * blockDim.x) st.strip_count[i] = 0;
}

// Inserts `strip` into the LDS table; returns the slot or kHashEmpty on overflow.
__device__ inline uint32_t strip_slot(uint32_t *keys, uint32_t strip) {
    uint32_t h = (strip * 2654435761u) & (kHashSlots - 1);
    for (uint32_t p = 0; p < 16; ++p) {
        uint32_t s = (h + p) & (kHashSlots - 1);
        uint32_t k = keys[s];
        if (k == strip) return s;
        if (k == kHashEmpty) {
            uint32_t old = atomicCAS(keys + s, kHashEmpty, strip);
            if (old == kHashEmpty || old == strip) return s;
        }
    }
    return kHashEmpty;
}

// Records of chunk `c` (valid only if it was handed out).
__device__ inline uint32_t chunk_records(const BinState &st, uint32_t c) {
    const uint32_t per_pool = st.pool_cap / kBinChunk;
    uint32_t pool = c / per_pool, local = c % per_pool;
    if (pool >= kBinPools || local * kBinChunk >= st.pool_next[pool]) return 0;
    uint32_t n = st.chunk_fill[c];
    return n > kBinChunk ? kBinChunk : n;
}

// Adds 1 per active lane to the LDS counter of its slot and returns a unique
// rank among all additions to that slot (the counter's value before this
// lane's addition).  When every active lane of the wave holds the same slot,
// one atomic serves the wave; otherwise each lane adds its own.
__device__ inline uint32_t slot_add(uint32_t *cnt, uint32_t slot, bool active) {
    uint64_t m = __ballot(active);
    if (!m) return 0;
    int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t ls = __shfl(slot, leader);
    if (__ballot(active && slot == ls) == m) {
        uint32_t lane = __lane_id();
        uint32_t before = 0;
        if ((int)lane == leader) before = atomicAdd(cnt + ls, (uint32_t)__popcll(m));
        before = __shfl(before, leader);
        return before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }
    return active ? atomicAdd(cnt + slot, 1u) : 0u;
}

// Chunk groups (kChunksPerBlock chunks) handed out so far, over all pools.
__device__ inline bool group_used(const BinState &st, uint32_t g) {
    const uint32_t per_pool = st.pool_cap / kBinChunk;
    uint32_t c0 = g * kChunksPerBlock;
    uint32_t pool = c0 / per_pool, local = c0 % per_pool;
    return pool < kBinPools && local * kBinChunk < st.pool_next[pool];
}

__global__ __launch_bounds__(256) void k_bins_count(BinState st) {
    __shared__ uint32_t keys[kHashSlots];
    __shared__ uint32_t cnt[kHashSlots];
    const uint32_t groups = kBinPools * (st.pool_cap / kBinChunk) / kChunksPerBlock;
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        if (!group_used(st, g)) continue;  // block-uniform
        for (uint32_t s = threadIdx.x; s < kHashSlots; s += blockDim.x) {
            keys[s] = kHashEmpty;
            cnt[s] = 0;
        }
        __syncthreads();
        const uint32_t c0 = g * kChunksPerBlock;
        for (uint32_t c = 0; c < kChunksPerBlock; ++c) {
            const uint32_t n = chunk_records(st, c0 + c), begin = (c0 + c) * kBinChunk;
            for (uint32_t i0 = 0; i0 < n; i0 += blockDim.x) {
                uint32_t i = i0 + threadIdx.x;
                bool act = i < n;
                uint32_t strip = act ? st.recs[begin + i].key >> kStripLog2 : 0;
                uint32_t s = act ? strip_slot(keys, strip) : kHashEmpty;
                slot_add(cnt, s, act && s != kHashEmpty);
                if (act && s == kHashEmpty) atomicAdd(st.strip_count + strip, 1u);
            }
        }
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < kHashSlots; s += blockDim.x)
            if (keys[s] != kHashEmpty) atomicAdd(st.strip_count + keys[s], cnt[s]);
        __syncthreads();
    }
}

// Exclusive scans of the strip record counts and of the accumulation parts.
__global__ __launch_bounds__(1024) void k_bins_scan(BinState st) {
    __shared__ uint32_t sa[1024], sp[1024];
    const uint32_t n = st.n_strips, T = blockDim.x;
    const uint32_t per = (n + T - 1) / T;
    const uint32_t b0 = threadIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
    uint32_t a = 0, p = 0;
    for (uint32_t i = b0; i < b1; ++i) {
        uint32_t c = st.strip_count[i];
        a += c;
        p += (c + kPartRecs - 1) / kPartRecs;
    }
    sa[threadIdx.x] = a;
    sp[threadIdx.x] = p;
    __syncthreads();
    for (uint32_t off = 1; off < T; off <<= 1) {  // inclusive Hillis-Steele scan
        uint32_t va = threadIdx.x >= off ? sa[threadIdx.x - off] : 0;
        uint32_t vp = threadIdx.x >= off ? sp[threadIdx.x - off] : 0;
        __syncthreads();
        sa[threadIdx.x] += va;
        sp[threadIdx.x] += vp;
        __syncthreads();
    }
    uint32_t ra = threadIdx.x ? sa[threadIdx.x - 1] : 0;
    uint32_t rp = threadIdx.x ? sp[threadIdx.x - 1] : 0;
    for (uint32_t i = b0; i < b1; ++i) {
        uint32_t c = st.strip_count[i];
        st.strip_off[i] = ra;
        st.strip_cur[i] = ra;
        st.part_off[i] = rp;
        ra += c;
        rp += (c + kPartRecs - 1) / kPartRecs;
    }
    if (threadIdx.x == T - 1) {
        st.strip_off[n] = sa[T - 1];
        st.part_off[n] = sp[T - 1];
    }
}

__global__ __launch_bounds__(256) void k_bins_scatter(BinState st) {
    __shared__ uint32_t keys[kHashSlots];
    __shared__ uint32_t cnt[kHashSlots];
    __shared__ uint32_t base[kHashSlots];
    const uint32_t groups = kBinPools * (st.pool_cap / kBinChunk) / kChunksPerBlock;
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        if (!group_used(st, g)) continue;
        for (uint32_t s = threadIdx.x; s < kHashSlots; s += blockDim.x) {
            keys[s] = kHashEmpty;
            cnt[s] = 0;
        }
        __syncthreads();
        const uint32_t c0 = g * kChunksPerBlock;
        for (uint32_t c = 0; c < kChunksPerBlock; ++c) {
            const uint32_t n = chunk_records(st, c0 + c), begin = (c0 + c) * kBinChunk;
            for (uint32_t i0 = 0; i0 < n; i0 += blockDim.x) {
                uint32_t i = i0 + threadIdx.x;
                bool act = i < n;
                uint32_t s = act ? strip_slot(keys, st.recs[begin + i].key >> kStripLog2) : kHashEmpty;
                slot_add(cnt, s, act && s != kHashEmpty);
            }
        }
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < kHashSlots; s += blockDim.x) {
            if (keys[s] != kHashEmpty) base[s] = atomicAdd(st.strip_cur + keys[s], cnt[s]);
            cnt[s] = 0;
        }
        __syncthreads();
        for (uint32_t c = 0; c < kChunksPerBlock; ++c) {
            const uint32_t n = chunk_records(st, c0 + c), begin = (c0 + c) * kBinChunk;
            for (uint32_t i0 = 0; i0 < n; i0 += blockDim.x) {
                uint32_t i = i0 + threadIdx.x;
                bool act = i < n;
                BinRec r;
                uint32_t strip = 0, s = kHashEmpty;
                if (act) {
                    r = st.recs[begin + i];
                    strip = r.key >> kStripLog2;
                    s = strip_slot(keys, strip);  // a lookup: inserted above, or overflowed
                }
                uint32_t rank = slot_add(cnt, s, act && s != kHashEmpty);
                if (act) {
                    uint32_t pos = s != kHashEmpty ? base[s] + rank : atomicAdd(st.strip_cur + strip, 1u);
                    st.sorted[pos] = r;
                }
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_bins_accumulate(BinState st, const uint32_t *job_shared) {
    __shared__ float acc[3 * kStripPix];
    __shared__ uint32_t s_part;
    const uint32_t total_parts = st.part_off[st.n_strips];
    while (true) {
        if (threadIdx.x == 0) s_part = atomicAdd(st.part_next, 1u);
        for (uint32_t k = threadIdx.x; k < 3 * kStripPix; k += blockDim.x) acc[k] = 0.0f;
        __syncthreads();
        const uint32_t p = s_part;
        if (p >= total_parts) break;
        // strip s: part_off[s] <= p < part_off[s + 1]
        uint32_t lo = 0, hi = st.n_strips;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (st.part_off[mid] <= p) lo = mid;
            else hi = mid;
        }
        const uint32_t s = lo;
        const uint32_t q = p - st.part_off[s];
        const uint32_t nparts = st.part_off[s + 1] - st.part_off[s];
        const uint32_t r0 = st.strip_off[s] + q * kPartRecs;
        uint32_t r1 = r0 + kPartRecs;
        if (r1 > st.strip_off[s + 1]) r1 = st.strip_off[s + 1];
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            BinRec r = st.sorted[i];
            uint32_t l = r.key & (kStripPix - 1);
            __hip_atomic_fetch_add(acc + 3 * l + 0, r.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc + 3 * l + 1, r.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc + 3 * l + 2, r.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();
        const uint32_t job = st.strip_job[s];
        const uint32_t tile = ((s << kStripLog2) - st.job_key[job]) >> kStripLog2;
        const uint32_t npix = st.job_npix[job], width = st.job_width[job];
        const uint32_t height = npix / width;
        const uint32_t tiles_x = (width + (1u << kTileLog2) - 1) >> kTileLog2;
        const uint32_t tx0 = (tile % tiles_x) << kTileLog2, ty0 = (tile / tiles_x) << kTileLog2;
        float *film = st.job_film[job];
        const bool exclusive = nparts == 1 && !job_shared[job];
        for (uint32_t k = threadIdx.x; k < kStripPix; k += blockDim.x) {
            uint32_t px = tx0 + (k & ((1u << kTileLog2) - 1)), py = ty0 + (k >> kTileLog2);
            float x = acc[3 * k], y = acc[3 * k + 1], z = acc[3 * k + 2];
            if (px >= width || py >= height || (x == 0.0f && y == 0.0f && z == 0.0f)) continue;
            uint32_t pix = py * width + px;
            if (exclusive) {
                film[3 * pix + 0] += x;
                film[3 * pix + 1] += y;
                film[3 * pix + 2] += z;
            } else {
                film_add(film, pix, mk(x, y, z));
            }
        }
        __syncthreads();
    }
}

hipError_t launch_bins_reset(const BinState &st, hipStream_t s) {
    uint32_t blocks = (st.n_strips + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    hipLaunchKernelGGL(k_bins_reset, dim3(blocks), dim3(256), 0, s, st);
    return hipGetLastError();
}

hipError_t launch_bins_resolve(const BinState &st, const uint32_t *job_shared, uint32_t acc_blocks,
                               hipStream_t s) {
    uint32_t groups = kBinPools * (st.pool_cap / kBinChunk) / kChunksPerBlock;
    uint32_t blocks = groups < 2 * acc_blocks ? groups : 2 * acc_blocks;  // grid-stride over used groups
    hipLaunchKernelGGL(k_bins_count, dim3(blocks), dim3(256), 0, s, st);
    hipLaunchKernelGGL(k_bins_scan, dim3(1), dim3(1024), 0, s, st);
    hipLaunchKernelGGL(k_bins_scatter, dim3(blocks), dim3(256), 0, s, st);
    hipLaunchKernelGGL(k_bins_accumulate, dim3(acc_blocks), dim3(256), 0, s, st, job_shared);
    return hipGetLastError();
}

}  // namespace spx
