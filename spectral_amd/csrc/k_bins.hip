// k_bins.hip -- the count / totals / scan / scatter / accumulate passes of
// binned photon deposits (spx_bins.hpp).
#include "spx_bins.hpp"
#include "spx_internal.hpp"

namespace spx {

constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;
constexpr uint32_t kBinThreads = 256;
constexpr uint32_t kRecsPerThread = kGroupRecs / kBinThreads;  // 16, held in registers
constexpr uint32_t kProbe = 16;  // count-pass probe bound; a longer chain overflows
static_assert(kGroupSlots <= kGroupRecs, "a group fills at most kGroupRecs slots");

__global__ __launch_bounds__(256) void k_bins_reset(BinState st) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    if (t < kBinPools) st.pool_next[t] = 0;
    if (t == 0) *st.part_next = 0;
    for (size_t i = t; i < st.zero_words; i += stride) st.sub_count[i] = 0;
}

__device__ inline uint32_t strip_hash(uint32_t strip) { return (strip * 2654435761u) & (kGroupSlots - 1); }

// Count pass: inserts `strip` within kProbe probes; kHashEmpty on overflow.
// A strip that fails once fails for the whole group (slots are never freed),
// so a strip's records in one group are either all tabled or all overflow.
__device__ inline uint32_t strip_slot(uint32_t *keys, uint32_t strip) {
    const uint32_t h = strip_hash(strip);
    for (uint32_t p = 0; p < kProbe; ++p) {
        uint32_t s = (h + p) & (kGroupSlots - 1);
        uint32_t k = keys[s];
        if (k == strip) return s;
        if (k == kHashEmpty) {
            uint32_t old = atomicCAS(keys + s, kHashEmpty, strip);
            if (old == kHashEmpty || old == strip) return s;
        }
    }
    return kHashEmpty;
}

// Scatter pass: the table holds only the group's listed strips (at most
// kGroupSlots), inserted without a probe bound, so a lookup ends at the key,
// at an empty slot, or after one full cycle.
__device__ inline uint32_t slot_insert(uint32_t *keys, uint32_t strip) {
    const uint32_t h = strip_hash(strip);
    for (uint32_t p = 0; p < kGroupSlots; ++p) {
        uint32_t s = (h + p) & (kGroupSlots - 1);
        uint32_t old = atomicCAS(keys + s, kHashEmpty, strip);
        if (old == kHashEmpty || old == strip) return s;
    }
    return 0;  // unreachable: a list never holds more strips than slots
}

__device__ inline uint32_t slot_find(const uint32_t *keys, uint32_t strip) {
    const uint32_t h = strip_hash(strip);
    for (uint32_t p = 0; p < kGroupSlots; ++p) {
        uint32_t s = (h + p) & (kGroupSlots - 1);
        uint32_t k = keys[s];
        if (k == strip) return s;
        if (k == kHashEmpty) break;
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
    uint32_t ls = __builtin_amdgcn_readlane(slot, leader);
    if (__ballot(active && slot == ls) == m) {
        uint32_t lane = __lane_id();
        uint32_t before = 0;
        if ((int)lane == leader) before = atomicAdd(cnt + ls, (uint32_t)__popcll(m));
        before = __builtin_amdgcn_readlane(before, leader);
        return before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    }
    return active ? atomicAdd(cnt + slot, 1u) : 0u;
}

// Chunk groups handed out so far, over all pools.
__device__ inline bool group_used(const BinState &st, uint32_t g) {
    const uint32_t per_pool = st.pool_cap / kBinChunk;
    uint32_t c0 = g * kChunksPerBlock;
    uint32_t pool = c0 / per_pool, local = c0 % per_pool;
    return pool < kBinPools && local * kBinChunk < st.pool_next[pool];
}

// Record counts of the group's chunks into s_n (thread < kChunksPerBlock).
__device__ inline void load_group_counts(const BinState &st, uint32_t c0, uint32_t *s_n) {
    if (threadIdx.x < kChunksPerBlock) s_n[threadIdx.x] = chunk_records(st, c0 + threadIdx.x);
}

// Slot k of this thread: record k * kBinThreads + threadIdx.x of the group,
// i.e. chunk k / 2, offset (k % 2) * kBinThreads + threadIdx.x.
__device__ inline bool group_rec(const uint32_t *s_n, uint32_t c0, uint32_t k, uint32_t &at) {
    constexpr uint32_t per_chunk = kBinChunk / kBinThreads;
    const uint32_t c = k / per_chunk, o = (k % per_chunk) * kBinThreads + threadIdx.x;
    at = (c0 + c) * kBinChunk + o;
    return o < s_n[c];
}

// Bit pattern of max(|x|, |y|, |z|): for non-negative floats (and +inf, NaN
// above it) the integer order is the numeric order.
__device__ inline uint32_t mag_bits(const BinRec &r) {
    uint32_t a = __float_as_uint(r.x) & 0x7FFFFFFFu, b = __float_as_uint(r.y) & 0x7FFFFFFFu,
             c = __float_as_uint(r.z) & 0x7FFFFFFFu;
    return max(a, max(b, c));
}

// Count pass: streams the keys only.  Each group's run of a strip is
// reserved padded to kRunAlign slots, so every run starts on a multiple of
// kRunAlign within its strip.
__global__ __launch_bounds__(kBinThreads) void k_bins_count(BinState st) {
    __shared__ uint32_t keys[kGroupSlots];
    __shared__ uint32_t cnt[kGroupSlots];
    __shared__ uint32_t s_n[kChunksPerBlock];
    __shared__ uint32_t s_m;
    const uint32_t groups = kBinPools * (st.pool_cap / kBinChunk) / kChunksPerBlock;
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        if (!group_used(st, g)) continue;  // block-uniform
        for (uint32_t s = threadIdx.x; s < kGroupSlots; s += blockDim.x) {
            keys[s] = kHashEmpty;
            cnt[s] = 0;
        }
        if (threadIdx.x == 0) s_m = 0;
        const uint32_t c0 = g * kChunksPerBlock;
        load_group_counts(st, c0, s_n);
        __syncthreads();
        // All keys first (16 loads in flight per thread), then the LDS work.
        uint32_t strip[kRecsPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kRecsPerThread; ++k) {
            uint32_t at;
            strip[k] = group_rec(s_n, c0, k, at) ? st.rkey[at] >> kStripLog2 : kHashEmpty;
        }
#pragma unroll
        for (uint32_t k = 0; k < kRecsPerThread; ++k) {
            const bool act = strip[k] != kHashEmpty;
            uint32_t sl = act ? strip_slot(keys, strip[k]) : kHashEmpty;
            slot_add(cnt, sl, act && sl != kHashEmpty);
            if (act && sl == kHashEmpty) atomicAdd(st.ovf_count + strip[k], 1u);
        }
        __syncthreads();
        const uint32_t sub = g % kSub;
        for (uint32_t s = threadIdx.x; s < kGroupSlots; s += blockDim.x) {
            const uint32_t key = keys[s];
            if (key == kHashEmpty) continue;
            const size_t i = (size_t)key * kSub + sub;
            uint32_t rel = atomicAdd(st.sub_count + i, run_pad(cnt[s]));
            uint32_t j = atomicAdd(&s_m, 1u);
            st.glist[(size_t)g * kGroupSlots + j] = make_uint2(key, rel);
        }
        __syncthreads();
        if (threadIdx.x == 0) st.gcount[g] = s_m;
    }
}

// Per strip: the slots it uses (padded runs, then its overflow records)
// and the sub-bucket sizes turned into offsets within the strip
// (sub-buckets 0..kSub-1, then the overflow bucket).
__global__ __launch_bounds__(256) void k_bins_totals(BinState st) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= st.n_strips) return;
    uint32_t c[kSub];
    const uint4 *sc = (const uint4 *)(st.sub_count + (size_t)s * kSub);
#pragma unroll
    for (uint32_t q = 0; q < kSub / 4; ++q) {
        uint4 v = sc[q];
        c[4 * q] = v.x;
        c[4 * q + 1] = v.y;
        c[4 * q + 2] = v.z;
        c[4 * q + 3] = v.w;
    }
    uint32_t run = 0;
    uint4 *so = (uint4 *)(st.sub_count + (size_t)s * kSub);
#pragma unroll
    for (uint32_t q = 0; q < kSub / 4; ++q) {
        uint4 o;
        o.x = run;
        run += c[4 * q];
        o.y = run;
        run += c[4 * q + 1];
        o.z = run;
        run += c[4 * q + 2];
        o.w = run;
        run += c[4 * q + 3];
        so[q] = o;
    }
    const uint32_t ovf = st.ovf_count[s];
    st.ovf_count[s] = run;
    st.strip_total[s] = run + ovf;
}

// Exclusive scans of the strips' slots (each padded to kRunAlign, so every
// strip starts aligned) and of the accumulation parts.
__global__ __launch_bounds__(1024) void k_bins_scan(BinState st) {
    __shared__ uint32_t sa[1024], sp[1024];
    const uint32_t n = st.n_strips, T = blockDim.x;
    const uint32_t per = (n + T - 1) / T;
    const uint32_t b0 = threadIdx.x * per < n ? threadIdx.x * per : n, b1 = b0 + per < n ? b0 + per : n;
    uint32_t a = 0, p = 0;
    for (uint32_t i = b0; i < b1; ++i) {
        uint32_t c = st.strip_total[i];
        a += run_pad(c);
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
        uint32_t c = st.strip_total[i];
        st.strip_off[i] = ra;
        st.part_off[i] = rp;
        ra += run_pad(c);
        rp += (c + kPartRecs - 1) / kPartRecs;
    }
    if (threadIdx.x == T - 1) {
        st.strip_off[n] = sa[T - 1];
        st.part_off[n] = sp[T - 1];
    }
    // Accumulator indices of the strips split over several parts.
    __syncthreads();
    uint32_t m = 0;
    for (uint32_t i = b0; i < b1; ++i) m += st.strip_total[i] > kPartRecs ? 1u : 0u;
    sa[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t off = 1; off < T; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? sa[threadIdx.x - off] : 0;
        __syncthreads();
        sa[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t rm = threadIdx.x ? sa[threadIdx.x - 1] : 0;
    for (uint32_t i = b0; i < b1; ++i) {
        uint32_t idx = kNoMacc;
        if (st.strip_total[i] > kPartRecs) {
            if (rm < st.max_multi) {
                idx = rm;
                st.macc_strip[rm] = i;
            }
            ++rm;
        }
        st.strip_macc[i] = idx;
    }
    if (threadIdx.x == T - 1) *st.macc_count = sa[T - 1] < st.max_multi ? sa[T - 1] : st.max_multi;
}

// Clears the accumulators of this round's split strips.
__global__ __launch_bounds__(256) void k_bins_macc_zero(BinState st) {
    const size_t n = (size_t)*st.macc_count * 3 * kStripPix;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        st.macc[i] = 0ull;
}

// Scatter pass: each record to its slot in strip order, the slack of every
// padded run filled with pad records, and each strip's largest magnitude
// (for the fixed-point scale) taken on the way.
__global__ __launch_bounds__(kBinThreads) void k_bins_scatter(BinState st) {
    __shared__ uint32_t keys[kGroupSlots];
    __shared__ uint32_t cnt[kGroupSlots];
    __shared__ uint32_t base[kGroupSlots];
    __shared__ uint32_t mx[kGroupSlots];
    __shared__ uint32_t s_n[kChunksPerBlock];
    const uint32_t groups = kBinPools * (st.pool_cap / kBinChunk) / kChunksPerBlock;
    for (uint32_t g = blockIdx.x; g < groups; g += gridDim.x) {
        if (!group_used(st, g)) continue;
        for (uint32_t s = threadIdx.x; s < kGroupSlots; s += blockDim.x) {
            keys[s] = kHashEmpty;
            cnt[s] = 0;
            mx[s] = 0;
        }
        const uint32_t c0 = g * kChunksPerBlock;
        load_group_counts(st, c0, s_n);
        __syncthreads();
        // The group's records stay in registers; their loads overlap the table build.
        BinRec r[kRecsPerThread];
        uint32_t live = 0;  // bit k: slot k holds a record
#pragma unroll
        for (uint32_t k = 0; k < kRecsPerThread; ++k) {
            uint32_t at;
            if (group_rec(s_n, c0, k, at)) {
                r[k].key = st.rkey[at];
                r[k].x = st.rval[3 * (size_t)at + 0];
                r[k].y = st.rval[3 * (size_t)at + 1];
                r[k].z = st.rval[3 * (size_t)at + 2];
                live |= 1u << k;
            } else {
                r[k].key = 0;
            }
        }
        const uint32_t m = st.gcount[g], sub = g % kSub;
        for (uint32_t j = threadIdx.x; j < m; j += blockDim.x) {
            const uint2 e = st.glist[(size_t)g * kGroupSlots + j];
            const uint32_t sl = slot_insert(keys, e.x);
            base[sl] = st.strip_off[e.x] + st.sub_count[(size_t)e.x * kSub + sub] + e.y;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kRecsPerThread; ++k) {
            const bool act = (live >> k) & 1u;
            const uint32_t strip = r[k].key >> kStripLog2;
            const uint32_t sl = act ? slot_find(keys, strip) : kHashEmpty;
            const uint32_t rank = slot_add(cnt, sl, act && sl != kHashEmpty);
            if (act) {
                uint32_t pos;
                if (sl != kHashEmpty) {
                    pos = base[sl] + rank;
                    atomicMax(mx + sl, mag_bits(r[k]));
                } else {
                    pos = st.strip_off[strip] + st.ovf_count[strip] + atomicAdd(st.ovf_cur + strip, 1u);
                    atomicMax(st.strip_max + strip, mag_bits(r[k]));
                }
                st.sorted[pos] = r[k];
            }
        }
        __syncthreads();
        // Pads after each run, and the runs' largest magnitudes.
        for (uint32_t j = threadIdx.x; j < m; j += blockDim.x) {
            const uint32_t strip = st.glist[(size_t)g * kGroupSlots + j].x;
            const uint32_t sl = slot_find(keys, strip);
            const uint32_t n = cnt[sl];
            BinRec pad;
            pad.key = kPadKey;
            pad.x = pad.y = pad.z = 0.0f;
            for (uint32_t q = n; q < run_pad(n); ++q) st.sorted[base[sl] + q] = pad;
            atomicMax(st.strip_max + strip, mx[sl]);
        }
        __syncthreads();
    }
}

// Scale exponent for a strip of `total` records whose largest magnitude has
// bit pattern `mb` (M < 2^(ex+1)): with b = bit length of total, every
// scaled value stays below 2^(63-b), so the strip's sum stays below 2^63.
// Resolution: 2^(ex-62+b), e.g. M * 2^-47 for a strip of 2^14 records.
__device__ inline int fixed_exponent(uint32_t mb, uint32_t total) {
    const int ex = (int)((mb >> 23) & 0xFFu) - 127;  // floor(log2 M) (-127 for subnormal M)
    const int b = 32 - __clz(total);
    return 62 - b - ex;
}

__device__ inline unsigned long long to_fixed(float v, int e) {
    return (unsigned long long)__float2ll_rn(ldexpf(v, e));
}

__device__ inline float from_fixed(unsigned long long a, int e) {
    return (float)ldexp((double)(long long)a, -e);
}

__global__ __launch_bounds__(256) void k_bins_accumulate(BinState st, const uint32_t *job_shared) {
    __shared__ unsigned long long acc[3 * kStripPix];  // fixed point, or fp32 in its first half
    __shared__ uint32_t s_part;
    float *accf = (float *)acc;
    const uint32_t total_parts = st.part_off[st.n_strips];
    while (true) {
        if (threadIdx.x == 0) s_part = atomicAdd(st.part_next, 1u);
        for (uint32_t k = threadIdx.x; k < 3 * kStripPix; k += blockDim.x) acc[k] = 0ull;
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
        if (r1 > st.strip_off[s] + st.strip_total[s]) r1 = st.strip_off[s] + st.strip_total[s];
        const uint32_t mb = st.strip_max[s];
        // Non-finite deposits keep fp32 sums, so inf / NaN reach the film.
        const uint32_t macc_i = nparts > 1 ? st.strip_macc[s] : kNoMacc;
        const bool fixed = mb < 0x7F800000u && (nparts == 1 || macc_i != kNoMacc);
        const int e = fixed_exponent(mb, st.strip_total[s]);
        constexpr uint32_t U = 8;  // loads in flight per thread
        for (uint32_t i0 = r0; i0 < r1; i0 += U * blockDim.x) {
            BinRec r[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                uint32_t i = i0 + u * blockDim.x + threadIdx.x;
                if (i < r1) r[u] = st.sorted[i];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                uint32_t i = i0 + u * blockDim.x + threadIdx.x;
                if (i >= r1 || r[u].key == kPadKey) continue;
                uint32_t l = r[u].key & (kStripPix - 1);
                if (fixed) {
                    atomicAdd(acc + 3 * l + 0, to_fixed(r[u].x, e));
                    atomicAdd(acc + 3 * l + 1, to_fixed(r[u].y, e));
                    atomicAdd(acc + 3 * l + 2, to_fixed(r[u].z, e));
                } else {
                    __hip_atomic_fetch_add(accf + 3 * l + 0, r[u].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(accf + 3 * l + 1, r[u].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(accf + 3 * l + 2, r[u].z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
        __syncthreads();
        const uint32_t job = st.strip_job[s];
        const uint32_t tile = ((s << kStripLog2) - st.job_key[job]) >> kStripLog2;
        const uint32_t npix = st.job_npix[job], width = st.job_width[job];
        const uint32_t height = npix / width;
        const uint32_t tiles_x = tiles_across(width, st.tile_wlog2);
        float *film = st.job_film[job];
        const bool exclusive = nparts == 1 && !job_shared[job] && !st.atomic_film;
        if (fixed && nparts > 1) {  // into the strip's accumulator; written by k_bins_macc_flush
            unsigned long long *dst = st.macc + (size_t)macc_i * 3 * kStripPix;
            for (uint32_t k = threadIdx.x; k < 3 * kStripPix; k += blockDim.x)
                if (acc[k]) atomicAdd(dst + k, acc[k]);
            __syncthreads();
            continue;
        }
        for (uint32_t k = threadIdx.x; k < kStripPix; k += blockDim.x) {
            uint32_t px, py;
            tile_pixel(tile, k, tiles_x, st.tile_wlog2, px, py);
            float x, y, z;
            if (fixed) {
                x = from_fixed(acc[3 * k], e);
                y = from_fixed(acc[3 * k + 1], e);
                z = from_fixed(acc[3 * k + 2], e);
            } else {
                x = accf[3 * k];
                y = accf[3 * k + 1];
                z = accf[3 * k + 2];
            }
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

// Film writes of the split strips: one owner per pixel.
__global__ __launch_bounds__(256) void k_bins_macc_flush(BinState st, const uint32_t *job_shared) {
    const size_t n = (size_t)*st.macc_count * kStripPix;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t m = (uint32_t)(i >> kStripLog2), k = (uint32_t)(i & (kStripPix - 1));
        const uint32_t s = st.macc_strip[m];
        const unsigned long long *a = st.macc + (size_t)m * 3 * kStripPix + 3 * k;
        if (!a[0] && !a[1] && !a[2]) continue;
        const int e = fixed_exponent(st.strip_max[s], st.strip_total[s]);
        const uint32_t job = st.strip_job[s];
        const uint32_t tile = ((s << kStripLog2) - st.job_key[job]) >> kStripLog2;
        const uint32_t npix = st.job_npix[job], width = st.job_width[job];
        uint32_t px, py;
        tile_pixel(tile, k, tiles_across(width, st.tile_wlog2), st.tile_wlog2, px, py);
        if (px >= width || py >= npix / width) continue;
        const float x = from_fixed(a[0], e), y = from_fixed(a[1], e), z = from_fixed(a[2], e);
        float *film = st.job_film[job];
        const uint32_t pix = py * width + px;
        if (job_shared[job] || st.atomic_film) {
            film_add(film, pix, mk(x, y, z));
        } else {
            film[3 * pix + 0] += x;
            film[3 * pix + 1] += y;
            film[3 * pix + 2] += z;
        }
    }
}

hipError_t launch_bins_reset(const BinState &st, hipStream_t s) {
    size_t blocks = (st.zero_words + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    hipLaunchKernelGGL(k_bins_reset, dim3((uint32_t)blocks), dim3(256), 0, s, st);
    return hipGetLastError();
}

hipError_t launch_bins_resolve(const BinState &st, const uint32_t *job_shared, uint32_t acc_blocks,
                               hipStream_t s) {
    uint32_t groups = kBinPools * (st.pool_cap / kBinChunk) / kChunksPerBlock;
    uint32_t blocks = groups < 2 * acc_blocks ? groups : 2 * acc_blocks;  // grid-stride over used groups
    hipLaunchKernelGGL(k_bins_count, dim3(blocks), dim3(kBinThreads), 0, s, st);
    hipLaunchKernelGGL(k_bins_totals, dim3(st.n_strips ? (st.n_strips + 255) / 256 : 1), dim3(256), 0, s, st);
    hipLaunchKernelGGL(k_bins_scan, dim3(1), dim3(1024), 0, s, st);
    hipLaunchKernelGGL(k_bins_macc_zero, dim3(acc_blocks), dim3(256), 0, s, st);
    hipLaunchKernelGGL(k_bins_scatter, dim3(blocks), dim3(kBinThreads), 0, s, st);
    hipLaunchKernelGGL(k_bins_accumulate, dim3(acc_blocks), dim3(256), 0, s, st, job_shared);
    hipLaunchKernelGGL(k_bins_macc_flush, dim3(acc_blocks), dim3(256), 0, s, st, job_shared);
    return hipGetLastError();
}

}  // namespace spx
