// spx_splat.hpp -- wave / block level film-splat aggregation in LDS.
//
// A block keeps an open-addressed hash of (pixel -> XYZ partial sum) in LDS.
// Deposits that land on the same pixel within the block are combined with
// LDS fp32 atomics (ds_add_f32); the table is drained to the film with
// global fp32 atomics (global_atomic_add_f32) when it is flushed.  A deposit
// that finds no free slot within its probes goes straight to the film, so the
// result never depends on the table size -- only the number of global
// atomics does.  Summation order changes float rounding only.
#pragma once

#include "spx_bins.hpp"
#include "spx_device.hpp"

namespace spx {

constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;

// PROBES: linear-probe length.  Draining to film atomics, a miss costs a
// global atomic, so the table probes far (8); draining to records, a miss
// costs only a record, and short probing is cheaper overall (2 measured
// fastest for the batched photon kernel: fewer divergent compare-and-swap
// rounds per deposit, slightly more records).
// COUNT_FILL: keep the occupied-slot counter that maybe_flush reads (tables
// drained only at fixed points skip it: one ballot and LDS atomic per insert).
template <int SLOTS, int PROBES = 8, bool COUNT_FILL = true>
struct LdsSplat {
    static_assert((SLOTS & (SLOTS - 1)) == 0, "power of two");
    uint32_t *keys;  // [SLOTS]
    float *vals;     // [3*SLOTS]
    uint32_t *fill;  // occupied-slot counter
    float *film;

    __device__ inline void init() {
        for (uint32_t k = threadIdx.x; k < (uint32_t)SLOTS; k += blockDim.x) {
            keys[k] = kEmptyKey;
            vals[3 * k] = vals[3 * k + 1] = vals[3 * k + 2] = 0.0f;
        }
        if (threadIdx.x == 0) *fill = 0;
    }

    __device__ inline void lds_add(uint32_t slot, f3 v) {
        __hip_atomic_fetch_add(vals + 3 * slot + 0, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(vals + 3 * slot + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(vals + 3 * slot + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }

    // Sums v into idx's slot; false when no slot is free within PROBES.
    // The occupied-slot counter takes one LDS atomic per wave, not per
    // inserting lane (all of a wave's inserts would hit that one address).
    __device__ inline bool try_add(uint32_t idx, f3 v) {
        uint32_t h = (idx * 2654435761u) & (uint32_t)(SLOTS - 1);
        int slot = -1;
        bool inserted = false;
        for (int p = 0; p < PROBES; ++p) {
            uint32_t s = (h + (uint32_t)p) & (uint32_t)(SLOTS - 1);
            if (PROBES <= 2) {
                // Short probes: one compare-and-swap per probe (it returns the
                // key when the slot is taken) instead of a read and then a
                // compare-and-swap, one LDS round trip fewer per new key.
                const uint32_t old = atomicCAS(keys + s, kEmptyKey, idx);
                if (old == kEmptyKey || old == idx) {
                    slot = (int)s;
                    inserted = old == kEmptyKey;
                    break;
                }
                continue;
            }
            uint32_t k = keys[s];
            if (k == idx) {
                slot = (int)s;
                break;
            }
            if (k == kEmptyKey) {
                uint32_t old = atomicCAS(keys + s, kEmptyKey, idx);
                if (old == kEmptyKey || old == idx) {
                    slot = (int)s;
                    inserted = old == kEmptyKey;
                    break;
                }
            }
        }
        if (COUNT_FILL) {
            const uint64_t ins = __ballot(inserted);
            if (ins && (int)__lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1)
                atomicAdd(fill, (uint32_t)__popcll(ins));
        }
        if (slot < 0) return false;
        lds_add((uint32_t)slot, v);
        return true;
    }

    __device__ inline void add(uint32_t idx, f3 v) {
        if (!try_add(idx, v)) film_add(film, idx, v);
    }

    // Whole block must call (contains barriers).  Hands every occupied slot
    // to emit(pixel, sum) and empties the table.
    template <class F>
    __device__ inline void drain(F emit) {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < (uint32_t)SLOTS; k += blockDim.x) {
            uint32_t key = keys[k];
            if (key != kEmptyKey) {
                emit(key, mk(vals[3 * k], vals[3 * k + 1], vals[3 * k + 2]));
                keys[k] = kEmptyKey;
                vals[3 * k] = vals[3 * k + 1] = vals[3 * k + 2] = 0.0f;
            }
        }
        if (threadIdx.x == 0) *fill = 0;
        __syncthreads();
    }

    // Drains to the film with global fp32 atomics.
    __device__ inline void flush() {
        drain([&](uint32_t key, f3 v) { film_add(film, key, v); });
    }

    // Whole block must call.  Flushes when more than half the slots are taken.
    __device__ inline void maybe_flush() {
        __syncthreads();
        uint32_t f = *fill;
        __syncthreads();  // every thread has read the same count
        if (f > (uint32_t)(SLOTS / 2)) flush();
    }
};

// The LDS hash in front of binned records (deposit mode 3): deposits on one
// film pixel within a block are summed in LDS first, so a beam's many
// deposits on few pixels become few records; the table drains to records,
// and a deposit that finds no free slot is appended as a record at once.
// Once a photon has scattered off a diffuse surface its deposits spread
// over the film and rarely meet another's, so they skip the table.
// The LDS fp32 sums' order varies between runs (last-bit differences only).
constexpr int kHashBinProbes = 2;

template <int SLOTS>
struct HashBinSink {
    LdsSplat<SLOTS, kHashBinProbes> hash;
    BinSink bins;
    bool spread = false;  // this lane's photon has scattered diffusely

    __device__ inline void add(uint32_t idx, f3 v) {
        if (spread || !hash.try_add(idx, v)) bins.add(idx, v);
    }
    __device__ inline void add_xy(uint32_t px, uint32_t py, uint32_t idx, f3 v) {
        if (spread || !hash.try_add(idx, v)) bins.add_xy(px, py, idx, v);
    }
    // Whole block must call.
    __device__ inline void flush() {
        hash.drain([&](uint32_t key, f3 v) { bins.add(key, v); });
    }
    // Whole block must call.  Drains when more than half the slots are taken.
    __device__ inline void maybe_flush() {
        __syncthreads();
        uint32_t f = *hash.fill;
        __syncthreads();
        if (f > (uint32_t)(SLOTS / 2)) flush();
    }
};

// The record sink of the batched photon kernel, called by every lane of the
// wave together (converged control flow, each lane with a deposit or
// none): the wave's open chunk and its fill are then the same on every lane
// and stay in scalar registers, so an append is a ballot, a prefix count
// and the lanes' stores -- no leader section, no LDS round trip.  HASH: the
// LDS table in front (deposit mode 3; a photon past a diffuse bounce skips
// it).  THREADS: the block size (the drain's trip count is uniform).
template <int SLOTS, bool HASH, int THREADS>
struct WaveSink {
    LdsSplat<SLOTS, kHashBinProbes, false> hash;
    uint32_t *rkey;
    float *rval;
    uint32_t *pool_next;
    uint32_t *chunk_fill;
    uint32_t pool_cap;
    uint32_t pool;                   // this block's first-choice pool
    uint32_t tile_wlog2;
    uint32_t chunk = kChunkEmpty;    // the wave's open chunk (first record index)
    uint32_t fill = 0;               // records written into it
    // the current item's job
    uint32_t key_base = 0, width = 1, tiles_x = 1;
    float *film = nullptr;           // fallback when every pool is full

    __device__ inline void job(uint32_t kb, uint32_t w, float *f) {
        key_base = kb;
        width = w;
        tiles_x = tiles_across(w, tile_wlog2);
        film = f;
    }

    __device__ inline void append(bool rec, uint32_t key, uint32_t idx, f3 v) {
        const uint64_t m = __ballot(rec);
        if (!m) return;
        const uint32_t cnt = (uint32_t)__popcll(m);
        if (chunk != kChunkFull && (chunk == kChunkEmpty || fill + cnt > kBinChunk)) {
            uint32_t got = 0;
            if (__lane_id() == 0) {
                if (chunk != kChunkEmpty) chunk_fill[chunk / kBinChunk] = fill;
                got = grab_chunk(pool_next, pool_cap, pool);
            }
            chunk = __builtin_amdgcn_readfirstlane(got);
            fill = 0;
        }
        if (chunk == kChunkFull) {
            if (rec) film_add(film, idx, v);
            return;
        }
        if (rec) {
            const uint32_t at = chunk + fill + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
            rkey[at] = key_base + key;
            rval[3 * (size_t)at + 0] = v.x;
            rval[3 * (size_t)at + 1] = v.y;
            rval[3 * (size_t)at + 2] = v.z;
        }
        fill += cnt;
    }

    // Every lane calls; d.ok = false for no deposit.
    __device__ inline void deposit(const Dep &d, bool spread) {
        bool rec = d.ok;
        if (HASH && d.ok && !spread) rec = !hash.try_add(d.idx, d.v);
        append(rec, tile_key_xy(d.px, d.py, tiles_x, tile_wlog2), d.idx, d.v);
    }

    // Whole block must call: the table's sums become records.
    __device__ inline void flush() {
        if (!HASH) return;
        __syncthreads();
        for (uint32_t k0 = 0; k0 < (uint32_t)SLOTS; k0 += THREADS) {
            const uint32_t k = k0 + threadIdx.x;
            const uint32_t key = hash.keys[k];
            const bool occ = key != kEmptyKey;
            f3 v = mk(0.f, 0.f, 0.f);
            if (occ) {
                v = mk(hash.vals[3 * k], hash.vals[3 * k + 1], hash.vals[3 * k + 2]);
                hash.keys[k] = kEmptyKey;
                hash.vals[3 * k] = hash.vals[3 * k + 1] = hash.vals[3 * k + 2] = 0.0f;
            }
            append(occ, occ ? tile_key(key, width, tiles_x, tile_wlog2) : 0u, key, v);
        }
        __syncthreads();
    }

    // Every wave calls this once at kernel end.
    __device__ inline void finish() {
        if (__lane_id() == 0 && chunk != kChunkEmpty && chunk != kChunkFull)
            chunk_fill[chunk / kBinChunk] = fill;
    }
};

template <int SLOTS>
__device__ inline void sink_photon_begin(HashBinSink<SLOTS> &s) {
    s.spread = false;
}
template <int SLOTS>
__device__ inline void sink_scattered(HashBinSink<SLOTS> &s, uint32_t kind) {
    if (kind == SPX_MAT_LAMBERTIAN) s.spread = true;
}

// A deposit at film pixel (px, py) = index idx: sinks that key records by
// film tile take the pixel itself.
__device__ inline void deposit_xy(BinSink &s, uint32_t px, uint32_t py, uint32_t idx, f3 v) {
    s.add_xy(px, py, idx, v);
}
template <int SLOTS>
__device__ inline void deposit_xy(HashBinSink<SLOTS> &s, uint32_t px, uint32_t py, uint32_t idx, f3 v) {
    s.add_xy(px, py, idx, v);
}

}  // namespace spx
