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
template <int SLOTS, int PROBES = 8>
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
        const uint64_t ins = __ballot(inserted);
        if (ins && (int)__lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1)
            atomicAdd(fill, (uint32_t)__popcll(ins));
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
