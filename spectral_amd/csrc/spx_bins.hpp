// spx_bins.hpp -- binned photon deposits: records instead of film atomics.
//
// Deposits are appended as records {key, X, Y, Z} (key = the job's key base
// + tile-ordered pixel) into chunked pools, one pool per XCD (a wave whose
// pool is full takes chunks from the others), keys and values in separate
// arrays (4 + 12 bytes), so the count pass streams only the keys.  Each
// wave owns one 512-record chunk at a time, so an append costs one LDS
// read-modify-write by the wave's leader and two coalesced stores per lane;
// a new chunk costs one global atomic per 512 records.  After the trace,
// the records are counted per strip (a 1024-pixel film tile, 2^tw x
// 2^(10-tw) with tw chosen per call -- 32x32 by default: spatially
// compact, so the few tiles a beam or caustic crosses collect its
// deposits), scanned, scattered into strip order -- each group's run of a
// strip padded to kRunAlign records, so that runs written by different
// blocks never share a memory segment -- and accumulated strip by strip in
// LDS as 64-bit fixed point scaled to the strip's largest deposit and
// record count (integer LDS atomics: order-independent, resolution finer
// than the fp32 rounding of a single add).  A strip's film pixels are
// written with plain stores by their one owner; strips split over several
// blocks first sum the blocks' fixed-point tiles with 64-bit integer
// atomics.  Films are therefore bit-reproducible unless two jobs share one
// (fp32 atomics between jobs) or a pool fills: a full pool falls back to
// direct film atomics, so results never depend on capacity.
#pragma once

#include "spx_device.hpp"

namespace spx {

constexpr uint32_t kBinChunk = 512;      // records per wave chunk
constexpr uint32_t kChunksPerBlock = 8;  // chunks per count / scatter block
constexpr uint32_t kStripLog2 = 10;      // 1024 film pixels per strip (one film tile)
constexpr uint32_t kTileLog2 = 5;        // default tile width log2 (32x32 tiles)
constexpr uint32_t kTileLog2Min = 2;     // tile shapes from 4x256 ...
constexpr uint32_t kTileLog2Max = 8;     // ... to 256x4
constexpr uint32_t kStripPix = 1u << kStripLog2;
constexpr uint32_t kBinPools = 8;        // one per XCD
constexpr uint32_t kPartRecs = 16384;    // records per accumulation work item
constexpr uint32_t kSub = 16;            // sub-counters per strip
constexpr uint32_t kGroupSlots = 1024;   // distinct strips per group tracked in LDS
constexpr uint32_t kGroupRecs = kChunksPerBlock * kBinChunk;  // records per group
constexpr uint32_t kNoMacc = 0xFFFFFFFFu;
// Sorted runs start on multiples of kRunAlign records (32 bytes, the memory
// write granule: unaligned runs cost read-modify-write traffic at their
// shared edges; 2, 4 and 8 records measured 29.5, 30.5 and 32.7 GB per
// 5e8-record round); the slack after a run holds pad records (key kPadKey),
// which the accumulation skips.
constexpr uint32_t kRunAlign = 2;
constexpr uint32_t kPadKey = 0xFFFFFFFFu;
constexpr uint32_t kChunkEmpty = 0xFFFFFFFFu;
constexpr uint32_t kChunkFull = 0xFFFFFFFEu;

struct __align__(16) BinRec {
    uint32_t key;
    float x, y, z;
};

__host__ __device__ constexpr uint32_t run_pad(uint32_t n) { return (n + kRunAlign - 1) & ~(kRunAlign - 1); }

// Sorted-array slots for `cap` records over `n_strips` strips: every group
// run and every strip can end with up to kRunAlign - 1 pads, and a group's
// 4096 records fall in at most kGroupSlots runs (cap / 4 over all groups).
__host__ __device__ constexpr size_t sorted_capacity(size_t cap, size_t n_strips) {
    return cap + (size_t)(kRunAlign - 1) * (cap / (kGroupRecs / kGroupSlots) + n_strips) + kRunAlign;
}
// Device bytes per record of a deposit round: key, value and the worst
// case of the padded sorted copy.
constexpr double kRoundBytesPerRecord =
    16.0 + 16.0 * (1.0 + (double)(kRunAlign - 1) * kGroupSlots / kGroupRecs);

// Device-side state of one deposit round (all pointers device memory).
// Records are handled in groups of kChunksPerBlock chunks.  The count pass
// tallies each group's strips in LDS and reserves the group's share of every
// strip in one of kSub sub-counters (g % kSub: the same strip's counters are
// spread so concurrent groups rarely meet on one address), leaving a list of
// {strip, offset in sub-bucket} per group for the scatter pass.
struct BinState {
    uint32_t *rkey;        // kBinPools * pool_cap record keys
    float *rval;           // 3 floats per record
    BinRec *sorted;        // sorted_capacity(pool records, n_strips) slots
    uint32_t *pool_next;   // [kBinPools] records taken per pool
    uint32_t *chunk_fill;  // [kBinPools * pool_cap / kBinChunk]
    // Zeroed each round, contiguous: sub_count | ovf_count | ovf_cur | strip_max.
    uint32_t *sub_count;   // [n_strips * kSub] padded run lengths; after the totals pass, offsets in the strip
    uint32_t *ovf_count;   // [n_strips] records of groups whose LDS table was full; then their offset
    uint32_t *ovf_cur;     // [n_strips] scatter cursor of those records
    uint32_t *strip_max;   // [n_strips] largest |x|,|y|,|z| bit pattern (the scatter pass)
    size_t zero_words;     // words from sub_count to the end of strip_max
    uint32_t *strip_total; // [n_strips] slots the strip uses (padded runs, then its overflow records)
    uint32_t *strip_off;   // [n_strips + 1] first slot (strips start on kRunAlign multiples)
    uint32_t *part_off;    // [n_strips + 1]
    uint32_t *part_next;   // [1]
    uint2 *glist;          // [groups * kGroupSlots] {strip, offset} per group
    uint32_t *gcount;      // [groups] list length
    // Strips split over several accumulation parts sum their parts' fixed-
    // point tiles here (64-bit integer atomics: exact, order-independent),
    // then one pass converts and writes the film.
    unsigned long long *macc;  // [max_multi * 3 * kStripPix]
    uint32_t *strip_macc;      // [n_strips] accumulator index, or kNoMacc
    uint32_t *macc_strip;      // [max_multi] strip of each accumulator
    uint32_t *macc_count;      // [1]
    uint32_t max_multi;
    const uint32_t *strip_job;   // [n_strips] owning job segment
    const uint32_t *job_key;     // [n_jobs] first key of each job
    const uint32_t *job_npix;    // [n_jobs] pixels of each job's film
    const uint32_t *job_width;   // [n_jobs] film width (tile grid)
    float *const *job_film;      // [n_jobs] film pointers
    uint32_t pool_cap;     // records per pool (multiple of kBinChunk)
    uint32_t n_strips;
    uint32_t tile_wlog2;   // tiles are 2^tile_wlog2 wide and 2^(10 - tile_wlog2) tall
    // Film writes of the accumulation are atomic adds: set while the next
    // round's trace kernel runs concurrently (its pool-overflow fallback adds
    // into the same films).
    uint32_t atomic_film;
};

__device__ inline uint32_t xcc_id() {
    // HW_REG_XCC_ID (hwreg 20), bits [3:0]: the XCD this wave runs on.
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xFu;
}

// Tiles across a film of `width` pixels for tiles 2^tw wide.
__host__ __device__ inline uint32_t tiles_across(uint32_t width, uint32_t tw) {
    return (width + (1u << tw) - 1) >> tw;
}

// Key of film pixel idx within its job: tile index * 1024 + pixel in tile,
// tiles 2^tw wide and 2^(10 - tw) tall, row-major in the film and inside.
__host__ __device__ inline uint32_t tile_key(uint32_t idx, uint32_t width, uint32_t tiles_x, uint32_t tw) {
    const uint32_t th = kStripLog2 - tw;
    uint32_t px = idx % width, py = idx / width;
    uint32_t tile = (py >> th) * tiles_x + (px >> tw);
    uint32_t local = ((py & ((1u << th) - 1)) << tw) | (px & ((1u << tw) - 1));
    return (tile << kStripLog2) | local;
}

// tile_key of pixel (px, py): no division.
__host__ __device__ inline uint32_t tile_key_xy(uint32_t px, uint32_t py, uint32_t tiles_x, uint32_t tw) {
    const uint32_t th = kStripLog2 - tw;
    const uint32_t tile = (py >> th) * tiles_x + (px >> tw);
    const uint32_t local = ((py & ((1u << th) - 1u)) << tw) | (px & ((1u << tw) - 1u));
    return (tile << kStripLog2) | local;
}

// Film pixel (px, py) of in-tile position k of tile `tile` (tile_key's inverse).
__host__ __device__ inline void tile_pixel(uint32_t tile, uint32_t k, uint32_t tiles_x, uint32_t tw, uint32_t &px,
                                           uint32_t &py) {
    const uint32_t th = kStripLog2 - tw;
    px = ((tile % tiles_x) << tw) + (k & ((1u << tw) - 1));
    py = ((tile / tiles_x) << th) + (k >> tw);
}

// A chunk of `pool` (this XCD's), else of another pool with room, else
// kChunkFull.  A pool that reads full is skipped, which bounds how far
// pool_next overshoots the capacity.  Kept out of line: the sink is inlined
// at every deposit site of the trace kernels.
__device__ __noinline__ uint32_t grab_chunk(uint32_t *pool_next, uint32_t pool_cap, uint32_t pool) {
    uint32_t got = atomicAdd(pool_next + pool, kBinChunk);
    if (got + kBinChunk <= pool_cap) return pool * pool_cap + got;
#pragma unroll 1
    for (uint32_t t = 1; t < kBinPools; ++t) {
        const uint32_t pl = (pool + t) % kBinPools;
        if (__hip_atomic_load(pool_next + pl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= pool_cap) continue;
        got = atomicAdd(pool_next + pl, kBinChunk);
        if (got + kBinChunk <= pool_cap) return pl * pool_cap + got;
    }
    return kChunkFull;
}

// A film sink that appends records.  `wstate` is LDS: {chunk, fill} per wave
// of the block, initialised to kChunkEmpty by the kernel.
struct BinSink {
    uint32_t *rkey;
    float *rval;
    uint32_t *pool_next;
    uint32_t *chunk_fill;
    uint32_t pool_cap;
    uint32_t pool;       // this block's first-choice pool (XCD id mod kBinPools)
    uint32_t key_base;
    float *film;         // fallback when the pool is exhausted
    uint32_t *wstate;
    uint32_t width;      // film width in pixels
    uint32_t tiles_x;    // tiles across the film
    uint32_t tile_wlog2; // tile width log2

    __device__ inline void add(uint32_t idx, f3 v) { append(tile_key(idx, width, tiles_x, tile_wlog2), idx, v); }
    // The same deposit when its pixel (px, py) is at hand: no division.
    __device__ inline void add_xy(uint32_t px, uint32_t py, uint32_t idx, f3 v) {
        append(tile_key_xy(px, py, tiles_x, tile_wlog2), idx, v);
    }

    __device__ inline void append(uint32_t key, uint32_t idx, f3 v) {
        uint64_t m = __ballot(1);
        uint32_t lane = __lane_id();
        uint32_t cnt = (uint32_t)__popcll(m);
        uint32_t pre = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t w = threadIdx.x >> 6;
        uint32_t base = 0;
        if ((int)lane == leader) {
            uint32_t c = wstate[2 * w], f = wstate[2 * w + 1];
            if (c != kChunkFull && (c == kChunkEmpty || f + cnt > kBinChunk)) {
                if (c != kChunkEmpty) chunk_fill[c / kBinChunk] = f;
                c = grab_chunk(pool_next, pool_cap, pool);
                f = 0;
            }
            if (c == kChunkFull) {
                base = kChunkFull;
            } else {
                base = c + f;
                f += cnt;
            }
            wstate[2 * w] = c;
            wstate[2 * w + 1] = f;
        }
        base = __builtin_amdgcn_readlane(base, leader);
        if (base == kChunkFull) {
            film_add(film, idx, v);
            return;
        }
        const uint32_t at = base + pre;
        rkey[at] = key_base + key;
        rval[3 * (size_t)at + 0] = v.x;
        rval[3 * (size_t)at + 1] = v.y;
        rval[3 * (size_t)at + 2] = v.z;
    }
};

// Every wave calls this once at kernel end (all 64 lanes present).
__device__ inline void bin_wave_finish(uint32_t *wstate, uint32_t *chunk_fill) {
    if (__lane_id() == 0) {
        uint32_t w = threadIdx.x >> 6;
        uint32_t c = wstate[2 * w];
        if (c != kChunkEmpty && c != kChunkFull) chunk_fill[c / kBinChunk] = wstate[2 * w + 1];
    }
}

}  // namespace spx
