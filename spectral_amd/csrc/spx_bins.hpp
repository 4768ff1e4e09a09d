// spx_bins.hpp -- binned photon deposits: records instead of film atomics.
//
// Deposits are appended as 16-byte records {key, X, Y, Z} (key = the job's
// key base + pixel) into chunked pools, one pool per XCD.  Each wave owns
// one 512-record chunk at a time, so an append costs one LDS read-modify-
// write by the wave's leader and one coalesced 16-byte store per lane; a new
// chunk costs one global atomic per 512 records.  After the trace, the
// records are counted per strip (a 32x32-pixel film tile: spatially
// compact, so the few tiles a beam or caustic crosses collect its deposits),
// scanned, scattered into strip order and accumulated strip by strip in LDS; a strip's film pixels are
// then written with plain stores (one owner) or, for strips split over
// several blocks, with one fp32 atomic per pixel and block.  A full pool
// falls back to direct film atomics, so results never depend on capacity.
#pragma once

#include "spx_device.hpp"

namespace spx {

constexpr uint32_t kBinChunk = 512;      // records per wave chunk
constexpr uint32_t kChunksPerBlock = 8;  // chunks per count / scatter block
constexpr uint32_t kStripLog2 = 10;      // 1024 film pixels per strip (a 32x32 tile)
constexpr uint32_t kTileLog2 = 5;        // tile side 32
constexpr uint32_t kStripPix = 1u << kStripLog2;
constexpr uint32_t kBinPools = 8;        // one per XCD
constexpr uint32_t kPartRecs = 16384;    // records per accumulation work item
constexpr uint32_t kChunkEmpty = 0xFFFFFFFFu;
constexpr uint32_t kChunkFull = 0xFFFFFFFEu;

struct __align__(16) BinRec {
    uint32_t key;
    float x, y, z;
};

// Device-side state of one deposit round (all pointers device memory).
struct BinState {
    BinRec *recs;          // kBinPools * pool_cap records
    BinRec *sorted;        // same size
    uint32_t *pool_next;   // [kBinPools] records taken per pool
    uint32_t *chunk_fill;  // [kBinPools * pool_cap / kBinChunk]
    uint32_t *strip_count; // [n_strips]
    uint32_t *strip_off;   // [n_strips + 1]
    uint32_t *strip_cur;   // [n_strips]
    uint32_t *part_off;    // [n_strips + 1]
    uint32_t *part_next;   // [1]
    const uint32_t *strip_job;   // [n_strips] owning job segment
    const uint32_t *job_key;     // [n_jobs] first key of each job
    const uint32_t *job_npix;    // [n_jobs] pixels of each job's film
    const uint32_t *job_width;   // [n_jobs] film width (tile grid)
    float *const *job_film;      // [n_jobs] film pointers
    uint32_t pool_cap;     // records per pool (multiple of kBinChunk)
    uint32_t n_strips;
};

__device__ inline uint32_t xcc_id() {
    // HW_REG_XCC_ID (hwreg 20), bits [3:0]: the XCD this wave runs on.
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xFu;
}

// Key of film pixel idx within its job: tile index * 1024 + pixel in tile.
__host__ __device__ inline uint32_t tile_key(uint32_t idx, uint32_t width) {
    uint32_t px = idx % width, py = idx / width;
    uint32_t tiles_x = (width + (1u << kTileLog2) - 1) >> kTileLog2;
    uint32_t tile = (py >> kTileLog2) * tiles_x + (px >> kTileLog2);
    uint32_t local = ((py & ((1u << kTileLog2) - 1)) << kTileLog2) | (px & ((1u << kTileLog2) - 1));
    return (tile << kStripLog2) | local;
}

// A film sink that appends records.  `wstate` is LDS: {chunk, fill} per wave
// of the block, initialised to kChunkEmpty by the kernel.
struct BinSink {
    BinRec *recs;
    uint32_t *pool_next;
    uint32_t *chunk_fill;
    uint32_t pool_cap;
    uint32_t pool;       // this block's pool (XCD id mod kBinPools)
    uint32_t key_base;
    float *film;         // fallback when the pool is exhausted
    uint32_t *wstate;
    uint32_t width;      // film width in pixels (the tile grid is ceil(width / 32) wide)

    __device__ inline void add(uint32_t idx, f3 v) {
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
                uint32_t got = atomicAdd(pool_next + pool, kBinChunk);
                c = (got + kBinChunk <= pool_cap) ? pool * pool_cap + got : kChunkFull;
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
        base = __shfl(base, leader);
        if (base == kChunkFull) {
            film_add(film, idx, v);
            return;
        }
        BinRec r;
        r.key = key_base + tile_key(idx, width);
        r.x = v.x;
        r.y = v.y;
        r.z = v.z;
        recs[base + pre] = r;
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
