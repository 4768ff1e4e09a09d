// spx_batch.hpp -- device job table layout of the batched engine.
#pragma once

#include "spx_internal.hpp"

namespace spx {

constexpr uint32_t kXcdQueues = 8;  // MI355X: 8 XCDs, blocks dealt round-robin

struct BatchJob {
    uint32_t kind;  // SPX_JOB_*
    uint32_t pad[3];
    CamParams cam;
    PhotonParams ph;
};

struct BatchItem {
    uint32_t job;
    uint32_t start;  // camera: first pixel; photons: first photon (relative to photon_base)
    uint32_t count;  // camera: pixels (<= 2048); photons: photons
    uint32_t cost;   // host-side estimate, for the longest-first order
};

struct BatchQueues {
    uint32_t next[kXcdQueues];   // claimed so far (device counters, zeroed per launch)
    uint32_t begin[kXcdQueues];  // first item of each queue
    uint32_t size[kXcdQueues];   // items in each queue
    uint32_t pad[8];
};

// camera: the camera kernel (items of camera jobs only); else the photon
// kernel for splat_mode (items of volume / lighttrace jobs only).  `blocks`
// counts blocks of 256 threads; with photon_threads 1024 the hash modes run
// blocks / 4 blocks of 1024 threads.
hipError_t launch_batch(bool camera, const BatchJob *jobs, const BatchItem *items, BatchQueues *q,
                        uint32_t blocks, uint32_t splat_mode, uint32_t photon_threads, const BinState &bins,
                        hipStream_t s);

}  // namespace spx
