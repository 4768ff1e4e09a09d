// spx_batch_host.hip -- host side of the batched engine: the launch schedule
// of a job table, its one-shot run (spx_batch_run), the same schedule
// captured once as a hipGraph and replayed (spx_batch_plan_*), and the
// autotuner that picks the engine's launch shape by timing (spx_autotune).
//
// A schedule is built on the host in three steps: every job becomes a
// BatchJob parameter block and (job, range) work items; photon items are
// split into deposit rounds bounded by the record capacity (binned deposits)
// or given film replicas (atomic deposits); jobs are dealt largest first to
// 8 XCD queues and the items ordered by round, queue and cost.  The device
// image of the schedule (queue heads, items, job table, replica jobs) is one
// contiguous block uploaded before the kernels run.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "spx_batch.hpp"
#include "spx_internal.hpp"

using namespace spx;

namespace spx {

hipError_t launch_queues_reset(BatchQueues *q, uint32_t n_rounds, size_t stride, hipStream_t s);

namespace {

struct Schedule {
    uint32_t splat_mode = 2, blocks = 0, n_rounds = 0, n_jobs = 0;
    bool binned = false;
    bool overlap = false;  // overlapped photon rounds (bin passes on a second stream)
    uint32_t nbuf = 1;     // record buffers the rounds cycle through (overlapped: 2 or 3)
    uint32_t cam_blocks = 0;  // > 0: the camera round runs concurrently, on its own stream
    uint32_t ph_threads = 256;  // photon-kernel block size (1024: one block per CU)
    std::vector<BatchJob> table;
    std::vector<BatchItem> ordered;
    std::vector<BatchQueues> qh;   // per round; round 0 = camera items
    std::vector<ReplicaJob> reps;  // src: offset (floats) into the replica buffer until bound
    size_t rep_total = 0;          // replica floats
    uint32_t rep_blocks = 0;
    BinJobMap map;
    size_t bin_want = 0;
    size_t qstride = 0, off_items = 0, off_jobs = 0, off_reps = 0, bytes = 0;
};

int build_batch_job(spx_ctx *ctx, const spx_job &jb, BatchJob &out, uint32_t idx) {
    char who[64];
    snprintf(who, sizeof who, "spx_batch_run job %u", idx);
    memset(&out, 0, sizeof out);
    out.kind = jb.kind;
    if (!jb.scene || jb.scene->ctx != ctx) return set_error(ctx, SPX_EINVAL, std::string(who) + ": bad scene");
    int rc;
    if (jb.kind == SPX_JOB_CAMERA) {
        if ((rc = check_film(ctx, jb.film, jb.width, jb.height, who))) return rc;
        if (!valid_opts(&jb.opts, true)) return set_error(ctx, SPX_EINVAL, std::string(who) + ": bad options");
        if (!finite3(jb.camera.origin) || !finite3(jb.camera.lower_left) || !finite3(jb.camera.horizontal) ||
            !finite3(jb.camera.vertical))
            return set_error(ctx, SPX_EINVAL, std::string(who) + ": non-finite camera");
        CamParams &p = out.cam;
        p.cam = jb.camera;
        p.w = jb.width; p.h = jb.height; p.spp = jb.spp; p.sample_base = jb.sample_base; p.seed = jb.seed;
        p.illuminant = jb.opts.illuminant; p.dispersion = jb.opts.dispersion;
        p.max_depth = jb.opts.max_depth; p.rr_depth = jb.opts.rr_depth;
        p.scene = jb.scene->dev;
        p.tbl = table_for(ctx, jb.opts.illuminant);
        p.film = jb.film;
        return SPX_OK;
    }
    if (jb.kind != SPX_JOB_VOLUME && jb.kind != SPX_JOB_LIGHTTRACE)
        return set_error(ctx, SPX_EINVAL, std::string(who) + ": unknown kind");
    PhotonParams &p = out.ph;
    if ((rc = photon_params(ctx, jb.scene, &jb.beam, &jb.opts, jb.n_photons, jb.photon_base, jb.seed, p, who)))
        return rc;
    if (jb.kind == SPX_JOB_LIGHTTRACE) {
        if ((rc = check_film(ctx, jb.film, jb.screen.width, jb.screen.height, who))) return rc;
        if ((rc = check_screen(ctx, jb.screen, who))) return rc;
        set_screen(p, jb.screen);
    } else {
        if ((rc = check_film(ctx, jb.film, jb.width, jb.height, who))) return rc;
        if (jb.estimator != SPX_EST_POINT && jb.estimator != SPX_EST_LEGACY)
            return set_error(ctx, SPX_EINVAL, std::string(who) + ": batches run the point or legacy estimator");
        const spx_vol_params &vp = jb.vol;
        if (!isfinite(vp.sigma_s) || !isfinite(vp.sigma_t) || !isfinite(vp.g) || !(vp.max_dist > 0.0f) ||
            !isfinite(vp.max_dist))
            return set_error(ctx, SPX_EINVAL, std::string(who) + ": bad volume parameters");
        if (!finite3(jb.camera.origin) || !finite3(jb.camera.lower_left) || !finite3(jb.camera.horizontal) ||
            !finite3(jb.camera.vertical))
            return set_error(ctx, SPX_EINVAL, std::string(who) + ": non-finite camera");
        make_proj(jb.camera, p.proj);
        p.vp = vp;
        p.w = jb.width;
        p.h = jb.height;
        p.estimator = jb.estimator;
        p.zbuf = jb.zbuf;
    }
    p.film = jb.film;
    return SPX_OK;
}

uint32_t photon_film_w(const spx_job &j) { return j.kind == SPX_JOB_LIGHTTRACE ? j.screen.width : j.width; }
uint32_t photon_film_h(const spx_job &j) { return j.kind == SPX_JOB_LIGHTTRACE ? j.screen.height : j.height; }

// Builds the schedule of a job table (validation included).
int plan_schedule(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, const spx_batch_cfg *cfg, Schedule &S) {
    const spx_tuning &T = ctx->tuning;
    uint32_t chunk_paths = cfg && cfg->chunk_paths ? cfg->chunk_paths : T.chunk_paths;
    uint32_t chunk_photons = cfg && cfg->chunk_photons ? cfg->chunk_photons : T.chunk_photons;
    const uint32_t cus = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256);
    S.blocks = cfg && cfg->blocks ? cfg->blocks : (T.batch_blocks ? T.batch_blocks : cus * 4u);
    // A table too small to give every block several items at the tuned size
    // (one job, a few served requests) gets smaller items, down to 256
    // paths / photons, so that the persistent grid is filled and the round's
    // tail is short: four camera items per block (one 512^2 x 64 spp job:
    // 17.4 against 16.2 G paths/s at two), two photon items (each ends in a
    // flush of the block's deposit hash).
    if (!(cfg && (cfg->chunk_paths || cfg->chunk_photons))) {
        uint64_t paths = 0, photons = 0;
        for (uint32_t j = 0; j < n_jobs; ++j) {
            if (jobs[j].kind == SPX_JOB_CAMERA) paths += (uint64_t)jobs[j].width * jobs[j].height * jobs[j].spp;
            else photons += jobs[j].n_photons;
        }
        // Such a table also keeps to three blocks per compute unit: with
        // four, its items shrink further (one 512^2 x 64 spp job: 0.94 ms at
        // 768 blocks, 1.06 at 1024), while large tables gain from the fourth
        // (the 512-job table's camera half: 67.4 -> 60.3 ms).
        const bool small = (paths && paths / (4ull * S.blocks) < chunk_paths) ||
                           (photons && photons / (2ull * S.blocks) < chunk_photons);
        if (small && !(cfg && cfg->blocks)) S.blocks = std::min(S.blocks, 3u * cus);
        auto fit = [&](uint32_t chunk, uint64_t units, uint64_t per_block) {
            uint64_t c = units / (per_block * S.blocks);
            c = c < 256 ? 256 : c;
            return (uint32_t)(c < chunk ? c : chunk);
        };
        chunk_paths = fit(chunk_paths, paths, 4);
        chunk_photons = fit(chunk_photons, photons, 2);
    }
    S.splat_mode = cfg ? cfg->splat_mode : T.splat_mode;
    if (S.splat_mode > 3) return set_error(ctx, SPX_EINVAL, "spx_batch_run: unknown splat mode");
    if (chunk_paths == 0 || chunk_photons == 0) return set_error(ctx, SPX_EINVAL, "spx_batch_run: zero item size");
    S.binned = S.splat_mode >= 2;
    S.n_jobs = n_jobs;
    S.table.assign(n_jobs, BatchJob{});
    S.map = BinJobMap(T.tile_wlog2);
    std::vector<BatchItem> items;
    std::vector<double> job_cost(n_jobs, 0.0);
    std::vector<std::pair<const float *, uint32_t>> films;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        int rc = build_batch_job(ctx, jobs[j], S.table[j], j);
        if (rc) return rc;
        films.push_back({jobs[j].film, jobs[j].kind});
        const spx_job &jb = jobs[j];
        if (jb.kind == SPX_JOB_CAMERA) {
            if (jb.spp == 0) continue;
            uint32_t npix = jb.width * jb.height;
            uint32_t per = chunk_paths / jb.spp;
            per = per < 1 ? 1 : (per > 2048 ? 2048 : per);
            for (uint32_t p0 = 0; p0 < npix; p0 += per) {
                uint32_t c = npix - p0 < per ? npix - p0 : per;
                uint64_t cost = (uint64_t)c * jb.spp;
                items.push_back({j, p0, c, (uint32_t)std::min<uint64_t>(cost, 0xFFFFFFFFu)});
                job_cost[j] += (double)cost;
            }
        } else {
            // 64-bit counter: n_photons near 2^32 must not wrap the loop
            for (uint64_t p0 = 0; p0 < jb.n_photons; p0 += chunk_photons) {
                uint32_t c = (uint32_t)std::min<uint64_t>(jb.n_photons - p0, chunk_photons);
                items.push_back({j, (uint32_t)p0, c, 2u * c});
                job_cost[j] += 2.0 * c;
            }
        }
    }
    // Camera films are written without atomics: no two jobs may share one.
    std::vector<std::pair<const float *, uint32_t>> sorted_films = films;
    std::sort(sorted_films.begin(), sorted_films.end());
    for (size_t k = 1; k < sorted_films.size(); ++k)
        if (sorted_films[k].first == sorted_films[k - 1].first &&
            (sorted_films[k].second == SPX_JOB_CAMERA || sorted_films[k - 1].second == SPX_JOB_CAMERA))
            return set_error(ctx, SPX_EINVAL, "spx_batch_run: a camera job shares its film with another job");

    // Round 0 holds the camera items (camera kernel); photon items follow in
    // rounds 1.. (photon kernel).  Binned deposits: key ranges per photon
    // job, and photon items split into rounds of at most kBinFill of the
    // record capacity (estimated deposits per photon from
    // records_per_photon; overflow falls back to atomics, so the estimate
    // only affects speed).
    std::vector<uint32_t> item_round(items.size(), 1);
    for (size_t k = 0; k < items.size(); ++k)
        if (jobs[items[k].job].kind == SPX_JOB_CAMERA) item_round[k] = 0;
    S.n_rounds = 2;
    if (S.binned) {
        std::map<const float *, uint32_t> first_owner;
        for (uint32_t j = 0; j < n_jobs; ++j) {
            if (jobs[j].kind == SPX_JOB_CAMERA) continue;
            uint32_t fw = photon_film_w(jobs[j]), fh = photon_film_h(jobs[j]);
            if (((uint64_t)S.map.n_strips + BinJobMap::film_strips(fw, fh, S.map.tile_wlog2)) * kStripPix >
                0xFFFFFFFFull)
                return set_error(ctx, SPX_ECAP, "spx_batch_run: photon films exceed 2^32 tile pixels");
            S.table[j].ph.key_base = S.map.add(jobs[j].film, fw, fh);
            S.table[j].ph.replicas = 1;
            S.table[j].ph.rep_stride = 0;
            auto it = first_owner.find(jobs[j].film);
            if (it != first_owner.end()) {
                S.map.shared.back() = 1;
                S.map.shared[it->second] = 1;
            } else {
                first_owner[jobs[j].film] = (uint32_t)S.map.film.size() - 1;
            }
        }
        const double budget = kBinFill * (double)((size_t)1 << T.bin_log2);
        double used = 0.0, most = 0.0;
        uint32_t r = 1;
        for (size_t k = 0; k < items.size(); ++k) {
            const spx_job &jb = jobs[items[k].job];
            if (jb.kind == SPX_JOB_CAMERA) continue;
            double est = (double)items[k].count * records_per_photon(jb.kind == SPX_JOB_LIGHTTRACE);
            if (used > 0.0 && used + est > budget) {
                ++r;
                used = 0.0;
            }
            used += est;
            most = std::max(most, used);
            item_round[k] = r;
        }
        S.n_rounds = r + 1;
        S.bin_want = (size_t)(most / kBinFill);
        S.overlap = T.overlap_bins != 0 && S.n_rounds > 2;
        S.nbuf = S.overlap ? std::min<uint32_t>(T.overlap_bins + 1u, (uint32_t)kRoundBufs) : 1u;
    }

    // Film replicas for the photon jobs (atomic deposit modes only).
    for (uint32_t j = 0; j < n_jobs && !S.binned; ++j) {
        if (jobs[j].kind == SPX_JOB_CAMERA || jobs[j].n_photons == 0) continue;
        size_t n = (size_t)3 * photon_film_w(jobs[j]) * photon_film_h(jobs[j]);
        uint32_t R = replica_count(ctx, n);
        if (R <= 1) continue;
        ReplicaJob rj;
        rj.dst = jobs[j].film;
        rj.src = (const float *)(uintptr_t)S.rep_total;  // offset, bound later
        rj.n = (uint32_t)n;
        rj.reps = R;
        rj.first_block = S.rep_blocks;
        rj.pad = j;
        S.reps.push_back(rj);
        S.rep_total += (size_t)R * n;
        S.rep_blocks += (uint32_t)((n + 1023) / 1024);
    }

    // Jobs to XCD queues, largest first onto the least-loaded queue.
    std::vector<uint32_t> order(n_jobs);
    for (uint32_t j = 0; j < n_jobs; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return job_cost[a] > job_cost[b]; });
    std::vector<uint32_t> job_queue(n_jobs, 0);
    double load[kXcdQueues] = {0};
    for (uint32_t j : order) {
        uint32_t best = 0;
        for (uint32_t q = 1; q < kXcdQueues; ++q)
            if (load[q] < load[best]) best = q;
        job_queue[j] = best;
        load[best] += job_cost[j];
    }
    std::vector<uint32_t> idx(items.size());
    for (size_t k = 0; k < items.size(); ++k) idx[k] = (uint32_t)k;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        if (item_round[a] != item_round[b]) return item_round[a] < item_round[b];
        uint32_t qa = job_queue[items[a].job], qb = job_queue[items[b].job];
        if (qa != qb) return qa < qb;
        return items[a].cost > items[b].cost;
    });
    S.ordered.resize(items.size());
    S.qh.assign(S.n_rounds, BatchQueues{});
    memset(S.qh.data(), 0, S.n_rounds * sizeof(BatchQueues));
    for (size_t k = 0; k < idx.size(); ++k) {
        S.ordered[k] = items[idx[k]];
        S.qh[item_round[idx[k]]].size[job_queue[items[idx[k]].job]]++;
    }
    uint32_t start = 0;
    for (uint32_t r = 0; r < S.n_rounds; ++r)
        for (uint32_t q = 0; q < kXcdQueues; ++q) {
            S.qh[r].begin[q] = start;
            start += S.qh[r].size[q];
        }
    uint32_t cam_items = 0, ph_items = 0;
    for (uint32_t q = 0; q < kXcdQueues; ++q) cam_items += S.qh[0].size[q];
    ph_items = (uint32_t)S.ordered.size() - cam_items;
    S.cam_blocks = T.camera_blocks && cam_items && ph_items ? T.camera_blocks : 0;
    S.ph_threads = (S.splat_mode == 1 || S.splat_mode == 3) && T.photon_threads == 1024 && S.blocks >= 4 ? 1024 : 256;
    S.qstride = a256(sizeof(BatchQueues));
    S.off_items = S.qstride * S.n_rounds;
    S.off_jobs = a256(S.off_items + S.ordered.size() * sizeof(BatchItem));
    S.off_reps = a256(S.off_jobs + S.table.size() * sizeof(BatchJob));
    S.bytes = S.off_reps + S.reps.size() * sizeof(ReplicaJob);
    return SPX_OK;
}

// Points the photon jobs' deposit films at their replicas in `replica`.
void bind_replicas(Schedule &S, float *replica) {
    for (ReplicaJob &rj : S.reps) {
        size_t off = (size_t)(uintptr_t)rj.src;
        rj.src = replica + off;
        PhotonParams &p = S.table[rj.pad].ph;
        p.film = replica + off;
        p.replicas = rj.reps;
        p.rep_stride = rj.n;
    }
}

// The device image: queue heads (claim counters zero), items, jobs, replica jobs.
void write_image(const Schedule &S, char *h) {
    memset(h, 0, S.bytes);
    for (uint32_t r = 0; r < S.n_rounds; ++r) memcpy(h + r * S.qstride, &S.qh[r], sizeof(BatchQueues));
    memcpy(h + S.off_items, S.ordered.data(), S.ordered.size() * sizeof(BatchItem));
    memcpy(h + S.off_jobs, S.table.data(), S.table.size() * sizeof(BatchJob));
    if (!S.reps.empty()) memcpy(h + S.off_reps, S.reps.data(), S.reps.size() * sizeof(ReplicaJob));
}

uint32_t round_items(const Schedule &S, uint32_t r) {
    uint32_t n = 0;
    for (uint32_t q = 0; q < kXcdQueues; ++q) n += S.qh[r].size[q];
    return n;
}

// Records a timing event pair around `launch` when the context's hook is on.
template <class F>
hipError_t timed(spx_ctx *ctx, bool timing, uint32_t kind, hipStream_t s, F launch) {
    if (!timing || !ctx->ktime_on) return launch();
    hipError_t e = hipSuccess;
    while (ctx->ktime_ev.size() < ctx->ktime_used + 2) {
        hipEvent_t ev;
        e = hipEventCreate(&ev);
        if (e != hipSuccess) return e;
        ctx->ktime_ev.push_back(ev);
    }
    if (ctx->ktime_kind.size() < ctx->ktime_ev.size() / 2) ctx->ktime_kind.resize(ctx->ktime_ev.size() / 2);
    e = hipEventRecord(ctx->ktime_ev[ctx->ktime_used], s);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(ctx->ktime_ev[ctx->ktime_used + 1], s);
    if (e == hipSuccess) {
        ctx->ktime_kind[ctx->ktime_used / 2] = kind;
        ctx->ktime_used += 2;
    }
    return e;
}

// Deposit-round buffers of one launch: bst[k] / shared[k] for k >= 1 are used
// only by overlapped schedules; `ev` holds 2 ordering events per photon round
// and `s2` is the second stream.
struct RoundBufs {
    BinState bst[kRoundBufs];
    const uint32_t *shared[kRoundBufs] = {};
    hipStream_t s2 = nullptr;
    hipEvent_t *ev = nullptr;   // overlapped rounds: 2 per photon round
    hipStream_t s3 = nullptr;   // concurrent camera round
    hipEvent_t *cev = nullptr;  // its fork and done events
};

// Photon rounds of a schedule (non-empty rounds after the camera round).
uint32_t photon_rounds(const Schedule &S) {
    uint32_t n = 0;
    for (uint32_t r = 1; r < S.n_rounds; ++r) n += round_items(S, r) ? 1 : 0;
    return n;
}

// The kernel sequence of a schedule whose image is at `d`: claim counters
// reset, replica clear, the camera round, then per photon round the bin
// reset, the trace kernel and the bin passes, and finally the replica sums.
// Overlapped schedules cycle through S.nbuf record buffers: round k's bin
// passes run on s2 while rounds k+1 .. k+nbuf-1 trace on s; round k+nbuf
// reuses k's buffer only after k's passes (and the buffer's reset behind
// them) end, and the passes stay in round order on s2 (they add into the
// same films).  Beside a photon kernel that holds most of the GPU the passes
// take longer than a round's trace, so with two buffers each round waited
// for the passes of the round before last (4-8 ms a round on the 512-job
// stream); three take that wait off.  Only kernel launches and event
// ordering (no memset or copy), so the sequence can be captured into a graph.
hipError_t enqueue(spx_ctx *ctx, const Schedule &S, char *d, const RoundBufs &B, float *replica, hipStream_t s,
                   bool timing) {
    const uint32_t n_photon_rounds = photon_rounds(S);
    hipError_t e = launch_queues_reset((BatchQueues *)d, S.n_rounds, S.qstride, s);
    if (e == hipSuccess && !S.reps.empty()) e = launch_fill(replica, 0.0f, S.rep_total, s);
    const uint32_t acc_blocks = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * 4u;
    auto trace_on = [&](uint32_t r, const BinState &bst, hipStream_t st, uint32_t blocks) {
        const bool camera = r == 0;
        return timed(ctx, timing, camera ? SPX_KT_CAMERA : SPX_KT_PHOTON, st, [&] {
            return launch_batch(camera, (const BatchJob *)(d + S.off_jobs), (const BatchItem *)(d + S.off_items),
                                (BatchQueues *)(d + r * S.qstride), blocks, S.splat_mode, S.ph_threads, bst, st);
        });
    };
    auto trace = [&](uint32_t r, const BinState &bst) { return trace_on(r, bst, s, S.blocks); };
    if (e == hipSuccess && S.cam_blocks) {  // fork the camera round
        e = hipEventRecord(B.cev[0], s);
        if (e == hipSuccess) e = hipStreamWaitEvent(B.s3, B.cev[0], 0);
        if (e == hipSuccess) e = trace_on(0, B.bst[0], B.s3, S.cam_blocks);
        if (e == hipSuccess) e = hipEventRecord(B.cev[1], B.s3);
    }
    uint32_t k = 0;  // photon rounds issued
    for (uint32_t r = 0; r < S.n_rounds && e == hipSuccess; ++r) {
        if (round_items(S, r) == 0) continue;
        if (r == 0) {
            if (!S.cam_blocks) e = trace(0, B.bst[0]);
            continue;
        }
        if (!S.overlap) {
            if (S.binned) e = launch_bins_reset(B.bst[0], s);
            if (e == hipSuccess) e = trace(r, B.bst[0]);
            if (e == hipSuccess && S.binned)
                e = timed(ctx, timing, SPX_KT_BINS, s,
                          [&] { return launch_bins_resolve(B.bst[0], B.shared[0], acc_blocks, s); });
            continue;
        }
        const uint32_t buf = k % S.nbuf;
        hipEvent_t traced = B.ev[2 * k], resolved = B.ev[2 * k + 1];
        if (k >= S.nbuf)  // the buffer's previous round resolved (and reset it)
            e = hipStreamWaitEvent(s, B.ev[2 * (k - S.nbuf) + 1], 0);
        else  // the buffer's first round of the launch
            e = launch_bins_reset(B.bst[buf], s);
        if (e == hipSuccess) e = trace(r, B.bst[buf]);
        if (e == hipSuccess) e = hipEventRecord(traced, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(B.s2, traced, 0);
        if (e == hipSuccess)
            e = timed(ctx, timing, SPX_KT_BINS, B.s2,
                      [&] { return launch_bins_resolve(B.bst[buf], B.shared[buf], acc_blocks, B.s2); });
        // The round that reuses the buffer finds it reset: the reset runs
        // behind the passes on s2, not on the trace stream's critical path.
        if (e == hipSuccess && k + S.nbuf < n_photon_rounds) e = launch_bins_reset(B.bst[buf], B.s2);
        if (e == hipSuccess) e = hipEventRecord(resolved, B.s2);
        ++k;
    }
    // The concurrent camera round's tail: once the photon rounds are traced,
    // more blocks join its grid in draining the camera queue (blocks that
    // find it empty exit at once), so the step does not wait on a camera
    // round left with its small grid alone.  Two blocks per compute unit, not
    // the full grid: the last rounds' bin passes still run on the second
    // stream, and persistent camera blocks on every slot starved them (one
    // scatter pass 24 ms instead of 5).  The 512-job mixed step on 256 CUs,
    // against the full grid of 768 blocks: 256, 512 and 1024 blocks gave
    // 244.5, 234.5 and 242.2 ms against 241.1 on one box; 384, 512 and 640
    // gave 223.7, 221.3 and 224.5 against 227.3 on another.
    const uint32_t cus = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256);
    if (e == hipSuccess && S.cam_blocks) e = trace_on(0, B.bst[0], s, std::min(S.blocks, 2u * cus));
    if (e == hipSuccess && S.overlap && k) e = hipStreamWaitEvent(s, B.ev[2 * (k - 1) + 1], 0);  // join
    if (e == hipSuccess && S.cam_blocks) e = hipStreamWaitEvent(s, B.cev[1], 0);
    if (e == hipSuccess && !S.reps.empty())
        e = launch_reduce_replicas((const ReplicaJob *)(d + S.off_reps), (uint32_t)S.reps.size(), S.rep_blocks, s);
    return e;
}

// `n` ordering events (no timing) in `pool`, created as needed.
hipError_t ensure_events(std::vector<hipEvent_t> &pool, size_t n) {
    while (pool.size() < n) {
        hipEvent_t ev;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
        pool.push_back(ev);
    }
    return hipSuccess;
}

int batch_stage(spx_ctx *ctx, size_t bytes, hipStream_t s) {
    if (ctx->batch_pending) {  // the previous staging copy may still read the pinned buffer
        hipError_t e = hipEventSynchronize(ctx->batch_done);
        if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_run wait");
        ctx->batch_pending = false;
    }
    if (!ctx->batch_done) {
        hipError_t e = hipEventCreateWithFlags(&ctx->batch_done, hipEventDisableTiming);
        if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_run event");
    }
    if (ctx->batch_host_bytes < bytes) {
        if (ctx->batch_host) hipHostFree(ctx->batch_host);
        ctx->batch_host = nullptr;
        ctx->batch_host_bytes = 0;
        hipError_t e = hipHostMalloc(&ctx->batch_host, bytes * 2, hipHostMallocDefault);
        if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_run pinned staging");
        ctx->batch_host_bytes = bytes * 2;
    }
    if (ctx->batch_dev_bytes < bytes) {
        hipError_t e = hipStreamSynchronize(s);  // earlier launches may still read the old table
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_run sync");
        if (ctx->batch_dev) hipFree(ctx->batch_dev);
        ctx->batch_dev = nullptr;
        ctx->batch_dev_bytes = 0;
        e = hipMalloc(&ctx->batch_dev, bytes * 2);
        ctx->buf_gen++;
        if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_run table");
        ctx->batch_dev_bytes = bytes * 2;
    }
    return SPX_OK;
}

// Identity of a schedule's layout: everything the captured kernels and the
// buffer sizes depend on, apart from the uploaded per-job values.
bool same_layout(const Schedule &a, const Schedule &b) {
    if (a.n_jobs != b.n_jobs || a.n_rounds != b.n_rounds || a.binned != b.binned || a.overlap != b.overlap ||
        a.splat_mode != b.splat_mode || a.cam_blocks != b.cam_blocks || a.ph_threads != b.ph_threads ||
        a.blocks != b.blocks || a.bytes != b.bytes || a.rep_total != b.rep_total || a.bin_want != b.bin_want ||
        a.ordered.size() != b.ordered.size() || a.reps.size() != b.reps.size())
        return false;
    if (memcmp(a.ordered.data(), b.ordered.data(), a.ordered.size() * sizeof(BatchItem)) != 0) return false;
    for (uint32_t r = 0; r < a.n_rounds; ++r)
        if (memcmp(a.qh[r].begin, b.qh[r].begin, sizeof a.qh[r].begin) != 0 ||
            memcmp(a.qh[r].size, b.qh[r].size, sizeof a.qh[r].size) != 0)
            return false;
    for (size_t k = 0; k < a.reps.size(); ++k)  // equal sizes in equal order: equal replica offsets
        if (a.reps[k].n != b.reps[k].n || a.reps[k].reps != b.reps[k].reps || a.reps[k].pad != b.reps[k].pad ||
            a.reps[k].first_block != b.reps[k].first_block)
            return false;
    for (uint32_t j = 0; j < a.n_jobs; ++j)
        if (a.table[j].kind != b.table[j].kind) return false;
    return a.map.n_strips == b.map.n_strips && a.map.key == b.map.key && a.map.shared == b.map.shared &&
           a.map.width == b.map.width && a.map.npix == b.map.npix && a.map.tile_wlog2 == b.map.tile_wlog2;
}

}  // namespace

int validate_job(spx_ctx *ctx, const spx_job &job) {
    BatchJob tmp;
    return build_batch_job(ctx, job, tmp, 0);
}

}  // namespace spx

struct spx_batch_plan {
    spx_ctx *ctx = nullptr;
    spx_tuning tuning;         // the context's tuning at creation
    bool has_cfg = false;
    spx_batch_cfg cfg;         // the creation's per-call overrides
    spx::Schedule S;
    char *dev = nullptr;   // schedule image
    char *host = nullptr;  // pinned copy of the image
    char *bins[kRoundBufs] = {};  // deposit round buffers (more than one: overlapped rounds)
    spx::BinLayout bl;
    spx::RoundBufs rb;
    float *replica = nullptr;
    hipStream_t cap = nullptr;   // capture stream
    hipStream_t cap2 = nullptr;  // second capture stream (overlapped bin passes)
    hipStream_t cap3 = nullptr;  // third capture stream (concurrent camera round)
    std::vector<hipEvent_t> ev;  // capture-time ordering events
    std::vector<hipEvent_t> cev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t uploaded = nullptr;  // after the last image upload from `host`
    bool upload_pending = false;
    size_t nodes = 0;
    bool shared = false;    // buffers are the context's (dev, host, bins, replica not owned)
    uint64_t gen = 0;       // the context's buffer generation at capture
};

static void plan_free(spx_batch_plan *p) {
    if (!p) return;
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    if (p->cap) hipStreamDestroy(p->cap);
    if (p->cap2) hipStreamDestroy(p->cap2);
    if (p->cap3) hipStreamDestroy(p->cap3);
    for (hipEvent_t ev : p->ev) hipEventDestroy(ev);
    for (hipEvent_t ev : p->cev) hipEventDestroy(ev);
    if (p->uploaded) hipEventDestroy(p->uploaded);
    if (!p->shared) {
        if (p->dev) hipFree(p->dev);
        for (char *b : p->bins)
            if (b) hipFree(b);
        if (p->replica) hipFree(p->replica);
        if (p->host) hipHostFree(p->host);
    }
    delete p;
}

// Stream-ordered upload of the job map into every deposit buffer of a plan.
static int plan_upload_maps(spx_batch_plan *p, hipStream_t s) {
    if (!p->S.binned) return SPX_OK;
    for (uint32_t k = 0; k < p->S.nbuf; ++k) {
        int rc = bins_upload_map(p->ctx, p->bins[k], p->bl, p->S.map, s);
        if (rc) return rc;
    }
    return SPX_OK;
}

// Uploads the plan's image and job maps on `s` (stream-ordered).
static int plan_upload(spx_batch_plan *p, hipStream_t s) {
    spx_ctx *ctx = p->ctx;
    if (p->shared) {
        if (p->gen != ctx->buf_gen)
            return set_error(ctx, SPX_EINVAL, "spx_batch_plan: the context's buffers moved; capture again");
        int rc = batch_stage(ctx, p->S.bytes, s);  // waits for the previous use of the staging copy
        if (rc) return rc;
        if (p->gen != ctx->buf_gen)
            return set_error(ctx, SPX_EINVAL, "spx_batch_plan: the context's buffers moved; capture again");
        write_image(p->S, (char *)ctx->batch_host);
        hipError_t e = hipMemcpyAsync(ctx->batch_dev, ctx->batch_host, p->S.bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(ctx->batch_done, s);
        if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_plan upload");
        ctx->batch_pending = true;
        return plan_upload_maps(p, s);
    }
    if (p->upload_pending) {
        hipError_t e = hipEventSynchronize(p->uploaded);
        if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_plan upload wait");
        p->upload_pending = false;
    }
    write_image(p->S, p->host);
    hipError_t e = hipMemcpyAsync(p->dev, p->host, p->S.bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(p->uploaded, s);
    if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_plan upload");
    p->upload_pending = true;
    return plan_upload_maps(p, s);
}

// Binds the plan's deposit buffers (already sized for p->bl) into p->rb.
static void plan_bind(spx_batch_plan *p) {
    memset(&p->rb.bst, 0, sizeof p->rb.bst);
    if (!p->S.binned) return;
    for (uint32_t k = 0; k < p->S.nbuf; ++k) {
        bins_bind(p->bins[k], p->bl, p->S.map, p->rb.bst[k], &p->rb.shared[k]);
        p->rb.bst[k].atomic_film = p->S.overlap ? 1u : 0u;
    }
}

// Captures the plan's kernel sequence into a graph (nothing to capture for
// a table of empty jobs).  Overlapped schedules fork onto a second capture
// stream and join back, so the graph holds both branches.
static int plan_capture(spx_batch_plan *p) {
    if (p->S.ordered.empty()) return SPX_OK;
    spx_ctx *ctx = p->ctx;
    hipError_t e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
    if (e == hipSuccess && p->S.overlap) {
        e = hipStreamCreateWithFlags(&p->cap2, hipStreamNonBlocking);
        if (e == hipSuccess) e = ensure_events(p->ev, 2 * (size_t)photon_rounds(p->S));
        p->rb.s2 = p->cap2;
        p->rb.ev = p->ev.data();
    }
    if (e == hipSuccess && p->S.cam_blocks) {
        e = hipStreamCreateWithFlags(&p->cap3, hipStreamNonBlocking);
        if (e == hipSuccess) e = ensure_events(p->cev, 2);
        p->rb.s3 = p->cap3;
        p->rb.cev = p->cev.data();
    }
    if (e == hipSuccess) e = hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t el = enqueue(ctx, p->S, p->dev, p->rb, p->replica, p->cap, false);
        e = hipStreamEndCapture(p->cap, &p->graph);
        if (el != hipSuccess) e = el;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphGetNodes(p->graph, nullptr, &p->nodes);
    return check_hip(ctx, e, "spx_batch_plan capture");
}

// The context's extra streams and ordering events for overlapped rounds and
// a concurrent camera round (events: 2 for the camera, then 2 per round).
static int ctx_overlap_streams(spx_ctx *ctx, const Schedule &S, RoundBufs &B) {
    if (!S.overlap && !S.cam_blocks) return SPX_OK;
    hipError_t e = hipSuccess;
    if (S.overlap && !ctx->aux) e = hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking);
    if (e == hipSuccess && S.cam_blocks && !ctx->aux_cam) e = hipStreamCreateWithFlags(&ctx->aux_cam, hipStreamNonBlocking);
    if (e == hipSuccess) e = ensure_events(ctx->sync_ev, 2 + 2 * (size_t)photon_rounds(S));
    if (e != hipSuccess) return check_hip(ctx, e, "overlap streams");
    B.s2 = ctx->aux;
    B.s3 = ctx->aux_cam;
    B.cev = ctx->sync_ev.data();
    B.ev = ctx->sync_ev.data() + 2;
    return SPX_OK;
}

// Sizes the context's deposit buffers for S (growing them may move them)
// and binds them into B.
static int ctx_bins(spx_ctx *ctx, const Schedule &S, hipStream_t s, RoundBufs &B) {
    memset(&B.bst, 0, sizeof B.bst);
    if (!S.binned) return SPX_OK;
    for (uint32_t k = 0; k < S.nbuf; ++k) {
        int rc = prepare_bins(ctx, S.map, S.bin_want, s, B.bst[k], &B.shared[k], k);
        if (rc) return rc;
        B.bst[k].atomic_film = S.overlap ? 1u : 0u;
    }
    return SPX_OK;
}

int spx::batch_plan_create_shared(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, hipStream_t s,
                                  spx_batch_plan **out) {
    *out = nullptr;
    spx_batch_plan *p = new spx_batch_plan();
    p->ctx = ctx;
    p->shared = true;
    p->tuning = ctx->tuning;
    int rc = plan_schedule(ctx, jobs, n_jobs, nullptr, p->S);
    Schedule &S = p->S;
    // grow the context's buffers first (this may move them), then bind
    RoundBufs tmp;
    if (rc == SPX_OK) rc = ctx_bins(ctx, S, s, tmp);
    if (rc == SPX_OK && !S.reps.empty()) rc = ensure_replicas(ctx, S.rep_total, s);
    if (rc == SPX_OK) rc = batch_stage(ctx, S.bytes, s);
    if (rc != SPX_OK) {
        plan_free(p);
        return rc;
    }
    if (S.binned) {
        p->bl = bins_layout(S.map, S.bin_want, ctx->tuning.bin_log2);
        for (int k = 0; k < kRoundBufs; ++k) p->bins[k] = (char *)ctx->bins[k];
    }
    plan_bind(p);  // the maps were uploaded by prepare_bins
    if (!S.reps.empty()) {
        p->replica = ctx->replica;
        bind_replicas(S, p->replica);
    }
    p->dev = (char *)ctx->batch_dev;
    p->gen = ctx->buf_gen;
    hipError_t e = hipEventCreateWithFlags(&p->uploaded, hipEventDisableTiming);
    if (e != hipSuccess) {
        rc = check_hip(ctx, e, "batch plan event");
        plan_free(p);
        return rc;
    }
    write_image(S, (char *)ctx->batch_host);
    e = hipMemcpyAsync(ctx->batch_dev, ctx->batch_host, S.bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(ctx->batch_done, s);
    if (e != hipSuccess) {
        rc = check_hip(ctx, e, "batch plan upload");
        plan_free(p);
        return rc;
    }
    ctx->batch_pending = true;
    if ((rc = plan_capture(p)) != SPX_OK) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return SPX_OK;
}

double spx::job_queue_bytes(const spx_ctx *ctx, const spx_job &job) {
    if (job.kind == SPX_JOB_CAMERA) return 0.0;
    if (ctx->tuning.splat_mode >= 2)  // records and their padded sorted copy
        return kRoundBytesPerRecord * (double)job.n_photons * records_per_photon(job.kind == SPX_JOB_LIGHTTRACE);
    size_t n = (size_t)3 * photon_film_w(job) * photon_film_h(job);
    uint32_t R = replica_count(ctx, n);
    return R > 1 ? (double)R * n * sizeof(float) : 0.0;
}

extern "C" {

int spx_batch_run(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, const spx_batch_cfg *cfg, void *stream) {
    if (!ctx) return SPX_EINVAL;
    if (!jobs && n_jobs) return set_error(ctx, SPX_EINVAL, "spx_batch_run: null job table");
    if (n_jobs == 0) return SPX_OK;
    hipStream_t s = (hipStream_t)stream;
    Schedule S;
    int rc = plan_schedule(ctx, jobs, n_jobs, cfg, S);
    if (rc) return rc;
    if (S.ordered.empty()) return SPX_OK;
    RoundBufs B;
    if ((rc = ctx_bins(ctx, S, s, B)) || (rc = ctx_overlap_streams(ctx, S, B))) return rc;
    if (!S.reps.empty()) {
        if ((rc = ensure_replicas(ctx, S.rep_total, s))) return rc;
        bind_replicas(S, ctx->replica);
    }
    if ((rc = batch_stage(ctx, S.bytes, s))) return rc;
    write_image(S, (char *)ctx->batch_host);
    char *d = (char *)ctx->batch_dev;
    hipError_t e = hipMemcpyAsync(d, ctx->batch_host, S.bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(ctx->batch_done, s);
    if (e != hipSuccess) return check_hip(ctx, e, "spx_batch_run upload");
    ctx->batch_pending = true;
    e = enqueue(ctx, S, d, B, ctx->replica, s, true);
    return check_hip(ctx, e, "spx_batch_run launch");
}

int spx_batch_plan_create(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, const spx_batch_cfg *cfg, void *stream,
                          spx_batch_plan **out) {
    if (!ctx || !out) return SPX_EINVAL;
    *out = nullptr;
    if (!jobs || n_jobs == 0) return set_error(ctx, SPX_EINVAL, "spx_batch_plan_create: empty job table");
    hipStream_t s = (hipStream_t)stream;
    spx_batch_plan *p = new spx_batch_plan();
    p->ctx = ctx;
    p->tuning = ctx->tuning;
    p->has_cfg = cfg != nullptr;
    if (cfg) p->cfg = *cfg;
    int rc = plan_schedule(ctx, jobs, n_jobs, cfg, p->S);
    if (rc) {
        plan_free(p);
        return rc;
    }
    Schedule &S = p->S;
    hipError_t e = hipMalloc(&p->dev, S.bytes);
    if (e == hipSuccess) e = hipHostMalloc(&p->host, S.bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->uploaded, hipEventDisableTiming);
    if (e == hipSuccess && S.binned) {
        p->bl = bins_layout(S.map, S.bin_want, ctx->tuning.bin_log2);
        for (uint32_t k = 0; k < S.nbuf && e == hipSuccess; ++k) e = hipMalloc(&p->bins[k], p->bl.total);
    }
    if (e == hipSuccess && !S.reps.empty()) {
        e = hipMalloc(&p->replica, S.rep_total * sizeof(float));
        if (e == hipSuccess) bind_replicas(S, p->replica);
    }
    if (e != hipSuccess) {
        rc = check_hip(ctx, e, "spx_batch_plan_create alloc");
        plan_free(p);
        return rc;
    }
    plan_bind(p);
    if ((rc = plan_upload(p, s))) {
        plan_free(p);
        return rc;
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess || (rc = plan_capture(p)) != SPX_OK) {
        if (e != hipSuccess) rc = check_hip(ctx, e, "spx_batch_plan_create sync");
        plan_free(p);
        return rc;
    }
    *out = p;
    return SPX_OK;
}

int spx_batch_plan_update(spx_batch_plan *p, const spx_job *jobs, uint32_t n_jobs, void *stream) {
    if (!p) return SPX_EINVAL;
    spx_ctx *ctx = p->ctx;
    if (!jobs || n_jobs != p->S.n_jobs) return set_error(ctx, SPX_EINVAL, "spx_batch_plan_update: job count differs");
    // Scheduled under the creation's tuning, whatever the context's is now.
    const spx_tuning keep = ctx->tuning;
    ctx->tuning = p->tuning;
    Schedule S;
    int rc = plan_schedule(ctx, jobs, n_jobs, p->has_cfg ? &p->cfg : nullptr, S);
    ctx->tuning = keep;
    if (rc) return rc;
    if (!same_layout(S, p->S))
        return set_error(ctx, SPX_EINVAL, "spx_batch_plan_update: the jobs need a different schedule");
    if (!S.reps.empty()) bind_replicas(S, p->replica);
    p->S = std::move(S);
    return plan_upload(p, (hipStream_t)stream);
}

int spx_batch_plan_launch(spx_batch_plan *p, void *stream) {
    if (!p) return SPX_EINVAL;
    if (p->S.ordered.empty()) return SPX_OK;
    return check_hip(p->ctx, hipGraphLaunch(p->exec, (hipStream_t)stream), "spx_batch_plan_launch");
}

int spx_batch_plan_info(const spx_batch_plan *p, uint32_t *rounds, uint32_t *graph_nodes) {
    if (!p) return SPX_EINVAL;
    if (rounds) {
        uint32_t n = 0;
        for (uint32_t r = 0; r < p->S.n_rounds; ++r) n += round_items(p->S, r) ? 1 : 0;
        *rounds = n;
    }
    if (graph_nodes) *graph_nodes = (uint32_t)p->nodes;
    return SPX_OK;
}

void spx_batch_plan_destroy(spx_batch_plan *p) {
    if (!p) return;
    hipSetDevice(p->ctx->device);
    hipDeviceSynchronize();
    plan_free(p);
}

int spx_autotune(spx_ctx *ctx, const spx_job *jobs, uint32_t n_jobs, uint32_t sweeps, void *stream, spx_tuning *out,
                 double *best_ms) {
    if (!ctx) return SPX_EINVAL;
    if (!jobs || n_jobs == 0 || !out || sweeps < 1 || sweeps > 4)
        return set_error(ctx, SPX_EINVAL, "spx_autotune: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    // Scratch films: one per distinct film of the table, same sizes and the
    // same sharing between jobs.
    std::map<float *, size_t> film_off;
    size_t floats = 0;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const spx_job &jb = jobs[j];
        size_t n = (size_t)3 * (jb.kind == SPX_JOB_LIGHTTRACE ? (size_t)jb.screen.width * jb.screen.height
                                                              : (size_t)jb.width * jb.height);
        if (!jb.film) return set_error(ctx, SPX_EINVAL, "spx_autotune: null film");
        if (film_off.count(jb.film)) continue;
        film_off[jb.film] = floats;
        floats += (n + 63) & ~(size_t)63;
    }
    float *scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&scratch, floats * sizeof(float));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        if (scratch) hipFree(scratch);
        if (e0) hipEventDestroy(e0);
        return check_hip(ctx, e, "spx_autotune setup");
    }
    std::vector<spx_job> tj(jobs, jobs + n_jobs);
    for (spx_job &jb : tj) jb.film = scratch + film_off[jb.film];
    const spx_tuning start = ctx->tuning;
    const bool ktime = ctx->ktime_on;
    ctx->ktime_on = false;
    int rc = SPX_OK;
    // Time of the table under tuning t (ms; +inf when it fails).
    auto eval = [&](const spx_tuning &t) -> double {
        ctx->tuning = t;
        double best = INFINITY;
        for (int k = 0; k < 3 && rc == SPX_OK; ++k) {
            hipError_t he = hipEventRecord(e0, s);
            if (he != hipSuccess) {
                rc = check_hip(ctx, he, "spx_autotune event");
                break;
            }
            int r = spx_batch_run(ctx, tj.data(), n_jobs, nullptr, stream);
            if (r != SPX_OK) {
                if (r == SPX_ENOMEM || r == SPX_EDEVICE) rc = r;
                return INFINITY;
            }
            float ms = 0.0f;
            he = hipEventRecord(e1, s);
            if (he == hipSuccess) he = hipEventSynchronize(e1);
            if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
            if (he != hipSuccess) {
                rc = check_hip(ctx, he, "spx_autotune timing");
                break;
            }
            if (k > 0) best = std::min(best, (double)ms);
        }
        return best;
    };
    const uint32_t cus = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256);
    struct Axis {
        uint32_t spx_tuning::*field;
        std::vector<uint32_t> values;
        bool binned_only;
    };
    // Launch-structure axes first (deposit mode, overlapped rounds, the
    // concurrent camera grid, then the persistent grid it shares the GPU
    // with), then the work-item and tile sizes.
    const std::vector<Axis> axes = {
        {&spx_tuning::splat_mode, {2, 3, 1}, false},
        {&spx_tuning::overlap_bins, {0, 1, 2}, true},
        {&spx_tuning::camera_blocks, {0, cus / 2, cus, cus * 2}, false},
        {&spx_tuning::batch_blocks, {cus * 2, cus * 3, cus * 4, cus * 8}, false},
        {&spx_tuning::chunk_photons, {2048, 4096, 8192, 16384, 32768}, false},
        {&spx_tuning::chunk_paths, {4096, 8192, 16384, 32768}, false},
        {&spx_tuning::tile_wlog2, {3, 4, 5, 6, 7}, true},
        {&spx_tuning::bin_log2, {27, 28, 29, 30}, true},
        {&spx_tuning::photon_threads, {256, 1024}, false},
    };
    spx_tuning best_t = start;
    if (best_t.batch_blocks == 0) best_t.batch_blocks = cus * 4;
    double best = eval(best_t);
    for (uint32_t sw = 0; sw < sweeps && rc == SPX_OK; ++sw) {
        bool improved = false;
        for (const Axis &ax : axes) {
            if (ax.binned_only && best_t.splat_mode < 2) continue;
            for (uint32_t v : ax.values) {
                if (best_t.*ax.field == v) continue;
                spx_tuning t = best_t;
                t.*ax.field = v;
                double ms = eval(t);
                if (rc != SPX_OK) break;
                if (ms < best * 0.99) {  // keep a change only for a clear gain
                    best = ms;
                    best_t = t;
                    improved = true;
                }
            }
            if (rc != SPX_OK) break;
        }
        if (!improved) break;
    }
    hipStreamSynchronize(s);
    hipFree(scratch);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ctx->ktime_on = ktime;
    ctx->tuning = rc == SPX_OK ? best_t : start;
    if (rc != SPX_OK) return rc;
    if (!isfinite(best)) return set_error(ctx, SPX_EINVAL, "spx_autotune: no candidate ran");
    *out = best_t;
    if (best_ms) *best_ms = best;
    return SPX_OK;
}

}  // extern "C"
