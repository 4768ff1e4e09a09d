// spx_serve.hip -- the serving loop: continuous batching of render requests.
//
// Requests are admitted, oldest first, when their footprint (film plus its
// share of the deposit queues: records or film replicas) fits under the
// memory cap; every active request's next frame
// runs in ONE batched launch per step, so requests that arrive while others
// are on the GPU join at the next step and their work overlaps on the
// device.  The step's schedule is captured as a batch plan (a hipGraph) and
// replayed, with the frames' RNG bases and cameras rewritten, for as long
// as the set of active requests stays the same; a new set is captured anew.  Films and RNG counters (frames_done) stay
// resident on the device between frames.  The reference has no server; its
// nearest analogue is the viewer's progressive loop (spectral-viewer
// main.rs:395-419), whose camera moves reset the accumulation
// (main.rs:370-391).
#include <math.h>
#include <string.h>

#include <chrono>
#include <algorithm>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "spx_internal.hpp"

using namespace spx;

namespace {

struct Req {
    spx_request r;
    uint32_t state = SPX_REQ_PENDING;
    uint32_t frames_done = 0;
    uint64_t footprint = 0;   // film + queue share
    uint64_t film_bytes = 0;
    uint64_t queue_bytes = 0; // deposit records (or film replicas) one frame needs
    double t_submit = 0, t_admit = 0, t_done = 0;
    float *film = nullptr;
    size_t film_floats = 0;
};

// Deposit queues are shared by the requests of a step and never exceed the
// round buffers' record capacity (binned deposits: records + padded sorted
// copy; one more buffer per overlapped round), so the requests' queue shares are
// charged up to that bound.
uint64_t queue_cap(const spx_ctx *ctx) {
    if (ctx->tuning.splat_mode < 2) return UINT64_MAX;
    return (uint64_t)(kRoundBytesPerRecord * (double)((uint64_t)1 << ctx->tuning.bin_log2)) *
           (ctx->tuning.overlap_bins + 1u);
}

size_t job_film_floats(const spx_job &j) {
    return j.kind == SPX_JOB_LIGHTTRACE ? (size_t)3 * j.screen.width * j.screen.height
                                        : (size_t)3 * j.width * j.height;
}

// Turns a camera basis by `deg` about the vertical axis through `c`.
spx_camera orbit_camera(const spx_camera &cam, const float c[3], double deg) {
    double a = deg * 3.14159265358979323846 / 180.0, ca = cos(a), sa = sin(a);
    auto rot_v = [&](const float *v, float *o) {
        double x = v[0], y = v[1], z = v[2];
        o[0] = (float)(ca * x + sa * z);
        o[1] = (float)y;
        o[2] = (float)(-sa * x + ca * z);
    };
    auto rot_p = [&](const float *p, float *o) {
        float d[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
        float r[3];
        rot_v(d, r);
        o[0] = r[0] + c[0];
        o[1] = r[1] + c[1];
        o[2] = r[2] + c[2];
    };
    spx_camera out;
    rot_p(cam.origin, out.origin);
    rot_p(cam.lower_left, out.lower_left);
    rot_v(cam.horizontal, out.horizontal);
    rot_v(cam.vertical, out.vertical);
    return out;
}

}  // namespace

struct spx_server {
    spx_ctx *ctx = nullptr;
    uint64_t cap = 0;
    uint64_t film_used = 0;   // films of admitted requests
    uint64_t queue_used = 0;  // their queue shares (charged up to queue_cap)
    uint64_t used(const spx_ctx *c) const { return film_used + std::min(queue_used, queue_cap(c)); }
    uint64_t next_id = 1;
    std::map<uint64_t, Req> reqs;
    std::deque<uint64_t> pending;
    std::vector<uint64_t> active;
    bool graphs = true;                 // steps replay captured plans
    spx_batch_plan *plan = nullptr;     // plan of the active set `plan_ids`
    std::vector<uint64_t> plan_ids;
    uint64_t replays = 0, captures = 0; // steps served by a replay / by a new capture
    std::chrono::steady_clock::time_point t0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static spx_camera frame_camera(const Req &q, uint32_t frame) {
    if (q.r.orbit_deg == 0.0f || frame == 0) return q.r.job.camera;
    return orbit_camera(q.r.job.camera, q.r.orbit_center, (double)q.r.orbit_deg * frame);
}

extern "C" {

int spx_server_create(spx_ctx *ctx, uint64_t mem_cap_bytes, spx_server **out) {
    if (!ctx || !out || mem_cap_bytes == 0) return set_error(ctx, SPX_EINVAL, "spx_server_create: bad arguments");
    spx_server *s = new spx_server();
    s->ctx = ctx;
    s->cap = mem_cap_bytes;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return SPX_OK;
}

void spx_server_destroy(spx_server *srv) {
    if (!srv) return;
    hipDeviceSynchronize();
    if (srv->plan) spx_batch_plan_destroy(srv->plan);
    for (auto &kv : srv->reqs)
        if (kv.second.film) hipFree(kv.second.film);
    delete srv;
}

double spx_server_clock(const spx_server *srv) { return srv ? srv->now() : 0.0; }

int spx_server_set_graphs(spx_server *srv, int enable) {
    if (!srv) return SPX_EINVAL;
    srv->graphs = enable != 0;
    return SPX_OK;
}

int spx_server_stats(const spx_server *srv, uint64_t *captures, uint64_t *replays) {
    if (!srv) return SPX_EINVAL;
    if (captures) *captures = srv->captures;
    if (replays) *replays = srv->replays;
    return SPX_OK;
}

uint64_t spx_server_bytes_in_use(const spx_server *srv) { return srv ? srv->used(srv->ctx) : 0; }

int spx_server_submit(spx_server *srv, const spx_request *req, uint64_t *id) {
    if (!srv || !req || !id) return SPX_EINVAL;
    spx_ctx *ctx = srv->ctx;
    if (req->frames == 0) return set_error(ctx, SPX_EINVAL, "spx_server_submit: zero frames");
    if (!isfinite(req->orbit_deg) || !isfinite(req->orbit_center[0]) || !isfinite(req->orbit_center[1]) ||
        !isfinite(req->orbit_center[2]))
        return set_error(ctx, SPX_EINVAL, "spx_server_submit: non-finite orbit");
    spx_job probe = req->job;
    static float dummy;
    probe.film = &dummy;  // validation only; the server allocates the film at admission
    int rc = validate_job(ctx, probe);
    if (rc) return rc;
    Req q;
    q.r = *req;
    q.film_floats = job_film_floats(req->job);
    q.film_bytes = q.film_floats * sizeof(float);
    q.queue_bytes = (uint64_t)job_queue_bytes(ctx, req->job);
    q.footprint = q.film_bytes + q.queue_bytes;
    if (q.film_bytes + std::min<uint64_t>(q.queue_bytes, queue_cap(ctx)) > srv->cap)
        return set_error(ctx, SPX_EBUSY, "spx_server_submit: footprint exceeds the cap");
    q.t_submit = srv->now();
    *id = srv->next_id++;
    srv->reqs[*id] = q;
    srv->pending.push_back(*id);
    return SPX_OK;
}

int spx_server_step(spx_server *srv, void *stream, uint32_t *n_active) {
    if (!srv) return SPX_EINVAL;
    spx_ctx *ctx = srv->ctx;
    hipStream_t s = (hipStream_t)stream;
    if (n_active) *n_active = 0;
    // Admission, oldest first; a request that does not fit waits while
    // smaller later ones may still be admitted.
    for (auto it = srv->pending.begin(); it != srv->pending.end();) {
        Req &q = srv->reqs[*it];
        if (srv->film_used + q.film_bytes + std::min(srv->queue_used + q.queue_bytes, queue_cap(ctx)) > srv->cap) {
            ++it;
            continue;
        }
        hipError_t e = hipMalloc(&q.film, q.film_floats * sizeof(float));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            q.film = nullptr;
            break;  // device full: retry at the next step
        }
        e = launch_fill(q.film, 0.0f, q.film_floats, s);
        if (e != hipSuccess) return check_hip(ctx, e, "spx_server_step clear");
        q.state = SPX_REQ_ACTIVE;
        q.t_admit = srv->now();
        srv->film_used += q.film_bytes;
        srv->queue_used += q.queue_bytes;
        srv->active.push_back(*it);
        it = srv->pending.erase(it);
    }
    if (srv->active.empty()) return SPX_OK;
    std::vector<spx_job> jobs;
    jobs.reserve(srv->active.size());
    for (uint64_t id : srv->active) {
        Req &q = srv->reqs[id];
        spx_job j = q.r.job;
        j.film = q.film;
        uint32_t f = q.frames_done;
        j.camera = frame_camera(q, f);
        if (j.kind == SPX_JOB_CAMERA) j.sample_base = q.r.job.sample_base + f * q.r.job.spp;
        else j.photon_base = q.r.job.photon_base + f * q.r.job.n_photons;
        if (q.r.orbit_deg != 0.0f && f > 0) {
            hipError_t e = launch_fill(q.film, 0.0f, q.film_floats, s);
            if (e != hipSuccess) return check_hip(ctx, e, "spx_server_step reset");
        }
        jobs.push_back(j);
    }
    int rc = SPX_OK;
    if (!srv->graphs) {
        rc = spx_batch_run(ctx, jobs.data(), (uint32_t)jobs.size(), nullptr, stream);
    } else {
        bool replayed = false;
        if (srv->plan && srv->plan_ids == srv->active) {
            rc = spx_batch_plan_update(srv->plan, jobs.data(), (uint32_t)jobs.size(), stream);
            if (rc == SPX_OK) {
                rc = spx_batch_plan_launch(srv->plan, stream);
                if (rc) return rc;
                replayed = true;
                srv->replays++;
            } else if (rc != SPX_EINVAL) {
                return rc;
            }
        }
        if (!replayed) {
            if (srv->plan) {
                spx_batch_plan_destroy(srv->plan);
                srv->plan = nullptr;
                srv->plan_ids.clear();
            }
            rc = batch_plan_create_shared(ctx, jobs.data(), (uint32_t)jobs.size(), s, &srv->plan);
            if (rc) return rc;
            srv->plan_ids = srv->active;
            srv->captures++;
            rc = spx_batch_plan_launch(srv->plan, stream);
        }
    }
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return check_hip(ctx, e, "spx_server_step sync");
    double t = srv->now();
    if (n_active) *n_active = (uint32_t)jobs.size();
    std::vector<uint64_t> still;
    for (uint64_t id : srv->active) {
        Req &q = srv->reqs[id];
        q.frames_done++;
        if (q.frames_done >= q.r.frames) {
            q.state = SPX_REQ_DONE;
            q.t_done = t;
        } else {
            still.push_back(id);
        }
    }
    srv->active.swap(still);
    return SPX_OK;
}

int spx_server_info(const spx_server *srv, uint64_t id, spx_request_info *out) {
    if (!srv || !out) return SPX_EINVAL;
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return SPX_EINVAL;
    const Req &q = it->second;
    out->state = q.state;
    out->frames_done = q.frames_done;
    out->footprint = q.footprint;
    out->t_submit = q.t_submit;
    out->t_admit = q.t_admit;
    out->t_done = q.t_done;
    return SPX_OK;
}

int spx_server_frame_camera(const spx_server *srv, uint64_t id, uint32_t frame, spx_camera *out) {
    if (!srv || !out) return SPX_EINVAL;
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end() || frame >= it->second.r.frames) return SPX_EINVAL;
    *out = frame_camera(it->second, frame);
    return SPX_OK;
}

int spx_server_film(spx_server *srv, uint64_t id, float *host_out, size_t floats) {
    if (!srv || !host_out) return SPX_EINVAL;
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end() || !it->second.film || floats != it->second.film_floats)
        return set_error(srv->ctx, SPX_EINVAL, "spx_server_film: unknown request, no film or size mismatch");
    hipError_t e = hipMemcpy(host_out, it->second.film, floats * sizeof(float), hipMemcpyDeviceToHost);
    return check_hip(srv->ctx, e, "spx_server_film");
}

int spx_server_release(spx_server *srv, uint64_t id) {
    if (!srv) return SPX_EINVAL;
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return SPX_EINVAL;
    Req &q = it->second;
    if (q.state == SPX_REQ_ACTIVE) return set_error(srv->ctx, SPX_EINVAL, "spx_server_release: request is active");
    if (q.state == SPX_REQ_PENDING) {
        for (auto p = srv->pending.begin(); p != srv->pending.end(); ++p)
            if (*p == id) {
                srv->pending.erase(p);
                break;
            }
    } else if (q.film) {
        hipError_t e = hipStreamSynchronize(nullptr);
        if (e == hipSuccess) e = hipFree(q.film);
        if (e != hipSuccess) return check_hip(srv->ctx, e, "spx_server_release");
        srv->film_used -= q.film_bytes;
        srv->queue_used -= q.queue_bytes;
    }
    srv->reqs.erase(it);
    return SPX_OK;
}

}  // extern "C"
