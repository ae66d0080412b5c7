// bm_serve.hip -- the request loop over one batch: rollout requests are
// queued with their arrival time, admitted into the batch while its particle,
// block and scene room and the HBM budget allow, stepped together, and
// retired one by one when their own substep count is reached, while later
// requests keep being admitted (continuous batching).  The reference has no
// serving (its solver holds one scene and runs on one thread).
//
// Scheduling, one bm_server_poll:
//   1. admit: walk the queue in arrival order; a request whose arrival has
//      come and that fits is added to the batch.  A later request may pass an
//      older one that does not fit, at most max_bypass times; then the queue
//      waits for the oldest.
//   2. step: the batch advances min(max_chunk, fewest remaining substeps of
//      the running requests) substeps, so that no request overshoots.
//   3. retire: requests with no substeps left are copied back to their
//      output buffers (and body), removed from the batch, and reported done;
//      a request whose scene raised a device error is retired as failed.
// With nothing running, a poll sleeps until the next arrival.
//
// Host-only code over the batch entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mpm_internal.hpp"

using namespace bm;

namespace {

struct Request {
    int64_t id;
    bm_request req;
    bm_body body;
    bool has_body;
    bm_request_info info;
    int32_t scene = -1;  // batch scene id while running
    int32_t left = 0;    // substeps still to run
    int64_t bytes = 0;   // budget charge while running
    int64_t blocks = 0;
};

int64_t blocks_of(const bm_scene_desc& d) {
    int64_t nb = 1;
    for (int a = 0; a < 3; ++a) nb *= (int64_t)((d.grid_dims[a] + 3) / 4);
    return nb;
}

}  // namespace

struct bm_server {
    bm_server_desc desc{};
    bm_batch_handle batch = nullptr;
    std::chrono::steady_clock::time_point t0;
    std::unordered_map<int64_t, Request> reqs;
    std::vector<int64_t> queue;    // queued ids, arrival order
    std::vector<int64_t> running;  // ids in the batch
    std::deque<int64_t> done;      // retired ids not yet reported
    int64_t next_id = 1;
    int64_t budget = 0, in_use = 0;
    int bypassed = 0;              // admissions past the current oldest waiting request
    int64_t oldest = 0;
    // seconds spent per activity: admission, stepping (launch to device
    // idle), retirement, waiting for arrivals; chunks run; substeps run
    double t_admit = 0.0, t_step = 0.0, t_retire = 0.0, t_idle = 0.0;
    int64_t chunks = 0, substeps = 0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static int check_server(bm_server_handle s) { return s ? BM_OK : fail(BM_ERR_INVALID_ARG, "null server"); }

static int64_t request_bytes(const bm_request& r) {
    int64_t b = 0;
    bm_batch_footprint(r.n, blocks_of(r.scene), 1, &b);
    return b - 256;
}

static void enqueue_sorted(bm_server_handle s, int64_t id) {
    const double a = s->reqs[id].info.arrival;
    auto it = std::upper_bound(s->queue.begin(), s->queue.end(), a,
                               [&](double t, int64_t other) { return t < s->reqs[other].info.arrival; });
    s->queue.insert(it, id);
}

// Step 1: admissions.
static int admit(bm_server_handle s) {
    const double t = s->now();
    for (size_t qi = 0; qi < s->queue.size();) {
        Request& R = s->reqs[s->queue[qi]];
        if (R.info.arrival > t) break;  // arrival order: nothing later has arrived either
        const bool is_oldest = qi == 0;
        if (is_oldest && s->oldest != R.id) {
            s->oldest = R.id;
            s->bypassed = 0;
        }
        if (!is_oldest && s->bypassed >= s->desc.max_bypass) break;
        int rc = BM_ERR_OOM;
        if (s->in_use + R.bytes <= s->budget) {
            rc = bm_batch_add_scene(s->batch, &R.req.scene, R.req.n, R.req.input, R.has_body ? &R.body : nullptr,
                                    &R.scene);
            if (rc == BM_ERR_HIP) return rc;
        }
        if (rc == BM_ERR_OOM) {  // no room now: wait (the request fits an empty server)
            ++qi;
            continue;
        }
        if (rc != BM_OK) {  // the request itself is bad (non-finite state, bad scene): it fails alone
            R.info.state = BM_REQ_FAILED;
            R.info.status = rc;
            R.info.finished = s->now();
            R.scene = -1;
            s->done.push_back(R.id);
            s->queue.erase(s->queue.begin() + (std::ptrdiff_t)qi);
            continue;
        }
        R.info.state = BM_REQ_RUNNING;
        R.info.admitted = s->now();
        s->in_use += R.bytes;
        s->running.push_back(R.id);
        s->queue.erase(s->queue.begin() + (std::ptrdiff_t)qi);
        if (!is_oldest) ++s->bypassed;
    }
    return BM_OK;
}

// Step 3: retirements (and device failures).
static int retire(bm_server_handle s) {
    for (size_t i = 0; i < s->running.size();) {
        Request& R = s->reqs[s->running[i]];
        int32_t code = BM_OK;
        if (int rc = bm_batch_scene_error(s->batch, R.scene, &code)) return rc;
        if (code == BM_OK && R.left > 0) {
            ++i;
            continue;
        }
        if (code == BM_OK) {
            if (R.req.output)
                if (int rc = bm_batch_download_scene(s->batch, R.scene, R.req.output)) return rc;
            if (R.has_body)
                if (int rc = bm_batch_get_body(s->batch, R.scene, &R.info.body, &R.info.reaction)) return rc;
            R.info.state = BM_REQ_DONE;
        } else {
            R.info.state = BM_REQ_FAILED;
            R.info.status = code;
        }
        if (int rc = bm_batch_remove_scene(s->batch, R.scene)) return rc;
        R.info.finished = s->now();
        s->in_use -= R.bytes;
        R.scene = -1;
        s->done.push_back(R.id);
        s->running.erase(s->running.begin() + (std::ptrdiff_t)i);
    }
    return BM_OK;
}

extern "C" {

int bm_server_create(const bm_server_desc* desc, bm_server_handle* out) {
    if (!desc || !out) return fail(BM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(desc->dt > 0.0f)) return fail(BM_ERR_INVALID_ARG, "dt must be > 0");
    if (desc->hbm_budget < 0 || desc->max_chunk < 0 || desc->max_bypass < 0)
        return fail(BM_ERR_INVALID_ARG, "negative server limit");
    auto* s = new bm_server();
    s->desc = *desc;
    if (s->desc.max_chunk == 0) s->desc.max_chunk = 32;
    if (s->desc.max_bypass == 0) s->desc.max_bypass = 8;
    if (int rc = bm_batch_create(&desc->batch, &s->batch)) {
        delete s;
        return rc;
    }
    int64_t arenas = 0;
    bm_batch_footprint(desc->batch.max_particles, desc->batch.max_blocks, desc->batch.max_scenes, &arenas);
    s->budget = desc->hbm_budget > 0 ? std::min(desc->hbm_budget, arenas) : arenas;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return BM_OK;
}

int bm_server_destroy(bm_server_handle s) {
    if (!s) return BM_OK;
    bm_batch_destroy(s->batch);
    delete s;
    return BM_OK;
}

int bm_server_clock(bm_server_handle s, double* seconds) {
    if (int rc = check_server(s)) return rc;
    if (!seconds) return fail(BM_ERR_INVALID_ARG, "null output");
    *seconds = s->now();
    return BM_OK;
}

int bm_server_batch(bm_server_handle s, bm_batch_handle* batch) {
    if (int rc = check_server(s)) return rc;
    if (!batch) return fail(BM_ERR_INVALID_ARG, "null output");
    *batch = s->batch;
    return BM_OK;
}

int bm_server_submit(bm_server_handle s, const bm_request* req, int64_t* id) {
    if (int rc = check_server(s)) return rc;
    if (!req || !id) return fail(BM_ERR_INVALID_ARG, "null argument");
    if (req->n <= 0 || !req->input) return fail(BM_ERR_INVALID_ARG, "a request needs particles");
    if (req->substeps <= 0) return fail(BM_ERR_INVALID_ARG, "substeps must be > 0");
    for (int a = 0; a < 3; ++a)
        if (req->scene.grid_dims[a] <= 0) return fail(BM_ERR_INVALID_ARG, "grid_dims must be > 0");
    Request R{};
    R.id = s->next_id++;
    R.req = *req;
    R.has_body = req->body != nullptr;
    if (R.has_body) R.body = *req->body;
    R.req.body = nullptr;
    R.left = req->substeps;
    R.bytes = request_bytes(*req);
    R.blocks = blocks_of(req->scene);
    // a request that cannot fit an empty server is refused at once
    if (R.bytes > s->budget || req->n > s->desc.batch.max_particles || R.blocks > s->desc.batch.max_blocks)
        return fail(BM_ERR_OOM, "request of %lld particles and %lld blocks exceeds the server's capacity",
                    (long long)req->n, (long long)R.blocks);
    R.info.state = BM_REQ_QUEUED;
    R.info.status = BM_OK;
    R.info.arrival = req->arrival >= 0.0 ? req->arrival : s->now();
    R.info.admitted = R.info.finished = -1.0;
    s->reqs.emplace(R.id, R);
    enqueue_sorted(s, R.id);
    *id = R.id;
    return BM_OK;
}

int bm_server_poll(bm_server_handle s, int64_t* done_ids, int32_t capacity, int32_t* n_done) {
    if (int rc = check_server(s)) return rc;
    if (capacity < 0 || (capacity > 0 && !done_ids)) return fail(BM_ERR_INVALID_ARG, "bad done-id buffer");
    double t = s->now();
    if (int rc = admit(s)) return rc;
    if (s->running.empty() && !s->queue.empty() && s->done.empty()) {
        // idle: wait for the next arrival, then admit it
        const double wait = s->reqs[s->queue.front()].info.arrival - s->now();
        const double t1 = s->now();
        if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        s->t_idle += s->now() - t1;
        if (int rc = admit(s)) return rc;
    }
    s->t_admit += s->now() - t;
    if (!s->running.empty()) {
        int32_t k = s->desc.max_chunk;
        for (int64_t id : s->running) k = std::min(k, s->reqs[id].left);
        t = s->now();
        if (int rc = bm_batch_step(s->batch, s->desc.dt, k)) return rc;
        int32_t code = 0;  // waits for the device (per-scene errors are read at retirement)
        if (int rc = bm_batch_scene_error(s->batch, s->reqs[s->running.front()].scene, &code)) return rc;
        s->t_step += s->now() - t;
        ++s->chunks;
        s->substeps += k;
        for (int64_t id : s->running) {
            s->reqs[id].left -= k;
            s->reqs[id].info.substeps_done += k;
        }
        t = s->now();
        if (int rc = retire(s)) return rc;
        s->t_retire += s->now() - t;
    }
    int32_t n = 0;
    while (n < capacity && !s->done.empty()) {
        done_ids[n++] = s->done.front();
        s->done.pop_front();
    }
    if (n_done) *n_done = n;
    return BM_OK;
}

int bm_server_info(bm_server_handle s, int64_t id, bm_request_info* info) {
    if (int rc = check_server(s)) return rc;
    if (!info) return fail(BM_ERR_INVALID_ARG, "null output");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(BM_ERR_INVALID_ARG, "no request %lld", (long long)id);
    *info = it->second.info;
    return BM_OK;
}

int bm_server_cancel(bm_server_handle s, int64_t id) {
    if (int rc = check_server(s)) return rc;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(BM_ERR_INVALID_ARG, "no request %lld", (long long)id);
    Request& R = it->second;
    if (R.info.state == BM_REQ_QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
    } else if (R.info.state == BM_REQ_RUNNING) {
        if (int rc = bm_batch_remove_scene(s->batch, R.scene)) return rc;
        s->in_use -= R.bytes;
        R.scene = -1;
        s->running.erase(std::find(s->running.begin(), s->running.end(), id));
    } else {
        return fail(BM_ERR_STATE, "request %lld has already finished", (long long)id);
    }
    R.info.state = BM_REQ_CANCELLED;
    R.info.finished = s->now();
    return BM_OK;
}

int bm_server_stats(bm_server_handle s, double* seconds, int64_t* counts) {
    if (int rc = check_server(s)) return rc;
    if (seconds) {
        seconds[0] = s->t_admit - s->t_idle;
        seconds[1] = s->t_step;
        seconds[2] = s->t_retire;
        seconds[3] = s->t_idle;
    }
    if (counts) {
        counts[0] = s->chunks;
        counts[1] = s->substeps;
    }
    return BM_OK;
}

int bm_server_load(bm_server_handle s, int32_t* queued, int32_t* running, int64_t* bytes_in_use, int64_t* budget) {
    if (int rc = check_server(s)) return rc;
    if (queued) *queued = (int32_t)s->queue.size();
    if (running) *running = (int32_t)s->running.size();
    if (bytes_in_use) *bytes_in_use = s->in_use;
    if (budget) *budget = s->budget;
    return BM_OK;
}

}  // extern "C"
