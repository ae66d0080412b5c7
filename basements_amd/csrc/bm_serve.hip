// bm_serve.hip -- the request loop over one batch: rollout requests are
// queued with their arrival time, admitted into the batch while its particle,
// block and scene room and the HBM budget allow, stepped together, and
// retired one by one when their own substep count is reached, while later
// requests keep being admitted (continuous batching).  The reference has no
// serving (its solver holds one scene and runs on one thread).
//
// Scheduling, one bm_server_poll:
//   1. wait for the chunk launched by the previous poll; meanwhile, requests
//      that have arrived but are still queued are transposed into pinned
//      memory and uploaded into the server's device staging area on a copy
//      stream, so that their admission is a device-to-device copy.  Then
//      retire: requests with no substeps left leave the batch -- their
//      particle state is first copied, in insertion order, into the staging
//      area -- and a request whose scene raised a device error is retired as
//      failed.
//   2. admit: walk the queue in arrival order; a request whose arrival has
//      come and that fits is added to the batch.  A later request may pass an
//      older one that does not fit, at most max_bypass times; then the queue
//      waits for the oldest.
//   3. step: the batch is launched for min(max_chunk, fewest remaining
//      substeps of the running requests) substeps, so that no request
//      overshoots; the poll does not wait for it.
//   4. while that chunk runs on the device: the retired requests' state
//      comes down to pinned host memory on the copy stream and is transposed
//      into their output buffers; then they are reported done.
// With nothing running or in flight, a poll sleeps until the next arrival.
// With overlap off (bm_server_set_overlap), every step waits for the device
// in the poll that launched it and admissions and retirements go through the
// batch's synchronous host paths.
//
// Host-only code over the batch entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_transpose.hpp"
#include "mpm_internal.hpp"
#include "trace.hpp"

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
    // input staged ahead of admission (overlap on): staging range, upload
    // event, mass quantum; or a non-finite input found while staging
    int64_t in_off = -1, in_len = 0;
    hipEvent_t in_ready = nullptr;
    int q_exp = 0;
    bool in_nonfinite = false;
};

// A retired request whose state is on its way to its output buffer.
struct Retiring {
    int64_t id;
    int64_t off, len, n, base;
    hipEvent_t landed;  // the download to pinned memory has finished
};

// First-fit ranges of the staging area, in 256-byte granules.
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;  // offset -> length
    static int64_t round(int64_t b) { return (b + 255) & ~(int64_t)255; }
    void init(int64_t size) {
        free_.clear();
        if (size > 0) free_[0] = size;
    }
    int64_t take(int64_t len) {
        len = round(len);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            const int64_t off = it->first, rest = it->second - len;
            free_.erase(it);
            if (rest > 0) free_[off + len] = rest;
            return off;
        }
        return -1;
    }
    void give(int64_t off, int64_t len) {
        len = round(len);
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Staging bytes of n particles: the component-major state, then the order.
int64_t soa_bytes(int64_t n) { return RangeAlloc::round((int64_t)sizeof(float) * PCOMP * n); }
int64_t stage_len(int64_t n) { return soa_bytes(n) + RangeAlloc::round((int64_t)sizeof(int) * n); }

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

    // overlap of staging with stepping
    bool overlap = true;
    hipStream_t cs = nullptr;          // copy stream (default priority)
    hipEvent_t chunk_done = nullptr;   // recorded after each launched chunk
    bool in_flight = false;            // a launched chunk not yet waited for
    // The chunk's kernels are enqueued by a launcher thread: the runtime
    // holds an enqueuing thread back while a few dozen dispatches are
    // outstanding, i.e. for most of a chunk, and the poll thread stages
    // meanwhile.  Guarded by mu: job_k (> 0: a chunk being enqueued), the
    // result of the last one, stop.
    int device = 0;
    std::thread launcher;
    std::mutex mu;
    std::condition_variable cv;
    int32_t job_k = 0;
    int job_rc = 0;
    std::string job_err;
    bool stop = false;
    char* dstage = nullptr;            // device staging area
    char* hstage = nullptr;            // pinned host staging, same offsets
    int64_t stage_bytes = 0;
    RangeAlloc ranges;
    std::vector<Retiring> retiring;

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

// Release a request's staged input (waits for its upload first).
static void drop_staged(bm_server_handle s, Request& R) {
    if (R.in_ready) {
        hipEventSynchronize(R.in_ready);
        hipEventDestroy(R.in_ready);
        R.in_ready = nullptr;
    }
    if (R.in_off >= 0) s->ranges.give(R.in_off, R.in_len);
    R.in_off = -1;
    R.in_len = 0;
}

// Admissions.
static int admit(bm_server_handle s) {
    BM_TRACE("server admit");
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
            if (R.in_nonfinite) {
                rc = fail(BM_ERR_NONFINITE, "particle state holds a non-finite value");
            } else if (R.in_off >= 0) {  // staged: the state is on the device already
                rc = batch_add_scene_staged(s->batch, &R.req.scene, R.req.n, R.q_exp,
                                            reinterpret_cast<const float*>(s->dstage + R.in_off), R.in_ready,
                                            R.has_body ? &R.body : nullptr, &R.scene);
                if (rc != BM_ERR_OOM) drop_staged(s, R);  // used (the batch stream has synced) or failed
            } else {
                rc = bm_batch_add_scene(s->batch, &R.req.scene, R.req.n, R.req.input,
                                        R.has_body ? &R.body : nullptr, &R.scene);
            }
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

// Retirements (and device failures).  With overlap on, a finished request's
// state goes to the staging area and starts down to pinned memory on the
// copy stream; finish_retirements writes its output later.
static int retire(bm_server_handle s) {
    BM_TRACE("server retire");
    for (size_t i = 0; i < s->running.size();) {
        Request& R = s->reqs[s->running[i]];
        int32_t code = BM_OK;
        if (int rc = bm_batch_scene_error(s->batch, R.scene, &code)) return rc;
        if (code == BM_OK && R.left > 0) {
            ++i;
            continue;
        }
        bool staged = false;
        Retiring T{R.id, -1, 0, 0, 0, nullptr};
        if (code == BM_OK) {
            if (R.req.output && s->overlap && s->stage_bytes > 0 && R.req.n > 0) {
                T.len = stage_len(R.req.n);
                T.off = s->ranges.take(T.len);
            }
            if (T.off >= 0) {
                if (int rc = batch_take_scene(s->batch, R.scene, reinterpret_cast<float*>(s->dstage + T.off),
                                              reinterpret_cast<int*>(s->dstage + T.off + soa_bytes(R.req.n)), &T.n,
                                              &T.base)) {
                    s->ranges.give(T.off, T.len);
                    return rc;
                }
                staged = true;
            } else if (R.req.output) {
                if (int rc = bm_batch_download_scene(s->batch, R.scene, R.req.output)) return rc;
            }
            if (R.has_body)
                if (int rc = bm_batch_get_body(s->batch, R.scene, &R.info.body, &R.info.reaction)) return rc;
        } else {
            R.info.state = BM_REQ_FAILED;
            R.info.status = code;
        }
        if (int rc = bm_batch_remove_scene(s->batch, R.scene)) return rc;  // syncs the batch stream
        R.info.finished = s->now();
        s->in_use -= R.bytes;
        R.scene = -1;
        s->running.erase(s->running.begin() + (std::ptrdiff_t)i);
        if (staged) {
            HIP_TRY(hipMemcpyAsync(s->hstage + T.off, s->dstage + T.off, (size_t)T.len, hipMemcpyDeviceToHost, s->cs));
            HIP_TRY(hipEventCreateWithFlags(&T.landed, hipEventDisableTiming));
            HIP_TRY(hipEventRecord(T.landed, s->cs));
            s->retiring.push_back(T);
        } else {
            if (code == BM_OK) R.info.state = BM_REQ_DONE;
            s->done.push_back(R.id);
        }
    }
    return BM_OK;
}

// Outputs of the staged retirements: wait for each download, transpose into
// the caller's buffers, report done.
static int finish_retirements(bm_server_handle s) {
    BM_TRACE("server finish retirements");
    for (Retiring& T : s->retiring) {
        Request& R = s->reqs[T.id];
        HIP_TRY(hipEventSynchronize(T.landed));
        hipEventDestroy(T.landed);
        const float* st = reinterpret_cast<const float*>(s->hstage + T.off);
        const int* order = reinterpret_cast<const int*>(s->hstage + T.off + soa_bytes(T.n));
        const bool ok = order_in_range(order, T.n, T.base);
        if (ok) particles_from_soa(st, T.n, nullptr, 0, R.req.output);
        s->ranges.give(T.off, T.len);
        if (ok) {
            R.info.state = BM_REQ_DONE;
        } else {
            R.info.state = BM_REQ_FAILED;
            R.info.status = fail(BM_ERR_HIP, "particle order corrupted");
        }
        s->done.push_back(T.id);
    }
    s->retiring.clear();
    return BM_OK;
}

// Stage the first queued request that has arrived and is not staged yet:
// transpose its input into pinned memory and upload it into the staging area
// on the copy stream.  *did: one was staged.  *next: arrival time of the
// first queued request still to come (infinity when none, or when the
// staging area is full).
static int stage_one(bm_server_handle s, bool* did, double* next) {
    BM_TRACE("server stage arrival");
    *did = false;
    *next = std::numeric_limits<double>::infinity();
    if (!s->overlap || s->stage_bytes == 0) return BM_OK;
    const double t = s->now();
    for (int64_t id : s->queue) {
        Request& R = s->reqs[id];
        if (R.info.arrival > t) {
            *next = R.info.arrival;
            return BM_OK;
        }
        if (R.in_off >= 0 || R.in_nonfinite) continue;
        const bm_particles* p = R.req.input;
        if (!p->x || !p->v || !p->F || !p->C || !p->m || !p->V0) continue;  // left to the batch's own checks
        const int64_t len = soa_bytes(R.req.n);
        const int64_t off = s->ranges.take(len);
        if (off < 0) return BM_OK;  // full: the rest are admitted the synchronous way
        float* st = reinterpret_cast<float*>(s->hstage + off);
        *did = true;
        if (!soa_from_particles(p->x, p->v, p->F, p->C, p->m, p->V0, R.req.n, st)) {
            s->ranges.give(off, len);
            R.in_nonfinite = true;
            return BM_OK;
        }
        R.q_exp = batch_quantum_exp(p->m, R.req.n);
        HIP_TRY(hipMemcpyAsync(s->dstage + off, st, (size_t)len, hipMemcpyHostToDevice, s->cs));
        HIP_TRY(hipEventCreateWithFlags(&R.in_ready, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(R.in_ready, s->cs));
        R.in_off = off;
        R.in_len = len;
        return BM_OK;
    }
    return BM_OK;
}

// The launcher thread: enqueue each chunk handed over (bm_batch_step, then
// the chunk's event on the batch stream) on the server's device.
static void launcher_main(bm_server_handle s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return s->stop || s->job_k > 0; });
        if (s->stop) return;
        const int32_t k = s->job_k;
        lk.unlock();
        int rc;
        {
            BM_TRACE("server enqueue chunk");
            rc = bm_batch_step(s->batch, s->desc.dt, k);
            if (rc == BM_OK && hipEventRecord(s->chunk_done, batch_stream(s->batch)) != hipSuccess)
                rc = fail(BM_ERR_HIP, "hipEventRecord failed after a chunk");
        }
        const std::string err = rc == BM_OK ? std::string() : std::string(bm_last_error());
        lk.lock();
        s->job_rc = rc;
        s->job_err = err;
        s->job_k = 0;
        s->cv.notify_all();
    }
}

// Launch a chunk of k substeps: handed to the launcher thread with overlap
// on, enqueued here otherwise.
static int launch_chunk(bm_server_handle s, int32_t k) {
    if (!s->overlap || !s->launcher.joinable()) {
        if (int rc = bm_batch_step(s->batch, s->desc.dt, k)) return rc;
        HIP_TRY(hipEventRecord(s->chunk_done, batch_stream(s->batch)));
        return BM_OK;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->job_k = k;
    s->job_rc = BM_OK;
    s->cv.notify_all();
    return BM_OK;
}

// True when no chunk is being enqueued; with block, waits for that.  A
// failure of the enqueuing thread is reported here, with its message.
static bool enqueued(bm_server_handle s, bool block, int* rc) {
    std::unique_lock<std::mutex> lk(s->mu);
    if (block) s->cv.wait(lk, [&] { return s->job_k == 0; });
    if (s->job_k != 0) return false;
    *rc = s->job_rc;
    if (s->job_rc != BM_OK) {
        *rc = fail(s->job_rc, "%s", s->job_err.c_str());
        s->job_rc = BM_OK;
    }
    return true;
}

// Wait for the chunk in flight, staging arrivals meanwhile: requests that
// have arrived are staged one at a time; with none waiting, the host checks
// the chunk every 50 us until the next arrival, or blocks on it when no
// arrival is due.
static int wait_chunk(bm_server_handle s) {
    BM_TRACE("server wait chunk");
    if (!s->in_flight) return BM_OK;
    const double t = s->now();
    for (;;) {
        int erc = BM_OK;
        const bool ready = enqueued(s, false, &erc);
        if (erc != BM_OK) {
            s->in_flight = false;
            return erc;
        }
        if (ready) {
            const hipError_t q = hipEventQuery(s->chunk_done);
            if (q == hipSuccess) break;
            if (q != hipErrorNotReady) HIP_TRY(q);
        }
        bool did = false;
        double next = 0.0;
        if (int rc = stage_one(s, &did, &next)) return rc;
        if (did) continue;
        if (!std::isfinite(next)) break;
        std::this_thread::sleep_for(std::chrono::duration<double>(std::min(std::max(next - s->now(), 0.0), 50e-6)));
    }
    int erc = BM_OK;
    enqueued(s, true, &erc);
    s->in_flight = false;
    if (erc != BM_OK) return erc;
    HIP_TRY(hipEventSynchronize(s->chunk_done));
    s->t_step += s->now() - t;
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
    // staging: room for the whole particle capacity once (state + order), in
    // device memory and in pinned host memory; the copy stream keeps the
    // default priority.  Without it, the synchronous paths serve.
    if (hipStreamCreateWithFlags(&s->cs, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->chunk_done, hipEventDisableTiming) != hipSuccess) {
        bm_server_destroy(s);
        return fail(BM_ERR_HIP, "server stream or event");
    }
    const int64_t want = stage_len(desc->batch.max_particles) + ((int64_t)1 << 20);
    void* d = nullptr;
    void* h = nullptr;
    if (hipMalloc(&d, (size_t)want) == hipSuccess && hipHostMalloc(&h, (size_t)want, hipHostMallocDefault) == hipSuccess) {
        s->dstage = static_cast<char*>(d);
        s->hstage = static_cast<char*>(h);
        s->stage_bytes = want;
        s->ranges.init(want);
    } else {
        (void)hipGetLastError();
        if (d) hipFree(d);
        s->overlap = false;
    }
    (void)hipGetDevice(&s->device);
    s->launcher = std::thread(launcher_main, s);
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return BM_OK;
}

int bm_server_destroy(bm_server_handle s) {
    if (!s) return BM_OK;
    if (s->launcher.joinable()) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->job_k == 0; });
            s->stop = true;
            s->cv.notify_all();
        }
        s->launcher.join();
    }
    if (s->cs) hipStreamSynchronize(s->cs);
    for (auto& kv : s->reqs)
        if (kv.second.in_ready) hipEventDestroy(kv.second.in_ready);
    for (Retiring& T : s->retiring) hipEventDestroy(T.landed);
    bm_batch_destroy(s->batch);
    if (s->chunk_done) hipEventDestroy(s->chunk_done);
    if (s->cs) hipStreamDestroy(s->cs);
    if (s->dstage) hipFree(s->dstage);
    if (s->hstage) hipHostFree(s->hstage);
    delete s;
    return BM_OK;
}

int bm_server_set_overlap(bm_server_handle s, int32_t on) {
    if (int rc = check_server(s)) return rc;
    if (on && s->stage_bytes == 0) return fail(BM_ERR_OOM, "no staging area (its allocation failed)");
    s->overlap = on != 0;
    if (int rc = wait_chunk(s)) return rc;  // the next poll starts from a settled device either way
    if (!on)
        for (int64_t id : s->queue) drop_staged(s, s->reqs[id]);
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
    if (int rc = wait_chunk(s)) return rc;  // the batch is the caller's until the next poll
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
    BM_TRACE("bm_server_poll");
    if (int rc = check_server(s)) return rc;
    if (capacity < 0 || (capacity > 0 && !done_ids)) return fail(BM_ERR_INVALID_ARG, "bad done-id buffer");
    // 1. the chunk in flight ends; finished requests leave the batch
    if (int rc = wait_chunk(s)) return rc;
    if (!s->running.empty()) {
        const double t = s->now();
        if (int rc = retire(s)) return rc;
        s->t_retire += s->now() - t;
    }
    // 2. admissions
    double t = s->now();
    if (int rc = admit(s)) return rc;
    if (s->running.empty() && !s->queue.empty() && s->done.empty() && s->retiring.empty()) {
        // idle: wait for the next arrival, then admit it
        const double wait = s->reqs[s->queue.front()].info.arrival - s->now();
        const double t1 = s->now();
        if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        s->t_idle += s->now() - t1;
        if (int rc = admit(s)) return rc;
    }
    s->t_admit += s->now() - t;
    // 3. the next chunk
    if (!s->running.empty()) {
        int32_t k = s->desc.max_chunk;
        for (int64_t id : s->running) k = std::min(k, s->reqs[id].left);
        t = s->now();
        if (int rc = launch_chunk(s, k)) return rc;
        s->in_flight = true;
        ++s->chunks;
        s->substeps += k;
        for (int64_t id : s->running) {
            s->reqs[id].left -= k;
            s->reqs[id].info.substeps_done += k;
        }
        s->t_step += s->now() - t;
    }
    // 4. while it runs: outputs of the retired requests (arrivals are staged
    // while the next poll waits for the chunk)
    t = s->now();
    if (int rc = finish_retirements(s)) return rc;
    (s->in_flight ? s->t_step : s->t_retire) += s->now() - t;
    if (!s->overlap && s->in_flight) {  // synchronous form: the chunk and its retirements end in this poll
        if (int rc = wait_chunk(s)) return rc;
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
    if (int rc = wait_chunk(s)) return rc;  // no chunk is being enqueued while the batch changes
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(BM_ERR_INVALID_ARG, "no request %lld", (long long)id);
    Request& R = it->second;
    if (R.info.state == BM_REQ_QUEUED) {
        drop_staged(s, R);
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
    } else if (R.info.state == BM_REQ_RUNNING && R.scene < 0) {
        return fail(BM_ERR_STATE, "request %lld has already finished (its output is being written)", (long long)id);
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
