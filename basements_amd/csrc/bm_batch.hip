// bm_batch.hip -- many scenes in shared device arrays (bm_batch_* of
// include/bm_mpm.h).
//
// Layout: scene s owns particles [p0_s, p0_s + n_s) of one particle arena and
// grid blocks [b0_s, b0_s + nb_s) of one grid arena (Domain::b0), packed in
// scene order.  A scene table (SceneDev), a block -> scene map and a chunk
// list (256 particles of one scene per workgroup) let every kernel of the fast
// substep run once for all scenes: the particle pass and the re-binning over
// the chunks, the cell deposit over the work units (each tagged with its
// scene), the grid update over the touched blocks, the couplers over
// (workgroup, scene).  Re-binning sorts by global block id, which keeps each
// scene's particles inside its range.  Nothing here syncs the host inside a
// substep, so the substeps can be captured in a HIP graph and replayed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_staging.hpp"
#include "mpm_internal.hpp"
#include "trace.hpp"

using namespace bm;

namespace {

struct SceneHost {
    int32_t id;
    bm_scene_desc desc;
    Domain d;
    int64_t p0, n, nb;
    int q_exp;
    bool has_body;
    int err = 0;  // device error bits raised by this scene (sticky)
};

__global__ void k_fill_scene(int* block_scene, int64_t b0, int64_t nb, int s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) block_scene[b0 + i] = s;
}
__global__ void k_iota_from(int* p, int64_t n, int64_t first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)(first + i);
}
__global__ void k_add_i32(int* p, int64_t n, int delta) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] += delta;
}
inline int grid256(int64_t n) { return (int)((n + 255) / 256); }

int quantum_exp_of(const float* m, int64_t n) {
    float mx = 0.0f;
    for (int64_t i = 0; i < n; ++i) mx = std::max(mx, m[i]);
    if (!(mx > 0.0f) || !std::isfinite(mx)) return 0;
    return std::ilogb(mx) - 22;  // as the single-scene handle: <= 2^23 - 1 quanta per particle
}

}  // namespace

struct bm_batch {
    bm_batch_desc desc{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int64_t capP = 0, capB = 0;
    int capS = 0;

    float* pbuf = nullptr;
    float* pbuf2 = nullptr;
    int* pid = nullptr;
    int* pid2 = nullptr;
    float* coef = nullptr;
    int* redo = nullptr;
    int* n_redo = nullptr;
    GridBuf grid[2] = {};
    uint8_t* bever = nullptr;
    uint8_t* bever_tmp = nullptr;
    int* d_err = nullptr;
    unsigned long long* d_counters = nullptr;
    BinScratch bs{};
    int* block_scene = nullptr;
    int4* chunks = nullptr;
    int64_t cap_chunks = 0;
    SceneDev* d_sc = nullptr;
    bm_body* d_bodies = nullptr;
    ColliderDev* d_cols = nullptr;
    int* d_has_body = nullptr;
    int* d_methods = nullptr;
    double* d_partials = nullptr;
    double* d_out7 = nullptr;

    std::vector<SceneHost> scenes;
    int32_t next_id = 1;
    int64_t P = 0, B = 0;
    int n_chunks = 0;

    int cur = 0;
    bool prepared = false, binned = false;
    float dt_prepared = 0.0f;
    int since_bin = 0;
    int unit_p = 640, sort_every = 16, deposit_groups = 4096;
    PinnedStage stage;  // host side of uploads and downloads

    // Captured substeps.  A replay starts and ends in the state it was
    // captured in (which particle buffer is current, which grid, re-binning
    // due), so one graph is kept per (particle buffer, grid) combination, and
    // per whether its last substep ends a call.
    bool use_graph = false;
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        const float* pbuf = nullptr;
        int cur = 0;
        bool last = false;
    } graphs[8];
    bool graph_valid = false;  // the slots match the current layout, dt and tuning
    float graph_dt = 0.0f;
    void drop_graphs() {
        for (GraphSlot& g : graphs) {
            if (g.exec) hipGraphExecDestroy(g.exec);
            g = GraphSlot{};
        }
    }

    SceneView view() const { return SceneView{d_sc, block_scene, chunks, (int)scenes.size()}; }
    PView pview() const { return PView{pbuf, capP, P}; }
    int n_units_bound() const { return (int)(B + (P + unit_p - 1) / unit_p + 1); }
    int find(int32_t id) const {
        for (size_t s = 0; s < scenes.size(); ++s)
            if (scenes[s].id == id) return (int)s;
        return -1;
    }
    void invalidate() {
        prepared = false;
        binned = false;
        graph_valid = false;
    }
};

static int check_batch(bm_batch_handle b) { return b ? BM_OK : fail(BM_ERR_INVALID_ARG, "null batch"); }

static Phys phys_of(const bm_scene_desc& sd) {
    Phys p;
    const float E = sd.young_modulus, nu = sd.poisson_ratio;
    p.mu = E / (2.0f * (1.0f + nu));
    p.lambda = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    p.alpha = sd.plastic_alpha;
    p.kohesion = (3.0f * p.lambda + 2.0f * p.mu) / (2.0f * p.mu);
    p.gx = sd.gravity[0];
    p.gy = sd.gravity[1];
    p.gz = sd.gravity[2];
    p.mu_floor = sd.floor_friction;
    p.walls = sd.walls;
    return p;
}

// Scene table, body flags, chunk list and block -> scene map from the host
// description of the scenes (after every change of the scene set).
static int upload_tables(bm_batch_handle b) {
    const int ns = (int)b->scenes.size();
    std::vector<SceneDev> sc((size_t)std::max(ns, 1));
    std::vector<int> has((size_t)std::max(ns, 1), 0), meth((size_t)std::max(ns, 1), 0);
    std::vector<int4> ch;
    for (int s = 0; s < ns; ++s) {
        const SceneHost& S = b->scenes[s];
        SceneDev& D = sc[s];
        D.d = S.d;
        D.ph = phys_of(S.desc);
        D.qscale = std::ldexp(1.0, S.q_exp);
        D.qscale_inv = std::ldexp(1.0, -S.q_exp);
        D.p0 = S.p0;
        D.n = S.n;
        has[s] = S.has_body ? 1 : 0;
        meth[s] = S.desc.coupler;
        for (int64_t off = 0; off < S.n; off += 256)
            ch.push_back(make_int4(s, (int)(S.p0 + off), (int)std::min<int64_t>(256, S.n - off), 0));
    }
    if ((int64_t)ch.size() > b->cap_chunks) return fail(BM_ERR_OOM, "chunk list overflow");
    b->n_chunks = (int)ch.size();
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (ns > 0) {
        HIP_TRY(hipMemcpyAsync(b->d_sc, sc.data(), sizeof(SceneDev) * ns, hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_has_body, has.data(), sizeof(int) * ns, hipMemcpyHostToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_methods, meth.data(), sizeof(int) * ns, hipMemcpyHostToDevice, b->stream));
    }
    if (!ch.empty())
        HIP_TRY(hipMemcpyAsync(b->chunks, ch.data(), sizeof(int4) * ch.size(), hipMemcpyHostToDevice, b->stream));
    for (int s = 0; s < ns; ++s) {
        const SceneHost& S = b->scenes[s];
        if (S.nb > 0)
            hipLaunchKernelGGL(k_fill_scene, dim3(grid256(S.nb)), dim3(256), 0, b->stream, b->block_scene, S.d.b0,
                               S.nb, s);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(b->stream));
    return BM_OK;
}

// Move the device error words into the scenes' sticky flags (the per-scene
// words are indexed by the current table order, so this runs before any
// change of that order).
static int batch_harvest_err(bm_batch_handle b) {
    HIP_TRY(hipStreamSynchronize(b->stream));
    int err = 0;
    HIP_TRY(hipMemcpy(&err, b->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        const int ns = (int)b->scenes.size();
        std::vector<int> w((size_t)ns + 1);
        HIP_TRY(hipMemcpy(w.data(), b->d_err, sizeof(int) * (ns + 1), hipMemcpyDeviceToHost));
        for (int s = 0; s < ns; ++s) b->scenes[s].err |= w[(size_t)s + 1];
        HIP_TRY(hipMemsetAsync(b->d_err, 0, sizeof(int) * (b->capS + 1), b->stream));
    }
    return BM_OK;
}

static int scene_err_code(int bits) {
    if (bits & 1) return BM_ERR_OUT_OF_DOMAIN;
    if (bits & 8) return BM_ERR_NONFINITE;
    return bits ? BM_ERR_HIP : BM_OK;
}

static int scene_fail(const SceneHost& S) {
    if (S.err & 1) return fail(BM_ERR_OUT_OF_DOMAIN, "scene %d: a particle stencil left the scene's grid box", S.id);
    if (S.err & 8) return fail(BM_ERR_NONFINITE, "scene %d: a particle's position or velocity became non-finite", S.id);
    return fail(BM_ERR_HIP, "scene %d: device error flag %d", S.id, S.err);
}

// Harvest, then fail if scene s (or, with s < 0, any scene) has raised an error.
static int batch_check_err(bm_batch_handle b, int s = -1) {
    if (int rc = batch_harvest_err(b)) return rc;
    for (int t = 0; t < (int)b->scenes.size(); ++t)
        if ((s < 0 || t == s) && b->scenes[t].err) return scene_fail(b->scenes[t]);
    return BM_OK;
}

// ------------------------------------------------------------------ substep
// with_vc = false when a G2P follows, which rewrites velocity and C.
static void batch_rebin(bm_batch_handle b, bool with_vc) {
    launch_rebin(b->pview(), b->pbuf2, b->pid, b->pid2, Domain{}, (int)b->B, b->unit_p, b->bs, b->stream, nullptr,
                 b->view(), b->n_chunks, with_vc);
    std::swap(b->pbuf, b->pbuf2);
    std::swap(b->pid, b->pid2);
    b->binned = true;
    b->since_bin = 0;
}

// clear = false: the grid update already cleared gout (and the redo count).
static void batch_pass(bm_batch_handle b, bool do_g2p, const GridBuf& gin, const GridBuf& gout, float dt,
                       bool store_vc, bool clear = true) {
    SplitBatch sb;
    sb.sv = b->view();
    sb.n_chunks = b->n_chunks;
    sb.n_units_dev = b->bs.n_units;
    sb.n_units_bound = b->n_units_bound();
    sb.deposit_groups = b->deposit_groups;
    // the clear before every pass also zeroes the redo counter; the pass
    // empties the output grid's block list
    if (clear) launch_clear(gout, (int)std::min<int64_t>(b->B, 1 << 30), b->stream, b->n_redo, true);
    launch_split(do_g2p, b->pview(), b->bs.units, 0, gin, gout, b->bever, Domain{}, Phys{}, dt, dt, 0.0, b->d_err,
                 b->d_counters, b->redo, b->n_redo, b->coef, b->stream, nullptr, nullptr, false, sb, store_vc);
}

// One coupled substep of every scene.  Host state only changes the way the
// device state does, so the same calls record a graph.  last: the last
// substep of a bm_batch_step call, whose G2P writes the particle velocity and
// C (the earlier ones leave them to it; only downloads read them).
static void batch_substep(bm_batch_handle b, float dt, bool last) {
    if (!b->binned) batch_rebin(b, true);
    if (!b->prepared || dt != b->dt_prepared) {
        batch_pass(b, false, b->grid[b->cur], b->grid[b->cur], dt, true);
        b->prepared = true;
        b->dt_prepared = dt;
    }
    // the update also clears the grid the G2P below deposits into
    launch_update(b->grid[b->cur], (int)b->B, Domain{}, Phys{}, 0.0, dt, b->stream, b->d_bodies, b->d_cols, LLONG_MAX,
                  b->d_err, b->view(), b->d_has_body, &b->grid[1 - b->cur], b->n_redo);
    launch_couple_batch(b->d_methods, b->d_has_body, b->d_bodies, b->d_cols, b->grid[b->cur], b->bever, b->view(), dt,
                        b->d_partials, b->d_out7, b->stream);
    if (b->since_bin >= b->sort_every) batch_rebin(b, false);
    const int nxt = 1 - b->cur;
    batch_pass(b, true, b->grid[b->cur], b->grid[nxt], dt, last, false);
    b->cur = nxt;
    b->since_bin += 1;
}

// The graph of 2 x sort_every substeps from the current state (re-binning
// due): two re-binnings and an even number of grid swaps bring buffers and
// grids back to where they started, so the graph can be replayed.  last: its
// final substep ends a bm_batch_step call (writes velocity and C).
static int graph_for_state(bm_batch_handle b, float dt, bool last, hipGraphExec_t* out) {
    if (!b->graph_valid || b->graph_dt != dt) {
        b->drop_graphs();
        b->graph_valid = true;
        b->graph_dt = dt;
    }
    for (auto& g : b->graphs)
        if (g.exec && g.pbuf == b->pbuf && g.cur == b->cur && g.last == last) {
            *out = g.exec;
            return BM_OK;
        }
    bm_batch::GraphSlot* slot = nullptr;
    for (auto& g : b->graphs)
        if (!g.exec) slot = &g;
    if (!slot) return fail(BM_ERR_HIP, "no free graph slot");
    const int G = 2 * b->sort_every;
    hipGraph_t g = nullptr;
    slot->pbuf = b->pbuf;
    slot->cur = b->cur;
    slot->last = last;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < G; ++i) batch_substep(b, dt, last && i == G - 1);
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) return fail(BM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    if (b->pbuf != slot->pbuf || b->cur != slot->cur) {
        hipGraphDestroy(g);
        return fail(BM_ERR_HIP, "captured substeps do not return to their starting state");
    }
    const hipError_t e2 = hipGraphInstantiate(&slot->exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e2 != hipSuccess) return fail(BM_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e2));
    *out = slot->exec;
    return BM_OK;
}

// Device bytes that bm_batch_create allocates per particle, per grid block
// and per scene of capacity (the unit list and chunk table rounded up).
static void batch_unit_bytes(int64_t& per_particle, int64_t& per_block, int64_t& per_scene) {
    per_particle = (int64_t)(sizeof(float) * PCOMP * 2      // particle state, double-buffered for re-binning
                             + sizeof(int) * 2              // insertion ids, double-buffered
                             + sizeof(float) * 16           // deposit records
                             + sizeof(int) * 3              // FP64 redo list, bin key, bin slot
                             + 2);                          // deposit units and chunk table
    per_block = (int64_t)(2 + sizeof(int)                   // touched flags, block -> scene
                          + 2 * (sizeof(long long) * 64 + sizeof(float4) * 64 + sizeof(float) * 3 * 64 +
                                 sizeof(uint32_t) + sizeof(int32_t))
                          + sizeof(int) * 4 + sizeof(int4)  // bin counts and starts, units
                          + 8);                             // scan scratch
    per_scene = (int64_t)(sizeof(SceneDev) + sizeof(bm_body) + sizeof(ColliderDev) + sizeof(int) * 3 +
                          sizeof(double) * (7 * COUPLE_WGS + 8) + sizeof(int4));
}

// Admission, first half: validation and placement at the arena ends
// (nothing is written yet).
static int place_scene(bm_batch_handle b, const bm_scene_desc* sd, int64_t n, int q_exp, const bm_body* body,
                       SceneHost& S) {
    if (!(sd->dx > 0.0f) || !std::isfinite(sd->dx)) return fail(BM_ERR_INVALID_ARG, "dx must be > 0");
    for (int a = 0; a < 3; ++a)
        if (sd->grid_dims[a] <= 0) return fail(BM_ERR_INVALID_ARG, "grid_dims must be > 0");
    if (body && (sd->coupler < BM_COUPLER_M0_DIRICHLET || sd->coupler > BM_COUPLER_M45_BSPLINE))
        return fail(BM_ERR_INVALID_ARG, "unknown coupler %d", sd->coupler);
    if (body && body->shape != BM_SHAPE_BOX && body->shape != BM_SHAPE_CYLINDER)
        return fail(BM_ERR_INVALID_ARG, "unknown shape %d", body->shape);
    S = SceneHost{};
    S.desc = *sd;
    Domain& d = S.d;
    d.ox = sd->grid_origin[0];
    d.oy = sd->grid_origin[1];
    d.oz = sd->grid_origin[2];
    d.nx = (sd->grid_dims[0] + 3) & ~3;
    d.ny = (sd->grid_dims[1] + 3) & ~3;
    d.nz = (sd->grid_dims[2] + 3) & ~3;
    d.nbx = d.nx / 4;
    d.nby = d.ny / 4;
    d.nbz = d.nz / 4;
    d.dx = sd->dx;
    d.inv_dx = 1.0f / sd->dx;
    d.b0 = b->B;
    S.nb = (int64_t)d.nbx * d.nby * d.nbz;
    if ((int)b->scenes.size() >= b->capS) return fail(BM_ERR_OOM, "batch scene capacity %d reached", b->capS);
    if (b->P + n > b->capP) return fail(BM_ERR_OOM, "batch particle capacity %lld exceeded", (long long)b->capP);
    if (b->B + S.nb > b->capB) return fail(BM_ERR_OOM, "batch block capacity %lld exceeded", (long long)b->capB);
    S.p0 = b->P;
    S.n = n;
    S.q_exp = q_exp;
    S.has_body = body != nullptr;
    return BM_OK;
}

// Admission, second half (the particles are on their way into the arena on
// the batch stream): body, reaction slot, tables.
static int commit_scene(bm_batch_handle b, SceneHost& S, const bm_body* body, int32_t* scene_id) {
    S.id = b->next_id++;
    const int s = (int)b->scenes.size();
    if (body) HIP_TRY(hipMemcpyAsync(b->d_bodies + s, body, sizeof(bm_body), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemsetAsync(b->d_out7 + (int64_t)s * 8, 0, sizeof(double) * 8, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));  // the caller's buffers and the staging are free again
    b->scenes.push_back(S);
    b->P += S.n;
    b->B += S.nb;
    b->invalidate();
    if (int rc = upload_tables(b)) return rc;
    *scene_id = S.id;
    return BM_OK;
}

namespace bm {

hipStream_t batch_stream(bm_batch_handle b) { return b->stream; }

int batch_quantum_exp(const float* m, int64_t n) { return n > 0 ? quantum_exp_of(m, n) : 0; }

int batch_add_scene_staged(bm_batch_handle b, const bm_scene_desc* sd, int64_t n, int q_exp, const float* dev_soa,
                           hipEvent_t ready, const bm_body* body, int32_t* scene_id) {
    BM_TRACE("batch_add_scene_staged");
    if (!b || !sd || !scene_id || n < 0 || (n > 0 && !dev_soa)) return fail(BM_ERR_INVALID_ARG, "null argument");
    SceneHost S{};
    if (int rc = place_scene(b, sd, n, q_exp, body, S)) return rc;
    if (n > 0) {
        if (ready) HIP_TRY(hipStreamWaitEvent(b->stream, ready, 0));
        HIP_TRY(hipMemcpy2DAsync(b->pbuf + S.p0, sizeof(float) * b->capP, dev_soa, sizeof(float) * n,
                                 sizeof(float) * n, PCOMP, hipMemcpyDeviceToDevice, b->stream));
        hipLaunchKernelGGL(k_iota_from, dim3(grid256(n)), dim3(256), 0, b->stream, b->pid + S.p0, n, S.p0);
        HIP_TRY(hipGetLastError());
    }
    return commit_scene(b, S, body, scene_id);
}

int batch_take_scene(bm_batch_handle b, int32_t id, float* dev_soa, int* dev_order, int64_t* n_out,
                     int64_t* base_out) {
    BM_TRACE("batch_take_scene");
    if (!b || !n_out || !base_out) return fail(BM_ERR_INVALID_ARG, "null argument");
    const int s = b->find(id);
    if (s < 0) return fail(BM_ERR_INVALID_ARG, "no scene %d", id);
    if (int rc = batch_check_err(b, s)) return rc;
    const SceneHost& S = b->scenes[s];
    *n_out = S.n;
    *base_out = S.p0;
    if (S.n == 0) return BM_OK;
    if (!dev_soa || !dev_order) return fail(BM_ERR_INVALID_ARG, "null staging");
    // as bm_batch_download_scene, with the copies staying on the device
    launch_unpermute(b->pview(), b->pbuf2, b->pid, S.p0, S.n, b->stream);
    HIP_TRY(hipMemcpyAsync(dev_order, b->pid + S.p0, sizeof(int) * S.n, hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipMemcpy2DAsync(dev_soa, sizeof(float) * S.n, b->pbuf2 + S.p0, sizeof(float) * b->capP,
                             sizeof(float) * S.n, PCOMP, hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipGetLastError());
    return BM_OK;
}

}  // namespace bm

extern "C" {

int bm_batch_footprint(int64_t particles, int64_t blocks, int32_t scenes, int64_t* bytes) {
    if (!bytes || particles < 0 || blocks < 0 || scenes < 0) return fail(BM_ERR_INVALID_ARG, "bad footprint query");
    int64_t pp, pb, ps;
    batch_unit_bytes(pp, pb, ps);
    *bytes = pp * particles + pb * blocks + ps * scenes + 256;
    return BM_OK;
}

int bm_batch_create(const bm_batch_desc* desc, bm_batch_handle* out) {
    if (!desc || !out) return fail(BM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (desc->max_particles <= 0 || desc->max_blocks <= 0 || desc->max_scenes <= 0)
        return fail(BM_ERR_INVALID_ARG, "batch capacities must be > 0");
    if (desc->max_particles > (int64_t)INT32_MAX - 1024 || desc->max_blocks > (int64_t)INT32_MAX / 64)
        return fail(BM_ERR_INVALID_ARG, "batch capacity too large");
    auto* b = new bm_batch();
    b->desc = *desc;
    b->capP = (desc->max_particles + 63) / 64 * 64;
    b->capB = desc->max_blocks;
    b->capS = desc->max_scenes;
    b->cap_chunks = (b->capP + 255) / 256 + b->capS;
    auto bail = [&](int code) {
        bm_batch_destroy(b);
        return code;
    };
    if (desc->stream) {
        b->stream = (hipStream_t)desc->stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(BM_ERR_HIP, "stream creation failed"));
        b->own_stream = true;
    }
    const int64_t cp = b->capP, cb = b->capB;
    auto A = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)) == hipSuccess; };
    bool ok = A((void**)&b->pbuf, sizeof(float) * PCOMP * cp) && A((void**)&b->pbuf2, sizeof(float) * PCOMP * cp) &&
              A((void**)&b->pid, sizeof(int) * cp) && A((void**)&b->pid2, sizeof(int) * cp) &&
              A((void**)&b->coef, sizeof(float) * 16 * cp) && A((void**)&b->redo, sizeof(int) * cp) &&
              A((void**)&b->n_redo, sizeof(int)) && A((void**)&b->bever, cb) && A((void**)&b->bever_tmp, cb) &&
              A((void**)&b->d_err, sizeof(int) * (b->capS + 1)) && A((void**)&b->d_counters, 2 * sizeof(unsigned long long)) &&
              A((void**)&b->block_scene, sizeof(int) * cb) && A((void**)&b->chunks, sizeof(int4) * b->cap_chunks) &&
              A((void**)&b->d_sc, sizeof(SceneDev) * b->capS) && A((void**)&b->d_bodies, sizeof(bm_body) * b->capS) &&
              A((void**)&b->d_cols, sizeof(ColliderDev) * b->capS) && A((void**)&b->d_has_body, sizeof(int) * b->capS) &&
              A((void**)&b->d_methods, sizeof(int) * b->capS) &&
              A((void**)&b->d_partials, sizeof(double) * 7 * COUPLE_WGS * b->capS) &&
              A((void**)&b->d_out7, sizeof(double) * 8 * b->capS);
    for (int k = 0; ok && k < 2; ++k) {
        GridBuf& g = b->grid[k];
        ok = A((void**)&g.mq, sizeof(long long) * 64 * cb) && A((void**)&g.mv, sizeof(float4) * 64 * cb) &&
             A((void**)&g.mom, sizeof(float) * 3 * 64 * cb) && A((void**)&g.bflag, sizeof(uint32_t) * cb) &&
             A((void**)&g.blist, sizeof(int32_t) * cb) && A((void**)&g.bcount, sizeof(int32_t));
        g.mstride = 64 * cb;
        if (ok) {
            ok = hipMemset(g.mq, 0, sizeof(long long) * 64 * cb) == hipSuccess &&
                 hipMemset(g.mv, 0, sizeof(float4) * 64 * cb) == hipSuccess &&
                 hipMemset(g.mom, 0, sizeof(float) * 3 * 64 * cb) == hipSuccess &&
                 hipMemset(g.bflag, 0, sizeof(uint32_t) * cb) == hipSuccess &&
                 hipMemset(g.bcount, 0, sizeof(int32_t)) == hipSuccess;
        }
    }
    BinScratch& bs = b->bs;
    ok = ok && A((void**)&bs.bcnt, sizeof(int) * cb) && A((void**)&bs.bstart, sizeof(int) * cb) &&
         A((void**)&bs.ucnt, sizeof(int) * cb) && A((void**)&bs.ustart, sizeof(int) * cb) &&
         A((void**)&bs.key, sizeof(int) * cp) && A((void**)&bs.slot, sizeof(int) * cp) &&
         A((void**)&bs.units, sizeof(int4) * (cb + cp / 16 + 2)) && A((void**)&bs.n_units, sizeof(int));
    if (ok) {
        bs.temp_bytes = scan_temp_bytes((int)cb);
        ok = A(&bs.temp, bs.temp_bytes);
    }
    if (!ok) return bail(fail(BM_ERR_OOM, "batch allocation"));
    if (hipMemset(b->bever, 0, cb) != hipSuccess || hipMemset(b->d_err, 0, sizeof(int) * (b->capS + 1)) != hipSuccess ||
        hipMemset(b->d_counters, 0, 2 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(b->d_out7, 0, sizeof(double) * 8 * b->capS) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return bail(fail(BM_ERR_HIP, "batch initialisation"));
    *out = b;
    return BM_OK;
}

int bm_batch_destroy(bm_batch_handle b) {
    if (!b) return BM_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    b->drop_graphs();
    for (void* p : {(void*)b->pbuf, (void*)b->pbuf2, (void*)b->pid, (void*)b->pid2, (void*)b->coef, (void*)b->redo,
                    (void*)b->n_redo, (void*)b->bever, (void*)b->bever_tmp, (void*)b->d_err, (void*)b->d_counters,
                    (void*)b->block_scene, (void*)b->chunks, (void*)b->d_sc, (void*)b->d_bodies, (void*)b->d_cols,
                    (void*)b->d_has_body, (void*)b->d_methods, (void*)b->d_partials, (void*)b->d_out7,
                    (void*)b->bs.bcnt, (void*)b->bs.bstart, (void*)b->bs.ucnt, (void*)b->bs.ustart, (void*)b->bs.key,
                    (void*)b->bs.slot, (void*)b->bs.units, (void*)b->bs.n_units, b->bs.temp})
        hipFree(p);
    for (GridBuf& g : b->grid) {
        hipFree(g.mq);
        hipFree(g.mv);
        hipFree(g.mom);
        hipFree(g.bflag);
        hipFree(g.blist);
        hipFree(g.bcount);
    }
    if (b->own_stream && b->stream) hipStreamDestroy(b->stream);
    delete b;
    return BM_OK;
}

int bm_batch_synchronize(bm_batch_handle b) {
    if (int rc = check_batch(b)) return rc;
    return batch_check_err(b);
}

int bm_batch_add_scene(bm_batch_handle b, const bm_scene_desc* sd, int64_t n, const bm_particles* p,
                       const bm_body* body, int32_t* scene_id) {
    BM_TRACE("bm_batch_add_scene");
    if (int rc = check_batch(b)) return rc;
    if (!sd || !scene_id || n < 0 || (n > 0 && !p)) return fail(BM_ERR_INVALID_ARG, "null argument");
    if (n > 0 && (!p->x || !p->v || !p->F || !p->C || !p->m || !p->V0))
        return fail(BM_ERR_INVALID_ARG, "null particle field");
    SceneHost S{};
    if (int rc = place_scene(b, sd, n, n > 0 ? quantum_exp_of(p->m, n) : 0, body, S)) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (n > 0) {
        // component-major pinned staging, then one pitched copy into the arena
        if (!b->stage.ensure(sizeof(float) * PCOMP * n)) return fail(BM_ERR_OOM, "pinned staging");
        float* st = b->stage.f();
        if (!soa_from_particles(p->x, p->v, p->F, p->C, p->m, p->V0, n, st))
            return fail(BM_ERR_NONFINITE, "particle state holds a non-finite value");
        HIP_TRY(hipMemcpy2DAsync(b->pbuf + S.p0, sizeof(float) * b->capP, st, sizeof(float) * n, sizeof(float) * n,
                                 PCOMP, hipMemcpyHostToDevice, b->stream));
        hipLaunchKernelGGL(k_iota_from, dim3(grid256(n)), dim3(256), 0, b->stream, b->pid + S.p0, n, S.p0);
        HIP_TRY(hipGetLastError());
    }
    return commit_scene(b, S, body, scene_id);
}

int bm_batch_remove_scene(bm_batch_handle b, int32_t id) {
    BM_TRACE("bm_batch_remove_scene");
    if (int rc = check_batch(b)) return rc;
    const int s = b->find(id);
    if (s < 0) return fail(BM_ERR_INVALID_ARG, "no scene %d", id);
    if (int rc = batch_harvest_err(b)) return rc;  // before the table order changes
    const SceneHost R = b->scenes[s];
    const int64_t tailP = b->P - (R.p0 + R.n), tailB = b->B - (R.d.b0 + R.nb);
    const size_t pitch = sizeof(float) * b->capP;
    // particles: the prefix and the shifted tail into the second buffer, then swap
    if (R.p0 > 0)
        HIP_TRY(hipMemcpy2DAsync(b->pbuf2, pitch, b->pbuf, pitch, sizeof(float) * R.p0, PCOMP,
                                 hipMemcpyDeviceToDevice, b->stream));
    if (tailP > 0)
        HIP_TRY(hipMemcpy2DAsync(b->pbuf2 + R.p0, pitch, b->pbuf + R.p0 + R.n, pitch, sizeof(float) * tailP, PCOMP,
                                 hipMemcpyDeviceToDevice, b->stream));
    if (R.p0 > 0)
        HIP_TRY(hipMemcpyAsync(b->pid2, b->pid, sizeof(int) * R.p0, hipMemcpyDeviceToDevice, b->stream));
    if (tailP > 0) {
        HIP_TRY(hipMemcpyAsync(b->pid2 + R.p0, b->pid + R.p0 + R.n, sizeof(int) * tailP, hipMemcpyDeviceToDevice,
                               b->stream));
        hipLaunchKernelGGL(k_add_i32, dim3(grid256(tailP)), dim3(256), 0, b->stream, b->pid2 + R.p0, tailP,
                           -(int)R.n);
    }
    std::swap(b->pbuf, b->pbuf2);
    std::swap(b->pid, b->pid2);
    // ever-touched flags move with their blocks; the grids start empty again
    if (tailB > 0) {
        HIP_TRY(hipMemcpyAsync(b->bever_tmp, b->bever + R.d.b0 + R.nb, tailB, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->bever + R.d.b0, b->bever_tmp, tailB, hipMemcpyDeviceToDevice, b->stream));
    }
    HIP_TRY(hipMemsetAsync(b->bever + b->B - R.nb, 0, R.nb, b->stream));
    for (GridBuf& g : b->grid) {
        HIP_TRY(hipMemsetAsync(g.mq, 0, sizeof(long long) * 64 * b->B, b->stream));
        HIP_TRY(hipMemsetAsync(g.mv, 0, sizeof(float4) * 64 * b->B, b->stream));
        for (int c = 0; c < 3; ++c)
            HIP_TRY(hipMemsetAsync(g.mom + c * g.mstride, 0, sizeof(float) * 64 * b->B, b->stream));
        HIP_TRY(hipMemsetAsync(g.bflag, 0, sizeof(uint32_t) * b->B, b->stream));
        HIP_TRY(hipMemsetAsync(g.bcount, 0, sizeof(int32_t), b->stream));
    }
    // bodies and reactions move down one slot
    const int ns = (int)b->scenes.size();
    if (s + 1 < ns) {
        std::vector<bm_body> bodies((size_t)(ns - s - 1));
        std::vector<double> out((size_t)(8 * (ns - s - 1)));
        HIP_TRY(hipMemcpyAsync(bodies.data(), b->d_bodies + s + 1, sizeof(bm_body) * bodies.size(),
                               hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipMemcpyAsync(out.data(), b->d_out7 + (int64_t)(s + 1) * 8, sizeof(double) * out.size(),
                               hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_bodies + s, bodies.data(), sizeof(bm_body) * bodies.size(), hipMemcpyHostToDevice,
                               b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_out7 + (int64_t)s * 8, out.data(), sizeof(double) * out.size(),
                               hipMemcpyHostToDevice, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (int t = s + 1; t < ns; ++t) {
        b->scenes[t].p0 -= R.n;
        b->scenes[t].d.b0 -= R.nb;
    }
    b->scenes.erase(b->scenes.begin() + s);
    b->P -= R.n;
    b->B -= R.nb;
    b->cur = 0;
    b->invalidate();
    return upload_tables(b);
}

int bm_batch_step(bm_batch_handle b, float dt, int32_t k) {
    BM_TRACE("bm_batch_step");
    if (int rc = check_batch(b)) return rc;
    if (k < 0) return fail(BM_ERR_INVALID_ARG, "k < 0");
    if (b->scenes.empty() || k == 0) return BM_OK;
    int32_t done = 0;
    if (b->use_graph) {
        const int G = 2 * b->sort_every;
        // bring the batch to the replay's starting point: deposited, re-binning due
        while (done < k && (!b->prepared || dt != b->dt_prepared || !b->binned || b->since_bin != b->sort_every)) {
            batch_substep(b, dt, done == k - 1);
            ++done;
        }
        for (; k - done >= G; done += G) {
            hipGraphExec_t ge = nullptr;
            if (int rc = graph_for_state(b, dt, k - done == G, &ge)) return rc;
            HIP_TRY(hipGraphLaunch(ge, b->stream));
        }
    }
    for (; done < k; ++done) batch_substep(b, dt, done == k - 1);
    HIP_TRY(hipGetLastError());
    return BM_OK;
}

int bm_batch_download_scene(bm_batch_handle b, int32_t id, bm_particles* p) {
    BM_TRACE("bm_batch_download_scene");
    if (int rc = check_batch(b)) return rc;
    if (!p) return fail(BM_ERR_INVALID_ARG, "null output");
    const int s = b->find(id);
    if (s < 0) return fail(BM_ERR_INVALID_ARG, "no scene %d", id);
    if (int rc = batch_check_err(b, s)) return rc;
    const SceneHost& S = b->scenes[s];
    const int64_t n = S.n;
    if (n == 0) return BM_OK;
    if (!b->stage.ensure((sizeof(float) * PCOMP + sizeof(int)) * n)) return fail(BM_ERR_OOM, "pinned staging");
    float* st = b->stage.f();
    int* order = reinterpret_cast<int*>(st + PCOMP * n);
    // slot p0 + i holds the scene's particle pid[p0 + i] (insertion index +
    // p0): the GPU puts them back in insertion order in the re-binning buffer
    // (free between re-binnings), so the host copy is a plain transpose
    launch_unpermute(b->pview(), b->pbuf2, b->pid, S.p0, n, b->stream);
    HIP_TRY(hipMemcpyAsync(order, b->pid + S.p0, sizeof(int) * n, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipMemcpy2DAsync(st, sizeof(float) * n, b->pbuf2 + S.p0, sizeof(float) * b->capP, sizeof(float) * n, PCOMP,
                             hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (!order_in_range(order, n, S.p0)) return fail(BM_ERR_HIP, "particle order corrupted");
    particles_from_soa(st, n, nullptr, 0, p);
    return BM_OK;
}

int bm_batch_get_body(bm_batch_handle b, int32_t id, bm_body* body, bm_reaction* r) {
    if (int rc = check_batch(b)) return rc;
    if (!body) return fail(BM_ERR_INVALID_ARG, "null output");
    const int s = b->find(id);
    if (s < 0) return fail(BM_ERR_INVALID_ARG, "no scene %d", id);
    if (!b->scenes[s].has_body) return fail(BM_ERR_STATE, "scene %d has no body", id);
    if (int rc = batch_check_err(b, s)) return rc;
    double o[7];
    HIP_TRY(hipMemcpy(body, b->d_bodies + s, sizeof(bm_body), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(o, b->d_out7 + (int64_t)s * 8, sizeof(double) * 7, hipMemcpyDeviceToHost));
    if (r) {
        for (int a = 0; a < 3; ++a) {
            r->force[a] = (float)o[a];
            r->torque[a] = (float)o[3 + a];
        }
        r->nodes_inside = (int32_t)std::llround(o[6]);
    }
    return BM_OK;
}

int bm_batch_scene_error(bm_batch_handle b, int32_t id, int32_t* code) {
    if (int rc = check_batch(b)) return rc;
    if (!code) return fail(BM_ERR_INVALID_ARG, "null output");
    const int s = b->find(id);
    if (s < 0) return fail(BM_ERR_INVALID_ARG, "no scene %d", id);
    if (int rc = batch_harvest_err(b)) return rc;
    *code = scene_err_code(b->scenes[s].err);
    return BM_OK;
}

int bm_batch_usage(bm_batch_handle b, int32_t* scenes, int64_t* particles, int64_t* blocks) {
    if (int rc = check_batch(b)) return rc;
    if (scenes) *scenes = (int32_t)b->scenes.size();
    if (particles) *particles = b->P;
    if (blocks) *blocks = b->B;
    return BM_OK;
}

int bm_batch_set_tuning(bm_batch_handle b, int32_t unit_particles, int32_t sort_every, int32_t deposit_groups) {
    if (int rc = check_batch(b)) return rc;
    if (deposit_groups > 0) b->deposit_groups = (deposit_groups + 7) / 8 * 8;
    if (unit_particles > 0) {
        if (unit_particles < 16 || unit_particles > UNIT_MAX)
            return fail(BM_ERR_INVALID_ARG, "unit_particles must be in [16, %d]", UNIT_MAX);
        b->unit_p = unit_particles;
    }
    if (sort_every > 0) b->sort_every = sort_every;
    b->binned = false;  // the units follow the unit size
    b->graph_valid = false;
    return BM_OK;
}

int bm_batch_set_graph(bm_batch_handle b, int32_t on) {
    if (int rc = check_batch(b)) return rc;
    b->use_graph = on != 0;
    return BM_OK;
}

int bm_batch_autotune(bm_batch_handle b, float dt, int32_t substeps, int32_t* unit_out, int32_t* sort_out,
                      int32_t* groups_out) {
    if (int rc = check_batch(b)) return rc;
    if (substeps <= 0 || !(dt > 0.0f)) return fail(BM_ERR_INVALID_ARG, "substeps and dt must be > 0");
    if (b->scenes.empty()) return fail(BM_ERR_STATE, "no scenes");
    HIP_TRY(hipStreamSynchronize(b->stream));
    // save the state the trial substeps change
    const int ns = (int)b->scenes.size();
    float* sp = nullptr;
    int* sid = nullptr;
    bm_body* sbody = nullptr;
    uint8_t* sever = nullptr;
    if (hipMalloc(&sp, sizeof(float) * PCOMP * std::max<int64_t>(b->P, 1)) != hipSuccess ||
        hipMalloc(&sid, sizeof(int) * std::max<int64_t>(b->P, 1)) != hipSuccess ||
        hipMalloc(&sbody, sizeof(bm_body) * ns) != hipSuccess ||
        hipMalloc(&sever, std::max<int64_t>(b->B, 1)) != hipSuccess) {
        hipFree(sp);
        hipFree(sid);
        hipFree(sbody);
        return fail(BM_ERR_OOM, "autotune state copy");
    }
    const size_t pitch = sizeof(float) * b->capP, spitch = sizeof(float) * std::max<int64_t>(b->P, 1);
    auto save = [&]() -> int {
        // stream-ordered: copies on the null stream would not be ordered
        // with the batch's (non-blocking) stream
        HIP_TRY(hipMemcpy2DAsync(sp, spitch, b->pbuf, pitch, sizeof(float) * b->P, PCOMP, hipMemcpyDeviceToDevice,
                                 b->stream));
        HIP_TRY(hipMemcpyAsync(sid, b->pid, sizeof(int) * b->P, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(sbody, b->d_bodies, sizeof(bm_body) * ns, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(sever, b->bever, b->B, hipMemcpyDeviceToDevice, b->stream));
        return BM_OK;
    };
    auto restore = [&]() -> int {
        HIP_TRY(hipStreamSynchronize(b->stream));
        HIP_TRY(hipMemcpy2DAsync(b->pbuf, pitch, sp, spitch, sizeof(float) * b->P, PCOMP, hipMemcpyDeviceToDevice,
                                 b->stream));
        HIP_TRY(hipMemcpyAsync(b->pid, sid, sizeof(int) * b->P, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_bodies, sbody, sizeof(bm_body) * ns, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->bever, sever, b->B, hipMemcpyDeviceToDevice, b->stream));
        b->invalidate();
        return BM_OK;
    };
    int rc = save();
    struct Cand {
        int unit, sort, groups;
        float ms;
    };
    std::vector<Cand> cands;
    for (int u : {512, 640})
        for (int so : {8, 16, 32})
            for (int g : {2048, 4096, 8192}) cands.push_back({u, so, g, 0.0f});
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == BM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(BM_ERR_HIP, "event creation");
    const bool graph = b->use_graph;
    b->use_graph = false;
    // time one candidate from the saved state over a whole number of
    // re-binning cycles (each re-binning falls inside the window)
    auto trial = [&](Cand& c) -> int {
        if (int r = restore()) return r;
        b->unit_p = c.unit;
        b->sort_every = c.sort;
        b->deposit_groups = c.groups;
        const int window = (std::max(substeps, c.sort) + c.sort - 1) / c.sort * c.sort;
        // the saved state is restored afterwards: no substep needs velocity and C
        for (int w = 0; w < 2; ++w) batch_substep(b, dt, false);
        hipEventRecord(e0, b->stream);
        for (int i = 0; i < window; ++i) batch_substep(b, dt, false);
        hipEventRecord(e1, b->stream);
        if (hipEventSynchronize(e1) != hipSuccess) return fail(BM_ERR_HIP, "autotune timing");
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        c.ms += ms / (float)window;
        return BM_OK;
    };
    // clocks settle over the first trials: one untimed pass, then a forward
    // and a backward pass so that drift weighs on every candidate alike
    if (rc == BM_OK) {
        Cand warm = cands[cands.size() / 2];
        rc = trial(warm);
    }
    for (size_t i = 0; rc == BM_OK && i < cands.size(); ++i) rc = trial(cands[i]);
    for (size_t i = cands.size(); rc == BM_OK && i-- > 0;) rc = trial(cands[i]);
    int bu = b->unit_p, bsrt = b->sort_every, bg = b->deposit_groups;
    float best = 1e30f;
    for (const Cand& c : cands)
        if (rc == BM_OK && c.ms < best) {
            best = c.ms;
            bu = c.unit;
            bsrt = c.sort;
            bg = c.groups;
        }
    b->use_graph = graph;
    if (rc == BM_OK) rc = restore();
    b->unit_p = bu;
    b->sort_every = bsrt;
    b->deposit_groups = bg;
    b->graph_valid = false;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(sp);
    hipFree(sid);
    hipFree(sbody);
    hipFree(sever);
    if (rc == BM_OK) rc = batch_check_err(b);
    if (rc == BM_OK) {
        if (unit_out) *unit_out = bu;
        if (sort_out) *sort_out = bsrt;
        if (groups_out) *groups_out = bg;
    }
    return rc;
}

}  // extern "C"
