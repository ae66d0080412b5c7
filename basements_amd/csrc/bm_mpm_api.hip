// bm_mpm_api.hip -- the C ABI of libbm_mpm.so (declared in include/bm_mpm.h).
//
// Host side of the handle: owns every device buffer, queues the substep
// kernels on its stream, converts the reference's POD structs, and turns HIP
// failures into status codes plus a last-error string.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_staging.hpp"
#include "mpm_internal.hpp"
#include "trace.hpp"

#include <atomic>

using namespace bm;

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

int bm::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

enum Phase { PH_EMPTY = 0, PH_P2G = 1, PH_UPDATED = 2 };

struct bm_mpm_solver {
    bm_mpm_desc desc;
    Domain dom;
    int64_t nblocks = 0, nnodes = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    float* pbuf = nullptr;  // 26 * cap floats
    int64_t cap = 0, n = 0;
    PinnedStage stage;      // host side of uploads and downloads

    GridBuf grid[2] = {};
    int cur = 0;
    uint8_t* bever = nullptr;
    int* d_err = nullptr;
    unsigned long long* d_scratch_u64 = nullptr;  // 2 words: mass total, max mass bits
    unsigned long long* d_fallbacks = nullptr;
    double* d_partials = nullptr;
    double* d_out7 = nullptr;
    // device body of the coupled loop
    bm_body* d_body = nullptr;
    ColliderDev* d_col = nullptr;
    int64_t body_blocks = 0;  // scan-range bound of the body's collider
    bool body_set = false;
    bool body_begun = false;  // the grid update already started the body's substep

    // MPMSolver parameters (mpm_solver.h:100-109 defaults)
    float E = 1.0e5f, nu = 0.3f, rho = 1500.0f, alpha = 0.40f, mu_floor = 0.4f;
    float g[3] = {0.0f, -9.81f, 0.0f};
    int walls = 0;

    int q_exp = 0;
    float max_mass = 0.0f;
    int phase = PH_EMPTY;

    // fast mode: binned particle order, double-buffered particles and grids
    bool fast = false;
    float* pbuf2 = nullptr;
    int* pid = nullptr;   // original (insertion) index of each stored particle
    int* pid2 = nullptr;
    int* redo = nullptr;    // particles redone with the FP64 SVD (one list per launch)
    int* n_redo = nullptr;
    float* coef = nullptr;  // deposit records of the two-kernel form (16 floats per particle)
    BinScratch bs{};
    int n_units = 0;
    bool binned = false;      // units describe the current particle order
    int since_bin = 0;        // substeps since the last re-binning
    bool next_cleared = false;  // the last grid update cleared the other grid for the next G2P
    int sort_every = 32;
    int unit_p = 640;
    int variant = 4;          // fast-kernel variant (see bm_mpm_set_tuning): two-kernel form
    bool prepared = false;    // grid[prep] already holds the next substep's P2G
    float dt_prepared = 0.0f;
    int prep = 0;

    // per-phase event timing
    struct Timed {
        int phase;
        hipEvent_t a, b;
    };
    int timing = 0;  // 1: every phase; 2: dispatch-recorded kernel events only; 3: the particle pass's only
    BodyEnd pending_end;  // a coupled substep's body end left to the next fast particle pass
    std::vector<hipEvent_t> ev_pool;
    std::vector<Timed> pending;
    double t_ms[BM_NUM_PHASES] = {};
    int64_t t_calls[BM_NUM_PHASES] = {};
    int open_phase = -1;
    hipEvent_t open_ev = nullptr;

    hipEvent_t take_event() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void begin(int ph) {
        if (!timing) return;
        open_phase = ph;
        open_ev = nullptr;
        if (timing != 1) return;  // the phase is known, no event between launches
        open_ev = take_event();
        if (open_ev) hipEventRecord(open_ev, stream);
    }
    void end() {
        if (!timing || open_phase < 0) return;
        if (open_ev) {
            hipEvent_t e = take_event();
            if (e) {
                hipEventRecord(e, stream);
                pending.push_back(Timed{open_phase, open_ev, e});
            }
        }
        open_phase = -1;
        open_ev = nullptr;
    }
    // Fold finished event pairs into the accumulators (stream must be idle).
    void harvest() {
        for (auto& t : pending) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
                t_ms[t.phase] += ms;
                t_calls[t.phase] += 1;
            }
            ev_pool.push_back(t.a);
            ev_pool.push_back(t.b);
        }
        pending.clear();
    }

    PView pview() const { return PView{pbuf, cap, n}; }
    Phys phys() const {
        Phys p;
        p.mu = E / (2.0f * (1.0f + nu));
        p.lambda = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
        p.alpha = alpha;
        p.kohesion = (3.0f * p.lambda + 2.0f * p.mu) / (2.0f * p.mu);
        p.gx = g[0];
        p.gy = g[1];
        p.gz = g[2];
        p.mu_floor = mu_floor;
        p.walls = walls;
        return p;
    }
    double qscale() const { return std::ldexp(1.0, q_exp); }
    double qscale_inv() const { return std::ldexp(1.0, -q_exp); }
};

// Mass quantum: 2^(ilogb(max mass) - 22) kg, so a particle of the largest
// mass is 2^22..2^23 - 1 quanta (its float mantissa exactly) and a node
// deposit resolves 2^-23 of a particle's mass -- FP32 resolution.  A particle
// gives one node at most (27/64) q_p + 13 quanta (the largest weight is
// (3/4)^3; the stencil centre's remainder exceeds its own share by at most
// the other 26 nodes' roundings of 1/2), so the sum over any 1,024 particles
// stays below 2^32: the fast path's LDS tiles (units of at most UNIT_MAX =
// 1,024 particles) accumulate node mass in one 32-bit word.  Products
// w * q_p need at most 24 + 23 bits, so the double-precision rounding of
// w * q_p is exact.
static int quantum_exp_for(float max_mass) {
    if (!(max_mass > 0.0f) || !std::isfinite(max_mass)) return 0;
    return std::ilogb(max_mass) - 22;
}

static int check_handle(bm_mpm_handle h) {
    if (!h) return fail(BM_ERR_INVALID_ARG, "null handle");
    return BM_OK;
}

static int alloc_grid(GridBuf& g, int64_t nnodes, int64_t nblocks, bool with_force, bool with_mom) {
    HIP_TRY(hipMalloc(&g.mq, sizeof(long long) * nnodes));
    HIP_TRY(hipMalloc(&g.mv, sizeof(float4) * nnodes));
    if (with_force) HIP_TRY(hipMalloc(&g.f, sizeof(float4) * nnodes));
    if (with_mom) {
        HIP_TRY(hipMalloc(&g.mom, sizeof(float) * 3 * nnodes));
        HIP_TRY(hipMemset(g.mom, 0, sizeof(float) * 3 * nnodes));
        g.mstride = nnodes;
    }
    HIP_TRY(hipMalloc(&g.bflag, sizeof(uint32_t) * nblocks));
    HIP_TRY(hipMalloc(&g.blist, sizeof(int32_t) * nblocks));
    HIP_TRY(hipMalloc(&g.bcount, sizeof(int32_t)));
    HIP_TRY(hipMemset(g.mq, 0, sizeof(long long) * nnodes));
    HIP_TRY(hipMemset(g.mv, 0, sizeof(float4) * nnodes));
    if (with_force) HIP_TRY(hipMemset(g.f, 0, sizeof(float4) * nnodes));
    HIP_TRY(hipMemset(g.bflag, 0, sizeof(uint32_t) * nblocks));
    HIP_TRY(hipMemset(g.bcount, 0, sizeof(int32_t)));
    return BM_OK;
}

static void free_grid(GridBuf& g) {
    hipFree(g.mq);
    hipFree(g.mv);
    hipFree(g.f);
    hipFree(g.mom);
    hipFree(g.bflag);
    hipFree(g.blist);
    hipFree(g.bcount);
    g = GridBuf{};
}

static int sync_and_check(bm_mpm_handle h) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    int err = 0;
    HIP_TRY(hipMemcpy(&err, h->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        HIP_TRY(hipMemsetAsync(h->d_err, 0, sizeof(int), h->stream));  // ordered with the handle's stream
        if (err & 1) return fail(BM_ERR_OUT_OF_DOMAIN, "a particle stencil left the grid domain");
        if (err & 8) return fail(BM_ERR_NONFINITE, "a particle's position or velocity became non-finite");
        return fail(BM_ERR_HIP, "device body collider exceeded its scan-range bound");
    }
    return BM_OK;
}

// ------------------------------------------------------------------ allocation-free coupler
// Scratch of bm_mpm_couple_noalloc: the coupler's per-workgroup partial
// reactions (COUPLE_WGS x 7 doubles) and their fixed-order sum (7 + 1
// doubles), 7,232 bytes per GPU, held until the process ends.  Every
// workgroup writes all 7 of its partials and the sum is written whole, so
// nothing needs clearing between calls; calls that overlap on one GPU share
// the array and race.
constexpr int NOALLOC_PARTIALS = COUPLE_WGS * 7;
constexpr int NOALLOC_DOUBLES = NOALLOC_PARTIALS + 8;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// The device address of `scratch` on each GPU (the host-side name is only a
// stand-in): looked up once per device, with that device current.
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::atomic<double*> g_scratch_ptr[NOALLOC_MAX_DEVICES];

static int noalloc_scratch(double** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(BM_ERR_NOT_SUPPORTED, "device %d beyond the %d the allocation-free coupler tracks", dev,
                    NOALLOC_MAX_DEVICES);
    double* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        void* addr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !addr)
            return fail(BM_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        p = static_cast<double*>(addr);
        g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
    *out = p;
    return BM_OK;
}

// The reaction in the caller's device memory: the FP64 sum rounded to FP32
// as bm_mpm_couple rounds it on the host (zero reaction when the sum is null).
__global__ void k_reaction_out(const double* out7, bm_reaction* r) {
    if (threadIdx.x != 0) return;
    for (int a = 0; a < 3; ++a) {
        r->force[a] = out7 ? (float)out7[a] : 0.0f;
        r->torque[a] = out7 ? (float)out7[3 + a] : 0.0f;
    }
    r->nodes_inside = out7 ? (int32_t)llround(out7[6]) : 0;
}

extern "C" {

int bm_abi_version(void) { return BM_MPM_ABI_VERSION; }
const char* bm_last_error(void) { return g_last_error.c_str(); }

int bm_mpm_create(const bm_mpm_desc* desc, bm_mpm_handle* out) {
    if (!desc || !out) return fail(BM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(desc->dx > 0.0f) || !std::isfinite(desc->dx)) return fail(BM_ERR_INVALID_ARG, "dx must be > 0");
    if (desc->max_particles < 0) return fail(BM_ERR_INVALID_ARG, "max_particles < 0");
    for (int a = 0; a < 3; ++a)
        if (desc->grid_dims[a] <= 0) return fail(BM_ERR_INVALID_ARG, "grid_dims must be > 0");
    if (desc->mode != BM_MODE_REFERENCE && desc->mode != BM_MODE_FAST)
        return fail(BM_ERR_INVALID_ARG, "unknown mode %d", desc->mode);

    // set-up of the allocation-free coupler's scratch on this device
    double* scratch_ptr = nullptr;
    if (int rc = noalloc_scratch(&scratch_ptr)) return rc;
    (void)scratch_ptr;

    auto* h = new bm_mpm_solver();
    h->desc = *desc;
    Domain& d = h->dom;
    d.ox = desc->grid_origin[0];
    d.oy = desc->grid_origin[1];
    d.oz = desc->grid_origin[2];
    d.nx = (desc->grid_dims[0] + 3) & ~3;
    d.ny = (desc->grid_dims[1] + 3) & ~3;
    d.nz = (desc->grid_dims[2] + 3) & ~3;
    d.nbx = d.nx / 4;
    d.nby = d.ny / 4;
    d.nbz = d.nz / 4;
    d.dx = desc->dx;
    d.inv_dx = 1.0f / desc->dx;  // SPGridCPU(dx): inv_dx = 1/dx (spgrid_cpu.h:182)
    d.b0 = 0;
    h->nblocks = (int64_t)d.nbx * d.nby * d.nbz;
    h->nnodes = h->nblocks * 64;
    if (h->nblocks > (int64_t)1 << 30) {
        delete h;
        return fail(BM_ERR_INVALID_ARG, "grid too large");
    }
    h->cap = desc->max_particles;

    int rc = BM_OK;
    auto bail = [&](int code) {
        bm_mpm_destroy(h);
        return code;
    };
    if (desc->stream) {
        h->stream = (hipStream_t)desc->stream;
    } else {
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(BM_ERR_HIP, "hipStreamCreate failed"));
        h->own_stream = true;
    }
    if (h->cap > 0 && hipMalloc(&h->pbuf, sizeof(float) * PCOMP * h->cap) != hipSuccess)
        return bail(fail(BM_ERR_OOM, "particle arrays (%lld particles)", (long long)h->cap));
    h->fast = desc->mode == BM_MODE_FAST;
    if (desc->sort_every > 0) h->sort_every = desc->sort_every;
    if ((rc = alloc_grid(h->grid[0], h->nnodes, h->nblocks, !h->fast, h->fast)) != BM_OK) return bail(BM_ERR_OOM);
    if (h->fast) {
        if ((rc = alloc_grid(h->grid[1], h->nnodes, h->nblocks, false, true)) != BM_OK) return bail(BM_ERR_OOM);
        if (h->cap > (int64_t)INT32_MAX) return bail(fail(BM_ERR_INVALID_ARG, "fast mode holds at most 2^31-1 particles"));
        if (h->nblocks > (int64_t)INT32_MAX / 64) return bail(fail(BM_ERR_INVALID_ARG, "grid too large"));
        const int64_t pc = std::max<int64_t>(h->cap, 1);
        BinScratch& b = h->bs;
        b.temp_bytes = scan_temp_bytes((int)h->nblocks);
        if (hipMalloc(&h->pbuf2, sizeof(float) * PCOMP * pc) != hipSuccess ||
            hipMalloc(&h->pid, sizeof(int) * pc) != hipSuccess || hipMalloc(&h->pid2, sizeof(int) * pc) != hipSuccess ||
            hipMalloc(&b.key, sizeof(int) * pc) != hipSuccess || hipMalloc(&b.slot, sizeof(int) * pc) != hipSuccess ||
            hipMalloc(&b.bcnt, sizeof(int) * h->nblocks) != hipSuccess ||
            hipMalloc(&b.bstart, sizeof(int) * h->nblocks) != hipSuccess ||
            hipMalloc(&b.ucnt, sizeof(int) * h->nblocks) != hipSuccess ||
            hipMalloc(&b.ustart, sizeof(int) * h->nblocks) != hipSuccess ||
            hipMalloc(&b.units, sizeof(int4) * (pc / 16 + h->nblocks + 1)) != hipSuccess ||
            hipMalloc(&b.n_units, sizeof(int)) != hipSuccess ||
            hipMalloc(&h->redo, sizeof(int) * pc) != hipSuccess || hipMalloc(&h->n_redo, sizeof(int)) != hipSuccess ||
            hipMalloc(&h->coef, sizeof(float) * 16 * ((pc + 63) / 64 * 64)) != hipSuccess ||
            hipMalloc(&b.temp, std::max<size_t>(b.temp_bytes, 16)) != hipSuccess)
            return bail(fail(BM_ERR_OOM, "fast-mode scratch"));
    }
    if (hipMalloc(&h->bever, h->nblocks) != hipSuccess) return bail(fail(BM_ERR_OOM, "block flags"));
    hipMemset(h->bever, 0, h->nblocks);
    if (hipMalloc(&h->d_partials, sizeof(double) * 7 * COUPLE_WGS) != hipSuccess)
        return bail(fail(BM_ERR_OOM, "coupler partials"));
    if (hipMalloc(&h->d_out7, sizeof(double) * 8) != hipSuccess) return bail(fail(BM_ERR_OOM, "coupler out"));
    if (hipMalloc(&h->d_body, sizeof(bm_body)) != hipSuccess) return bail(fail(BM_ERR_OOM, "device body"));
    if (hipMalloc(&h->d_col, sizeof(ColliderDev)) != hipSuccess) return bail(fail(BM_ERR_OOM, "device collider"));
    hipMemset(h->d_out7, 0, sizeof(double) * 8);
    if (hipMalloc(&h->d_err, sizeof(int)) != hipSuccess) return bail(fail(BM_ERR_OOM, "err flag"));
    if (hipMalloc(&h->d_scratch_u64, sizeof(unsigned long long) * 2) != hipSuccess) return bail(fail(BM_ERR_OOM, "scratch"));
    if (hipMalloc(&h->d_fallbacks, 2 * sizeof(unsigned long long)) != hipSuccess) return bail(fail(BM_ERR_OOM, "counter"));
    hipMemset(h->d_err, 0, sizeof(int));
    hipMemset(h->d_fallbacks, 0, 2 * sizeof(unsigned long long));
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(BM_ERR_HIP, "device sync after allocation"));
    *out = h;
    return BM_OK;
}

int bm_mpm_destroy(bm_mpm_handle h) {
    if (!h) return BM_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    h->harvest();
    for (hipEvent_t e : h->ev_pool) hipEventDestroy(e);
    hipFree(h->pbuf);
    free_grid(h->grid[0]);
    free_grid(h->grid[1]);
    hipFree(h->bever);
    hipFree(h->d_partials);
    hipFree(h->d_out7);
    hipFree(h->d_body);
    hipFree(h->d_col);
    hipFree(h->d_err);
    hipFree(h->d_scratch_u64);
    hipFree(h->d_fallbacks);
    hipFree(h->pbuf2);
    hipFree(h->pid);
    hipFree(h->pid2);
    hipFree(h->redo);
    hipFree(h->n_redo);
    hipFree(h->coef);
    hipFree(h->bs.key);
    hipFree(h->bs.slot);
    hipFree(h->bs.bcnt);
    hipFree(h->bs.bstart);
    hipFree(h->bs.ucnt);
    hipFree(h->bs.ustart);
    hipFree(h->bs.units);
    hipFree(h->bs.n_units);
    hipFree(h->bs.temp);
    if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
    delete h;
    return BM_OK;
}

int bm_mpm_synchronize(bm_mpm_handle h) {
    if (int rc = check_handle(h)) return rc;
    return sync_and_check(h);
}

int bm_mpm_set_material(bm_mpm_handle h, float E, float nu, float rho) {
    if (int rc = check_handle(h)) return rc;
    if (E != h->E || nu != h->nu || rho != h->rho) h->prepared = false;  // the deposited stress is stale
    h->E = E;
    h->nu = nu;
    if (rho > 1e-12f && h->rho > 1e-12f && rho != h->rho && h->n > 0) {
        const float rescale = h->rho / rho;
        launch_scale_v0(h->pview(), rescale, h->stream);
        HIP_TRY(hipGetLastError());
    }
    h->rho = rho;
    return BM_OK;
}
int bm_mpm_set_gravity(bm_mpm_handle h, float gx, float gy, float gz) {
    if (int rc = check_handle(h)) return rc;
    h->g[0] = gx;
    h->g[1] = gy;
    h->g[2] = gz;
    return BM_OK;
}
int bm_mpm_set_plastic_alpha(bm_mpm_handle h, float a) {
    if (int rc = check_handle(h)) return rc;
    h->alpha = a;
    return BM_OK;
}
int bm_mpm_set_floor_friction(bm_mpm_handle h, float mu) {
    if (int rc = check_handle(h)) return rc;
    h->mu_floor = mu;
    return BM_OK;
}
int bm_mpm_set_walls(bm_mpm_handle h, int32_t thickness) {
    if (int rc = check_handle(h)) return rc;
    if (thickness < 0) return fail(BM_ERR_INVALID_ARG, "wall thickness < 0");
    h->walls = thickness;
    return BM_OK;
}

// Copy host SoA (3n/9n interleaved per particle) into component-major device arrays.
// Particles [off, off + n) from the caller's layout: component-major pinned
// staging (host threads), then one pitched copy of all components.
static int upload_range(bm_mpm_handle h, int64_t off, int64_t n, const float* x, const float* v, const float* F,
                        const float* Cm, const float* m, const float* V0) {
    if (!h->stage.ensure(sizeof(float) * PCOMP * n)) return fail(BM_ERR_OOM, "pinned staging");
    float* st = h->stage.f();
    if (!soa_from_particles(x, v, F, Cm, m, V0, n, st))
        return fail(BM_ERR_NONFINITE, "particle state holds a non-finite value");
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy2DAsync(h->pbuf + off, sizeof(float) * h->cap, st, sizeof(float) * n, sizeof(float) * n, PCOMP,
                             hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return BM_OK;
}

static void refresh_quantum(bm_mpm_handle h) { h->q_exp = quantum_exp_for(h->max_mass); }

// The particle arrays were rewritten from outside: the stored order is the
// caller's order again and nothing precomputed from the old state is valid.
// `first` is the first slot whose identity is new (slots before it keep
// theirs: add_particles appends behind the binned particles).
static void particles_changed(bm_mpm_handle h, int64_t first = 0) {
    h->phase = PH_EMPTY;
    h->prepared = false;
    h->binned = false;
    if (h->fast && h->n > first) launch_iota(h->pid + first, h->n - first, first, h->stream);
}

int bm_mpm_add_particles(bm_mpm_handle h, int64_t n, const float* pos, const float* vel, const float* mass) {
    if (int rc = check_handle(h)) return rc;
    if (n < 0 || (n > 0 && (!pos || !vel || !mass))) return fail(BM_ERR_INVALID_ARG, "bad particle arrays");
    if (h->n + n > h->cap) return fail(BM_ERR_OOM, "capacity %lld exceeded", (long long)h->cap);
    if (n == 0) return BM_OK;
    std::vector<float> F((size_t)n * 9, 0.0f), Cm((size_t)n * 9, 0.0f), V0((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        F[i * 9 + 0] = F[i * 9 + 4] = F[i * 9 + 8] = 1.0f;
        V0[i] = (h->rho > 1e-12f) ? (mass[i] / h->rho) : 1.0f;
        h->max_mass = std::max(h->max_mass, mass[i]);
    }
    if (int rc = upload_range(h, h->n, n, pos, vel, F.data(), Cm.data(), mass, V0.data())) return rc;
    const int64_t first = h->n;
    h->n += n;
    refresh_quantum(h);
    particles_changed(h, first);
    return BM_OK;
}

int bm_mpm_upload_particles(bm_mpm_handle h, int64_t n, const bm_particles* p) {
    BM_TRACE("bm_mpm_upload_particles");
    if (int rc = check_handle(h)) return rc;
    if (n < 0 || !p) return fail(BM_ERR_INVALID_ARG, "bad particle arrays");
    if (n > h->cap) return fail(BM_ERR_OOM, "capacity %lld exceeded", (long long)h->cap);
    if (n > 0 && (!p->x || !p->v || !p->F || !p->C || !p->m || !p->V0))
        return fail(BM_ERR_INVALID_ARG, "null particle field");
    h->max_mass = 0.0f;
    for (int64_t i = 0; i < n; ++i) h->max_mass = std::max(h->max_mass, p->m[i]);
    if (n > 0)
        if (int rc = upload_range(h, 0, n, p->x, p->v, p->F, p->C, p->m, p->V0)) return rc;
    h->n = n;
    refresh_quantum(h);
    particles_changed(h);
    return BM_OK;
}

int bm_mpm_download_particles(bm_mpm_handle h, bm_particles* p) {
    BM_TRACE("bm_mpm_download_particles");
    if (int rc = check_handle(h)) return rc;
    if (!p) return fail(BM_ERR_INVALID_ARG, "null output");
    if (int rc = sync_and_check(h)) return rc;
    const int64_t n = h->n;
    if (n == 0) return BM_OK;
    if (!h->stage.ensure((sizeof(float) * PCOMP + sizeof(int)) * n)) return fail(BM_ERR_OOM, "pinned staging");
    float* st = h->stage.f();
    int* order = reinterpret_cast<int*>(st + PCOMP * n);
    // fast mode stores particles binned by block (slot i holds particle
    // pid[i]): the GPU puts them back in insertion order in the re-binning
    // buffer (free between re-binnings), so the host copy is a plain transpose
    const float* src = h->pbuf;
    if (h->fast) {
        launch_unpermute(h->pview(), h->pbuf2, h->pid, 0, n, h->stream);
        HIP_TRY(hipMemcpyAsync(order, h->pid, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
        src = h->pbuf2;
    }
    HIP_TRY(hipMemcpy2DAsync(st, sizeof(float) * n, src, sizeof(float) * h->cap, sizeof(float) * n, PCOMP,
                             hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->fast && !order_in_range(order, n, 0)) return fail(BM_ERR_HIP, "particle order corrupted");
    particles_from_soa(st, n, nullptr, 0, p);
    return BM_OK;
}

int bm_mpm_num_particles(bm_mpm_handle h, int64_t* n) {
    if (int rc = check_handle(h)) return rc;
    if (!n) return fail(BM_ERR_INVALID_ARG, "null output");
    *n = h->n;
    return BM_OK;
}

int bm_mpm_particles_device(bm_mpm_handle h, bm_particles_device* out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(BM_ERR_INVALID_ARG, "null output");
    out->x = h->pbuf + PX * h->cap;
    out->v = h->pbuf + PV * h->cap;
    out->F = h->pbuf + PF * h->cap;
    out->C = h->pbuf + PC * h->cap;
    out->m = h->pbuf + PM * h->cap;
    out->V0 = h->pbuf + PV0 * h->cap;
    out->capacity = h->cap;
    return BM_OK;
}

int bm_mpm_commit_device_particles(bm_mpm_handle h, int64_t n) {
    if (int rc = check_handle(h)) return rc;
    if (n < 0 || n > h->cap) return fail(BM_ERR_INVALID_ARG, "count %lld outside [0, capacity]", (long long)n);
    h->n = n;
    unsigned int bits = 0;
    HIP_TRY(hipMemsetAsync(h->d_scratch_u64, 0, sizeof(unsigned long long) * 2, h->stream));
    launch_max_mass(h->pview(), (unsigned int*)h->d_scratch_u64, h->stream);
    HIP_TRY(hipMemcpyAsync(&bits, h->d_scratch_u64, sizeof(unsigned int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::memcpy(&h->max_mass, &bits, sizeof(float));
    refresh_quantum(h);
    particles_changed(h);
    return BM_OK;
}

// ------------------------------------------------------------------ substep
// Fast mode: re-bin the particles by grid block (swaps the particle buffers).
// with_vc = false when a G2P follows, which rewrites velocity and C.
static int rebin(bm_mpm_handle h, bool with_vc = true) {
    BM_TRACE("rebin");
    h->begin(BM_PHASE_SORT);
    const int rc = launch_rebin(h->pview(), h->pbuf2, h->pid, h->pid2, h->dom, (int)h->nblocks, h->unit_p, h->bs,
                                h->stream, &h->n_units, SceneView{}, 0, with_vc);
    h->end();
    if (rc != 0) return fail(BM_ERR_HIP, "particle re-binning failed: %s", hipGetErrorString(hipGetLastError()));
    std::swap(h->pbuf, h->pbuf2);
    std::swap(h->pid, h->pid2);
    h->binned = true;
    h->since_bin = 0;
    return BM_OK;
}

// The G2P form of a single scene (bm_mpm_set_tuning bits 4 and 5, else by
// the number of units).
static bool g2p_tiled(bm_mpm_handle h) {
    if (h->variant & 16) return false;
    if (h->variant & 32) return true;
    return h->n_units >= G2P_TILE_MIN_UNITS;
}

// The fast substep's particle kernels: the per-particle G2P (or the deposit
// record of the current state when do_g2p = false) and the per-cell deposit.
// store_vc = false skips the velocity and C stores of the G2P.
static void fast_kernels(bm_mpm_handle h, bool do_g2p, const GridBuf& gin, const GridBuf& gout, float dt_g2p,
                         float dt_fold, bool store_vc = true) {
    // Time each kernel on its own: the particle pass under the caller's phase
    // (G2P, or P2G for a deposit of the current state) and the cell deposit
    // under P2G, each with events recorded by its own dispatch; the small
    // bookkeeping launches between them are OTHER.
    struct Ctx {
        bm_mpm_handle h;
        int adv_phase;
    } ctx{h, h->open_phase};
    auto mark = [](void* p, int stage, hipEvent_t* a, hipEvent_t* b) {
        auto* c = (Ctx*)p;
        bm_mpm_handle hh = c->h;
        if (!hh->timing || c->adv_phase < 0) return;
        if (stage == SPLIT_REDO) {
            hh->begin(BM_PHASE_OTHER);
            return;
        }
        hh->end();
        if (hh->timing == 3 && stage == SPLIT_DEPOSIT) return;  // level 3: the particle pass only
        *a = hh->take_event();
        *b = hh->take_event();
        if (*a && *b)
            hh->pending.push_back(bm_mpm_solver::Timed{stage == SPLIT_ADVANCE ? c->adv_phase : BM_PHASE_P2G, *a, *b});
        else
            *a = *b = nullptr;
    };
    if (ctx.adv_phase >= 0) h->open_phase = BM_PHASE_OTHER;  // the opening event now starts OTHER
    launch_split(do_g2p, h->pview(), h->bs.units, h->n_units, gin, gout, h->bever, h->dom, h->phys(), dt_g2p, dt_fold,
                 h->qscale_inv(), h->d_err, h->d_fallbacks, h->redo, h->n_redo, h->coef, h->stream, mark, &ctx,
                 (h->variant & 8) != 0, SplitBatch{}, store_vc, g2p_tiled(h), do_g2p ? h->pending_end : BodyEnd{});
    if (do_g2p) h->pending_end = BodyEnd{};
}

static int p2g_fast(bm_mpm_handle h, float dt) {
    if (h->prepared && dt == h->dt_prepared) {  // deposited by the previous fused G2P
        h->cur = h->prep;
        h->prepared = false;
        h->phase = PH_P2G;
        return BM_OK;
    }
    if (!h->binned)
        if (int rc = rebin(h)) return rc;
    GridBuf& g = h->grid[h->cur];
    h->begin(BM_PHASE_CLEAR);
    launch_clear(g, (int)std::min<int64_t>(h->nblocks, 1 << 30), h->stream, h->n_redo, true);
    h->end();
    h->begin(BM_PHASE_P2G);
    fast_kernels(h, false, g, g, dt, dt);
    h->end();
    HIP_TRY(hipGetLastError());
    h->prepared = false;
    h->phase = PH_P2G;
    return BM_OK;
}

// store_vc = false inside a multi-substep call: a later G2P of the same call
// writes velocity and C (the re-binning before a G2P leaves them behind).
static int g2p_fast(bm_mpm_handle h, float dt, bool store_vc) {
    // the current grid survives the re-binning: only the particle order changes
    if (!h->binned || h->since_bin >= h->sort_every)
        if (int rc = rebin(h, false)) return rc;
    const int nxt = 1 - h->cur;
    if (!h->next_cleared) {
        h->begin(BM_PHASE_CLEAR);
        launch_clear(h->grid[nxt], (int)std::min<int64_t>(h->nblocks, 1 << 30), h->stream, h->n_redo,
                     true);
        h->end();
    }
    h->next_cleared = false;
    h->begin(BM_PHASE_G2P);
    fast_kernels(h, true, h->grid[h->cur], h->grid[nxt], dt, dt, store_vc);
    h->end();
    HIP_TRY(hipGetLastError());
    h->prepared = true;
    h->dt_prepared = dt;
    h->prep = nxt;
    h->since_bin += 1;
    return BM_OK;
}

int bm_mpm_p2g(bm_mpm_handle h, float dt) {
    BM_TRACE("bm_mpm_p2g");
    (void)dt;  // p2g_single ignores dt (mpm_solver.h:168)
    if (int rc = check_handle(h)) return rc;
    if (h->fast) return p2g_fast(h, dt);
    GridBuf& g = h->grid[h->cur];
    h->begin(BM_PHASE_CLEAR);
    launch_clear(g, (int)std::min<int64_t>(h->nblocks, 1 << 30), h->stream);
    h->end();
    h->begin(BM_PHASE_P2G);
    launch_p2g_ref(h->pview(), g, h->bever, h->dom, h->phys(), h->qscale_inv(), h->d_err, h->stream);
    h->end();
    HIP_TRY(hipGetLastError());
    h->phase = PH_P2G;
    return BM_OK;
}

// with_body: the update kernel also starts the device body's coupled substep
// (the caller couples the body next).
static int grid_update_impl(bm_mpm_handle h, float dt, bool with_body) {
    if (h->phase != PH_P2G) return fail(BM_ERR_STATE, "grid_update needs a preceding p2g");
    // two-kernel fast form: the update also clears the grid the next G2P
    // deposits into (and the FP64 redo count), saving that launch
    const bool clear_next = h->fast;
    const GridBuf* nxt = clear_next ? &h->grid[1 - h->cur] : nullptr;
    int* zw = clear_next ? h->n_redo : nullptr;
    h->begin(BM_PHASE_UPDATE);
    if (with_body)
        launch_update(h->grid[h->cur], (int)h->nblocks, h->dom, h->phys(), h->qscale(), dt, h->stream, h->d_body,
                      h->d_col, h->body_blocks, h->d_err, SceneView{}, nullptr, nxt, zw);
    else
        launch_update(h->grid[h->cur], (int)h->nblocks, h->dom, h->phys(), h->qscale(), dt, h->stream, nullptr,
                      nullptr, 0, nullptr, SceneView{}, nullptr, nxt, zw);
    h->end();
    HIP_TRY(hipGetLastError());
    h->next_cleared = clear_next;
    h->phase = PH_UPDATED;
    h->body_begun = with_body;
    return BM_OK;
}

int bm_mpm_grid_update(bm_mpm_handle h, float dt) {
    BM_TRACE("bm_mpm_grid_update");
    if (int rc = check_handle(h)) return rc;
    return grid_update_impl(h, dt, false);
}

// last = false: not the last G2P of the caller's call (see g2p_fast).
static int g2p_impl(bm_mpm_handle h, float dt, bool last) {
    if (int rc = check_handle(h)) return rc;
    if (h->phase != PH_UPDATED) return fail(BM_ERR_STATE, "g2p needs an updated grid");
    if (h->fast) return g2p_fast(h, dt, last);
    h->begin(BM_PHASE_G2P);
    launch_g2p_ref(h->pview(), h->grid[h->cur], h->dom, h->phys(), dt, h->d_err, h->stream);
    h->end();
    HIP_TRY(hipGetLastError());
    return BM_OK;
}

int bm_mpm_g2p(bm_mpm_handle h, float dt) {
    BM_TRACE("bm_mpm_g2p");
    return g2p_impl(h, dt, true);
}

static int step_impl(bm_mpm_handle h, float dt, bool last) {
    if (int rc = bm_mpm_p2g(h, dt)) return rc;
    if (int rc = bm_mpm_grid_update(h, dt)) return rc;
    return g2p_impl(h, dt, last);
}

int bm_mpm_step(bm_mpm_handle h, float dt) {
    BM_TRACE("bm_mpm_step");
    return step_impl(h, dt, true);
}

int bm_mpm_step_n(bm_mpm_handle h, float dt, int32_t k) {
    BM_TRACE("bm_mpm_step_n");
    if (k < 0) return fail(BM_ERR_INVALID_ARG, "k < 0");
    for (int32_t i = 0; i < k; ++i)
        if (int rc = step_impl(h, dt, i == k - 1)) return rc;
    return BM_OK;
}

// ------------------------------------------------------------------ couplers
static ColliderDev make_collider(bm_mpm_handle h, const bm_collider* b) {
    return collider_build(b->center, b->quat, b->half_extents, b->lin_vel, b->ang_vel, b->friction, b->mass, b->shape,
                          h->dom);
}

int bm_mpm_couple(bm_mpm_handle h, int32_t method, const bm_collider* body, float dt, bm_reaction* out) {
    BM_TRACE("bm_mpm_couple");
    if (int rc = check_handle(h)) return rc;
    if (!body || !out) return fail(BM_ERR_INVALID_ARG, "null argument");
    if (method < BM_COUPLER_M0_DIRICHLET || method > BM_COUPLER_M45_BSPLINE)
        return fail(BM_ERR_INVALID_ARG, "unknown coupler %d", method);
    std::memset(out, 0, sizeof(*out));
    if (!(dt > 0.0f)) return BM_OK;
    if (h->phase != PH_UPDATED) return fail(BM_ERR_STATE, "couple needs an updated grid");
    ColliderDev c = make_collider(h, body);
    h->begin(BM_PHASE_COUPLE);
    launch_couple(method, c, h->grid[h->cur], h->bever, h->dom, dt, h->d_partials, h->d_out7, h->stream);
    h->end();
    HIP_TRY(hipGetLastError());
    double r[7];
    HIP_TRY(hipMemcpyAsync(r, h->d_out7, sizeof(double) * 7, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int a = 0; a < 3; ++a) {
        out->force[a] = (float)r[a];
        out->torque[a] = (float)r[3 + a];
    }
    out->nodes_inside = (int32_t)std::llround(r[6]);
    return BM_OK;
}

int bm_mpm_couple_noalloc(bm_mpm_handle h, int32_t method, const bm_collider* body, float dt, bm_reaction* reaction) {
    if (int rc = check_handle(h)) return rc;
    if (!body || !reaction) return fail(BM_ERR_INVALID_ARG, "null argument");
    if (method < BM_COUPLER_M0_DIRICHLET || method > BM_COUPLER_M45_BSPLINE)
        return fail(BM_ERR_INVALID_ARG, "unknown coupler %d", method);
    // the partials fit the scratch for every input: one launch always runs
    // exactly COUPLE_WGS workgroups (the scan range only changes their work)
    static_assert(NOALLOC_PARTIALS >= COUPLE_WGS * 7, "partials exceed the scratch");
    double* sp = nullptr;
    if (int rc = noalloc_scratch(&sp)) return rc;
    if (!(dt > 0.0f)) {  // zero reaction, as bm_mpm_couple (mpm_rigid_coupler.cpp:62)
        hipLaunchKernelGGL(k_reaction_out, dim3(1), dim3(64), 0, h->stream, (const double*)nullptr, reaction);
        HIP_TRY(hipGetLastError());
        return BM_OK;
    }
    if (h->phase != PH_UPDATED) return fail(BM_ERR_STATE, "couple needs an updated grid");
    ColliderDev c = make_collider(h, body);
    h->begin(BM_PHASE_COUPLE);
    launch_couple(method, c, h->grid[h->cur], h->bever, h->dom, dt, sp, sp + NOALLOC_PARTIALS, h->stream);
    hipLaunchKernelGGL(k_reaction_out, dim3(1), dim3(64), 0, h->stream, (const double*)(sp + NOALLOC_PARTIALS),
                       reaction);
    h->end();
    HIP_TRY(hipGetLastError());
    return BM_OK;
}

// ------------------------------------------------------------------ device body
int bm_mpm_set_body(bm_mpm_handle h, const bm_body* body) {
    if (int rc = check_handle(h)) return rc;
    if (!body) return fail(BM_ERR_INVALID_ARG, "null body");
    if (body->shape != BM_SHAPE_BOX && body->shape != BM_SHAPE_CYLINDER)
        return fail(BM_ERR_INVALID_ARG, "unknown shape %d", body->shape);
    for (int a = 0; a < 3; ++a)
        if (!(body->half_extents[a] >= 0.0f) || !std::isfinite(body->half_extents[a]))
            return fail(BM_ERR_INVALID_ARG, "half extents must be finite and >= 0");
    const int64_t blocks = couple_blocks_bound(body->half_extents, h->dom);
    HIP_TRY(hipMemcpyAsync(h->d_body, body, sizeof(bm_body), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->body_blocks = blocks;
    h->body_set = true;
    return BM_OK;
}

int bm_mpm_get_body(bm_mpm_handle h, bm_body* body, bm_reaction* reaction) {
    if (int rc = check_handle(h)) return rc;
    if (!body) return fail(BM_ERR_INVALID_ARG, "null body");
    if (!h->body_set) return fail(BM_ERR_STATE, "no device body set");
    double r[7];
    HIP_TRY(hipMemcpyAsync(body, h->d_body, sizeof(bm_body), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(r, h->d_out7, sizeof(double) * 7, hipMemcpyDeviceToHost, h->stream));
    if (int rc = sync_and_check(h)) return rc;
    if (reaction) {
        for (int a = 0; a < 3; ++a) {
            reaction->force[a] = (float)r[a];
            reaction->torque[a] = (float)r[3 + a];
        }
        reaction->nodes_inside = (int32_t)std::llround(r[6]);
    }
    return BM_OK;
}

// defer_end: the body's end of substep is left to the next fast particle
// pass (h->pending_end), which runs it without a launch of its own.
static int couple_body_impl(bm_mpm_handle h, int32_t method, float dt, bm_reaction* out, bool defer_end) {
    if (method < BM_COUPLER_M0_DIRICHLET || method > BM_COUPLER_M45_BSPLINE)
        return fail(BM_ERR_INVALID_ARG, "unknown coupler %d", method);
    if (!h->body_set) return fail(BM_ERR_STATE, "no device body set");
    if (h->phase != PH_UPDATED) return fail(BM_ERR_STATE, "couple needs an updated grid");
    h->begin(BM_PHASE_COUPLE);
    h->pending_end = launch_couple_body(method, h->d_body, h->d_col, h->grid[h->cur], h->bever, h->dom, dt,
                                        h->body_blocks, h->d_partials, h->d_out7, h->d_err, h->body_begun, h->stream,
                                        defer_end && !out);
    h->body_begun = false;
    h->end();
    HIP_TRY(hipGetLastError());
    if (!out) return BM_OK;
    double r[7];
    HIP_TRY(hipMemcpyAsync(r, h->d_out7, sizeof(double) * 7, hipMemcpyDeviceToHost, h->stream));
    if (int rc = sync_and_check(h)) return rc;
    for (int a = 0; a < 3; ++a) {
        out->force[a] = (float)r[a];
        out->torque[a] = (float)r[3 + a];
    }
    out->nodes_inside = (int32_t)std::llround(r[6]);
    return BM_OK;
}

int bm_mpm_couple_body(bm_mpm_handle h, int32_t method, float dt, bm_reaction* out) {
    BM_TRACE("bm_mpm_couple_body");
    if (int rc = check_handle(h)) return rc;
    return couple_body_impl(h, method, dt, out, false);
}

int bm_mpm_step_coupled(bm_mpm_handle h, int32_t method, float dt, int32_t k) {
    BM_TRACE("bm_mpm_step_coupled");
    if (int rc = check_handle(h)) return rc;
    if (k < 0) return fail(BM_ERR_INVALID_ARG, "k < 0");
    if (method < BM_COUPLER_M0_DIRICHLET || method > BM_COUPLER_M45_BSPLINE)
        return fail(BM_ERR_INVALID_ARG, "unknown coupler %d", method);
    if (!h->body_set) return fail(BM_ERR_STATE, "no device body set");
    for (int32_t i = 0; i < k; ++i) {
        if (int rc = bm_mpm_p2g(h, dt)) return rc;
        if (int rc = grid_update_impl(h, dt, true)) return rc;
        // fast mode: the body's end of substep rides in the particle pass
        if (int rc = couple_body_impl(h, method, dt, nullptr, h->fast)) return rc;
        const int rc = g2p_impl(h, dt, i == k - 1);
        if (h->pending_end.body) {  // not taken by a particle pass (or the pass failed first)
            launch_body_end(h->pending_end, h->stream);
            h->pending_end = BodyEnd{};
        }
        if (rc) return rc;
    }
    return BM_OK;
}

// ------------------------------------------------------------------ grid access
int bm_mpm_download_grid(bm_mpm_handle h, const int32_t org[3], const int32_t dims[3], int64_t* mass_q, float* mass,
                         float* vel, float* force, uint8_t* active, uint8_t* alloc) {
    if (int rc = check_handle(h)) return rc;
    if (!org || !dims) return fail(BM_ERR_INVALID_ARG, "null box");
    for (int a = 0; a < 3; ++a)
        if (dims[a] < 0) return fail(BM_ERR_INVALID_ARG, "negative box");
    const int64_t tot = (int64_t)dims[0] * dims[1] * dims[2];
    if (tot == 0) return BM_OK;
    long long* d_mq = nullptr;
    float *d_mass = nullptr, *d_vel = nullptr, *d_force = nullptr;
    uint8_t *d_act = nullptr, *d_alloc = nullptr;
    int rc = BM_OK;
    auto cleanup = [&]() {
        hipFree(d_mq);
        hipFree(d_mass);
        hipFree(d_vel);
        hipFree(d_force);
        hipFree(d_act);
        hipFree(d_alloc);
    };
    if ((mass_q && hipMalloc(&d_mq, 8 * tot) != hipSuccess) || (mass && hipMalloc(&d_mass, 4 * tot) != hipSuccess) ||
        (vel && hipMalloc(&d_vel, 12 * tot) != hipSuccess) || (force && hipMalloc(&d_force, 12 * tot) != hipSuccess) ||
        (active && hipMalloc(&d_act, tot) != hipSuccess) || (alloc && hipMalloc(&d_alloc, tot) != hipSuccess)) {
        cleanup();
        return fail(BM_ERR_OOM, "grid download staging");
    }
    const GridBuf& g = h->grid[h->cur];
    launch_grid_gather(g, h->bever, h->dom, org, dims, h->qscale(), h->phase == PH_UPDATED ? 1 : 0, d_mq, d_mass,
                       d_vel, d_force, d_act, d_alloc, h->stream);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && mass_q) e = hipMemcpy(mass_q, d_mq, 8 * tot, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mass) e = hipMemcpy(mass, d_mass, 4 * tot, hipMemcpyDeviceToHost);
    if (e == hipSuccess && vel) e = hipMemcpy(vel, d_vel, 12 * tot, hipMemcpyDeviceToHost);
    if (e == hipSuccess && force) e = hipMemcpy(force, d_force, 12 * tot, hipMemcpyDeviceToHost);
    if (e == hipSuccess && active) e = hipMemcpy(active, d_act, tot, hipMemcpyDeviceToHost);
    if (e == hipSuccess && alloc) e = hipMemcpy(alloc, d_alloc, tot, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) rc = fail(BM_ERR_HIP, "grid download: %s", hipGetErrorString(e));
    return rc;
}

int bm_mpm_upload_grid(bm_mpm_handle h, const int32_t org[3], const int32_t dims[3], const float* mass,
                       const float* vel, const uint8_t* active) {
    if (int rc = check_handle(h)) return rc;
    if (!org || !dims || !mass || !vel || !active) return fail(BM_ERR_INVALID_ARG, "null argument");
    for (int a = 0; a < 3; ++a)
        if (dims[a] < 0) return fail(BM_ERR_INVALID_ARG, "negative box");
    const int64_t tot = (int64_t)dims[0] * dims[1] * dims[2];
    GridBuf& g = h->grid[h->cur];
    if (h->phase == PH_EMPTY) launch_clear(g, (int)h->nblocks, h->stream);
    if (tot > 0) {
        float *d_mass = nullptr, *d_vel = nullptr;
        uint8_t* d_act = nullptr;
        if (hipMalloc(&d_mass, 4 * tot) != hipSuccess || hipMalloc(&d_vel, 12 * tot) != hipSuccess ||
            hipMalloc(&d_act, tot) != hipSuccess) {
            hipFree(d_mass);
            hipFree(d_vel);
            hipFree(d_act);
            return fail(BM_ERR_OOM, "grid upload staging");
        }
        hipError_t e = hipMemcpyAsync(d_mass, mass, 4 * tot, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_vel, vel, 12 * tot, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_act, active, tot, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) {
            launch_grid_scatter(g, h->bever, h->dom, org, dims, h->qscale_inv(), d_mass, d_vel, d_act, h->stream);
            e = hipStreamSynchronize(h->stream);
        }
        hipFree(d_mass);
        hipFree(d_vel);
        hipFree(d_act);
        if (e != hipSuccess) return fail(BM_ERR_HIP, "grid upload: %s", hipGetErrorString(e));
    }
    h->prepared = false;
    h->phase = PH_UPDATED;
    return BM_OK;
}

int bm_mpm_grid_settle(bm_mpm_handle h, float dvy) {
    if (int rc = check_handle(h)) return rc;
    if (h->phase != PH_UPDATED) return fail(BM_ERR_STATE, "grid_settle needs an updated grid");
    launch_grid_settle(h->grid[h->cur], (int)std::min<int64_t>(h->nblocks, 1 << 30), dvy, h->stream);
    HIP_TRY(hipGetLastError());
    return BM_OK;
}

int bm_mpm_grid_mass_total(bm_mpm_handle h, int64_t* total) {
    if (int rc = check_handle(h)) return rc;
    if (!total) return fail(BM_ERR_INVALID_ARG, "null output");
    unsigned long long v = 0;
    HIP_TRY(hipMemsetAsync(h->d_scratch_u64, 0, sizeof(unsigned long long), h->stream));
    launch_mass_total(h->grid[h->cur], h->d_scratch_u64, h->stream);
    HIP_TRY(hipMemcpyAsync(&v, h->d_scratch_u64, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    if (int rc = sync_and_check(h)) return rc;
    *total = (int64_t)v;
    return BM_OK;
}

int bm_mpm_mass_quantum_exp(bm_mpm_handle h, int32_t* e) {
    if (int rc = check_handle(h)) return rc;
    if (!e) return fail(BM_ERR_INVALID_ARG, "null output");
    *e = h->q_exp;
    return BM_OK;
}

int bm_mpm_stats(bm_mpm_handle h, int64_t* active_blocks, int64_t* fp64_fallbacks) {
    if (int rc = check_handle(h)) return rc;
    if (int rc = sync_and_check(h)) return rc;
    if (active_blocks) {
        int c = 0;
        HIP_TRY(hipMemcpy(&c, h->grid[h->cur].bcount, sizeof(int), hipMemcpyDeviceToHost));
        *active_blocks = c;
    }
    if (fp64_fallbacks) {
        unsigned long long f = 0;
        HIP_TRY(hipMemcpy(&f, h->d_fallbacks, sizeof(f), hipMemcpyDeviceToHost));
        *fp64_fallbacks = (int64_t)f;
    }
    return BM_OK;
}

int bm_mpm_set_tuning(bm_mpm_handle h, int32_t unit_particles, int32_t sort_every, int32_t variant) {
    if (int rc = check_handle(h)) return rc;
    if (unit_particles > UNIT_MAX || (unit_particles > 0 && unit_particles < 16))
        return fail(BM_ERR_INVALID_ARG, "unit_particles must be in [16, %d]", UNIT_MAX);
    if (variant > 63 || (variant >= 0 && (variant & 48) == 48))
        return fail(BM_ERR_INVALID_ARG, "unknown variant %d", variant);
    if (variant >= 0 && !(variant & 4))
        return fail(BM_ERR_NOT_SUPPORTED, "variant %d: the single-kernel G2P+P2G form was retired (bit 2 must be set)",
                    variant);
    if (unit_particles > 0 && unit_particles != h->unit_p) {
        h->unit_p = unit_particles;
        h->binned = false;  // units must be rebuilt
    }
    if (sort_every > 0) h->sort_every = sort_every;
    if (variant >= 0) h->variant = variant;
    return BM_OK;
}

int bm_mpm_set_timing(bm_mpm_handle h, int32_t on) {
    if (int rc = check_handle(h)) return rc;
    if (on < 0 || on > 3) return fail(BM_ERR_INVALID_ARG, "timing level %d not in 0..3", on);
    h->timing = on;
    return BM_OK;
}

int bm_mpm_read_timing(bm_mpm_handle h, int32_t reset, double* ms, int64_t* launches) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->harvest();
    for (int i = 0; i < BM_NUM_PHASES; ++i) {
        if (ms) ms[i] = h->t_ms[i];
        if (launches) launches[i] = h->t_calls[i];
        if (reset) {
            h->t_ms[i] = 0.0;
            h->t_calls[i] = 0;
        }
    }
    return BM_OK;
}

}  // extern "C"
