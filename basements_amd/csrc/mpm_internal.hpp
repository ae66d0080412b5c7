// mpm_internal.hpp -- handle layout and kernel launchers shared by the
// translation units of libbm_mpm.so (host side only sees these declarations).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bm_mpm.h"
#include "particle_layout.hpp"

namespace bm {

// Record the last-error string of the C ABI and return code.
int fail(int code, const char* fmt, ...);
#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(BM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                 \
    } while (0)

// One node buffer of the dense block-linear grid.  Node (i,j,k) of the domain
// lives in block ((i>>2)*NBY + (j>>2))*NBZ + (k>>2), slot ((i&3)*4+(j&3))*4+(k&3)
// -- the reference's in-block order (spgrid_cpu.h:205-218).
struct GridBuf {
    long long* mq;    // fixed-point node mass
    float4* mv;       // xyz: momentum (after P2G, reference mode) / velocity (after update); w: node mass after update
    float4* f;        // xyz: stress force (reference mode)
    // Fast mode: the P2G momentum, one array per component (component c of
    // node n at mom[c * mstride + n]).  The deposit's three atomic adds per
    // node then land on three different cache lines, not on one 16-byte slot
    // (measured: the flush of the cell deposit 54 -> 19 us on the wheel bed);
    // the grid update reads it and writes the velocity into mv.
    float* mom;
    int64_t mstride;
    uint32_t* bflag;  // 1 while the block is in blist
    int32_t* blist;   // blocks touched since the last clear
    int32_t* bcount;  // length of blist
};

struct Domain {
    int ox, oy, oz;      // node index of the lowest corner
    int nx, ny, nz;      // nodes per axis (multiples of 4)
    int nbx, nby, nbz;   // blocks per axis
    float dx, inv_dx;
    int64_t b0;          // first block of this box in the grid arrays (0 for a lone scene)
};

// Particle SoA view: component c of particle i at p[c * cap + i]
// (component indices in particle_layout.hpp).

struct PView {
    float* p;
    int64_t cap;
    int64_t n;
};

struct Phys {
    float mu, lambda, alpha, kohesion;
    float gx, gy, gz;
    float mu_floor;
    int walls;
};

__host__ __device__ inline int64_t block_of(const Domain& d, int lx, int ly, int lz) {
    return d.b0 + ((int64_t)(lx >> 2) * d.nby + (ly >> 2)) * d.nbz + (lz >> 2);
}
// Block coordinates of a block of the grid arrays inside its box.
// 32-bit division: a box holds fewer than 2^31 blocks (64-bit division is a
// long software sequence on the GPU).
__host__ __device__ inline void block_coords(const Domain& d, int64_t blk, int& bx, int& by, int& bz) {
    const int l = (int)(blk - d.b0);
    const int q = l / d.nbz;
    bz = l - q * d.nbz;
    bx = q / d.nby;
    by = q - bx * d.nby;
}
__host__ __device__ inline int64_t node_of(const Domain& d, int lx, int ly, int lz) {
    return block_of(d, lx, ly, lz) * 64 + ((lx & 3) * 4 + (ly & 3)) * 4 + (lz & 3);
}

// Collider in the form the coupler kernel wants: rotation matrices built
// from the quaternion (matrix3.h:432-448), the |R| he world AABB
// (mpm_rigid_coupler.cpp:70-76) and the block range to scan.
struct ColliderDev {
    float c[3];
    float Rl[9];  // local -> world
    float Rw[9];  // world -> local (transpose)
    float he[3];
    float vl[3], wa[3];
    float friction;
    float mass;          // body mass (the M4.5 couplers spread it over the inside nodes)
    float mb;            // M1 per-node body mass
    float lo[3], hi[3];  // world AABB of the body
    int shape;
    int b0[3];  // first block (domain-local block coordinates) of the scanned range
    int nb[3];  // blocks in the range per axis
};

// RigidColliderState -> ColliderDev, on the host for bm_mpm_couple and on the
// device for the device body: the same FP32 operations, no contraction, so
// both give the same bits.
__host__ __device__ inline ColliderDev collider_build(const float center[3], const float q[4], const float he[3],
                                                      const float vl[3], const float wa[3], float friction,
                                                      float mass, int shape, const Domain& d) {
#pragma clang fp contract(off)
    ColliderDev c{};
    const float w = q[0], x = q[1], y = q[2], z = q[3];
    // Matrix3::from_quaternion, matrix3.h:432-448
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y,
                wz = w * z;
    const float R[9] = {1.0f - 2.0f * (yy + zz), 2.0f * (xy - wz),        2.0f * (xz + wy),
                        2.0f * (xy + wz),        1.0f - 2.0f * (xx + zz), 2.0f * (yz - wx),
                        2.0f * (xz - wy),        2.0f * (yz + wx),        1.0f - 2.0f * (xx + yy)};
    for (int i = 0; i < 9; ++i) c.Rl[i] = R[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c.Rw[i * 3 + j] = R[j * 3 + i];
    for (int a = 0; a < 3; ++a) {
        c.c[a] = center[a];
        c.he[a] = he[a];
        c.vl[a] = vl[a];
        c.wa[a] = wa[a];
    }
    c.friction = friction;
    c.mass = mass;
    c.shape = shape;
    // M1 per-node body mass from the box volume, mpm_rigid_coupler_pic.h:104-108
    const float dx = d.dx;
    const float vol = 8.0f * he[0] * he[1] * he[2];
    const int approx = (int)roundf(vol / (dx * dx * dx));
    c.mb = mass / (float)(approx > 1 ? approx : 1);
    // |R| he world AABB, mpm_rigid_coupler.cpp:70-76
    for (int a = 0; a < 3; ++a) {
        const float hw = fabsf(R[a * 3 + 0]) * c.he[0] + fabsf(R[a * 3 + 1]) * c.he[1] + fabsf(R[a * 3 + 2]) * c.he[2];
        c.lo[a] = c.c[a] - hw;
        c.hi[a] = c.c[a] + hw;
    }
    // block range: every block whose node-centre box may overlap the AABB,
    // widened by one block; the kernel applies the reference's exact cull.
    const int org[3] = {d.ox, d.oy, d.oz};
    const int nbv[3] = {d.nbx, d.nby, d.nbz};
    for (int a = 0; a < 3; ++a) {
        const double lo_n = floor((double)c.lo[a] / dx - 0.5) - org[a];
        const double hi_n = ceil((double)c.hi[a] / dx - 0.5) - org[a];
        long long blo = (long long)floor(lo_n / 4.0) - 1;
        long long bhi = (long long)floor(hi_n / 4.0) + 1;
        blo = blo > 0 ? blo : 0;
        bhi = bhi < nbv[a] - 1 ? bhi : nbv[a] - 1;
        c.b0[a] = (int)blo;
        c.nb[a] = (bhi >= blo) ? (int)(bhi - blo + 1) : 0;
    }
    return c;
}

// Blocks of the coupler's scan range for any pose of a body with half
// extents he: each axis of the |R| he box is at most |he| long each way.
inline int64_t couple_blocks_bound(const float he[3], const Domain& d) {
    const double r = sqrt((double)he[0] * he[0] + (double)he[1] * he[1] + (double)he[2] * he[2]);
    const double nodes = ceil(2.0 * r / (double)d.dx) + 4.0;
    const int64_t per = (int64_t)(nodes / 4.0) + 5;
    const int nbv[3] = {d.nbx, d.nby, d.nbz};
    int64_t tot = 1;
    for (int a = 0; a < 3; ++a) tot *= per < nbv[a] ? per : nbv[a];
    return tot;
}

// One scene of a batch: what the kernels need that differs per scene.
struct SceneDev {
    Domain d;  // the scene's node box; d.b0 = its first block in the shared grid arrays
    Phys ph;
    double qscale, qscale_inv;  // mass quantum and its inverse
    int64_t p0, n;              // the scene's particle range
};
// Scene lookup of a launch over a batch.  All null for a lone scene: then
// the Domain / Phys passed by value apply.
struct SceneView {
    const SceneDev* sc;      // scene table
    const int* block_scene;  // scene of each block of the grid arrays
    const int4* chunks;      // (scene, first particle, count, 0): one per workgroup of the particle kernels
    int n_scenes;
};

// Device error flags: err[0] collects every code; a batch also keeps one word
// per scene at err[1 + scene] so that a failing scene names itself.
__device__ __forceinline__ void flag_error(int* err, int scene, int code) {
    atomicOr(err, code);
    if (scene >= 0) atomicOr(err + 1 + scene, code);
}
// Error codes: 1 a stencil left the box, 4 a collider outgrew its scan
// bound, 8 a particle's new position or velocity is not finite.
__device__ __forceinline__ void check_finite(int* err, int scene, float x, float y, float z, float vx, float vy,
                                             float vz) {
    if (!isfinite(((x + y) + (z + vx)) + (vy + vz))) flag_error(err, scene, 8);
}

// ------------------------------------------------------------- launchers
// Zero the blocks on g's list and empty the list; also zeroes *zero_word
// (a counter of the next kernels) when given.  defer_reset: the list length
// is left for the next kernel to zero (the fast particle pass does).
void launch_clear(const GridBuf& g, int nblocks, hipStream_t s, int* zero_word = nullptr, bool defer_reset = false);
void launch_p2g_ref(const PView& P, const GridBuf& g, uint8_t* bever, const Domain& d, const Phys& ph,
                    double qscale_inv, int* err, hipStream_t s);
// Grid update; with a device body also the start of its coupled substep
// (applied loads, collider), as k_body_begin.
// Over a batch (sv.sc set) each block takes its scene's box and physics, and
// body[s] / col[s] are the bodies of the scenes with has_body[s] != 0.
// clear_next: also clear that grid's listed blocks (as launch_clear with
// defer_reset, for a grid without a force array) and zero *zero_word.
void launch_update(const GridBuf& g, int nblocks, const Domain& d, const Phys& ph, double qscale, float dt,
                   hipStream_t s, bm_body* body = nullptr, ColliderDev* col = nullptr, int64_t max_blocks = 0,
                   int* err = nullptr, SceneView sv = SceneView{}, const int* has_body = nullptr,
                   const GridBuf* clear_next = nullptr, int* zero_word = nullptr);
void launch_g2p_ref(const PView& P, const GridBuf& g, const Domain& d, const Phys& ph, float dt, int* err,
                    hipStream_t s);
// The couplers run a fixed number of workgroups over the scanned blocks, so
// the FP64 partial sums -- and the reaction -- do not depend on how the scan
// range was bounded (host-built or device-built collider).
constexpr int COUPLE_WGS = 128;
// Coupler workgroups per scene of a batch: a tool's scan range is a few
// blocks, so most of COUPLE_WGS would find nothing to do in every scene.
constexpr int COUPLE_BATCH_WGS = 16;
// A device body's end of substep (body_device.hpp: sum the coupler's partial
// reactions, integrate) handed to the particle pass, which runs it in the
// first workgroup of its FP64 redo launch.  body == nullptr: none.
struct BodyEnd {
    bm_body* body = nullptr;
    const double* partials = nullptr;
    int nparts = 0;
    double* out7 = nullptr;
    float dt = 0.0f;
};

void launch_couple(int method, const ColliderDev& b, const GridBuf& g, uint8_t* bever, const Domain& d, float dt,
                   double* partials, double* out7, hipStream_t s);
// Coupled substep of the device body: applied loads, collider from the body
// pose, coupler over at most max_blocks blocks (a bound from
// couple_blocks_bound), reaction into the body, integration.  err bit 2 is
// set if the collider's range exceeded max_blocks (then nothing is coupled).
// begun = true: the start was already done (fused into launch_update).
// defer_end: the body's end of substep is not launched but returned, for the
// next particle pass to run (launch_split); otherwise it is launched here.
BodyEnd launch_couple_body(int method, bm_body* body, ColliderDev* col, const GridBuf& g, uint8_t* bever,
                           const Domain& d, float dt, int64_t max_blocks, double* partials, double* out7, int* err,
                           bool begun, hipStream_t s, bool defer_end = false);
void launch_body_end(const BodyEnd& be, hipStream_t s);
// The device bodies of a batch (scene s: bodies[s], cols[s] built by the
// update, methods[s]); partials hold COUPLE_WGS * 7 doubles and out7 8
// doubles per scene.
void launch_couple_batch(const int* methods, const int* has_body, bm_body* bodies, const ColliderDev* cols,
                         const GridBuf& g, uint8_t* bever, SceneView sv, float dt, double* partials, double* out7,
                         hipStream_t s);
void launch_grid_settle(const GridBuf& g, int nblocks, float dvy, hipStream_t s);
void launch_grid_gather(const GridBuf& g, uint8_t* bever, const Domain& d, const int* org, const int* dims,
                        double qscale, int updated, long long* o_mq, float* o_mass, float* o_vel, float* o_force,
                        uint8_t* o_act, uint8_t* o_alloc, hipStream_t s);
void launch_grid_scatter(const GridBuf& g, uint8_t* bever, const Domain& d, const int* org, const int* dims,
                         double qscale_inv, const float* i_mass, const float* i_vel, const uint8_t* i_act,
                         hipStream_t s);
void launch_mass_total(const GridBuf& g, unsigned long long* out, hipStream_t s);
void launch_scale_v0(const PView& P, float rescale, hipStream_t s);
void launch_max_mass(const PView& P, unsigned int* out, hipStream_t s);

// ------------------------------------------------------------- fast path
// Scratch of the particle re-binning (all device memory, sized at create).
struct BinScratch {
    int* bcnt;      // particles per block
    int* bstart;    // exclusive scan of bcnt
    int* ucnt;      // units per block
    int* ustart;    // exclusive scan of ucnt
    int* key;       // block of each particle
    int* slot;      // rank of each particle inside its block
    int4* units;    // (block, first particle, count, 0)
    int* n_units;   // device copy of the unit count
    void* temp;     // hipcub scan scratch
    size_t temp_bytes;
};
size_t scan_temp_bytes(int n);
// Over a batch (sv.chunks set) k_bin runs one workgroup per chunk in its
// scene's box.  n_units_host = nullptr skips the read-back (no host sync).
// with_vc = false does not carry the velocity and C components over: only
// for a re-binning that a G2P follows (it rewrites them before they are read).
int launch_rebin(const PView& P, float* pdst, const int* pid_src, int* pid_dst, const Domain& d, int nblocks,
                 int unit_p, BinScratch& bs, hipStream_t s, int* n_units_host, SceneView sv = SceneView{},
                 int n_chunks = 0, bool with_vc = true);
void launch_iota(int* p, int64_t n, int64_t first, hipStream_t s);
// Slots [base, base + n) of P into insertion order in dst (same pitch):
// dst[c * cap + pid[base + i]] = P[c * cap + base + i].
void launch_unpermute(const PView& P, float* dst, const int* pid, int64_t base, int64_t n, hipStream_t s);
// Largest unit the fast kernels accept: a particle of at most 2^23 - 1 mass
// quanta gives a node at most (27/64) q_p + 13 of them, so 1,024 particles
// keep every LDS node-mass sum below 2^32 (1,213 would).  The default unit
// (640) holds a whole 4^3 block of the 8-per-cell bed with the drift of a
// re-binning period: blocks of 512 + a few particles split into two units
// at 512, the second nearly empty but staging and flushing a tile of its own.
constexpr int UNIT_MAX = 1024;
// The fast substep's two kernels: per-particle G2P writing a 16-float
// deposit record per particle into `coef`, then the per-cell block P2G.
// `mark(ctx, stage, start, stop)` is called on the host right before each
// stage's launch (the timing hook of the handle); for the two main kernels it
// may hand back a pair of events that the dispatch itself records.
enum { SPLIT_ADVANCE = 0, SPLIT_REDO = 1, SPLIT_DEPOSIT = 2 };
// A batch: the particle kernels run one workgroup per chunk, and the deposit
// kernel reads the unit count on the device with n_units_bound workgroups.
struct SplitBatch {
    SceneView sv{};
    int n_chunks = 0;
    const int* n_units_dev = nullptr;
    int n_units_bound = 0;
    int deposit_groups = 4096;  // fixed grid of the deposit when the count is on the device
};
// store_vc = false: the G2P does not write the particle velocity and C.  Only
// a download and a deposit of the current state (do_g2p = false) read them,
// and the deposit record carries what the next P2G needs, so a multi-substep
// call stores them on its last substep only.
void launch_split(bool do_g2p, const PView& P, const int4* units, int n_units, const GridBuf& gin,
                  const GridBuf& gout, uint8_t* bever, const Domain& d, const Phys& ph, float dt_g2p, float dt_fold,
                  double qscale_inv, int* err, unsigned long long* counters, int* redo, int* n_redo, float* coef,
                  hipStream_t s, void (*mark)(void*, int, hipEvent_t*, hipEvent_t*) = nullptr,
                  void* ctx = nullptr, bool two_pass = false, const SplitBatch& batch = SplitBatch{},
                  bool store_vc = true, bool g2p_tile = true, const BodyEnd& body_end = BodyEnd{});

// Units below which a single scene's G2P runs per particle over the flat
// particle range: with few units the workgroups of the staged-tile G2P do not
// hide its staging latency.  Measured per substep on dragged-tool scenes: the
// per-particle form faster by 3 us from 2e4 to 5e5 particles, by 1 us at 1e6
// (3,240 blocks), even at 2e6 (5,808 blocks); the staged form 15 us faster on
// the 4.1e6-particle wheel bed (11,264 blocks).
constexpr int G2P_TILE_MIN_UNITS = 6144;

// Batch entry points of the request loop (bm_batch.hip): admission of a scene
// whose component-major particle state (PCOMP rows of n floats) is already on
// the device, or on its way there behind `ready`; and the copy of a scene's
// state, in insertion order, into device staging on the batch stream, with
// nothing waiting (the scene stays in the batch).
hipStream_t batch_stream(bm_batch_handle b);
int batch_quantum_exp(const float* m, int64_t n);
int batch_add_scene_staged(bm_batch_handle b, const bm_scene_desc* sd, int64_t n, int q_exp, const float* dev_soa,
                           hipEvent_t ready, const bm_body* body, int32_t* scene_id);
// base_out: the scene's first arena slot; the staged order entries are
// insertion index + base.
int batch_take_scene(bm_batch_handle b, int32_t id, float* dev_soa, int* dev_order, int64_t* n_out,
                     int64_t* base_out);

}  // namespace bm
