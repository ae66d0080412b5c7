// mpm_split.hip -- the two-kernel form of the fast substep (tuning variant
// bit 2): a G2P kernel over the re-binning units (each unit's grid nodes
// staged in LDS, one lane per particle) that also writes each particle's
// deposit coefficients, then a per-block P2G kernel that ranks the particles
// by their *final* stencil cell and deposits with ordered LDS
// read-modify-writes.
//
// Compared with a fused G2P+P2G kernel it moves 128 more bytes per particle
// (the deposit record written and read back) but neither kernel holds the
// other's registers: the G2P kernel needs 8 KiB of LDS and no deposit
// accumulators, and the P2G kernel no SVD.
//
// Deposit record: 16 floats per particle as four float4 quads
//   (x, y, z, q_p)  (m v, A00)  (A01, A02, A10, A11)  (A12, A20, A21, A22)
// stored quad-major within groups of 64 particles (quad q of particle i at
// float4 index ((i / 64) * 4 + q) * 64 + i % 64): a wave writes each quad as
// 1 KiB of contiguous memory, and the P2G kernel reads a particle with four
// wide loads.
// with q_p the particle's mass quanta (an integer < 2^23, exact in FP32) and
// A = m C + dt K row-major, K = -4/dx^2 V0 tau.

#include <hip/hip_ext.h>

#include <algorithm>

#include "body_device.hpp"
#include "mpm_fast_device.hpp"

namespace bm {

constexpr int NCOEF = 12;
constexpr int SUNIT_MAX = 1024;   // largest unit (see UNIT_MAX)

__device__ __forceinline__ int stix(int ix, int iy, int iz) { return ((ix * TE + iy) * TE) + ((iz + 4 * (iy & 1)) & 7); }
__device__ __forceinline__ void stile_coords(int k, int& ix, int& iy, int& iz) {
    ix = k >> 6;
    iy = (k >> 3) & 7;
    iz = ((k & 7) - 4 * (iy & 1)) & 7;
}

__device__ __forceinline__ int64_t rec_at(int64_t i, int q) { return ((i >> 6) * 4 + q) * 64 + (i & 63); }
__device__ __forceinline__ void store_rec(float4* rec, int64_t i, float x, float y, float z, float qp,
                                          const v3<float>& mvel, const m3<float>& A) {
    rec[rec_at(i, 0)] = make_float4(x, y, z, qp);
    rec[rec_at(i, 1)] = make_float4(mvel.x, mvel.y, mvel.z, A.a[0]);
    rec[rec_at(i, 2)] = make_float4(A.a[1], A.a[2], A.a[3], A.a[4]);
    rec[rec_at(i, 3)] = make_float4(A.a[5], A.a[6], A.a[7], A.a[8]);
}
struct DepositRec {
    float x, y, z;
    float qp;
    v3<float> mvel;
    m3<float> A;
};
__device__ __forceinline__ DepositRec load_rec(const float4* __restrict__ rec, int64_t i) {
    const float4 r0 = rec[rec_at(i, 0)], r1 = rec[rec_at(i, 1)], r2 = rec[rec_at(i, 2)], r3 = rec[rec_at(i, 3)];
    DepositRec d;
    d.x = r0.x;
    d.y = r0.y;
    d.z = r0.z;
    d.qp = r0.w;
    d.mvel = mk(r1.x, r1.y, r1.z);
    d.A = m3<float>{{r1.w, r2.x, r2.y, r2.z, r2.w, r3.x, r3.y, r3.z, r3.w}};
    return d;
}
// The same record with its first quad (position, quanta) from LDS copies.
__device__ __forceinline__ DepositRec load_rec_lds(const float4* __restrict__ rec, int64_t i, const float* px,
                                                   const float* py, const float* pz, const float* pq, int k) {
    const float4 r1 = rec[rec_at(i, 1)], r2 = rec[rec_at(i, 2)], r3 = rec[rec_at(i, 3)];
    DepositRec d;
    d.x = px[k];
    d.y = py[k];
    d.z = pz[k];
    d.qp = pq[k];
    d.mvel = mk(r1.x, r1.y, r1.z);
    d.A = m3<float>{{r1.w, r2.x, r2.y, r2.z, r2.w, r3.x, r3.y, r3.z, r3.w}};
    return d;
}
__device__ __forceinline__ float quanta_of(float m, double qscale_inv) { return (float)rint((double)m * qscale_inv); }

__global__ void k_set_zero(int* p) { *p = 0; }

// Workgroups are dealt to the 8 XCDs round-robin (w % 8).  This bijection of
// [0, G) gives XCD x a contiguous run of logical indices, so the particles of
// neighbouring blocks -- which gather the same grid nodes -- are served from
// one XCD's L2 instead of being spread over all eight.
__device__ __forceinline__ int xcd_slot(int w, int G) {
    const int x = w & 7, q = w >> 3, base = G >> 3, rem = G & 7;
    return x * base + min(x, rem) + q;
}

// The state a G2P reads of particle i: position, F, mass and volume.
struct G2PIn {
    float x, y, z, m, V0;
    m3<float> F;
};
__device__ __forceinline__ G2PIn g2p_load(const PView& P, int64_t i) {
    G2PIn q;
    q.x = P.p[(PX + 0) * P.cap + i];
    q.y = P.p[(PX + 1) * P.cap + i];
    q.z = P.p[(PX + 2) * P.cap + i];
#pragma unroll
    for (int c = 0; c < 9; ++c) q.F.a[c] = P.p[(PF + c) * P.cap + i];
    q.m = P.p[PM * P.cap + i];
    q.V0 = P.p[PV0 * P.cap + i];
    return q;
}

// G2P + plasticity + write-back + deposit coefficients of particle i from its
// loaded state, with GATHER(s, lx, ly, lz, bad, vel, C) the G2P gather of its
// stencil (base lx, ly, lz in box nodes; `bad` when the stencil leaves the
// box, in which case the gather must still read in bounds).
template <class GatherFn>
__device__ __forceinline__ void g2p_particle(int64_t i, const G2PIn& in, const PView& P, const Domain& d,
                                             const Phys& ph, const FastArgs& fa, double qscale_inv, float4* rec,
                                             int* err, int scene, GatherFn gather) {
    const float inv_dx = d.inv_dx;
    float x = in.x, y = in.y, z = in.z;
    m3<float> F = in.F;
    const float m = in.m, V0 = in.V0;
    v3<float> vel, lns;
    m3<float> Cm, Us;
    Stencil s;
    make_stencil(x, y, z, inv_dx, s);
    const int lx = s.b[0] - d.ox, ly = s.b[1] - d.oy, lz = s.b[2] - d.oz;
    // A stencil outside the box is gathered in bounds and rejected after the
    // gather, so no branch separates the loads above from the gather's and
    // all of them are in flight together.
    const bool bad = !stencil_ok(d, lx, ly, lz);
    gather(s, lx, ly, lz, bad, vel, Cm);
    if (bad) {
        flag_error(err, scene, 1);
        store_rec(rec, i, x, y, z, 0.0f, mk(0.f, 0.f, 0.f), m3<float>{{0, 0, 0, 0, 0, 0, 0, 0, 0}});
        return;
    }
    const float dt = fa.dt_g2p;
    x += vel.x * dt;
    y += vel.y * dt;
    z += vel.z * dt;
    m3<float> A;
#pragma unroll
    for (int c = 0; c < 9; ++c) A.a[c] = ((c % 4 == 0) ? 1.0f : 0.0f) + Cm.a[c] * dt;
    const m3<float> Ftr = mmul(A, F);
    m3<float> V;
    v3<float> sg;
    if (!svd_fast(Ftr, Us, sg, V)) {
        fa.redo[atomicAdd(fa.n_redo, 1)] = (int)i;  // redone from scratch in FP64
        return;
    }
    const Material M{ph.mu, ph.lambda, ph.alpha, ph.kohesion};
    const v3<float> sn = dp_fast(sg, M, lns);
    F = recompose(Us, sn, V);
    if (y < 0.0f) {
        y = 0.0f;
        vel.y = 0.0f;
    }
    check_finite(err, scene, x, y, z, vel.x, vel.y, vel.z);
    P.p[(PX + 0) * P.cap + i] = x;
    P.p[(PX + 1) * P.cap + i] = y;
    P.p[(PX + 2) * P.cap + i] = z;
#pragma unroll
    for (int c = 0; c < 9; ++c) P.p[(PF + c) * P.cap + i] = F.a[c];
    if (fa.store_vc) {
        P.p[(PV + 0) * P.cap + i] = vel.x;
        P.p[(PV + 1) * P.cap + i] = vel.y;
        P.p[(PV + 2) * P.cap + i] = vel.z;
#pragma unroll
        for (int c = 0; c < 9; ++c) P.p[(PC + c) * P.cap + i] = Cm.a[c];
    }
    v3<float> mvel;
    m3<float> Aff;
    p2g_coeffs(m, V0, vel, Cm, Us, lns, ph, inv_dx, fa.dt_fold, mvel, Aff);
    store_rec(rec, i, x, y, z, quanta_of(m, qscale_inv), mvel, Aff);
}

// The G2P gather straight from the block-linear grid (27 global loads).
struct GlobalGather {
    const float4* __restrict__ g;
    Domain d;
    __device__ __forceinline__ void operator()(const Stencil& s, int lx, int ly, int lz, bool bad, v3<float>& vel,
                                               m3<float>& Cm) const {
        const NodeAxes na(d, bad ? 0 : lx, bad ? 0 : ly, bad ? 0 : lz);
        gather_sep(s, d.dx, 4.0f * d.inv_dx * d.inv_dx, [&](int a, int b, int c) { return g[na(a, b, c)]; }, vel, Cm);
    }
};

// One lane per particle over the whole particle range: the G2P (DO_G2P), or
// the deposit record of the current state (velocity and C as stored).
template <bool DO_G2P>
__global__ void __launch_bounds__(256) k_advance(PView P, GridBuf gin, Domain d, Phys ph, FastArgs fa, float4* rec,
                                                 int* err, SceneView sv) {
    // the clear before this pass left the output grid's block list to be
    // emptied here, before the deposit kernel appends to it
    if (blockIdx.x == 0 && threadIdx.x == 0 && fa.list_reset) *fa.list_reset = 0;
    const int wg = xcd_slot((int)blockIdx.x, (int)gridDim.x);
    int64_t i = (int64_t)wg * blockDim.x + threadIdx.x;
    int scene = -1;
    if (sv.chunks) {  // a batch: one scene per workgroup
        const int4 ch = sv.chunks[wg];
        if ((int)threadIdx.x >= ch.z) return;
        i = ch.y + threadIdx.x;
        scene = ch.x;
        const SceneDev& S = sv.sc[ch.x];
        d = S.d;
        ph = S.ph;
        fa.qscale_inv = S.qscale_inv;
    } else if (i >= P.n) {
        return;
    }
    if (DO_G2P) {
        // every load of the particle is issued before the gather's (the state
        // stores may alias P.p, so later loads would wait for the whole G2P)
        g2p_particle(i, g2p_load(P, i), P, d, ph, fa, fa.qscale_inv, rec, err, scene, GlobalGather{gin.mv, d});
        return;
    }
    const float inv_dx = d.inv_dx;
    const float m = P.p[PM * P.cap + i], V0 = P.p[PV0 * P.cap + i];
    const float x = P.p[(PX + 0) * P.cap + i], y = P.p[(PX + 1) * P.cap + i], z = P.p[(PX + 2) * P.cap + i];
    m3<float> F;
#pragma unroll
    for (int c = 0; c < 9; ++c) F.a[c] = P.p[(PF + c) * P.cap + i];
    const v3<float> vel = mk(P.p[(PV + 0) * P.cap + i], P.p[(PV + 1) * P.cap + i], P.p[(PV + 2) * P.cap + i]);
    m3<float> Cm, Us, V;
#pragma unroll
    for (int c = 0; c < 9; ++c) Cm.a[c] = P.p[(PC + c) * P.cap + i];
    v3<float> sg;
    if (!svd_fast(F, Us, sg, V)) {
        fa.redo[atomicAdd(fa.n_redo, 1)] = (int)i;
        return;
    }
    const v3<float> lns = mk(flog(fmaxf(sg.x, 1e-6f)), flog(fmaxf(sg.y, 1e-6f)), flog(fmaxf(sg.z, 1e-6f)));
    v3<float> mvel;
    m3<float> Aff;
    p2g_coeffs(m, V0, vel, Cm, Us, lns, ph, inv_dx, fa.dt_fold, mvel, Aff);
    store_rec(rec, i, x, y, z, quanta_of(m, fa.qscale_inv), mvel, Aff);
}

// The G2P over the re-binning units: a unit's particles share one 4^3-node
// block (the block of their stencil base at the last re-binning), so their
// stencils lie in the 8^3 nodes around it (the block, 2 stencil layers and 1
// drift layer on each side).  The workgroup stages those nodes' velocities
// in LDS once and every particle gathers its 27 nodes from there at constant
// offsets from one address; a particle that drifted further gathers from the
// grid.  One staged tile replaces ~27 x 12 B of L1 traffic per particle with
// 8^3 x 16 B per unit.  Units are dealt XCD-contiguously (neighbouring units
// stage overlapping nodes, which then come from one L2).  With the unit count
// on the device only (a batch) a fixed grid walks the units.
__global__ void __launch_bounds__(256) k_advance_tile(PView P, GridBuf gin, Domain d, Phys ph, FastArgs fa, float4* rec,
                                                      int* err, SceneView sv, const int4* __restrict__ units) {
    __shared__ float4 tile[TNODES];
    if (blockIdx.x == 0 && threadIdx.x == 0 && fa.list_reset) *fa.list_reset = 0;
    const int nu = fa.n_units_dev ? *fa.n_units_dev : fa.n_units;
    for (int u = xcd_slot((int)blockIdx.x, (int)gridDim.x); u < nu; u += gridDim.x) {
        const int4 U = units[u];
        Domain dd = d;
        Phys pp = ph;
        double qsi = fa.qscale_inv;
        int scene = -1;
        if (sv.sc) {  // a batch: the unit's scene
            const SceneDev& S = sv.sc[U.w];
            dd = S.d;
            pp = S.ph;
            qsi = S.qscale_inv;
            scene = U.w;
        }
        int bx, by, bz;
        block_coords(dd, U.x, bx, by, bz);
        const int tx0 = bx * 4 - 1, ty0 = by * 4 - 1, tz0 = bz * 4 - 1;
        __syncthreads();  // the previous unit's gathers are done with the tile
        for (int k = threadIdx.x; k < TNODES; k += blockDim.x) {
            const int lx = tx0 + (k >> 6), ly = ty0 + ((k >> 3) & 7), lz = tz0 + (k & 7);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lx >= 0 && ly >= 0 && lz >= 0 && lx < dd.nx && ly < dd.ny && lz < dd.nz) v = gin.mv[node_of(dd, lx, ly, lz)];
            tile[k] = v;
        }
        __syncthreads();
        const float4* gm = gin.mv;
        auto gather = [&](const Stencil& s, int lx, int ly, int lz, bool bad, v3<float>& vel, m3<float>& Cm) {
            const int qx = lx - tx0, qy = ly - ty0, qz = lz - tz0;
            const float kC = 4.0f * dd.inv_dx * dd.inv_dx;
            if ((unsigned)qx <= TE - 3 && (unsigned)qy <= TE - 3 && (unsigned)qz <= TE - 3) {
                const float4* t = tile + (qx * TE + qy) * TE + qz;
                gather_sep(s, dd.dx, kC, [&](int a, int b, int c) { return t[(a * TE + b) * TE + c]; }, vel, Cm);
            } else {
                GlobalGather{gm, dd}(s, lx, ly, lz, bad, vel, Cm);
            }
        };
        for (int j = threadIdx.x; j < U.z; j += blockDim.x) {
            const int64_t i = (int64_t)U.y + j;
            g2p_particle(i, g2p_load(P, i), P, dd, pp, fa, qsi, rec, err, scene, gather);
        }
    }
}

// The FP64 pass for the particles k_advance listed: same outputs.
template <bool DO_G2P>
__global__ void __launch_bounds__(64) k_advance_redo(PView P, GridBuf gin, Domain d0, Phys ph0, FastArgs fa, float4* rec,
                                                     int* err, unsigned long long* counters, SceneView sv,
                                                     BodyEnd be) {
    // the coupled substep's body end rides in the first workgroup (it only
    // has to precede the next grid update)
    if (be.body && blockIdx.x == 0) body_end(be.body, be.partials, be.nparts, be.out7, be.dt);
    const int cnt = *fa.n_redo;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < cnt; r += gridDim.x * blockDim.x) {
        const int64_t i = fa.redo[r];
        Domain d = d0;
        Phys ph = ph0;
        int scene = -1;
        if (sv.sc) {  // the scene whose particle range holds i
            int lo = 0, hi = sv.n_scenes - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sv.sc[mid].p0 <= i) lo = mid;
                else hi = mid - 1;
            }
            d = sv.sc[lo].d;
            ph = sv.sc[lo].ph;
            fa.qscale_inv = sv.sc[lo].qscale_inv;
            scene = lo;
        }
        const float dx = d.dx, inv_dx = d.inv_dx;
        float x = P.p[(PX + 0) * P.cap + i], y = P.p[(PX + 1) * P.cap + i], z = P.p[(PX + 2) * P.cap + i];
        m3<float> F;
#pragma unroll
        for (int c = 0; c < 9; ++c) F.a[c] = P.p[(PF + c) * P.cap + i];
        v3<float> vel, lns;
        m3<float> Cm, Us;
        if (DO_G2P) {
            Stencil s;
            make_stencil(x, y, z, inv_dx, s);
            const int lx = s.b[0] - d.ox, ly = s.b[1] - d.oy, lz = s.b[2] - d.oz;
            gather(s, dx, 4.0f * inv_dx * inv_dx,
                   [&](int a, int b, int c) { return gin.mv[node_of(d, lx + a, ly + b, lz + c)]; }, vel, Cm);
            const float dt = fa.dt_g2p;
            x += vel.x * dt;
            y += vel.y * dt;
            z += vel.z * dt;
            m3<float> A;
#pragma unroll
            for (int c = 0; c < 9; ++c) A.a[c] = ((c % 4 == 0) ? 1.0f : 0.0f) + Cm.a[c] * dt;
            const m3<float> Ftr = mmul(A, F);
            m3<float> V;
            v3<float> sg;
            svd_guarded(Ftr, Us, sg, V);
            const Material M{ph.mu, ph.lambda, ph.alpha, ph.kohesion};
            const v3<float> sn = dp_project(sg, M, &lns);
            F = recompose(Us, sn, V);
            if (y < 0.0f) {
                y = 0.0f;
                vel.y = 0.0f;
            }
            check_finite(err, sv.sc ? scene : -1, x, y, z, vel.x, vel.y, vel.z);
            P.p[(PX + 0) * P.cap + i] = x;
            P.p[(PX + 1) * P.cap + i] = y;
            P.p[(PX + 2) * P.cap + i] = z;
            P.p[(PV + 0) * P.cap + i] = vel.x;
            P.p[(PV + 1) * P.cap + i] = vel.y;
            P.p[(PV + 2) * P.cap + i] = vel.z;
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                P.p[(PF + c) * P.cap + i] = F.a[c];
                P.p[(PC + c) * P.cap + i] = Cm.a[c];
            }
        } else {
            vel = mk(P.p[(PV + 0) * P.cap + i], P.p[(PV + 1) * P.cap + i], P.p[(PV + 2) * P.cap + i]);
#pragma unroll
            for (int c = 0; c < 9; ++c) Cm.a[c] = P.p[(PC + c) * P.cap + i];
            m3<float> V;
            v3<float> sg;
            svd_guarded(F, Us, sg, V);
            lns = mk(logf(fmaxf(sg.x, 1e-6f)), logf(fmaxf(sg.y, 1e-6f)), logf(fmaxf(sg.z, 1e-6f)));
        }
        v3<float> mvel;
        m3<float> Aff;
        const float m = P.p[PM * P.cap + i];
        p2g_coeffs(m, P.p[PV0 * P.cap + i], vel, Cm, Us, lns, ph, inv_dx, fa.dt_fold, mvel, Aff);
        store_rec(rec, i, x, y, z, quanta_of(m, fa.qscale_inv), mvel, Aff);
        atomicAdd(&counters[0], 1ull);
    }
}

// P2G of one unit (one wave) with per-cell register accumulation.  The
// unit's particles are counting-sorted by the cell of their final stencil
// base; lane c then owns cell c of the block, runs through that cell's
// particles accumulating all 27 node deposits (momentum and mass quanta) in
// registers -- they share their 27 nodes -- and adds the totals into the
// LDS tile with one ordered read-modify-write chain.  For a fixed stencil
// offset the 64 lanes address 64 different nodes, and program order across
// offsets keeps neighbouring cells' shared nodes consistent.
// Units of at most SPOS particles keep the positions and quanta the count
// phase loaded for the accumulation (x, y, z in the tile's space, which the
// chain only needs afterwards): one scattered 16-byte load per particle less.
constexpr int SPOS = 640;
struct CellSmem {
    union {
        float4 pt[TNODES];   // momentum xyz; w = mass quanta (uint, < 2^32 for <= UNIT_MAX particles)
        float pos[3][SPOS];  // before the chain: the unit's particle positions
    };
    float pq[SPOS];
    unsigned short ord[SUNIT_MAX];
    unsigned short rank[SUNIT_MAX];
    unsigned char cell[SUNIT_MAX];
    unsigned short left[SUNIT_MAX];
    int cnt[64];
    int start[64];
    int nleft;
    int tb[27];
};

// Deposits of slabs [A0, A1) of the stencils of one cell's particles,
// accumulated in registers (node n = ((a - A0) * 3 + b) * 3 + c).  The
// centre (1, 1, 1) takes each particle's remaining quanta.  With the whole
// stencil in range the centre's quanta are formed once per cell, after the
// loop: the cell's total quanta less the other 26 nodes' sums (every particle
// deposits exactly q_p, so the per-particle remainders add up to that).  With
// a part of the stencil, the quanta of the other slabs are formed per particle
// when slab 1 is in range.  Same arithmetic as scatter() for the nodes it
// covers.
template <int A0, int A1>
__device__ __forceinline__ void accum_slabs(const float4* __restrict__ rec, const int4& U, const unsigned short* ord,
                                            int s0, int mine, float dx, float inv_dx, v3<float>* acc,
                                            unsigned* accm, const CellSmem* lp) {
#pragma unroll
    for (int k = 0; k < (A1 - A0) * 9; ++k) {
        acc[k] = mk(0.f, 0.f, 0.f);
        accm[k] = 0u;
    }
    constexpr bool CENTRE = A0 <= 1 && 1 < A1;
    constexpr bool FULL = A0 == 0 && A1 == 3;
    constexpr int NC = ((1 - A0) * 3 + 1) * 3 + 1;  // the centre's register slot (when CENTRE)
    unsigned qsum = 0u;                              // the cell's quanta (FULL)
    // the next particle's record is fetched while this one is deposited
    auto fetch = [&](int k) {
        const int pi = ord[s0 + k];
        return lp ? load_rec_lds(rec, (int64_t)U.y + pi, lp->pos[0], lp->pos[1], lp->pos[2], lp->pq, pi)
                  : load_rec(rec, (int64_t)U.y + pi);
    };
    DepositRec nxt;
    if (mine > 0) nxt = fetch(0);
    for (int k = 0; k < mine; ++k) {
        const DepositRec r = nxt;
        if (k + 1 < mine) nxt = fetch(k + 1);
        Stencil s;
        make_stencil(r.x, r.y, r.z, inv_dx, s);
        const v3<float>& mvel = r.mvel;
        const m3<float>& Aff = r.A;
        const float qp = r.qp;
        int rest = 0;
        if (CENTRE && !FULL) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (a >= A0 && a < A1) continue;
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int c = 0; c < 3; ++c) rest += mass_quanta(s.w[0][a] * s.w[1][b] * s.w[2][c], qp);
            }
        }
        if (FULL) qsum += (unsigned)qp;
        // node offsets ((b + a) - f - 1/2) dx formed as ((a - fr) - 1/2) dx, fr = f - b: the same
        // values (fr is exact, and so are a - fr and the subtraction of 1/2, all below 2 in magnitude)
        const float fr0 = s.f[0] - (float)s.b[0], fr1 = s.f[1] - (float)s.b[1], fr2 = s.f[2] - (float)s.b[2];
#pragma unroll
        for (int a = A0; a < A1; ++a) {
            const float ddx = (((float)a - fr0) - 0.5f) * dx;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const float ddy = (((float)b - fr1) - 0.5f) * dx;
                const float px = mvel.x + Aff.a[0] * ddx + Aff.a[1] * ddy;
                const float py = mvel.y + Aff.a[3] * ddx + Aff.a[4] * ddy;
                const float pz = mvel.z + Aff.a[6] * ddx + Aff.a[7] * ddy;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float ddz = (((float)c - fr2) - 0.5f) * dx;
                    const float w = s.w[0][a] * s.w[1][b] * s.w[2][c];
                    const v3<float> mom =
                        mk(w * (px + Aff.a[2] * ddz), w * (py + Aff.a[5] * ddz), w * (pz + Aff.a[8] * ddz));
                    const int n = ((a - A0) * 3 + b) * 3 + c;
                    acc[n] = acc[n] + mom;
                    if (a == 1 && b == 1 && c == 1) continue;  // the centre's quanta: below
                    const int dq = mass_quanta(w, qp);
                    if (!FULL) rest += dq;
                    accm[n] += (unsigned)dq;
                }
            }
        }
        if (CENTRE && !FULL) accm[NC] += (unsigned)((int)qp - rest);
    }
    if (FULL) {  // modulo 2^32: the centre's true sum, a node sum, is below 2^32 (UNIT_MAX)
        unsigned others = 0u;
#pragma unroll
        for (int n = 0; n < 27; ++n)
            if (n != NC) others += accm[n];
        accm[NC] = qsum - others;
    }
}

// Add the register sums of slabs [A0, A1) into the LDS tile: for a fixed
// stencil offset the 64 lanes (cells) address 64 different nodes, and program
// order across offsets keeps neighbouring cells' shared nodes consistent.
template <int A0, int A1>
__device__ __forceinline__ void chain_slabs(float4* pt, int lane, const v3<float>* acc, const unsigned* accm) {
    const int rx = (lane >> 4) + 1, ry = ((lane >> 2) & 3) + 1, rz = (lane & 3) + 1;
#pragma unroll
    for (int a = A0; a < A1; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int n = ((a - A0) * 3 + b) * 3 + c;
                const int k = stix(rx + a, ry + b, rz + c);
                float4 q = pt[k];
                q.x += acc[n].x;
                q.y += acc[n].y;
                q.z += acc[n].z;
                q.w = __uint_as_float(__float_as_uint(q.w) + accm[n]);
                pt[k] = q;
                __asm__ volatile("" ::: "memory");
            }
}

// One unit's deposit.  NPASS = 1: all 27 nodes in one pass over the cell's
// particles.  NPASS = 2: slabs 0-1, then slab 2 in a second pass (fewer live
// registers, more waves per SIMD, the particle data read twice).
template <int NPASS>
__device__ __forceinline__ void p2g_unit(int u, CellSmem& S, const float4* __restrict__ rec,
                                         const int4* __restrict__ units, const GridBuf& gout, uint8_t* bever,
                                         Domain d, const FastArgs& fa, int* err, unsigned long long* counters,
                                         const SceneView& sv) {
    const int4 U = units[u];
    if (sv.sc) d = sv.sc[U.w].d;  // a batch: the unit's scene
    const int blk = U.x;
    int bx, by, bz;
    block_coords(d, blk, bx, by, bz);
    const int tx0 = bx * 4 - 1, ty0 = by * 4 - 1, tz0 = bz * 4 - 1;
    const int lane = threadIdx.x;
    const float dx = d.dx, inv_dx = d.inv_dx;

    // one pass with positions in LDS (they share the tile's space until the chain)
    const bool lds_pos = NPASS == 1 && U.z <= SPOS;
    if (!lds_pos)
        for (int k = lane; k < TNODES; k += 64) S.pt[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    S.cnt[lane] = 0;
    if (lane < 27) S.tb[lane] = 0;
    if (lane == 0) S.nleft = 0;
    __syncthreads();

    // ---- count the particles per cell of the block (a unit holds at most
    // SUNIT_MAX = 16 x 64 particles: all of a lane's positions are in flight at once)
    float4 pos[SUNIT_MAX / 64];
#pragma unroll
    for (int j = 0; j < SUNIT_MAX / 64; ++j) {
        const int pi = lane + 64 * j;
        if (pi < U.z) pos[j] = rec[rec_at((int64_t)U.y + pi, 0)];
    }
#pragma unroll
    for (int j = 0; j < SUNIT_MAX / 64; ++j) {
        const int pi = lane + 64 * j;
        if (pi >= U.z) break;
        const float4 r0 = pos[j];
        if (lds_pos) {
            S.pos[0][pi] = r0.x;
            S.pos[1][pi] = r0.y;
            S.pos[2][pi] = r0.z;
            S.pq[pi] = r0.w;
        }
        Stencil s;
        make_stencil(r0.x, r0.y, r0.z, inv_dx, s);
        const int lx = s.b[0] - d.ox, ly = s.b[1] - d.oy, lz = s.b[2] - d.oz;
        const int cx = lx - tx0 - 1, cy = ly - ty0 - 1, cz = lz - tz0 - 1;
        // a stencil that leaves the box (a particle the G2P just moved into
        // the box's last two cells) goes to the checked path below: the tile
        // flush writes every node it holds
        if (cx >= 0 && cy >= 0 && cz >= 0 && cx < 4 && cy < 4 && cz < 4 && stencil_ok(d, lx, ly, lz)) {
            const int c = (cx * 4 + cy) * 4 + cz;
            S.cell[pi] = (unsigned char)c;
            S.rank[pi] = (unsigned short)atomicAdd(&S.cnt[c], 1);
        } else {
            S.cell[pi] = 255;
            S.left[atomicAdd(&S.nleft, 1)] = (unsigned short)pi;
        }
    }
    __syncthreads();
    // exclusive scan of the 64 counts across the wave
    const int mine = S.cnt[lane];
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    S.start[lane] = incl - mine;
    __syncthreads();
    for (int pi = lane; pi < U.z; pi += 64) {
        const int c = S.cell[pi];
        if (c != 255) S.ord[S.start[c] + S.rank[pi]] = (unsigned short)pi;
    }
    __syncthreads();

    // ---- per-cell accumulation in registers
    const int s0 = S.start[lane];
    if (NPASS == 1) {
        v3<float> acc[27];
        unsigned accm[27];
        accum_slabs<0, 3>(rec, U, S.ord, s0, mine, dx, inv_dx, acc, accm, lds_pos ? &S : nullptr);
        if (lds_pos) {  // every lane is done with the positions: the tile takes their space
            __syncthreads();
            for (int k = lane; k < TNODES; k += 64) S.pt[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            __syncthreads();
        }
        if (mine > 0) chain_slabs<0, 3>(S.pt, lane, acc, accm);
    } else {
        v3<float> acc[18];
        unsigned accm[18];
        accum_slabs<0, 2>(rec, U, S.ord, s0, mine, dx, inv_dx, acc, accm, nullptr);
        if (mine > 0) chain_slabs<0, 2>(S.pt, lane, acc, accm);
        accum_slabs<2, 3>(rec, U, S.ord, s0, mine, dx, inv_dx, acc, accm, nullptr);
        if (mine > 0) chain_slabs<2, 3>(S.pt, lane, acc, accm);
    }
    __syncthreads();

    // ---- particles outside the block's cells (drifted since the last
    // re-binning) or with a stencil outside the box (flagged, not deposited):
    // LDS atomics inside the tile, global atomics outside
    int off = 0;
    const int nl = S.nleft;
    for (int k = lane; k < nl; k += 64) {
        const DepositRec r = load_rec(rec, (int64_t)U.y + S.left[k]);
        Stencil s;
        make_stencil(r.x, r.y, r.z, inv_dx, s);
        const v3<float>& mvel = r.mvel;
        const m3<float>& Aff = r.A;
        const float qp = r.qp;
        const int lx = s.b[0] - d.ox, ly = s.b[1] - d.oy, lz = s.b[2] - d.oz;
        if (!stencil_ok(d, lx, ly, lz)) {
            flag_error(err, sv.sc ? U.w : -1, 1);
            continue;
        }
        const int qx = lx - tx0, qy = ly - ty0, qz = lz - tz0;
        if (qx >= 0 && qy >= 0 && qz >= 0 && qx <= TE - 3 && qy <= TE - 3 && qz <= TE - 3) {
            scatter(s, dx, qp, mvel, Aff, [&](int a, int b, int c, float w, int dq, v3<float> mom) {
                float* q = (float*)&S.pt[stix(qx + a, qy + b, qz + c)];
                atomicAdd(q + 0, mom.x);
                atomicAdd(q + 1, mom.y);
                atomicAdd(q + 2, mom.z);
                atomicAdd((unsigned*)(q + 3), (unsigned)dq);
            });
        } else {
            ++off;
            scatter(s, dx, qp, mvel, Aff, [&](int a, int b, int c, float w, int dq, v3<float> mom) {
                if (w != 0.0f) deposit_global(gout, bever, d, lx + a, ly + b, lz + c, dq, mom);
            });
        }
    }
    if (off) atomicAdd(&counters[1], (unsigned long long)off);
    __syncthreads();

    // ---- flush the tile into the next grid
    for (int k = lane; k < TNODES; k += 64) {
        const float4 q = S.pt[k];
        const unsigned m = __float_as_uint(q.w);
        if (m == 0u) continue;
        int ix, iy, iz;
        stile_coords(k, ix, iy, iz);
        const int lx = tx0 + ix, ly = ty0 + iy, lz = tz0 + iz;
        const int64_t nid = node_of(d, lx, ly, lz);
        atomicAdd((unsigned long long*)&gout.mq[nid], (unsigned long long)m);
        atomicAdd(gout.mom + nid, q.x);
        atomicAdd(gout.mom + gout.mstride + nid, q.y);
        atomicAdd(gout.mom + 2 * gout.mstride + nid, q.z);
        S.tb[(((lx >> 2) - bx + 1) * 3 + ((ly >> 2) - by + 1)) * 3 + ((lz >> 2) - bz + 1)] = 1;
    }
    __syncthreads();
    if (lane < 27 && S.tb[lane]) {
        const int qx = lane / 9, qy = (lane / 3) % 3, qz = lane % 3;
        touch_block_f(gout, bever, (int)(d.b0 + ((int64_t)(bx + qx - 1) * d.nby + (by + qy - 1)) * d.nbz + (bz + qz - 1)));
    }
}

template <int NPASS>
__global__ void __launch_bounds__(64) k_p2g_cell(const float4* __restrict__ rec,
                                                 const int4* __restrict__ units, GridBuf gout, uint8_t* bever,
                                                 Domain d, FastArgs fa, int* err, unsigned long long* counters,
                                                 SceneView sv) {
    __shared__ CellSmem S;
    // Units in list order, dealt round-robin to the XCDs: the neighbouring
    // units that run at the same time then sit on different XCDs, so their
    // flushes of shared halo nodes do not queue on the same atomics (measured
    // faster than giving each XCD a contiguous share of the list).
    if (!fa.n_units_dev) {  // one workgroup per unit
        const int u = blockIdx.x;
        if (u >= fa.n_units) return;
        p2g_unit<NPASS>(u, S, rec, units, gout, bever, d, fa, err, counters, sv);
        return;
    }
    // The unit count is known on the device only (a batch): a fixed grid
    // walks the units.
    const int nu = *fa.n_units_dev;
    for (int u = blockIdx.x; u < nu; u += gridDim.x) {
        p2g_unit<NPASS>(u, S, rec, units, gout, bever, d, fa, err, counters, sv);
        __syncthreads();  // the next unit reuses the LDS
    }
}


// G2P (or the coefficients of the current state when !do_g2p), the FP64
// pass for the particles it listed, then the ranked P2G.
void launch_split(bool do_g2p, const PView& P, const int4* units, int n_units, const GridBuf& gin,
                  const GridBuf& gout, uint8_t* bever, const Domain& d, const Phys& ph, float dt_g2p, float dt_fold,
                  double qscale_inv, int* err, unsigned long long* counters, int* redo, int* n_redo, float* coef,
                  hipStream_t s, void (*mark)(void*, int, hipEvent_t*, hipEvent_t*), void* ctx, bool two_pass,
                  const SplitBatch& batch, bool store_vc, bool g2p_tile, const BodyEnd& be) {
    const bool batched = batch.sv.chunks != nullptr;
    if (batched ? batch.n_chunks <= 0 : P.n <= 0) {  // no particle pass to empty the output grid's block list
        hipLaunchKernelGGL(k_set_zero, dim3(1), dim3(1), 0, s, gout.bcount);
        launch_body_end(be, s);
        return;
    }
    float4* rec = (float4*)coef;  // 16 floats per particle (whole groups of 64: see rec_at)
    FastArgs fa{dt_g2p, dt_fold, n_units, qscale_inv, redo, n_redo, gout.bcount, batch.n_units_dev, store_vc ? 1 : 0};
    // n_redo was zeroed by the grid clear that precedes every fast pass
    const int g = batched ? batch.n_chunks : (int)((P.n + 255) / 256);
    const SceneView sv = batch.sv;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (mark) mark(ctx, SPLIT_ADVANCE, &e0, &e1);
    // with timing events the launch records them around the dispatch itself.
    // The G2P runs over the units (LDS-staged node tiles): one workgroup per
    // unit, or a fixed grid walking a count known on the device only.
    const int tg = batch.n_units_dev ? std::min(batch.n_units_bound, batch.deposit_groups) : n_units;
    if (do_g2p && g2p_tile && tg > 0) {
        if (e0)
            hipExtLaunchKernelGGL(k_advance_tile, dim3(tg), dim3(256), 0, s, e0, e1, 0, P, gin, d, ph, fa, rec, err,
                                  sv, units);
        else
            hipLaunchKernelGGL(k_advance_tile, dim3(tg), dim3(256), 0, s, P, gin, d, ph, fa, rec, err, sv, units);
    } else if (e0) {
        if (do_g2p)
            hipExtLaunchKernelGGL(k_advance<true>, dim3(g), dim3(256), 0, s, e0, e1, 0, P, gin, d, ph, fa, rec, err,
                                  sv);
        else
            hipExtLaunchKernelGGL(k_advance<false>, dim3(g), dim3(256), 0, s, e0, e1, 0, P, gin, d, ph, fa, rec, err,
                                  sv);
    } else if (do_g2p) {
        hipLaunchKernelGGL(k_advance<true>, dim3(g), dim3(256), 0, s, P, gin, d, ph, fa, rec, err, sv);
    } else {
        hipLaunchKernelGGL(k_advance<false>, dim3(g), dim3(256), 0, s, P, gin, d, ph, fa, rec, err, sv);
    }
    if (mark) mark(ctx, SPLIT_REDO, nullptr, nullptr);
    if (do_g2p)
        hipLaunchKernelGGL(k_advance_redo<true>, dim3(32), dim3(64), 0, s, P, gin, d, ph, fa, rec, err, counters, sv,
                           be);
    else
        hipLaunchKernelGGL(k_advance_redo<false>, dim3(32), dim3(64), 0, s, P, gin, d, ph, fa, rec, err, counters, sv,
                           be);
    e0 = e1 = nullptr;
    if (mark) mark(ctx, SPLIT_DEPOSIT, &e0, &e1);
    // a batch walks the device-counted units with a fixed grid
    const int nu = batch.n_units_dev ? std::min(batch.n_units_bound, batch.deposit_groups) : n_units;
    if (nu > 0) {
        const int nwg = ((nu + 7) / 8) * 8;
        auto kern = two_pass ? k_p2g_cell<2> : k_p2g_cell<1>;
        if (e0)
            hipExtLaunchKernelGGL(kern, dim3(nwg), dim3(64), 0, s, e0, e1, 0, (const float4*)rec, units, gout, bever,
                                  d, fa, err, counters, sv);
        else
            hipLaunchKernelGGL(kern, dim3(nwg), dim3(64), 0, s, (const float4*)rec, units, gout, bever, d, fa, err,
                               counters, sv);
    }
}

}  // namespace bm
