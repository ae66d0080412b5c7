// mpm_ref_kernels.hip -- grid bookkeeping, the reference-order substep
// kernels, and the M0 / M1 couplers.
//
// Built with -ffp-contract=off so every a*b+c is rounded the way the
// reference's scalar C++ rounds it (mpm_solver.h, svd.h compiled without FMA).
// The grid is dense over a fixed domain and block-linear (4^3-node blocks, the
// reference's block size, spgrid_cpu.h:12-15).  A block list per grid buffer
// records the blocks P2G touched so that the clear and the grid update visit
// only those blocks; a persistent per-block "ever touched" flag reproduces the
// reference's never-freed block set (spgrid_cpu.h:46-65), which decides the
// node set M1 may claim.

#include "mpm_device.hpp"
#include "mpm_internal.hpp"
#include "body_device.hpp"

namespace bm {

// --------------------------------------------------------------- bookkeeping
// Zero every node of the blocks listed in g and drop their flags.  One wave
// per block; grid-stride over the device-side count.  The list length is
// zeroed by a later kernel (k_reset_count or the fast particle pass).
// zero_word, when given, is a counter of a later kernel zeroed on the way.
__global__ void __launch_bounds__(64) k_clear_blocks(GridBuf g, int* zero_word) {
    const int cnt = *g.bcount;
    for (int b = blockIdx.x; b < cnt; b += gridDim.x) {
        const int blk = g.blist[b];
        const int64_t nid = (int64_t)blk * 64 + threadIdx.x;
        g.mq[nid] = 0;
        g.mv[nid] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (g.f) g.f[nid] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (g.mom) g.mom[nid] = g.mom[g.mstride + nid] = g.mom[2 * g.mstride + nid] = 0.f;
        if (threadIdx.x == 0) g.bflag[blk] = 0u;
    }
    if (zero_word && blockIdx.x == 0 && threadIdx.x == 0) *zero_word = 0;
}

__global__ void k_reset_count(GridBuf g) { *g.bcount = 0; }

// Record that block blk holds data of this grid use.
__device__ __forceinline__ void touch_block(const GridBuf& g, uint8_t* bever, int blk) {
    if (g.bflag[blk] == 0u && atomicExch(&g.bflag[blk], 1u) == 0u) {
        const int slot = atomicAdd(g.bcount, 1);
        g.blist[slot] = blk;
    }
    if (bever && bever[blk] == 0) bever[blk] = 1;
}

__device__ __forceinline__ bool stencil_in_domain(const Domain& d, const Stencil& s) {
    const int lx = s.b[0] - d.ox, ly = s.b[1] - d.oy, lz = s.b[2] - d.oz;
    return lx >= 0 && ly >= 0 && lz >= 0 && lx + 2 < d.nx && ly + 2 < d.ny && lz + 2 < d.nz;
}

__device__ __forceinline__ m3<float> load_m3(const PView& P, int comp, int64_t i) {
    m3<float> r;
#pragma unroll
    for (int c = 0; c < 9; ++c) r.a[c] = P.p[(comp + c) * P.cap + i];
    return r;
}
__device__ __forceinline__ void store_m3(const PView& P, int comp, int64_t i, const m3<float>& r) {
#pragma unroll
    for (int c = 0; c < 9; ++c) P.p[(comp + c) * P.cap + i] = r.a[c];
}

// --------------------------------------------------------------- P2G (reference order)
// mpm_solver.h:168-248 with the fixed-point mass deposit.  One particle per
// lane; node accumulation with device-scope float / int64 atomics.
__global__ void __launch_bounds__(256) k_p2g_ref(PView P, GridBuf g, uint8_t* bever, Domain d, Phys ph,
                                                 double qscale_inv, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const float x = P.p[(PX + 0) * P.cap + i], y = P.p[(PX + 1) * P.cap + i], z = P.p[(PX + 2) * P.cap + i];
    Stencil s;
    make_stencil(x, y, z, d.inv_dx, s);
    if (!stencil_in_domain(d, s)) {
        atomicOr(err, 1);
        return;
    }
    const v3<float> vel = mk(P.p[(PV + 0) * P.cap + i], P.p[(PV + 1) * P.cap + i], P.p[(PV + 2) * P.cap + i]);
    const m3<float> F = load_m3(P, PF, i);
    const m3<float> Cm = load_m3(P, PC, i);
    const float mass = P.p[PM * P.cap + i];
    const float V0 = P.p[PV0 * P.cap + i];

    m3<float> U, V;
    v3<float> sg;
    svd_ref(F, U, sg, V);
    sg.x = fmaxf(sg.x, 1e-6f);
    sg.y = fmaxf(sg.y, 1e-6f);
    sg.z = fmaxf(sg.z, 1e-6f);
    const m3<float> tau = hencky_from_svd(U, sg, ph.mu, ph.lambda);
    const float ff = -4.0f * d.inv_dx * d.inv_dx * V0;
    m3<float> K;
#pragma unroll
    for (int c = 0; c < 9; ++c) K.a[c] = tau.a[c] * ff;

    // integer quanta of this particle; the stencil centre takes the remainder
    const float qp = (float)rint((double)mass * qscale_inv);
    long long rest = (long long)qp;
#pragma unroll
    for (int a = 0; a < 27; ++a) {
        if (a == 13) continue;
        const float w = s.w[0][a / 9] * s.w[1][(a / 3) % 3] * s.w[2][a % 3];
        rest -= mass_quanta(w, qp);
    }

    const int lx0 = s.b[0] - d.ox, ly0 = s.b[1] - d.oy, lz0 = s.b[2] - d.oz;
    int last_blk = -1;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float w = s.w[0][a] * s.w[1][b] * s.w[2][c];
                if (w == 0.0f) continue;
                const v3<float> dd = mk(((float)(s.b[0] + a) - s.f[0] - 0.5f) * d.dx,
                                        ((float)(s.b[1] + b) - s.f[1] - 0.5f) * d.dx,
                                        ((float)(s.b[2] + c) - s.f[2] - 0.5f) * d.dx);
                const v3<float> aff = mvec(Cm, dd);
                const float mc = w * mass;
                const v3<float> mom = (vel + aff) * mc;
                const v3<float> fo = mvec(K, dd) * w;
                const int64_t nid = node_of(d, lx0 + a, ly0 + b, lz0 + c);
                const long long dq = (a == 1 && b == 1 && c == 1) ? rest : mass_quanta(w, qp);
                atomicAdd((unsigned long long*)&g.mq[nid], (unsigned long long)dq);
                float* nm = (float*)&g.mv[nid];
                atomicAdd(nm + 0, mom.x);
                atomicAdd(nm + 1, mom.y);
                atomicAdd(nm + 2, mom.z);
                float* nf = (float*)&g.f[nid];
                atomicAdd(nf + 0, fo.x);
                atomicAdd(nf + 1, fo.y);
                atomicAdd(nf + 2, fo.z);
                const int blk = (int)(nid >> 6);
                if (blk != last_blk) {
                    touch_block(g, bever, blk);
                    last_blk = blk;
                }
            }
}

__device__ void body_begin(bm_body* body, ColliderDev* col, const Domain& d, long long max_blocks, int* err,
                           int scene = -1);

// --------------------------------------------------------------- grid update
// mpm_solver.h:250-291 plus the optional domain walls.  A force buffer is
// folded in when present (reference mode); in fast mode P2G already folded
// the force into the momentum.  With a device body, one lane also starts the
// body's coupled substep (body_begin), which touches no node.

__global__ void __launch_bounds__(64) k_update(GridBuf g, Domain d0, Phys ph0, double qscale0, float dt, bm_body* body,
                                               ColliderDev* col, long long max_blocks, int* err, SceneView sv,
                                               const int* has_body, GridBuf clr, int* zero_word) {
    // the other grid's clear rides along (what k_clear_blocks does, block
    // list length left to the next particle pass)
    if (clr.mq) {
        const int cc = *clr.bcount;
        for (int b = blockIdx.x; b < cc; b += gridDim.x) {
            const int blk = clr.blist[b];
            const int64_t nid = (int64_t)blk * 64 + threadIdx.x;
            clr.mq[nid] = 0;
            clr.mv[nid] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (clr.mom) clr.mom[nid] = clr.mom[clr.mstride + nid] = clr.mom[2 * clr.mstride + nid] = 0.f;
            if (threadIdx.x == 0) clr.bflag[blk] = 0u;
        }
        if (zero_word && blockIdx.x == 0 && threadIdx.x == 0) *zero_word = 0;
    }
    if (body && threadIdx.x == 0) {
        if (!sv.sc) {
            if (blockIdx.x == 0) body_begin(body, col, d0, max_blocks, err);
        } else {  // one body per scene that has one (the launch has at least n_scenes workgroups)
            const int s = blockIdx.x;
            if (s < sv.n_scenes && has_body[s]) body_begin(body + s, col + s, sv.sc[s].d, max_blocks, err, s);
        }
    }
    const int cnt = *g.bcount;
    for (int b = blockIdx.x; b < cnt; b += gridDim.x) {
        const int blk = g.blist[b];
        Domain d = d0;
        Phys ph = ph0;
        double qscale = qscale0;
        if (sv.block_scene) {
            const SceneDev& S = sv.sc[sv.block_scene[blk]];
            d = S.d;
            ph = S.ph;
            qscale = S.qscale;
        }
        const int64_t nid = (int64_t)blk * 64 + threadIdx.x;
        const long long mq = g.mq[nid];
        if (mq == 0) continue;  // inactive node (no particle weight)
        // the deposit's momentum: per-component arrays (fast mode) or mv
        float4 q = g.mom ? make_float4(g.mom[nid], g.mom[g.mstride + nid], g.mom[2 * g.mstride + nid], 0.f)
                         : g.mv[nid];
        const float m = (float)((double)mq * qscale);
        q.w = m;
        if (m > 1e-6f) {
            v3<float> mom = mk(q.x, q.y, q.z);
            if (g.f) {
                const float4 fo = g.f[nid];
                mom = mom + mk(fo.x, fo.y, fo.z) * dt;
            }
            v3<float> v = divg(mom, m);
            v = v + mk(ph.gx * dt, ph.gy * dt, ph.gz * dt);
            int bx, by, bz;
            block_coords(d, blk, bx, by, bz);
            const int t = threadIdx.x;
            const int gj = d.oy + by * 4 + ((t >> 2) & 3);
            const float node_y = ((float)gj + 0.5f) * d.dx;
            if (node_y < d.dx && v.y < 0.0f) {
                const float vn = -v.y;
                v.y = 0.0f;
                const float tx = v.x, tz = v.z;
                const float tm = sqrtf(tx * tx + tz * tz);
                const float budget = ph.mu_floor * vn;
                if (tm > 1e-12f) {
                    const float sc = fmaxf(0.0f, (tm - budget) / tm);
                    v.x = tx * sc;
                    v.z = tz * sc;
                }
            }
            if (ph.walls > 0) {
                const int lx = bx * 4 + (t >> 4), ly = by * 4 + ((t >> 2) & 3), lz = bz * 4 + (t & 3);
                const int W = ph.walls;
                if ((lx < W && v.x < 0.f) || (lx >= d.nx - W && v.x > 0.f)) v.x = 0.f;
                if ((ly < W && v.y < 0.f) || (ly >= d.ny - W && v.y > 0.f)) v.y = 0.f;
                if ((lz < W && v.z < 0.f) || (lz >= d.nz - W && v.z > 0.f)) v.z = 0.f;
            }
            q.x = v.x;
            q.y = v.y;
            q.z = v.z;
        }
        g.mv[nid] = q;
    }
}

// --------------------------------------------------------------- G2P (reference order)
// mpm_solver.h:293-348 and apply_plasticity :114-166.
__global__ void __launch_bounds__(256) k_g2p_ref(PView P, GridBuf g, Domain d, Phys ph, float dt, int* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    float x = P.p[(PX + 0) * P.cap + i], y = P.p[(PX + 1) * P.cap + i], z = P.p[(PX + 2) * P.cap + i];
    Stencil s;
    make_stencil(x, y, z, d.inv_dx, s);
    if (!stencil_in_domain(d, s)) {
        atomicOr(err, 1);
        return;
    }
    const int lx0 = s.b[0] - d.ox, ly0 = s.b[1] - d.oy, lz0 = s.b[2] - d.oz;
    const float kC = 4.0f * d.inv_dx * d.inv_dx;
    v3<float> v = mk(0.f, 0.f, 0.f);
    m3<float> Cm;
#pragma unroll
    for (int c = 0; c < 9; ++c) Cm.a[c] = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float w = s.w[0][a] * s.w[1][b] * s.w[2][c];
                if (w == 0.0f) continue;
                const float4 q = g.mv[node_of(d, lx0 + a, ly0 + b, lz0 + c)];
                const v3<float> dd = mk(((float)(s.b[0] + a) - s.f[0] - 0.5f) * d.dx,
                                        ((float)(s.b[1] + b) - s.f[1] - 0.5f) * d.dx,
                                        ((float)(s.b[2] + c) - s.f[2] - 0.5f) * d.dx);
                const v3<float> wv = mk(q.x, q.y, q.z) * w;
                v = v + wv;
                const float wvv[3] = {wv.x, wv.y, wv.z}, ddv[3] = {dd.x, dd.y, dd.z};
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) Cm.a[r * 3 + cc] = Cm.a[r * 3 + cc] + (wvv[r] * ddv[cc]) * kC;
            }
    x = x + v.x * dt;
    y = y + v.y * dt;
    z = z + v.z * dt;

    const m3<float> F = load_m3(P, PF, i);
    m3<float> A;
#pragma unroll
    for (int c = 0; c < 9; ++c) A.a[c] = ((c % 4 == 0) ? 1.0f : 0.0f) + Cm.a[c] * dt;
    const m3<float> Ftr = mmul(A, F);
    m3<float> U, V;
    v3<float> sg;
    svd_ref(Ftr, U, sg, V);
    const Material M{ph.mu, ph.lambda, ph.alpha, ph.kohesion};
    const v3<float> sn = dp_project(sg, M);
    const m3<float> Fn = recompose(U, sn, V);

    if (y < 0.0f) {
        y = 0.0f;
        v.y = 0.0f;
    }
    check_finite(err, -1, x, y, z, v.x, v.y, v.z);
    P.p[(PX + 0) * P.cap + i] = x;
    P.p[(PX + 1) * P.cap + i] = y;
    P.p[(PX + 2) * P.cap + i] = z;
    P.p[(PV + 0) * P.cap + i] = v.x;
    P.p[(PV + 1) * P.cap + i] = v.y;
    P.p[(PV + 2) * P.cap + i] = v.z;
    store_m3(P, PF, i, Fn);
    store_m3(P, PC, i, Cm);
}

// --------------------------------------------------------------- couplers
__device__ __forceinline__ bool shape_inside_d(const ColliderDev& b, v3<float> l) {
    if (b.shape == 1) {
        const float R = b.he[0], L = b.he[1];
        const float r2 = l.x * l.x + l.z * l.z;
        return fabsf(l.y) <= L && r2 <= R * R;
    }
    return fabsf(l.x) <= b.he[0] && fabsf(l.y) <= b.he[1] && fabsf(l.z) <= b.he[2];
}
__device__ __forceinline__ v3<float> shape_normal_d(const ColliderDev& b, v3<float> l) {
    if (b.shape == 1) {
        const float R = b.he[0], L = b.he[1];
        const float r = sqrtf(l.x * l.x + l.z * l.z);
        const float ad = L - fabsf(l.y), rd = R - r;
        if (ad < rd) return mk(0.f, l.y >= 0.f ? 1.f : -1.f, 0.f);
        if (r > 1e-9f) return mk(l.x / r, 0.f, l.z / r);
        return mk(1.f, 0.f, 0.f);
    }
    const float ddx = b.he[0] - fabsf(l.x), ddy = b.he[1] - fabsf(l.y), ddz = b.he[2] - fabsf(l.z);
    if (ddx <= ddy && ddx <= ddz) return mk(l.x >= 0.f ? 1.f : -1.f, 0.f, 0.f);
    if (ddy <= ddz) return mk(0.f, l.y >= 0.f ? 1.f : -1.f, 0.f);
    return mk(0.f, 0.f, l.z >= 0.f ? 1.f : -1.f);
}

// Cell-overlap fraction of a node inside the body (shape_cell_overlap_fraction,
// mpm_rigid_coupler.h:139-158).
__device__ __forceinline__ float overlap_fraction_d(const ColliderDev& b, v3<float> l) {
    auto clamp01 = [](float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); };
    if (b.shape == 1) {
        const float R = b.he[0], L = b.he[1];
        const float r = sqrtf(l.x * l.x + l.z * l.z);
        const float axial = 1.0f - fabsf(l.y) / fmaxf(L, 1e-9f);
        const float radial = 1.0f - r / fmaxf(R, 1e-9f);
        return clamp01(axial) * clamp01(radial);
    }
    const float ux = clamp01(1.0f - fabsf(l.x) / fmaxf(b.he[0], 1e-9f));
    const float uy = clamp01(1.0f - fabsf(l.y) / fmaxf(b.he[1], 1e-9f));
    const float uz = clamp01(1.0f - fabsf(l.z) / fmaxf(b.he[2], 1e-9f));
    return ux * uy * uz;
}
// 1D quadratic B-spline, m4_5_hybrid.h:44-49.
__device__ __forceinline__ float bspline1d_d(float xi) {
    const float a = fabsf(xi);
    if (a < 0.5f) return 0.75f - a * a;
    if (a < 1.5f) {
        const float t = 1.5f - a;
        return 0.5f * t * t;
    }
    return 0.0f;
}
// Coulomb tangent blend toward the body's tangential velocity (m4_5_hybrid.h:224-231).
__device__ __forceinline__ v3<float> friction_tangent_d(v3<float> v, v3<float> vb, v3<float> n, float friction) {
    const float vn = dot(v, n);
    const v3<float> vt = v - n * vn;
    const v3<float> vtb = vb - n * dot(vb, n);
    const v3<float> vtl = vt + (vtb - vt) * fminf(1.0f, friction);
    return n * vn + vtl;
}

// The scanned range is cut into chunks of 4 grid blocks; workgroup w of the
// fixed COUPLE_WGS takes chunks w, w + COUPLE_WGS, ...  Lane t handles node
// t & 63 of block t >> 6 of each chunk and accumulates its reaction
// contribution (7 values) in FP64; the workgroup reduces them in a fixed tree
// order and writes one partial per workgroup.
//
// Methods 2-4 (the M4.5 ladder) are the two-pass sweep of m4_5_hybrid.h:104-205
// over the allocated blocks within the shape's world AABB (tight for a
// cylinder, shape_world_aabb_half): pass 1, the sum S of the overlap fractions
// of the inside nodes, is taken by every workgroup over the whole range in the
// same fixed order (the range is a body's neighbourhood: small), so all agree
// on it without another launch; pass 2 gives each inside node m_b = M frac / S,
// the mass-weighted average velocity (+ the tangent blend for 3, + a B-spline
// weight on m_b for 4), mass old + m_b, and the reaction (v' - v_b) m_b / dt.
__device__ __forceinline__ void couple_wg(int method, const ColliderDev& b, const GridBuf& g, const uint8_t* bever,
                                          const Domain& d, float dt, double* partials) {
    __shared__ double red[7][256];
    const int t = threadIdx.x;
    const int64_t total_blocks = (int64_t)b.nb[0] * b.nb[1] * b.nb[2];
    const float dx = d.dx;
    // culling box: |R| he (M0, M1, M4.5 box), the cylinder's own for M4.5
    float alo[3], ahi[3];
    for (int a = 0; a < 3; ++a) {
        alo[a] = b.lo[a];
        ahi[a] = b.hi[a];
    }
    if (method >= 2 && b.shape == 1) {
        const float R = b.he[0], L = b.he[1];
        for (int a = 0; a < 3; ++a) {
            const float ai = b.Rl[a * 3 + 1];  // the axis: column 1 of R_l2w
            const float sq = 1.0f - ai * ai;
            const float hw = fabsf(ai) * L + (sq > 0.0f ? sqrtf(sq) : 0.0f) * R;
            alo[a] = b.c[a] - hw;
            ahi[a] = b.c[a] + hw;
        }
    }
    // the node (ln of block rb of the range): world position, block, culled
    auto locate = [&](int64_t rb, int ln, int& blk, int& ox, int& oy, int& oz) -> bool {
        const int rbz = (int)(rb % b.nb[2]);
        const int rby = (int)((rb / b.nb[2]) % b.nb[1]);
        const int rbx = (int)(rb / ((int64_t)b.nb[2] * b.nb[1]));
        const int bx = b.b0[0] + rbx, by = b.b0[1] + rby, bz = b.b0[2] + rbz;
        blk = (int)(d.b0 + ((int64_t)bx * d.nby + by) * d.nbz + bz);
        // global node index of the block origin
        ox = d.ox + bx * 4;
        oy = d.oy + by * 4;
        oz = d.oz + bz * 4;
        // block cull, mpm_rigid_coupler.cpp:82-91 (m4_5_hybrid.h:128-135 alike)
        const float blx = ((float)ox + 0.5f) * dx, bly = ((float)oy + 0.5f) * dx, blz = ((float)oz + 0.5f) * dx;
        const float bhx = ((float)(ox + 3) + 0.5f) * dx, bhy = ((float)(oy + 3) + 0.5f) * dx,
                    bhz = ((float)(oz + 3) + 0.5f) * dx;
        bool culled = (bhx < alo[0] || blx > ahi[0]) || (bhy < alo[1] || bly > ahi[1]) ||
                      (bhz < alo[2] || blz > ahi[2]);
        // M1 and the M4.5 sweep see only blocks that exist in the sparse grid (ever touched)
        if (method >= 1 && bever[blk] == 0) culled = true;
        (void)ln;
        return culled;
    };
    auto tree = [&](int k) {  // fixed-order sum of red[k][0..255] into red[k][0]
        for (int st = 128; st > 0; st >>= 1) {
            if (t < st) red[k][t] += red[k][t + st];
            __syncthreads();
        }
    };
    m3<float> Rw, Rl;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        Rw.a[c] = b.Rw[c];
        Rl.a[c] = b.Rl[c];
    }
    // ---- M4.5 pass 1: S = sum of overlap fractions (every workgroup, whole range)
    float mpu = 0.0f;  // body mass per unit overlap fraction
    bool sweep = method < 2;
    if (method >= 2) {
        double fs = 0.0;
        for (int64_t rb = t >> 6; rb < total_blocks; rb += 4) {
            int blk, ox, oy, oz;
            const int ln = t & 63;
            if (locate(rb, ln, blk, ox, oy, oz)) continue;
            const int li = ln >> 4, lj = (ln >> 2) & 3, lk = ln & 3;
            const v3<float> xn = mk(((float)(ox + li) + 0.5f) * dx, ((float)(oy + lj) + 0.5f) * dx,
                                    ((float)(oz + lk) + 0.5f) * dx);
            const v3<float> loc = mvec(Rw, xn - mk(b.c[0], b.c[1], b.c[2]));
            if (shape_inside_d(b, loc)) fs += (double)overlap_fraction_d(b, loc);
        }
        red[0][t] = fs;
        __syncthreads();
        tree(0);
        const float S = (float)red[0][0];
        __syncthreads();
        sweep = !(S < 1e-9f);  // m4_5_hybrid.h:149: no overlap, no coupling
        if (sweep) mpu = b.mass / S;
    }
    double c7[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t rb = (int64_t)blockIdx.x * 4 + (t >> 6); sweep && rb < total_blocks; rb += (int64_t)gridDim.x * 4) {
        int blk, ox, oy, oz;
        const int ln = t & 63;
        const bool culled = locate(rb, ln, blk, ox, oy, oz);
        const int li = ln >> 4, lj = (ln >> 2) & 3, lk = ln & 3;
        const int64_t nid = (int64_t)blk * 64 + ln;
        if (culled) continue;
        float4 q = g.mv[nid];
        const bool active = q.w > 0.0f;
        const bool visit = (method == 0) ? (active && q.w > 1e-8f) : true;
        if (!visit) continue;
        const v3<float> xn = mk(((float)(ox + li) + 0.5f) * dx, ((float)(oy + lj) + 0.5f) * dx,
                                ((float)(oz + lk) + 0.5f) * dx);
        const v3<float> r = xn - mk(b.c[0], b.c[1], b.c[2]);
        const v3<float> loc = mvec(Rw, r);
        if (!shape_inside_d(b, loc)) continue;
        const v3<float> vb = mk(b.vl[0], b.vl[1], b.vl[2]) + cross(mk(b.wa[0], b.wa[1], b.wa[2]), r);
        if (method == 0) {
            // mpm_rigid_coupler.cpp:108-138
            const v3<float> nl = shape_normal_d(b, loc);
            const v3<float> v0 = mk(q.x, q.y, q.z);
            v3<float> n = mvec(Rl, nl);
            const float nlen = sqrtf(dot(n, n));
            n = (nlen < 1e-6f) ? mk(0.f, 0.f, 0.f) : divg(n, nlen);
            const float vnn = dot(v0, n), vnb = dot(vb, n);
            const float vnc = fmaxf(vnn, vnb);
            const v3<float> vtn = v0 - n * vnn;
            const v3<float> vtb = vb - n * vnb;
            const v3<float> vtnew = vtn + (vtb - vtn) * fminf(1.0f, b.friction);
            const v3<float> va = n * vnc + vtnew;
            const v3<float> dp = (va - v0) * q.w;
            q.x = va.x;
            q.y = va.y;
            q.z = va.z;
            g.mv[nid] = q;
            const v3<float> fo = divg(dp, dt);
            const v3<float> to = divg(cross(r, dp), dt);
            c7[0] += -(double)fo.x;
            c7[1] += -(double)fo.y;
            c7[2] += -(double)fo.z;
            c7[3] += -(double)to.x;
            c7[4] += -(double)to.y;
            c7[5] += -(double)to.z;
            c7[6] += 1.0;
        } else if (method == 1) {
            // mpm_rigid_coupler_pic.h:145-177
            v3<float> va;
            if (!active || q.w < 1e-9f) {
                q = make_float4(vb.x, vb.y, vb.z, b.mb);
                va = vb;
                touch_block(g, nullptr, blk);  // cleared with the rest next substep
            } else {
                const v3<float> pold = mk(q.x, q.y, q.z) * q.w;
                const v3<float> padd = vb * b.mb;
                const float mn = q.w + b.mb;
                va = divg(pold + padd, mn);
                q = make_float4(va.x, va.y, va.z, mn);
            }
            g.mv[nid] = q;
            const v3<float> dF = (va - vb) * (b.mb / dt);
            const v3<float> to = cross(r, dF);
            c7[0] += dF.x;
            c7[1] += dF.y;
            c7[2] += dF.z;
            c7[3] += to.x;
            c7[4] += to.y;
            c7[5] += to.z;
            c7[6] += 1.0;
        } else {
            // M4.5 pass 2, m4_5_hybrid.h:152-198 and the layer rules :237-305
            const float m_b = mpu * overlap_fraction_d(b, loc);
            if (!(m_b > 0.0f)) continue;
            const bool was = active && q.w >= 1e-9f;  // kInactiveNodeMassEpsilon
            const float om = was ? q.w : 0.0f;
            const v3<float> ov = was ? mk(q.x, q.y, q.z) : mk(0.f, 0.f, 0.f);
            float mbe = m_b;
            v3<float> va = vb;
            bool blend = method >= 3;
            if (method == 4) {  // B-spline weight on the slice (layer C)
                const float w = bspline1d_d((xn.x - b.c[0]) / dx) * bspline1d_d((xn.y - b.c[1]) / dx) *
                                bspline1d_d((xn.z - b.c[2]) / dx);
                if (w < 1e-6f) {
                    mbe = 0.0f;
                    blend = false;  // v' = v_b, no mass, no reaction
                } else {
                    mbe = m_b * w;
                }
            }
            if (blend || method == 2) {
                va = was ? divg(ov * om + vb * mbe, om + mbe) : vb;  // mass_weighted_avg
                if (blend) {
                    v3<float> n = mvec(Rl, shape_normal_d(b, loc));
                    const float nlen = sqrtf(dot(n, n));
                    n = (nlen < 1e-6f) ? mk(0.f, 0.f, 0.f) : divg(n, nlen);
                    va = friction_tangent_d(va, vb, n, b.friction);
                }
            }
            if (!active) touch_block(g, nullptr, blk);  // claimed, cleared with the rest next substep
            q = make_float4(va.x, va.y, va.z, om + mbe);
            g.mv[nid] = q;
            const v3<float> dF = (va - vb) * (mbe / dt);
            const v3<float> to = cross(r, dF);
            c7[0] += dF.x;
            c7[1] += dF.y;
            c7[2] += dF.z;
            c7[3] += to.x;
            c7[4] += to.y;
            c7[5] += to.z;
            c7[6] += 1.0;
        }
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) red[k][t] = c7[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < 7; ++k) red[k][t] += red[k][t + s];
        }
        __syncthreads();
    }
    if (t < 7) partials[(int64_t)blockIdx.x * 7 + t] = red[t][0];
}
__global__ void __launch_bounds__(256) k_couple(int method, ColliderDev b, GridBuf g, uint8_t* bever, Domain d,
                                                float dt, double* partials) {
    couple_wg(method, b, g, bever, d, dt, partials);
}

// Sum of the per-workgroup partials in a fixed order (lane t takes partials
// t, t+64, ...; then a pairwise tree over the 64 lanes), so the reaction is
// deterministic run to run; the loads of one lane are independent.
__global__ void __launch_bounds__(64) k_reduce_partials(const double* partials, int nparts, double* out7) {
    __shared__ double red[7][64];
    const int t = threadIdx.x;
    double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = t; i < nparts; i += 64)
#pragma unroll
        for (int c = 0; c < 7; ++c) s[c] += partials[(int64_t)i * 7 + c];
#pragma unroll
    for (int c = 0; c < 7; ++c) red[c][t] = s[c];
    __syncthreads();
    for (int w = 32; w >= 1; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int c = 0; c < 7; ++c) red[c][t] += red[c][t + w];
        __syncthreads();
    }
    if (t < 7) out7[t] = red[t][0];
}

// --------------------------------------------------------------- grid I/O
__global__ void k_grid_gather(GridBuf g, uint8_t* bever, Domain d, int org0, int org1, int org2, int n0, int n1,
                              int n2, double qscale, int updated, long long* o_mq, float* o_mass, float* o_vel,
                              float* o_force, uint8_t* o_act, uint8_t* o_alloc) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)n0 * n1 * n2;
    if (idx >= tot) return;
    const int k = (int)(idx % n2), j = (int)((idx / n2) % n1), i = (int)(idx / ((int64_t)n2 * n1));
    const int lx = org0 + i - d.ox, ly = org1 + j - d.oy, lz = org2 + k - d.oz;
    long long mq = 0;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f), f = make_float4(0.f, 0.f, 0.f, 0.f);
    uint8_t al = 0;
    if (lx >= 0 && ly >= 0 && lz >= 0 && lx < d.nx && ly < d.ny && lz < d.nz) {
        const int64_t nid = node_of(d, lx, ly, lz);
        mq = g.mq[nid];
        q = g.mv[nid];
        if (!updated && g.mom) {  // the deposit's momentum (fast mode)
            q.x = g.mom[nid];
            q.y = g.mom[g.mstride + nid];
            q.z = g.mom[2 * g.mstride + nid];
        }
        if (g.f) f = g.f[nid];
        al = bever ? bever[nid >> 6] : 0;
    }
    const float mass = updated ? q.w : (float)((double)mq * qscale);
    if (o_mq) o_mq[idx] = mq;
    if (o_mass) o_mass[idx] = mass;
    if (o_vel) {
        o_vel[3 * idx] = q.x;
        o_vel[3 * idx + 1] = q.y;
        o_vel[3 * idx + 2] = q.z;
    }
    if (o_force) {
        o_force[3 * idx] = f.x;
        o_force[3 * idx + 1] = f.y;
        o_force[3 * idx + 2] = f.z;
    }
    if (o_act) o_act[idx] = (mq != 0 || (updated && q.w != 0.0f)) ? 1 : 0;
    if (o_alloc) o_alloc[idx] = al;
}

__global__ void k_grid_scatter(GridBuf g, uint8_t* bever, Domain d, int org0, int org1, int org2, int n0, int n1,
                               int n2, double qscale_inv, const float* i_mass, const float* i_vel,
                               const uint8_t* i_act) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)n0 * n1 * n2;
    if (idx >= tot) return;
    if (!i_act[idx]) return;
    const int k = (int)(idx % n2), j = (int)((idx / n2) % n1), i = (int)(idx / ((int64_t)n2 * n1));
    const int lx = org0 + i - d.ox, ly = org1 + j - d.oy, lz = org2 + k - d.oz;
    if (!(lx >= 0 && ly >= 0 && lz >= 0 && lx < d.nx && ly < d.ny && lz < d.nz)) return;
    const int64_t nid = node_of(d, lx, ly, lz);
    const float m = i_mass[idx];
    g.mv[nid] = make_float4(i_vel[3 * idx], i_vel[3 * idx + 1], i_vel[3 * idx + 2], m);
    long long mq = (long long)rint((double)m * qscale_inv);
    g.mq[nid] = mq;
    touch_block(g, bever, (int)(nid >> 6));
}

// bench_compare_all.cpp:117-126: v_y += dvy, clamped at 0, on active nodes
// (mass > 0 after the update; w holds the node mass).
__global__ void __launch_bounds__(64) k_grid_settle(GridBuf g, float dvy) {
    const int cnt = *g.bcount;
    for (int b = blockIdx.x; b < cnt; b += gridDim.x) {
        const int64_t nid = (int64_t)g.blist[b] * 64 + threadIdx.x;
        float4 q = g.mv[nid];
        if (!(q.w > 0.0f)) continue;
        q.y = q.y + dvy;
        if (q.y < 0.0f) q.y = 0.0f;
        g.mv[nid] = q;
    }
}
void launch_grid_settle(const GridBuf& g, int nblocks, float dvy, hipStream_t s) {
    const int grid = nblocks < 4096 ? (nblocks > 0 ? nblocks : 1) : 4096;
    hipLaunchKernelGGL(k_grid_settle, dim3(grid), dim3(64), 0, s, g, dvy);
}

__global__ void __launch_bounds__(256) k_mass_total(GridBuf g, unsigned long long* out) {
    const int cnt = *g.bcount;
    const int64_t tot = (int64_t)cnt * 64;
    unsigned long long s = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < tot; k += (int64_t)gridDim.x * blockDim.x) {
        const int blk = g.blist[k >> 6];
        s += (unsigned long long)g.mq[(int64_t)blk * 64 + (k & 63)];
    }
    atomicAdd(out, s);
}

// V0 *= rescale for every particle (set_material density change, mpm_solver.h:45-49).
__global__ void k_scale_v0(PView P, float rescale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.n) P.p[PV0 * P.cap + i] *= rescale;
}

// Largest particle mass, as float bits (masses are non-negative).
__global__ void k_max_mass(PView P, unsigned int* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned int v = 0;
    if (i < P.n) v = __float_as_uint(fmaxf(P.p[PM * P.cap + i], 0.0f));
    atomicMax(out, v);
}

// ================================================================ launchers
static inline int grid_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

void launch_clear(const GridBuf& g, int nblocks, hipStream_t s, int* zero_word, bool defer_reset) {
    const int grid = nblocks < 8192 ? (nblocks > 0 ? nblocks : 1) : 8192;
    hipLaunchKernelGGL(k_clear_blocks, dim3(grid), dim3(64), 0, s, g, zero_word);
    if (!defer_reset) hipLaunchKernelGGL(k_reset_count, dim3(1), dim3(1), 0, s, g);
}
void launch_p2g_ref(const PView& P, const GridBuf& g, uint8_t* bever, const Domain& d, const Phys& ph,
                    double qscale_inv, int* err, hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_p2g_ref, dim3(grid_for(P.n, 256)), dim3(256), 0, s, P, g, bever, d, ph, qscale_inv, err);
}
void launch_update(const GridBuf& g, int nblocks, const Domain& d, const Phys& ph, double qscale, float dt,
                   hipStream_t s, bm_body* body, ColliderDev* col, int64_t max_blocks, int* err, SceneView sv,
                   const int* has_body, const GridBuf* clear_next, int* zero_word) {
    int grid = nblocks < 16384 ? (nblocks > 0 ? nblocks : 1) : 16384;
    if (grid < sv.n_scenes) grid = sv.n_scenes;
    GridBuf clr{};
    if (clear_next) clr = *clear_next;
    hipLaunchKernelGGL(k_update, dim3(grid), dim3(64), 0, s, g, d, ph, qscale, dt, body, col, (long long)max_blocks,
                       err, sv, has_body, clr, zero_word);
}
void launch_g2p_ref(const PView& P, const GridBuf& g, const Domain& d, const Phys& ph, float dt, int* err,
                    hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_g2p_ref, dim3(grid_for(P.n, 256)), dim3(256), 0, s, P, g, d, ph, dt, err);
}
// ------------------------------------------------------------ device body
// Start of the body's coupled substep (bench_compare_all.cpp:642-658): the
// applied loads go into the accumulators (RigidBody::apply_force /
// apply_torque, rigid_body.h:168-182), then the collider is built from the
// body's pose and velocities.
__device__ void body_begin(bm_body* body, ColliderDev* col, const Domain& d, long long max_blocks, int* err,
                           int scene) {
    bm_body b = *body;
    for (int a = 0; a < 3; ++a) {
        b.force[a] = b.force[a] + b.applied_force[a];
        b.torque[a] = b.torque[a] + b.applied_torque[a];
    }
    ColliderDev c = collider_build(b.position, b.orientation, b.half_extents, b.linear_velocity, b.angular_velocity,
                                   b.friction, b.mass, b.shape, d);
    if ((long long)c.nb[0] * c.nb[1] * c.nb[2] > max_blocks) {
        flag_error(err, scene, 4);
        c.nb[0] = c.nb[1] = c.nb[2] = 0;
    }
    for (int a = 0; a < 3; ++a) {
        body->force[a] = b.force[a];
        body->torque[a] = b.torque[a];
    }
    *col = c;
}
__global__ void k_body_begin(bm_body* body, ColliderDev* col, Domain d, long long max_blocks, int* err) {
    if (threadIdx.x == 0) body_begin(body, col, d, max_blocks, err);
}

__global__ void __launch_bounds__(64) k_body_end(bm_body* body, const double* partials, int nparts, double* out7,
                                                 float dt) {
    body_end(body, partials, nparts, out7, dt);
}

// The device body's coupler: the collider was built on the device.
__global__ void __launch_bounds__(256) k_couple_body(int method, const ColliderDev* bp, GridBuf g, uint8_t* bever,
                                                     Domain d, float dt, double* partials) {
    const ColliderDev b = *bp;
    couple_wg(method, b, g, bever, d, dt, partials);
}

// The couplers of a batch: blockIdx.y is the scene; scenes without a body
// return at once.  Partials and reaction per scene.
__global__ void __launch_bounds__(256) k_couple_batch(const int* methods, const int* has_body, const ColliderDev* cols,
                                                      GridBuf g, uint8_t* bever, SceneView sv, float dt,
                                                      double* partials) {
    const int s = blockIdx.y;
    if (!has_body[s]) return;
    const ColliderDev b = cols[s];
    couple_wg(methods[s], b, g, bever, sv.sc[s].d, dt, partials + (int64_t)s * COUPLE_BATCH_WGS * 7);
}
__global__ void __launch_bounds__(64) k_body_end_batch(bm_body* bodies, const int* has_body, const double* partials,
                                                       double* out7, float dt) {
    const int s = blockIdx.x;
    if (!has_body[s]) return;
    body_end(bodies + s, partials + (int64_t)s * COUPLE_BATCH_WGS * 7, COUPLE_BATCH_WGS, out7 + (int64_t)s * 8, dt);
}
void launch_couple_batch(const int* methods, const int* has_body, bm_body* bodies, const ColliderDev* cols,
                         const GridBuf& g, uint8_t* bever, SceneView sv, float dt, double* partials, double* out7,
                         hipStream_t s) {
    if (sv.n_scenes <= 0 || !(dt > 0.0f)) return;
    hipLaunchKernelGGL(k_couple_batch, dim3(COUPLE_BATCH_WGS, sv.n_scenes), dim3(256), 0, s, methods, has_body, cols, g,
                       bever, sv, dt, partials);
    hipLaunchKernelGGL(k_body_end_batch, dim3(sv.n_scenes), dim3(64), 0, s, bodies, has_body,
                       (const double*)partials, out7, dt);
}

BodyEnd launch_couple_body(int method, bm_body* body, ColliderDev* col, const GridBuf& g, uint8_t* bever,
                           const Domain& d, float dt, int64_t max_blocks, double* partials, double* out7, int* err,
                           bool begun, hipStream_t s, bool defer_end) {
    if (!begun) hipLaunchKernelGGL(k_body_begin, dim3(1), dim3(64), 0, s, body, col, d, (long long)max_blocks, err);
    const int nparts = (dt > 0.0f) ? COUPLE_WGS : 0;  // dt <= 0: zero reaction
    if (nparts > 0)
        hipLaunchKernelGGL(k_couple_body, dim3(nparts), dim3(256), 0, s, method, (const ColliderDev*)col, g, bever, d,
                           dt, partials);
    const BodyEnd be{body, partials, nparts, out7, dt};
    if (defer_end) return be;
    launch_body_end(be, s);
    return BodyEnd{};
}
void launch_body_end(const BodyEnd& be, hipStream_t s) {
    if (be.body)
        hipLaunchKernelGGL(k_body_end, dim3(1), dim3(64), 0, s, be.body, be.partials, be.nparts, be.out7, be.dt);
}

void launch_couple(int method, const ColliderDev& b, const GridBuf& g, uint8_t* bever, const Domain& d, float dt,
                   double* partials, double* out7, hipStream_t s) {
    hipLaunchKernelGGL(k_couple, dim3(COUPLE_WGS), dim3(256), 0, s, method, b, g, bever, d, dt, partials);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, s, partials, COUPLE_WGS, out7);
}
void launch_grid_gather(const GridBuf& g, uint8_t* bever, const Domain& d, const int* org, const int* dims,
                        double qscale, int updated, long long* o_mq, float* o_mass, float* o_vel, float* o_force,
                        uint8_t* o_act, uint8_t* o_alloc, hipStream_t s) {
    const int64_t tot = (int64_t)dims[0] * dims[1] * dims[2];
    if (tot == 0) return;
    hipLaunchKernelGGL(k_grid_gather, dim3(grid_for(tot, 256)), dim3(256), 0, s, g, bever, d, org[0], org[1], org[2],
                       dims[0], dims[1], dims[2], qscale, updated, o_mq, o_mass, o_vel, o_force, o_act, o_alloc);
}
void launch_grid_scatter(const GridBuf& g, uint8_t* bever, const Domain& d, const int* org, const int* dims,
                         double qscale_inv, const float* i_mass, const float* i_vel, const uint8_t* i_act,
                         hipStream_t s) {
    const int64_t tot = (int64_t)dims[0] * dims[1] * dims[2];
    if (tot == 0) return;
    hipLaunchKernelGGL(k_grid_scatter, dim3(grid_for(tot, 256)), dim3(256), 0, s, g, bever, d, org[0], org[1],
                       org[2], dims[0], dims[1], dims[2], qscale_inv, i_mass, i_vel, i_act);
}
void launch_mass_total(const GridBuf& g, unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mass_total, dim3(1024), dim3(256), 0, s, g, out);
}
void launch_scale_v0(const PView& P, float rescale, hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_scale_v0, dim3(grid_for(P.n, 256)), dim3(256), 0, s, P, rescale);
}
void launch_max_mass(const PView& P, unsigned int* out, hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_max_mass, dim3(grid_for(P.n, 256)), dim3(256), 0, s, P, out);
}

}  // namespace bm
