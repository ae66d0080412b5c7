// mpm_fast_device.hpp -- per-particle device routines of the fast substep:
// fast-math Jacobi SVD, Drucker-Prager return map, the 27-node G2P gather
// and P2G scatter, and the folded-momentum coefficients.  Included by the
// fast-path translation units only (FMA contraction allowed there).
#pragma once

#include "mpm_device.hpp"
#include "mpm_internal.hpp"

namespace bm {

constexpr int TE = 8;  // LDS tile edge in nodes: the block's 4 + 2 stencil layers + 1 drift layer each side
constexpr int TNODES = TE * TE * TE;

__device__ __forceinline__ void touch_block_f(const GridBuf& g, uint8_t* bever, int blk) {
    if (g.bflag[blk] == 0u && atomicExch(&g.bflag[blk], 1u) == 0u) {
        const int slot = atomicAdd(g.bcount, 1);
        g.blist[slot] = blk;
    }
    if (bever && bever[blk] == 0) bever[blk] = 1;
}

// --------------------------------------------------------------- fast per-particle math
// Reciprocal / square-root through the hardware approximations (<= 1 ulp);
// the reference-order path keeps IEEE division in its own translation unit.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// The reference's cyclic Jacobi (svd.h:18-108: pairs (0,1),(0,2),(1,2), skip
// |a_ij| < 1e-6, <= 5 sweeps, stop at max |a_ij| < 1e-5) with fast reciprocal
// and square root.  Returns false when U = F V / sigma has lost orthogonality
// (|u_i . u_j| > 2e-4), which is how near-coincident singular values or an
// unconverged sweep show up; the caller then redoes the particle in FP64.
__device__ __forceinline__ bool svd_fast(const m3<float>& F, m3<float>& U, v3<float>& S, m3<float>& V) {
    m3<float> A = mtmul(F, F);
#pragma unroll
    for (int i = 0; i < 9; ++i) V.a[i] = (i % 4 == 0) ? 1.0f : 0.0f;
    for (int sweep = 0; sweep < 5; ++sweep) {
        float max_off = 0.0f;
#pragma unroll
        for (int pr = 0; pr < 3; ++pr) {
            const int i = pr < 2 ? 0 : 1;
            const int j = pr == 0 ? 1 : 2;
            const int k = 3 - i - j;
            const float aij = A.a[i * 3 + j], aii = A.a[i * 3 + i], ajj = A.a[j * 3 + j];
            max_off = fmaxf(max_off, fabsf(aij));
            if (fabsf(aij) < 1e-6f) continue;
            const float tau = (ajj - aii) * (0.5f * frcp(aij));
            const float t = copysignf(frcp(fabsf(tau) + fsqrt(1.0f + tau * tau)), tau);
            const float c = frsq(1.0f + t * t);
            const float s = c * t;
            const float nii = c * c * aii - 2.0f * s * c * aij + s * s * ajj;
            const float njj = s * s * aii + 2.0f * s * c * aij + c * c * ajj;
            A.a[i * 3 + i] = nii;
            A.a[j * 3 + j] = njj;
            A.a[i * 3 + j] = 0.0f;
            A.a[j * 3 + i] = 0.0f;
            const float aik = A.a[i * 3 + k], ajk = A.a[j * 3 + k];
            A.a[i * 3 + k] = c * aik - s * ajk;
            A.a[k * 3 + i] = A.a[i * 3 + k];
            A.a[j * 3 + k] = s * aik + c * ajk;
            A.a[k * 3 + j] = A.a[j * 3 + k];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const float vi = V.a[r * 3 + i], vj = V.a[r * 3 + j];
                V.a[r * 3 + i] = c * vi - s * vj;
                V.a[r * 3 + j] = s * vi + c * vj;
            }
        }
        if (max_off < 1e-5f) break;
    }
    S = mk(fsqrt(fmaxf(0.0f, A.a[0])), fsqrt(fmaxf(0.0f, A.a[4])), fsqrt(fmaxf(0.0f, A.a[8])));
    const m3<float> FV = mmul(F, V);
    const float sv[3] = {S.x, S.y, S.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (sv[c] > 1e-6f) {
            const float r = frcp(sv[c]);
            U.a[c] = FV.a[c] * r;
            U.a[3 + c] = FV.a[3 + c] * r;
            U.a[6 + c] = FV.a[6 + c] * r;
        } else {
            U.a[c] = (c == 0) ? 1.0f : 0.0f;
            U.a[3 + c] = (c == 1) ? 1.0f : 0.0f;
            U.a[6 + c] = (c == 2) ? 1.0f : 0.0f;
        }
    }
    const v3<float> u0 = mk(U.a[0], U.a[3], U.a[6]), u1 = mk(U.a[1], U.a[4], U.a[7]), u2 = mk(U.a[2], U.a[5], U.a[8]);
    const float o = fmaxf(fabsf(dot(u0, u1)), fmaxf(fabsf(dot(u0, u2)), fabsf(dot(u1, u2))));
    return o <= 2e-4f;
}

// Drucker-Prager return map (mpm_solver.h:125-162) with fast log/exp; returns
// the projected singular values and their logarithms.
__device__ __forceinline__ v3<float> dp_fast(v3<float> s, const Material& M, v3<float>& lns) {
    const v3<float> e = mk(flog(fmaxf(s.x, 1e-6f)), flog(fmaxf(s.y, 1e-6f)), flog(fmaxf(s.z, 1e-6f)));
    const float tr = e.x + e.y + e.z;
    const float t3 = tr * (1.0f / 3.0f);
    v3<float> eh = mk(e.x - t3, e.y - t3, e.z - t3);
    const float H = fsqrt(dot(eh, eh));
    if (tr > 0.0f) {
        lns = mk(0.0f, 0.0f, 0.0f);
        return mk(1.0f, 1.0f, 1.0f);
    }
    const float limit = -M.alpha * M.kohesion * tr;
    if (H > limit) {
        eh = eh * (limit * frcp(H));
        lns = mk(eh.x + t3, eh.y + t3, eh.z + t3);
        return mk(fexp(lns.x), fexp(lns.y), fexp(lns.z));
    }
    lns = e;
    return mk(fmaxf(s.x, 1e-6f), fmaxf(s.y, 1e-6f), fmaxf(s.z, 1e-6f));
}

struct FastArgs {
    float dt_g2p;   // substep ending in this G2P
    float dt_fold;  // substep whose force is folded into the deposited momentum
    int n_units;
    double qscale_inv;
    int* redo;      // particles whose SVD needs the FP64 pass
    int* n_redo;
    int* list_reset;  // block-list length of the output grid, zeroed by the particle pass (two-kernel form)
    const int* n_units_dev;  // unit count on the device (batches: no host read-back), else n_units
    int store_vc = 1;        // the G2P writes the particle velocity and C (see launch_split)
};

__device__ __forceinline__ bool stencil_ok(const Domain& d, int lx, int ly, int lz) {
    return lx >= 0 && ly >= 0 && lz >= 0 && lx + 2 < d.nx && ly + 2 < d.ny && lz + 2 < d.nz;
}

// G2P gather of one particle (mpm_solver.h:307-338) from a node source:
// NODE(a, b, c) returns the float4 of stencil node (a, b, c).
template <class NodeFn>
__device__ __forceinline__ void gather(const Stencil& s, float dx, float kC, NodeFn node, v3<float>& vel,
                                       m3<float>& Cm) {
    vel = mk(0.f, 0.f, 0.f);
#pragma unroll
    for (int c = 0; c < 9; ++c) Cm.a[c] = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ddx = ((float)(s.b[0] + a) - s.f[0] - 0.5f) * dx;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float ddy = ((float)(s.b[1] + b) - s.f[1] - 0.5f) * dx;
            const float wab = s.w[0][a] * s.w[1][b];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float ddz = ((float)(s.b[2] + c) - s.f[2] - 0.5f) * dx;
                const float w = wab * s.w[2][c];
                const float4 q = node(a, b, c);
                const float wx = q.x * w, wy = q.y * w, wz = q.z * w;
                vel.x += wx;
                vel.y += wy;
                vel.z += wz;
                Cm.a[0] += wx * ddx;
                Cm.a[1] += wx * ddy;
                Cm.a[2] += wx * ddz;
                Cm.a[3] += wy * ddx;
                Cm.a[4] += wy * ddy;
                Cm.a[5] += wy * ddz;
                Cm.a[6] += wz * ddx;
                Cm.a[7] += wz * ddy;
                Cm.a[8] += wz * ddz;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) Cm.a[c] *= kC;
}

// The same gather factored by axis: the quadratic B-spline weight and the
// node offset d are products / sums of per-axis terms, so
//   v = sum_a wx_a sum_b wy_b sum_c wz_c q,  C[:,z] ~ sum_a wx_a sum_b wy_b sum_c (wz d_z)_c q, ...
// takes 6 FMAs per node plus 9 per (a, b) row and 12 per a plane (279 in
// all) instead of 15 per node.
template <class NodeFn>
__device__ __forceinline__ void gather_sep(const Stencil& s, float dx, float kC, NodeFn node, v3<float>& vel,
                                           m3<float>& Cm) {
    float wd[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int k = 0; k < 3; ++k) wd[ax][k] = s.w[ax][k] * (((float)(s.b[ax] + k) - s.f[ax] - 0.5f) * dx);
    v3<float> V = mk(0.f, 0.f, 0.f), Cx = V, Cy = V, Cz = V;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v3<float> Y0 = mk(0.f, 0.f, 0.f), Y1 = Y0, Yd = Y0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            v3<float> Z0 = mk(0.f, 0.f, 0.f), Z1 = Z0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float4 q = node(a, b, c);
                const float wz = s.w[2][c], wdz = wd[2][c];
                Z0 = mk(Z0.x + wz * q.x, Z0.y + wz * q.y, Z0.z + wz * q.z);
                Z1 = mk(Z1.x + wdz * q.x, Z1.y + wdz * q.y, Z1.z + wdz * q.z);
            }
            const float wy = s.w[1][b], wdy = wd[1][b];
            Y0 = mk(Y0.x + wy * Z0.x, Y0.y + wy * Z0.y, Y0.z + wy * Z0.z);
            Y1 = mk(Y1.x + wy * Z1.x, Y1.y + wy * Z1.y, Y1.z + wy * Z1.z);
            Yd = mk(Yd.x + wdy * Z0.x, Yd.y + wdy * Z0.y, Yd.z + wdy * Z0.z);
        }
        const float wx = s.w[0][a], wdx = wd[0][a];
        V = mk(V.x + wx * Y0.x, V.y + wx * Y0.y, V.z + wx * Y0.z);
        Cx = mk(Cx.x + wdx * Y0.x, Cx.y + wdx * Y0.y, Cx.z + wdx * Y0.z);
        Cy = mk(Cy.x + wx * Yd.x, Cy.y + wx * Yd.y, Cy.z + wx * Yd.z);
        Cz = mk(Cz.x + wx * Y1.x, Cz.y + wx * Y1.y, Cz.z + wx * Y1.z);
    }
    vel = V;
    Cm.a[0] = Cx.x * kC;
    Cm.a[1] = Cy.x * kC;
    Cm.a[2] = Cz.x * kC;
    Cm.a[3] = Cx.y * kC;
    Cm.a[4] = Cy.y * kC;
    Cm.a[5] = Cz.y * kC;
    Cm.a[6] = Cx.z * kC;
    Cm.a[7] = Cy.z * kC;
    Cm.a[8] = Cz.z * kC;
}

// Node index of (lx+a, ly+b, lz+c) as a sum of per-axis terms (the
// block-linear layout is separable): nid = X[a] + Y[b] + Z[c].
struct NodeAxes {
    int X[3], Y[3], Z[3];
    __device__ __forceinline__ NodeAxes(const Domain& d, int lx, int ly, int lz) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            X[k] = (int)(d.b0 * 64) + (((lx + k) >> 2) * d.nby * d.nbz) * 64 + ((lx + k) & 3) * 16;
            Y[k] = (((ly + k) >> 2) * d.nbz) * 64 + ((ly + k) & 3) * 4;
            Z[k] = ((lz + k) >> 2) * 64 + ((lz + k) & 3);
        }
    }
    __device__ __forceinline__ int operator()(int a, int b, int c) const { return X[a] + Y[b] + Z[c]; }
};

// P2G scatter of one particle's mass quanta and folded momentum
// w (m v + (m C + dt K) d) into 27 nodes; SINK(a, b, c, dq, mom) deposits.
template <class SinkFn>
__device__ __forceinline__ void scatter(const Stencil& s, float dx, float qp, const v3<float>& mvel,
                                        const m3<float>& Aff, SinkFn sink) {
    // the stencil centre (1,1,1) is deposited last and takes the remainder,
    // so each other node's quanta are used as soon as they are formed
    int acc = 0;
    v3<float> mom_c;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float ddx = ((float)(s.b[0] + a) - s.f[0] - 0.5f) * dx;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float ddy = ((float)(s.b[1] + b) - s.f[1] - 0.5f) * dx;
            const float px = mvel.x + Aff.a[0] * ddx + Aff.a[1] * ddy;
            const float py = mvel.y + Aff.a[3] * ddx + Aff.a[4] * ddy;
            const float pz = mvel.z + Aff.a[6] * ddx + Aff.a[7] * ddy;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float ddz = ((float)(s.b[2] + c) - s.f[2] - 0.5f) * dx;
                const float w = s.w[0][a] * s.w[1][b] * s.w[2][c];
                const v3<float> mom = mk(w * (px + Aff.a[2] * ddz), w * (py + Aff.a[5] * ddz), w * (pz + Aff.a[8] * ddz));
                if (a == 1 && b == 1 && c == 1) {
                    mom_c = mom;
                    continue;
                }
                const int dq = mass_quanta(w, qp);
                acc += dq;
                sink(a, b, c, w, dq, mom);
            }
        }
    }
    sink(1, 1, 1, s.w[0][1] * s.w[1][1] * s.w[2][1], (int)qp - acc, mom_c);
}

__device__ __forceinline__ void deposit_global(const GridBuf& g, uint8_t* bever, const Domain& d, int lx, int ly,
                                               int lz, int dq, v3<float> mom) {
    const int64_t nid = node_of(d, lx, ly, lz);
    atomicAdd((unsigned long long*)&g.mq[nid], (unsigned long long)(long long)dq);
    atomicAdd(g.mom + nid, mom.x);
    atomicAdd(g.mom + g.mstride + nid, mom.y);
    atomicAdd(g.mom + 2 * g.mstride + nid, mom.z);
    touch_block_f(g, bever, (int)(nid >> 6));
}

// Folded-momentum coefficients of the post-G2P state: m v and m C + dt K,
// K = -4/dx^2 V0 tau with tau the Kirchhoff stress of (U, ln sigma).
__device__ __forceinline__ void p2g_coeffs(float mass, float V0, const v3<float>& vel, const m3<float>& Cm,
                                           const m3<float>& U, const v3<float>& lns, const Phys& ph, float inv_dx,
                                           float dt_fold, v3<float>& mvel, m3<float>& Aff) {
    const m3<float> tau = hencky_from_log(U, lns, ph.mu, ph.lambda);
    const float ff = -4.0f * inv_dx * inv_dx * V0 * dt_fold;
#pragma unroll
    for (int c = 0; c < 9; ++c) Aff.a[c] = mass * Cm.a[c] + ff * tau.a[c];
    mvel = vel * mass;
}


}  // namespace bm
