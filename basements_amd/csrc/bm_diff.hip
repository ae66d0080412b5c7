// bm_diff.hip -- differentiable rollout of the soil substep: the reference
// algorithm's P2G / grid update / M0 tool / G2P with plasticity, in FP32 or
// FP64, and its hand-written adjoint with checkpoint-and-recompute.
//
// The reference has no backward pass (its autodiff tape records nothing,
// include/basements/physics/mpm/mpm_autodiff_tape.h:63-92); this follows the
// forward algorithm of mpm_solver.h:114-348 and the M0 clamp of
// src/physics/coupling/mpm_rigid_coupler.cpp:95-138 and differentiates it by
// hand.  The loss is the mean squared difference between each particle's
// final displacement and a target displacement; its gradient is taken with
// respect to the tool velocity, the Drucker-Prager alpha (friction angle),
// the Lame parameters (Young's modulus) and the cone factor k.
//
// Forward form (per substep, particle state s = x, v, C, F; m, V0 fixed):
//   P2G   m_i = sum_p w m,  q_i = sum_p w (m (v + C d) + dt K d),
//         K = -(4/dx^2) V0 tau(F),  tau = U diag(2 mu ln s + lam tr ln s) U^T,
//         d = x_i - x_p (node i at i*dx)
//   grid  u = q/m + g dt (m > 1e-6), floor Coulomb rule on the bottom layer,
//         M0 clamp on nodes inside the tool (m > 1e-8); lighter nodes keep q
//   G2P   v' = sum w u,  C' = (4/dx^2) sum w u d^T,  x' = x + dt v',
//         F' = proj((I + dt C') F), floor clamp of x'.y
// The dense node box replaces the reference's unbounded grid; the per-node
// mass is a plain float sum (the product path's fixed-point mass is not
// differentiable); the SVD is the same cyclic Jacobi run to a tighter
// tolerance so that U and V are accurate enough to differentiate.
//
// Adjoint of a function G(F) = U diag(h(sigma)) V^T (both the plastic
// projection and P(F) = tau F^-T = U diag(tau_k / sigma_k) V^T have this
// form): with M = U^T Gbar V,
//   Fbar = U K V^T,  K_kk = sum_i M_ii dh_i/dsigma_k,
//   K_ij = M_ij (D + S)/2 + M_ji (D - S)/2   (i != j),
//   D = (h_j - h_i) / (sigma_j - sigma_i),  S = (h_j + h_i) / (sigma_j + sigma_i),
// with D taken as its limit dh_j/dsigma_j - dh_i/dsigma_j when the two
// singular values coincide (the first-order SVD perturbation of
// Papadopoulo & Lourakis 2000, rearranged so that nothing divides by
// sigma_j^2 - sigma_i^2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "mpm_device.hpp"
#include "mpm_internal.hpp"
#include "trace.hpp"

namespace bm {
namespace dif {

enum { QX = 0, QV = 3, QC = 6, QF = 15, QCOMP = 24 };

template <class R> struct PS {
    R* p;
    int64_t cap;
    __device__ R& at(int c, int64_t i) const { return p[c * cap + i]; }
};

struct Box {
    int ox, oy, oz, nx, ny, nz;
};

// Parameters of one rollout; tc is the tool centre of the substep at hand
// (par_at), tc0 its value at time 0.
template <class R> struct Par {
    R dx, inv_dx, mu, lam, alpha, koh, mu_floor, g[3];
    int ref_svd;  // the reference's SVD stopping rule (forward comparisons only)
    int tool;
    R tc[3], he[3], Rl[9], Rw[9], vt[3], friction;
    double tc0[3], vt0[3];
};

// ------------------------------------------------------------------ rollouts
// A handle runs its rollouts side by side: each launch has one row of
// workgroups (blockIdx.y) per rollout.  Rollout r's particle states sit
// QCOMP * cap values after rollout r-1's in every state slot (the component
// stride stays cap), its grid values (v values per node) at r * v * nn, its
// gradient accumulators at r * GN; the particles' m, V0, x(0) and target are
// one copy that every rollout reads.
template <class R> __device__ __forceinline__ PS<R> roll(PS<R> s) {
    s.p += (int64_t)blockIdx.y * QCOMP * s.cap;
    return s;
}
__device__ __forceinline__ int64_t roll_off(int64_t per_rollout) { return (int64_t)blockIdx.y * per_rollout; }

// This workgroup's rollout's parameters at time t: the tool centre is
// tc0 + vt0 t, in double and unfused, so that every rollout and both
// precisions round the same centre.
template <class R> __device__ __forceinline__ Par<R> par_at(const Par<R>* __restrict__ Pv, double t) {
#pragma clang fp contract(off)
    Par<R> P = Pv[blockIdx.y];
    if (P.tool) {
#pragma unroll
        for (int a = 0; a < 3; ++a) P.tc[a] = (R)(P.tc0[a] + P.vt0[a] * t);
    }
    return P;
}

// Gradient accumulators (double, device): tool velocity (3), alpha, koh, mu, lam.
enum { GVT = 0, GALPHA = 3, GKOH = 4, GMU = 5, GLAM = 6, GLOSS = 7, GN = 8 };

template <class R> __device__ __forceinline__ int64_t node_index(const Box& b, int i, int j, int k) {
    const int a = i - b.ox, c = j - b.oy, e = k - b.oz;
    if (a < 0 || c < 0 || e < 0 || a >= b.nx || c >= b.ny || e >= b.nz) return -1;
    return ((int64_t)a * b.ny + c) * b.nz + e;
}

__device__ __forceinline__ int64_t box_nodes(const Box& b) { return (int64_t)b.nx * b.ny * b.nz; }

template <class R> struct St {
    int b[3];
    R w[3][3], dw[3][3];  // weights and their derivatives along each axis (per unit x)
};
template <class R> __device__ __forceinline__ void stencil(const R x[3], R inv_dx, St<R>& s) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const R f = x[a] * inv_dx - R(0.5);
        s.b[a] = (int)floor(f);
        const R fr = f - (R)s.b[a];
        s.w[a][0] = R(0.5) * (R(1) - fr) * (R(1) - fr);
        s.w[a][1] = R(0.75) - (fr - R(0.5)) * (fr - R(0.5));
        s.w[a][2] = R(0.5) * fr * fr;
        s.dw[a][0] = -(R(1) - fr) * inv_dx;
        s.dw[a][1] = R(-2) * (fr - R(0.5)) * inv_dx;
        s.dw[a][2] = fr * inv_dx;
    }
}

template <class R> __device__ __forceinline__ m3<R> load_m(const PS<R>& s, int c0, int64_t i) {
    m3<R> m;
#pragma unroll
    for (int c = 0; c < 9; ++c) m.a[c] = s.at(c0 + c, i);
    return m;
}
template <class R> __device__ __forceinline__ void store_m(const PS<R>& s, int c0, int64_t i, const m3<R>& m) {
#pragma unroll
    for (int c = 0; c < 9; ++c) s.at(c0 + c, i) = m.a[c];
}
template <class R> __device__ __forceinline__ m3<R> transpose(const m3<R>& a) {
    m3<R> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = a.a[j * 3 + i];
    return r;
}
template <class R> __device__ __forceinline__ m3<R> mmt(const m3<R>& p, const m3<R>& q) {  // p q^T
    return mmul(p, transpose(q));
}

// SVD accurate enough to differentiate: the reference's cyclic Jacobi with a
// relative stopping test (and more sweeps in FP64); with ref, the
// reference's own constants (skip 1e-6, stop 1e-5, 5 sweeps; svd.h:24-107).
template <class R> __device__ __forceinline__ void svd_d(const m3<R>& F, m3<R>& U, v3<R>& S, m3<R>& V, int ref) {
    if (ref) {
        jacobi_svd<R, 5>(F, U, S, V, R(1e-6), R(1e-5));
        return;
    }
    R fro = R(0);
#pragma unroll
    for (int c = 0; c < 9; ++c) fro += F.a[c] * F.a[c];
    if (sizeof(R) == 8)
        jacobi_svd<R, 12>(F, U, S, V, R(1e-300), R(1e-14) * fro);
    else
        jacobi_svd<R, 8>(F, U, S, V, R(1e-30), R(1e-7) * fro);
}

// h and dh/dsigma of the Drucker-Prager return map (mpm_solver.h:114-166):
// case 0 tension (h = 1), 1 elastic (h = sigma), 2 projected onto the cone.
template <class R> struct DPh {
    v3<R> h;
    R J[9];             // J[k*3+j] = dh_k / dsigma_j
    R dalpha[3], dkoh[3];  // dh_k / dalpha, dh_k / dkoh
};
template <class R> __device__ __forceinline__ DPh<R> dp_h(v3<R> s, R alpha, R koh) {
    DPh<R> o;
    const R sv[3] = {fmax(s.x, R(1e-6)), fmax(s.y, R(1e-6)), fmax(s.z, R(1e-6))};
    const R e[3] = {log(sv[0]), log(sv[1]), log(sv[2])};
    const R tr = e[0] + e[1] + e[2];
    const R eh[3] = {e[0] - tr / R(3), e[1] - tr / R(3), e[2] - tr / R(3)};
    const R H = sqrt(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
#pragma unroll
    for (int c = 0; c < 9; ++c) o.J[c] = R(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) o.dalpha[k] = o.dkoh[k] = R(0);
    if (tr > R(0)) {
        o.h = mk(R(1), R(1), R(1));
        return o;
    }
    const R limit = -alpha * koh * tr;
    if (!(H > limit)) {
        o.h = s;
        o.J[0] = o.J[4] = o.J[8] = R(1);
        return o;
    }
    const R sc = limit / H;
    R hk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) hk[k] = exp(sc * eh[k] + tr / R(3));
    o.h = mk(hk[0], hk[1], hk[2]);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const R dsc = -alpha * koh / H - sc * eh[j] / (H * H);  // d sc / d e_j
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const R de = sc * ((k == j ? R(1) : R(0)) - R(1) / R(3)) + eh[k] * dsc + R(1) / R(3);
            o.J[k * 3 + j] = hk[k] * de / sv[j];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o.dalpha[k] = hk[k] * eh[k] * (-koh * tr / H);
        o.dkoh[k] = hk[k] * eh[k] * (-alpha * tr / H);
    }
    return o;
}

// h = tau_k / sigma_k of the Hencky-Kirchhoff stress (mpm_solver.h:186-217)
// with its sigma, mu and lambda derivatives.
template <class R> struct Hh {
    v3<R> h;
    R J[9];
    R dmu[3], dlam[3];
};
template <class R> __device__ __forceinline__ Hh<R> hencky_h(v3<R> s, R mu, R lam) {
    Hh<R> o;
    const R sv[3] = {fmax(s.x, R(1e-6)), fmax(s.y, R(1e-6)), fmax(s.z, R(1e-6))};
    const R e[3] = {log(sv[0]), log(sv[1]), log(sv[2])};
    const R tr = e[0] + e[1] + e[2];
    R hk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R tau = R(2) * mu * e[k] + lam * tr;
        hk[k] = tau / sv[k];
        o.dmu[k] = R(2) * e[k] / sv[k];
        o.dlam[k] = tr / sv[k];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const R dtau = (k == j ? R(2) * mu : R(0)) / sv[j] + lam / sv[j];
            o.J[k * 3 + j] = (dtau - (k == j ? hk[k] : R(0))) / sv[k];
        }
    }
    o.h = mk(hk[0], hk[1], hk[2]);
    return o;
}

template <class R> __device__ __forceinline__ m3<R> usv(const m3<R>& U, v3<R> h, const m3<R>& V) {
    m3<R> D;
#pragma unroll
    for (int c = 0; c < 9; ++c) D.a[c] = R(0);
    D.a[0] = h.x;
    D.a[4] = h.y;
    D.a[8] = h.z;
    return mmt(mmul(U, D), V);
}

// Fbar of G(F) = U diag(h) V^T given Gbar; returns M = U^T Gbar V too.
template <class R>
__device__ __forceinline__ m3<R> usv_adjoint(const m3<R>& U, v3<R> s, const m3<R>& V, v3<R> h, const R J[9],
                                             const m3<R>& Gbar, m3<R>& M) {
    M = mmul(mtmul(U, Gbar), V);
    const R sv[3] = {s.x, s.y, s.z}, hv[3] = {h.x, h.y, h.z};
    const R thr = sizeof(R) == 8 ? R(1e-7) : R(1e-3);
    m3<R> K;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R acc = R(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) acc += M.a[i * 4] * J[i * 3 + k];
        K.a[k * 4] = acc;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (i == j) continue;
            const R ds = sv[j] - sv[i];
            R D;
            if (fabs(ds) > thr * (sv[i] + sv[j]))
                D = (hv[j] - hv[i]) / ds;
            else
                D = R(0.5) * ((J[j * 3 + j] - J[i * 3 + j]) + (J[i * 3 + i] - J[j * 3 + i]));
            const R S = (hv[j] + hv[i]) / (sv[j] + sv[i]);
            K.a[i * 3 + j] = M.a[i * 3 + j] * R(0.5) * (D + S) + M.a[j * 3 + i] * R(0.5) * (D - S);
        }
    return mmt(mmul(U, K), V);
}

// Kirchhoff stress tau = P F^T with P = U diag(h) V^T.
template <class R> __device__ __forceinline__ m3<R> stress(const m3<R>& F, R mu, R lam, int ref) {
    m3<R> U, V;
    v3<R> s;
    svd_d(F, U, s, V, ref);
    const Hh<R> H = hencky_h(s, mu, lam);
    return mmt(usv(U, H.h, V), F);
}

// ------------------------------------------------------------------ wave sums
template <class R> __device__ __forceinline__ void wave_add(double* dst, R v) {
    double x = (double)v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0 && x != 0.0) atomicAdd(dst, x);
}

// ------------------------------------------------------------------ lane runs
// Upload groups a cell's particles into consecutive lanes, and lanes with one
// stencil base walking the 27 nodes in one order add to the same node at each
// step.  A run = the consecutive lanes of an aligned 8-lane group with the
// same base: its lanes sum their values in registers (DPP row shifts, which
// stay inside the 16-lane row that holds the group) and only the run's first
// lane issues the atomic, so a wave's LDS atomics carry about one lane in
// four instead of every lane.
struct LaneRun {
    bool head;  // the run's first lane (every dead lane is a run of its own)
    int rem;    // lanes after this one in its run (0..7)
};
__device__ __forceinline__ LaneRun lane_run(bool live, const int b[3]) {
    const int lane = (int)(threadIdx.x & 63);
    const int px = __shfl_up(b[0], 1), py = __shfl_up(b[1], 1), pz = __shfl_up(b[2], 1);
    const bool plive = __shfl_up(live ? 1 : 0, 1) != 0;
    LaneRun r;
    r.head = !live || (lane & 7) == 0 || !plive || px != b[0] || py != b[1] || pz != b[2];
    const unsigned long long cont = __ballot(!r.head);
    const unsigned long long after = lane == 63 ? 0ull : cont >> (lane + 1);
    r.rem = __builtin_ctzll(~after);
    return r;
}
// The value of lane + N of the same 16-lane row (DPP row_shl:N).
template <int N> __device__ __forceinline__ float lane_above(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + N, 0xf, 0xf, true));
}
template <int N> __device__ __forceinline__ double lane_above(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u & 0xffffffffull), 0x100 + N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x100 + N, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// Segmented suffix sum inside the run: the head gets the run's total
// (runs are at most 8 lanes, so three doubling steps).
template <class R> __device__ __forceinline__ R run_sum(R v, int rem) {
    R t = lane_above<1>(v);
    v += rem >= 1 ? t : R(0);
    t = lane_above<2>(v);
    v += rem >= 2 ? t : R(0);
    t = lane_above<4>(v);
    v += rem >= 4 ? t : R(0);
    return v;
}

// ------------------------------------------------------------------ scatter tile
// The scatter kernels run one wave per workgroup over consecutive particles,
// which sit close together: the wave's deposits are summed in an LDS tile
// over the nodes its stencils cover and flushed with one global atomic per
// node and value, instead of 27 contended global atomics per particle.  A
// wave whose nodes do not fit the tile deposits directly.
constexpr int STILE = 1024;

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

struct TileBox {
    int lo[3], ex[3];
    bool on;
    __device__ __forceinline__ int at(int i, int j, int k) const {
        return ((i - lo[0]) * ex[1] + (j - lo[1])) * ex[2] + (k - lo[2]);
    }
};
// The box of nodes the live lanes' stencils (base .. base + 2) cover.
__device__ __forceinline__ TileBox tile_box(bool live, const int base[3]) {
    TileBox t;
    int vol = 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        t.lo[a] = wave_min_i(live ? base[a] : INT_MAX);
        const int hi = wave_max_i(live ? base[a] + 2 : INT_MIN);
        t.ex[a] = hi - t.lo[a] + 1;
        vol = (t.ex[a] > 0 && vol <= STILE) ? vol * t.ex[a] : STILE + 1;
    }
    t.on = vol <= STILE;
    return t;
}

// NV values per node from the tile to global memory (dst[v] + NV * id + ...):
// out-of-box nodes raise the domain error.
template <class R, int NV>
__device__ __forceinline__ void tile_flush(const TileBox& t, R (*tl)[STILE], const Box& b, R* g0, R* g1, int* err) {
    const int vol = t.ex[0] * t.ex[1] * t.ex[2];
    for (int e = threadIdx.x; e < vol; e += 64) {
        bool any = false;
#pragma unroll
        for (int v = 0; v < NV; ++v) any = any || tl[v][e] != R(0);
        if (!any) continue;
        const int k = e % t.ex[2], j = (e / t.ex[2]) % t.ex[1], i = e / (t.ex[2] * t.ex[1]);
        const int64_t id = node_index<R>(b, t.lo[0] + i, t.lo[1] + j, t.lo[2] + k);
        if (id < 0) {
            if (err) atomicOr(err, 1);
            continue;
        }
        if (g0) atomicAdd(g0 + id, tl[0][e]);
#pragma unroll
        for (int v = (g0 ? 1 : 0); v < NV; ++v) atomicAdd(g1 + 3 * id + (v - (g0 ? 1 : 0)), tl[v][e]);
    }
}

// ------------------------------------------------------------------ forward
template <class R>
__global__ void __launch_bounds__(64) k_p2g(PS<R> s, const R* m, const R* V0, int64_t n, Box b,
                                            const Par<R>* Pv, double t, R dt, R* gm, R* gq, int* err, const R* ks) {
    __shared__ R tl[4][STILE];
    const Par<R> P = par_at(Pv, t);
    s = roll(s);
    gm += roll_off(4 * box_nodes(b));
    gq += roll_off(4 * box_nodes(b));
    if (ks) ks += roll_off(9 * s.cap);
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p < n;
    R x[3] = {R(0), R(0), R(0)};
    St<R> st{};
    if (live) {
        x[0] = s.at(QX, p);
        x[1] = s.at(QX + 1, p);
        x[2] = s.at(QX + 2, p);
        stencil(x, P.inv_dx, st);
    }
    const TileBox tb = tile_box(live, st.b);
    const LaneRun run = lane_run(live, st.b);
    if (tb.on) {
        for (int e = threadIdx.x; e < STILE; e += 64) tl[0][e] = tl[1][e] = tl[2][e] = tl[3][e] = R(0);
        __syncthreads();
    }
    if (live) {  // a run's lanes are all live
        const v3<R> v = mk(s.at(QV, p), s.at(QV + 1, p), s.at(QV + 2, p));
        const m3<R> C = load_m(s, QC, p), F = load_m(s, QF, p);
        const R mp = m[p];
        const R kf = -R(4) * P.inv_dx * P.inv_dx * V0[p];
        m3<R> K;
        if (ks) {
#pragma unroll
            for (int c = 0; c < 9; ++c) K.a[c] = ks[c * s.cap + p];
        } else {
            K = stress(F, P.mu, P.lam, P.ref_svd);
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) K.a[c] *= kf;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R w = st.w[0][i] * st.w[1][j] * st.w[2][k];
                    const int ni = st.b[0] + i, nj = st.b[1] + j, nk = st.b[2] + k;
                    const v3<R> d = mk((R)ni * P.dx - x[0], (R)nj * P.dx - x[1], (R)nk * P.dx - x[2]);
                    const v3<R> a = (v + mvec(C, d)) * mp + mvec(K, d) * dt;
                    const R dm = run_sum(w * mp, run.rem), dqx = run_sum(w * a.x, run.rem),
                            dqy = run_sum(w * a.y, run.rem), dqz = run_sum(w * a.z, run.rem);
                    if (!run.head) continue;
                    if (tb.on) {
                        const int e = tb.at(ni, nj, nk);
                        atomicAdd(&tl[0][e], dm);
                        atomicAdd(&tl[1][e], dqx);
                        atomicAdd(&tl[2][e], dqy);
                        atomicAdd(&tl[3][e], dqz);
                        continue;
                    }
                    const int64_t id = node_index<R>(b, ni, nj, nk);
                    if (id < 0) {
                        atomicOr(err, 1);
                        continue;
                    }
                    atomicAdd(gm + id, dm);
                    atomicAdd(gq + 3 * id + 0, dqx);
                    atomicAdd(gq + 3 * id + 1, dqy);
                    atomicAdd(gq + 3 * id + 2, dqz);
                }
    }
    if (tb.on) {
        __syncthreads();
        tile_flush<R, 4>(tb, tl, b, gm, gq, err);
    }
}

// The tool's probe at a node: inside flag and outward world normal.
template <class R> __device__ __forceinline__ bool tool_probe(const Par<R>& P, v3<R> xn, v3<R>& n) {
    const v3<R> r = xn - mk(P.tc[0], P.tc[1], P.tc[2]);
    m3<R> Rw, Rl;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        Rw.a[c] = P.Rw[c];
        Rl.a[c] = P.Rl[c];
    }
    const v3<R> l = mvec(Rw, r);
    if (!(fabs(l.x) <= P.he[0] && fabs(l.y) <= P.he[1] && fabs(l.z) <= P.he[2])) return false;
    const R ddx = P.he[0] - fabs(l.x), ddy = P.he[1] - fabs(l.y), ddz = P.he[2] - fabs(l.z);
    v3<R> nl;
    if (ddx <= ddy && ddx <= ddz)
        nl = mk(l.x >= R(0) ? R(1) : R(-1), R(0), R(0));
    else if (ddy <= ddz)
        nl = mk(R(0), l.y >= R(0) ? R(1) : R(-1), R(0));
    else
        nl = mk(R(0), R(0), l.z >= R(0) ? R(1) : R(-1));
    n = mvec(Rl, nl);
    const R len = sqrt(dot(n, n));
    n = len < R(1e-6) ? mk(R(0), R(0), R(0)) : n * (R(1) / len);
    return true;
}

// Node velocity through update, floor and M0; the intermediate values the
// adjoint needs are recomputed from (m, q) by the same routine.
template <class R> struct NodeFwd {
    bool upd, floor_on, floor_tiny, floor_zero, m0, m0_node_branch;
    v3<R> u0, u1, u2, n;
    R sc, tm, vn, fric;
};
template <class R>
__device__ __forceinline__ NodeFwd<R> node_forward(R m, v3<R> q, int j_node, v3<R> xn_c, const Par<R>& P, R dt) {
    NodeFwd<R> o{};
    o.upd = m > R(1e-6);
    if (o.upd) {
        o.u0 = q * (R(1) / m) + mk(P.g[0] * dt, P.g[1] * dt, P.g[2] * dt);
        o.u1 = o.u0;
        const R ny = ((R)j_node + R(0.5)) * P.dx;
        if (ny < P.dx && o.u0.y < R(0)) {
            o.floor_on = true;
            o.vn = -o.u0.y;
            o.u1.y = R(0);
            o.tm = sqrt(o.u0.x * o.u0.x + o.u0.z * o.u0.z);
            const R budget = P.mu_floor * o.vn;
            if (o.tm > R(1e-12)) {
                const R raw = (o.tm - budget) / o.tm;
                o.floor_zero = !(raw > R(0));
                o.sc = o.floor_zero ? R(0) : raw;
                o.u1.x = o.u0.x * o.sc;
                o.u1.z = o.u0.z * o.sc;
            } else {
                o.floor_tiny = true;
            }
        }
    } else {
        o.u0 = o.u1 = q;
    }
    o.u2 = o.u1;
    if (P.tool && m > R(1e-8) && tool_probe(P, xn_c, o.n)) {
        o.m0 = true;
        o.fric = fmin(R(1), P.friction);
        const v3<R> vb = mk(P.vt[0], P.vt[1], P.vt[2]);
        const R vnn = dot(o.u1, o.n), vnb = dot(vb, o.n);
        o.m0_node_branch = vnn >= vnb;
        const R vnc = o.m0_node_branch ? vnn : vnb;
        const v3<R> vtn = o.u1 - o.n * vnn, vtb = vb - o.n * vnb;
        o.u2 = o.n * vnc + vtn + (vtb - vtn) * o.fric;
    }
    return o;
}

// Each thread owns one node and reads only that node's (m, q), so it may
// clear them for the next P2G once read (clear_mq: the grid's last reader
// this substep).
template <class R>
__global__ void __launch_bounds__(256) k_update(Box b, const Par<R>* Pv, double t, R dt, R* gm, R* gq, R* gu,
                                                int clear_mq) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = box_nodes(b);
    if (id >= nn) return;
    const Par<R> P = par_at(Pv, t);
    gm += roll_off(4 * nn);
    gq += roll_off(4 * nn);
    gu += roll_off(3 * nn);
    const R m = gm[id];
    const v3<R> q = mk(gq[3 * id], gq[3 * id + 1], gq[3 * id + 2]);
    if (clear_mq) {
        gm[id] = R(0);
        gq[3 * id] = gq[3 * id + 1] = gq[3 * id + 2] = R(0);
    }
    if (m == R(0)) {
        gu[3 * id] = gu[3 * id + 1] = gu[3 * id + 2] = R(0);
        return;
    }
    const int k = (int)(id % b.nz), j = (int)((id / b.nz) % b.ny), i = (int)(id / ((int64_t)b.nz * b.ny));
    const int gi = b.ox + i, gj = b.oy + j, gk = b.oz + k;
    const v3<R> xc = mk(((R)gi + R(0.5)) * P.dx, ((R)gj + R(0.5)) * P.dx, ((R)gk + R(0.5)) * P.dx);
    const NodeFwd<R> f = node_forward(m, q, gj, xc, P, dt);
    gu[3 * id] = f.u2.x;
    gu[3 * id + 1] = f.u2.y;
    gu[3 * id + 2] = f.u2.z;
}

// ------------------------------------------------------------------ kept factors
// The SVD U diag(s) V^T of each particle's trial F at one substep, kept by
// the backward's recompute of a segment for both adjoint kernels: the G2P
// adjoint of that substep factors the same trial F, and the P2G adjoint of
// the next substep factors the projected F = U diag(h(s)) V^T, whose SVD is
// (U, h(s), V).  Component c of particle p at f[c * cap + p]; a rollout's
// factors FCOMP * cap values after the previous rollout's.
constexpr int FCOMP = 21;
template <class R>
__device__ __forceinline__ void store_fac(R* f, int64_t cap, int64_t p, const m3<R>& U, const v3<R>& s,
                                          const m3<R>& V) {
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        f[c * cap + p] = U.a[c];
        f[(12 + c) * cap + p] = V.a[c];
    }
    f[9 * cap + p] = s.x;
    f[10 * cap + p] = s.y;
    f[11 * cap + p] = s.z;
}
template <class R>
__device__ __forceinline__ void load_fac(const R* f, int64_t cap, int64_t p, m3<R>& U, v3<R>& s, m3<R>& V) {
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        U.a[c] = f[c * cap + p];
        V.a[c] = f[(12 + c) * cap + p];
    }
    s = mk(f[9 * cap + p], f[10 * cap + p], f[11 * cap + p]);
}

// G2P forward quantities of one particle (shared by the forward kernel and
// the adjoint's recomputation).
template <class R> struct G2PFwd {
    St<R> st;
    v3<R> vn;
    m3<R> Cn, Ftr, U, V;
    v3<R> s;
    DPh<R> dp;
    v3<R> xn;
    bool clamped;
};
template <class R>
__device__ __forceinline__ void g2p_forward(const R x[3], const m3<R>& F, const Box& b, const Par<R>& P, R dt,
                                            const R* gu, G2PFwd<R>& o, int* err, const R* fac = nullptr,
                                            int64_t cap = 0, int64_t p = 0) {
    stencil(x, P.inv_dx, o.st);
    const R kC = R(4) * P.inv_dx * P.inv_dx;
    o.vn = mk(R(0), R(0), R(0));
#pragma unroll
    for (int c = 0; c < 9; ++c) o.Cn.a[c] = R(0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                const R w = o.st.w[0][i] * o.st.w[1][j] * o.st.w[2][k];
                const int ni = o.st.b[0] + i, nj = o.st.b[1] + j, nk = o.st.b[2] + k;
                const int64_t id = node_index<R>(b, ni, nj, nk);
                if (id < 0) {
                    if (err) atomicOr(err, 1);
                    continue;
                }
                const v3<R> u = mk(gu[3 * id], gu[3 * id + 1], gu[3 * id + 2]);
                const v3<R> d = mk((R)ni * P.dx - x[0], (R)nj * P.dx - x[1], (R)nk * P.dx - x[2]);
                o.vn = o.vn + u * w;
                const v3<R> wu = u * (w * kC);
                o.Cn.a[0] += wu.x * d.x; o.Cn.a[1] += wu.x * d.y; o.Cn.a[2] += wu.x * d.z;
                o.Cn.a[3] += wu.y * d.x; o.Cn.a[4] += wu.y * d.y; o.Cn.a[5] += wu.y * d.z;
                o.Cn.a[6] += wu.z * d.x; o.Cn.a[7] += wu.z * d.y; o.Cn.a[8] += wu.z * d.z;
            }
    o.xn = mk(x[0] + dt * o.vn.x, x[1] + dt * o.vn.y, x[2] + dt * o.vn.z);
    m3<R> A = o.Cn;
#pragma unroll
    for (int c = 0; c < 9; ++c) A.a[c] *= dt;
    A.a[0] += R(1);
    A.a[4] += R(1);
    A.a[8] += R(1);
    o.Ftr = mmul(A, F);
    if (fac)
        load_fac(fac, cap, p, o.U, o.s, o.V);  // the factors the segment's recompute kept
    else
        svd_d(o.Ftr, o.U, o.s, o.V, P.ref_svd);
    o.dp = dp_h(o.s, P.alpha, P.koh);
    o.clamped = o.xn.y < R(0);
}

// fac (optional): the trial F's factors of this substep, kept for the
// backward; out.p == nullptr: the factors alone (the segment's last substep,
// whose new state is the next checkpoint, already stored).
template <class R>
__global__ void __launch_bounds__(256) k_g2p(PS<R> s, PS<R> out, int64_t n, Box b, const Par<R>* Pv, double t,
                                             R dt, const R* gu, int* err, R* ks, R* fac) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Par<R> P = par_at(Pv, t);
    s = roll(s);
    gu += roll_off(3 * box_nodes(b));
    const R x[3] = {s.at(QX, p), s.at(QX + 1, p), s.at(QX + 2, p)};
    const m3<R> F = load_m(s, QF, p);
    G2PFwd<R> f;
    g2p_forward(x, F, b, P, dt, gu, f, err);
    if (fac) store_fac(fac + roll_off(FCOMP * s.cap), s.cap, p, f.U, f.s, f.V);
    if (!out.p) return;
    out = roll(out);
    if (ks) ks += roll_off(9 * out.cap);
    v3<R> xn = f.xn, vn = f.vn;
    if (f.clamped) {
        xn.y = R(0);
        vn.y = R(0);
    }
    out.at(QX, p) = xn.x;
    out.at(QX + 1, p) = xn.y;
    out.at(QX + 2, p) = xn.z;
    out.at(QV, p) = vn.x;
    out.at(QV + 1, p) = vn.y;
    out.at(QV + 2, p) = vn.z;
    store_m(out, QC, p, f.Cn);
    store_m(out, QF, p, usv(f.U, f.dp.h, f.V));
    if (ks) {
        // the new F is U diag(s') V^T (s' = f.dp.h), so the Kirchhoff stress the
        // next deposit needs, (U diag(hk) V^T) F^T, is U diag(hk s') U^T: no
        // second SVD of the same F
        const Hh<R> H = hencky_h(f.dp.h, P.mu, P.lam);
        const m3<R> K = usv(f.U, mk(H.h.x * f.dp.h.x, H.h.y * f.dp.h.y, H.h.z * f.dp.h.z), f.U);
#pragma unroll
        for (int c = 0; c < 9; ++c) ks[c * out.cap + p] = K.a[c];
    }
}

// The Kirchhoff stress of each particle's F (component c of particle p at
// ks[c * cap + p]): the first deposit after a checkpoint reads it.
template <class R>
__global__ void __launch_bounds__(256) k_stress(PS<R> s, int64_t n, const Par<R>* Pv, R* ks) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Par<R> P = par_at(Pv, 0.0);
    s = roll(s);
    ks += roll_off(9 * s.cap);
    const m3<R> K = stress(load_m(s, QF, p), P.mu, P.lam, P.ref_svd);
#pragma unroll
    for (int c = 0; c < 9; ++c) ks[c * s.cap + p] = K.a[c];
}

// ------------------------------------------------------------------ loss
// L = (1/N) sum_p |x_p(T) - x_p(0) - d_p|^2; seeds xbar of the final state.
template <class R>
__global__ void __launch_bounds__(256) k_loss(PS<R> fin, const R* x0, const R* tgt, int64_t n, PS<R> bar,
                                              double* acc) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    fin = roll(fin);
    bar = roll(bar);
    acc += roll_off(GN);
    R l = R(0);
    if (p < n) {
        const R inv = R(1) / (R)n;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const R r = fin.at(QX + a, p) - x0[3 * p + a] - tgt[3 * p + a];
            l += r * r * inv;
            bar.at(QX + a, p) = R(2) * r * inv;
        }
#pragma unroll
        for (int c = QV; c < QCOMP; ++c) bar.at(c, p) = R(0);
    }
    wave_add(acc + GLOSS, l);
}

// ------------------------------------------------------------------ adjoint
// G2P: from the adjoint of the new state (ain) to the adjoint of x and F of
// the old state (aout; its v and C parts are zeroed for the P2G adjoint) and
// the adjoint of the node velocities (gub, accumulated).
template <class R>
__global__ void __launch_bounds__(64) k_g2p_adj(PS<R> s, PS<R> ain, PS<R> aout, int64_t n, Box b,
                                                const Par<R>* Pv, double t, R dt, const R* gu, R* gub,
                                                double* grad, const R* fac) {
    __shared__ R tl[3][STILE];
    const Par<R> P = par_at(Pv, t);
    s = roll(s);
    ain = roll(ain);
    aout = roll(aout);
    gu += roll_off(3 * box_nodes(b));
    gub += roll_off(3 * box_nodes(b));
    grad += roll_off(GN);
    if (fac) fac += roll_off(FCOMP * s.cap);
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    R galpha = R(0), gkoh = R(0);
    int base[3] = {0, 0, 0};
    if (p < n) {
        const R xs[3] = {s.at(QX, p), s.at(QX + 1, p), s.at(QX + 2, p)};
        St<R> st;
        stencil(xs, P.inv_dx, st);
        base[0] = st.b[0];
        base[1] = st.b[1];
        base[2] = st.b[2];
    }
    const TileBox tb = tile_box(p < n, base);
    const LaneRun run = lane_run(p < n, base);
    if (tb.on) {
        for (int e = threadIdx.x; e < STILE; e += 64) tl[0][e] = tl[1][e] = tl[2][e] = R(0);
        __syncthreads();
    }
    if (p < n) {  // a run's lanes are all live
        const R x[3] = {s.at(QX, p), s.at(QX + 1, p), s.at(QX + 2, p)};
        const m3<R> F = load_m(s, QF, p);
        G2PFwd<R> f;
        g2p_forward(x, F, b, P, dt, gu, f, (int*)nullptr, fac, s.cap, p);
        v3<R> xb = mk(ain.at(QX, p), ain.at(QX + 1, p), ain.at(QX + 2, p));
        v3<R> vb = mk(ain.at(QV, p), ain.at(QV + 1, p), ain.at(QV + 2, p));
        m3<R> Cb = load_m(ain, QC, p);
        const m3<R> Fb_new = load_m(ain, QF, p);
        if (f.clamped) {  // x'.y and v'.y were set to constants
            xb.y = R(0);
            vb.y = R(0);
        }
        // F' = proj(Ftr)
        m3<R> M;
        const m3<R> Ftrb = usv_adjoint(f.U, f.s, f.V, f.dp.h, f.dp.J, Fb_new, M);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            galpha += M.a[k * 4] * f.dp.dalpha[k];
            gkoh += M.a[k * 4] * f.dp.dkoh[k];
        }
        // Ftr = (I + dt C') F
        m3<R> A = f.Cn;
#pragma unroll
        for (int c = 0; c < 9; ++c) A.a[c] *= dt;
        A.a[0] += R(1);
        A.a[4] += R(1);
        A.a[8] += R(1);
        const m3<R> Fb = mtmul(A, Ftrb);
        const m3<R> dC = mmt(Ftrb, F);
#pragma unroll
        for (int c = 0; c < 9; ++c) Cb.a[c] += dt * dC.a[c];
        // x' = x + dt v'
        vb = vb + xb * dt;
        v3<R> xo = xb;
        const R kC = R(4) * P.inv_dx * P.inv_dx;
        // one node order for every lane: a run's lanes add to the same node
        // at each step (lane runs above)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R w = f.st.w[0][i] * f.st.w[1][j] * f.st.w[2][k];
                    const v3<R> gw = mk(f.st.dw[0][i] * f.st.w[1][j] * f.st.w[2][k],
                                        f.st.w[0][i] * f.st.dw[1][j] * f.st.w[2][k],
                                        f.st.w[0][i] * f.st.w[1][j] * f.st.dw[2][k]);
                    const int ni = f.st.b[0] + i, nj = f.st.b[1] + j, nk = f.st.b[2] + k;
                    const int64_t id = node_index<R>(b, ni, nj, nk);
                    if (id < 0) continue;
                    const v3<R> u = mk(gu[3 * id], gu[3 * id + 1], gu[3 * id + 2]);
                    const v3<R> d = mk((R)ni * P.dx - x[0], (R)nj * P.dx - x[1], (R)nk * P.dx - x[2]);
                    const v3<R> Cd = mvec(Cb, d);
                    // v' = sum w u;  C' = kC sum w u d^T
                    const v3<R> ub = (vb + Cd * kC) * w;
                    // (a run shares its base, so the whole run skipped an
                    // out-of-box node together above)
                    const R ubx = run_sum(ub.x, run.rem), uby = run_sum(ub.y, run.rem),
                            ubz = run_sum(ub.z, run.rem);
                    if (run.head) {
                        if (tb.on) {
                            const int e = tb.at(ni, nj, nk);
                            atomicAdd(&tl[0][e], ubx);
                            atomicAdd(&tl[1][e], uby);
                            atomicAdd(&tl[2][e], ubz);
                        } else {
                            atomicAdd(gub + 3 * id + 0, ubx);
                            atomicAdd(gub + 3 * id + 1, uby);
                            atomicAdd(gub + 3 * id + 2, ubz);
                        }
                    }
                    const R su = dot(u, vb) + kC * dot(u, Cd);
                    const v3<R> CTu = mk(Cb.a[0] * u.x + Cb.a[3] * u.y + Cb.a[6] * u.z,
                                         Cb.a[1] * u.x + Cb.a[4] * u.y + Cb.a[7] * u.z,
                                         Cb.a[2] * u.x + Cb.a[5] * u.y + Cb.a[8] * u.z);
                    xo = xo + gw * su - CTu * (kC * w);
                }
        aout.at(QX, p) = xo.x;
        aout.at(QX + 1, p) = xo.y;
        aout.at(QX + 2, p) = xo.z;
        aout.at(QV, p) = aout.at(QV + 1, p) = aout.at(QV + 2, p) = R(0);
#pragma unroll
        for (int c = 0; c < 9; ++c) aout.at(QC + c, p) = R(0);
        store_m(aout, QF, p, Fb);
    }
    if (tb.on) {
        __syncthreads();
        tile_flush<R, 3>(tb, tl, b, (R*)nullptr, gub, (int*)nullptr);
    }
    wave_add(grad + GALPHA, galpha);
    wave_add(grad + GKOH, gkoh);
}

// Grid: from the node-velocity adjoint to the adjoint of (m, q); the tool
// velocity's adjoint from the M0 nodes.
template <class R>
__global__ void __launch_bounds__(256) k_update_adj(Box b, const Par<R>* Pv, double t, R dt, R* gm, R* gq,
                                                    R* gub, R* gqb, R* gmb, double* grad, int clear_mq) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = box_nodes(b);
    const Par<R> P = par_at(Pv, t);
    gm += roll_off(4 * nn);
    gq += roll_off(4 * nn);
    gub += roll_off(3 * nn);
    gqb += roll_off(3 * nn);
    gmb += roll_off(nn);
    grad += roll_off(GN);
    v3<R> vtb = mk(R(0), R(0), R(0));
    if (id < nn) {
        // the substep's last reader of (m, q) and of the velocity adjoint:
        // each thread reads its own node and leaves it cleared for the next
        // P2G / G2P adjoint that accumulates into it; (m, q) that the forward
        // kept (clear_mq = 0) stay for a later backward
        const R m = gm[id];
        const v3<R> q = mk(gq[3 * id], gq[3 * id + 1], gq[3 * id + 2]);
        const v3<R> u2b_in = mk(gub[3 * id], gub[3 * id + 1], gub[3 * id + 2]);
        if (clear_mq) {
            gm[id] = R(0);
            gq[3 * id] = gq[3 * id + 1] = gq[3 * id + 2] = R(0);
        }
        gub[3 * id] = gub[3 * id + 1] = gub[3 * id + 2] = R(0);
        gqb[3 * id] = gqb[3 * id + 1] = gqb[3 * id + 2] = R(0);
        gmb[id] = R(0);
        if (m != R(0)) {
            const int k = (int)(id % b.nz), j = (int)((id / b.nz) % b.ny), i = (int)(id / ((int64_t)b.nz * b.ny));
            const int gj = b.oy + j;
            const v3<R> xc = mk(((R)(b.ox + i) + R(0.5)) * P.dx, ((R)gj + R(0.5)) * P.dx,
                                ((R)(b.oz + k) + R(0.5)) * P.dx);
            const NodeFwd<R> f = node_forward(m, q, gj, xc, P, dt);
            v3<R> u2b = u2b_in;
            v3<R> u1b = u2b;
            if (f.m0) {
                const R nb = dot(u2b, f.n);
                const v3<R> tan = u2b - f.n * nb;
                u1b = tan * (R(1) - f.fric) + (f.m0_node_branch ? f.n * nb : mk(R(0), R(0), R(0)));
                vtb = tan * f.fric + (f.m0_node_branch ? mk(R(0), R(0), R(0)) : f.n * nb);
            }
            v3<R> u0b = u1b;
            if (f.floor_on) {
                if (f.floor_tiny) {
                    u0b = mk(u1b.x, R(0), u1b.z);
                } else if (f.floor_zero) {
                    u0b = mk(R(0), R(0), R(0));
                } else {
                    const R gsc = u1b.x * f.u0.x + u1b.z * f.u0.z;  // dL / d sc
                    const R t3 = f.tm * f.tm * f.tm;
                    u0b.x = u1b.x * f.sc + gsc * P.mu_floor * f.vn * f.u0.x / t3;
                    u0b.z = u1b.z * f.sc + gsc * P.mu_floor * f.vn * f.u0.z / t3;
                    u0b.y = gsc * P.mu_floor / f.tm;
                }
            }
            if (f.upd) {
                const R im = R(1) / m;
                gqb[3 * id] = u0b.x * im;
                gqb[3 * id + 1] = u0b.y * im;
                gqb[3 * id + 2] = u0b.z * im;
                gmb[id] = -dot(u0b, q) * im * im;
            } else {
                gqb[3 * id] = u0b.x;
                gqb[3 * id + 1] = u0b.y;
                gqb[3 * id + 2] = u0b.z;
            }
        }
    }
    wave_add(grad + GVT + 0, vtb.x);
    wave_add(grad + GVT + 1, vtb.y);
    wave_add(grad + GVT + 2, vtb.z);
}

// P2G: from the adjoint of (m, q) to the adjoint of x, v, C, F (added to
// aout) and of mu and lambda.
// facp (optional): the previous substep's kept factors (U, s, V of its trial
// F): this state's F is U diag(h(s)) V^T, so its SVD is (U, h(s), V).
template <class R>
__global__ void __launch_bounds__(256) k_p2g_adj(PS<R> s, const R* m, const R* V0, int64_t n, Box b,
                                                 const Par<R>* Pv, double t, R dt, const R* gqb, const R* gmb,
                                                 PS<R> aout, double* grad, const R* facp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const Par<R> P = par_at(Pv, t);
    s = roll(s);
    aout = roll(aout);
    gqb += roll_off(3 * box_nodes(b));
    gmb += roll_off(box_nodes(b));
    grad += roll_off(GN);
    R gmu = R(0), glam = R(0);
    if (p < n) {
        const R x[3] = {s.at(QX, p), s.at(QX + 1, p), s.at(QX + 2, p)};
        const v3<R> v = mk(s.at(QV, p), s.at(QV + 1, p), s.at(QV + 2, p));
        const m3<R> C = load_m(s, QC, p), F = load_m(s, QF, p);
        const R mp = m[p];
        const R kf = -R(4) * P.inv_dx * P.inv_dx * V0[p];
        m3<R> U, V;
        v3<R> sg;
        if (facp) {
            v3<R> sp;
            load_fac(facp + roll_off(FCOMP * s.cap), s.cap, p, U, sp, V);
            sg = dp_h(sp, P.alpha, P.koh).h;
        } else {
            svd_d(F, U, sg, V, P.ref_svd);
        }
        const Hh<R> H = hencky_h(sg, P.mu, P.lam);
        const m3<R> Pm = usv(U, H.h, V);
        m3<R> K = mmt(Pm, F);
#pragma unroll
        for (int c = 0; c < 9; ++c) K.a[c] *= kf;
        St<R> st;
        stencil(x, P.inv_dx, st);
        v3<R> xb = mk(R(0), R(0), R(0)), vb = xb;
        m3<R> Cb, Kb;
#pragma unroll
        for (int c = 0; c < 9; ++c) Cb.a[c] = Kb.a[c] = R(0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R w = st.w[0][i] * st.w[1][j] * st.w[2][k];
                    const v3<R> gw = mk(st.dw[0][i] * st.w[1][j] * st.w[2][k],
                                        st.w[0][i] * st.dw[1][j] * st.w[2][k],
                                        st.w[0][i] * st.w[1][j] * st.dw[2][k]);
                    const int ni = st.b[0] + i, nj = st.b[1] + j, nk = st.b[2] + k;
                    const int64_t id = node_index<R>(b, ni, nj, nk);
                    if (id < 0) continue;
                    const v3<R> Qb = mk(gqb[3 * id], gqb[3 * id + 1], gqb[3 * id + 2]);
                    const R Mb = gmb[id];
                    const v3<R> d = mk((R)ni * P.dx - x[0], (R)nj * P.dx - x[1], (R)nk * P.dx - x[2]);
                    const v3<R> a = (v + mvec(C, d)) * mp + mvec(K, d) * dt;
                    vb = vb + Qb * (w * mp);
                    const v3<R> wq = Qb * w;
                    // C += w m Qb d^T ; K += dt w Qb d^T
                    Cb.a[0] += wq.x * d.x * mp; Cb.a[1] += wq.x * d.y * mp; Cb.a[2] += wq.x * d.z * mp;
                    Cb.a[3] += wq.y * d.x * mp; Cb.a[4] += wq.y * d.y * mp; Cb.a[5] += wq.y * d.z * mp;
                    Cb.a[6] += wq.z * d.x * mp; Cb.a[7] += wq.z * d.y * mp; Cb.a[8] += wq.z * d.z * mp;
                    Kb.a[0] += wq.x * d.x * dt; Kb.a[1] += wq.x * d.y * dt; Kb.a[2] += wq.x * d.z * dt;
                    Kb.a[3] += wq.y * d.x * dt; Kb.a[4] += wq.y * d.y * dt; Kb.a[5] += wq.y * d.z * dt;
                    Kb.a[6] += wq.z * d.x * dt; Kb.a[7] += wq.z * d.y * dt; Kb.a[8] += wq.z * d.z * dt;
                    // weights and d depend on x
                    const v3<R> CTq = mk(C.a[0] * Qb.x + C.a[3] * Qb.y + C.a[6] * Qb.z,
                                         C.a[1] * Qb.x + C.a[4] * Qb.y + C.a[7] * Qb.z,
                                         C.a[2] * Qb.x + C.a[5] * Qb.y + C.a[8] * Qb.z);
                    const v3<R> KTq = mk(K.a[0] * Qb.x + K.a[3] * Qb.y + K.a[6] * Qb.z,
                                         K.a[1] * Qb.x + K.a[4] * Qb.y + K.a[7] * Qb.z,
                                         K.a[2] * Qb.x + K.a[5] * Qb.y + K.a[8] * Qb.z);
                    xb = xb + gw * (dot(Qb, a) + Mb * mp) - (CTq * mp + KTq * dt) * w;
                }
        // K = kf tau,  tau = P F^T
        m3<R> taub = Kb;
#pragma unroll
        for (int c = 0; c < 9; ++c) taub.a[c] *= kf;
        m3<R> Fb = mtmul(taub, Pm);  // tau^T-bar P  (d tau = P dF^T)
        const m3<R> Pb = mmul(taub, F);
        m3<R> M;
        const m3<R> Fb2 = usv_adjoint(U, sg, V, H.h, H.J, Pb, M);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gmu += M.a[k * 4] * H.dmu[k];
            glam += M.a[k * 4] * H.dlam[k];
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) Fb.a[c] += Fb2.a[c];
        aout.at(QX, p) += xb.x;
        aout.at(QX + 1, p) += xb.y;
        aout.at(QX + 2, p) += xb.z;
        aout.at(QV, p) += vb.x;
        aout.at(QV + 1, p) += vb.y;
        aout.at(QV + 2, p) += vb.z;
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            aout.at(QC + c, p) += Cb.a[c];
            aout.at(QF + c, p) += Fb.a[c];
        }
    }
    wave_add(grad + GMU, gmu);
    wave_add(grad + GLAM, glam);
}

// The two per-particle SVD maps of the rollout and their vector-Jacobian
// products on their own, for checking the adjoint in isolation: map 0 the
// Drucker-Prager return map F -> U h(sigma) V^T, map 1 the first
// Piola-Kirchhoff stress P(F) = U diag(tau_k / sigma_k) V^T.
template <class R>
__global__ void __launch_bounds__(64) k_map_vjp(int map, int64_t n, const double* F, const double* Gb, Par<R> P,
                                                double* G, double* Fb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    m3<R> Fm, Gbar;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        Fm.a[c] = (R)F[9 * i + c];
        Gbar.a[c] = (R)Gb[9 * i + c];
    }
    m3<R> U, V;
    v3<R> s;
    svd_d(Fm, U, s, V, 0);
    v3<R> h;
    R J[9];
    if (map == 0) {
        const DPh<R> d = dp_h(s, P.alpha, P.koh);
        h = d.h;
#pragma unroll
        for (int c = 0; c < 9; ++c) J[c] = d.J[c];
    } else {
        const Hh<R> d = hencky_h(s, P.mu, P.lam);
        h = d.h;
#pragma unroll
        for (int c = 0; c < 9; ++c) J[c] = d.J[c];
    }
    const m3<R> Gm = usv(U, h, V);
    m3<R> M;
    const m3<R> Fbar = usv_adjoint(U, s, V, h, J, Gbar, M);
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        G[9 * i + c] = (double)Gm.a[c];
        Fb[9 * i + c] = (double)Fbar.a[c];
    }
}

// Grid-stride: the kept grids of a long rollout exceed the 2^32 threads a
// launch may have.
template <class R> __global__ void k_zero(R* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = R(0);
}

}  // namespace dif
}  // namespace bm

using namespace bm;
using namespace bm::dif;

namespace {

inline int blocks_for(int64_t n) { return (int)((n + 255) / 256); }
// Workgroups of a grid-stride fill (k_zero) of n values.
inline int fill_blocks(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 1 << 16); }
// Particle kernels run 64 lanes per workgroup: a rollout of a few 10^4
// particles then still reaches every CU (each lane's work -- SVD, stress,
// 27-node scatter -- is long, so the launch is latency bound).
constexpr int PWG = 64;
inline int pblocks(int64_t n) { return (int)((n + PWG - 1) / PWG); }

}  // namespace

struct bm_diff {
    bm_diff_desc desc{};
    int nr = 1;  // rollouts run side by side
    std::vector<bm_diff_params> prm;  // per rollout
    std::vector<char> prm_set;
    void* pard = nullptr;  // the rollouts' Par<R>, uploaded by each forward
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // the kept-grid backward recomputes the next segment on a second stream
    // (default priority) while the current one is walked: a second set of
    // segment states and factors, and events that order their reuse
    hipStream_t side = nullptr;
    hipEvent_t ev_rec[2] = {nullptr, nullptr}, ev_walk[2] = {nullptr, nullptr};
    void *seg2 = nullptr, *fac2 = nullptr;
    int64_t seg2_slots = 0;
    bool fp64 = false;
    int K = 16;
    Box box{};
    int64_t cap = 0, n = 0, nn = 0;
    size_t rs = 4;  // bytes per scalar
    // device buffers (scalar type per precision)
    void *m = nullptr, *V0 = nullptr, *x0 = nullptr, *tgt = nullptr, *init = nullptr;
    // gm: per rollout the node masses (nn) then momenta (3 nn), gq = gm + nn
    void *gm = nullptr, *gu = nullptr, *gub = nullptr, *gqb = nullptr, *gmb = nullptr;
    void *ck = nullptr, *seg = nullptr, *abar = nullptr, *abar2 = nullptr;
    // the backward's grids of one segment: (m, q) and u of each of its substeps
    void *sgmq = nullptr, *sgu = nullptr;
    // every substep's grid (m, q and u) of the last forward, kept for the
    // backward when it fits grid_budget bytes (then the backward recomputes
    // only the particle states); all_grids: the last forward kept them
    void *agmq = nullptr, *agu = nullptr;
    int64_t all_slots = 0, grid_budget = int64_t(4) << 30;
    bool all_grids = false;
    // per-particle Kirchhoff stress handed from each G2P to the next P2G
    // (unused with the reference's SVD stopping rule, whose comparisons
    // factor every F as the reference does)
    void* ks = nullptr;
    // the backward's SVD factors of each substep of the segment it walks
    // (FCOMP values per particle and substep)
    void* fac = nullptr;
    int64_t ck_slots = 0, seg_slots = 0, grid_slots = 0, fac_slots = 0;
    // device slot -> the caller's particle index (upload reorders, download restores)
    std::vector<int64_t> order;
    double* grad = nullptr;
    int* err = nullptr;
    // last forward
    float dt = 0.0f;
    int substeps = 0;
    bool have_forward = false;
};

static int check_diff(bm_diff_handle h) { return h ? BM_OK : fail(BM_ERR_INVALID_ARG, "null diff handle"); }

static int check_params(const bm_diff_params* p) {
    if (!p) return fail(BM_ERR_INVALID_ARG, "null params");
    if (!(p->young_modulus > 0.0) || !(p->poisson_ratio > -1.0 && p->poisson_ratio < 0.5))
        return fail(BM_ERR_INVALID_ARG, "bad elastic constants");
    if (p->has_tool && !(p->tool_half_extents[0] > 0 && p->tool_half_extents[1] > 0 && p->tool_half_extents[2] > 0))
        return fail(BM_ERR_INVALID_ARG, "tool half extents must be > 0");
    return BM_OK;
}

// Every rollout has its parameters.
static bool all_params_set(const bm_diff_handle h) {
    return std::all_of(h->prm_set.begin(), h->prm_set.end(), [](char c) { return c != 0; });
}

// The gradient of one rollout's loss from its accumulators (g: GN values):
// the solver's constants (alpha, koh, mu, lambda) chained to the caller's
// parameters (friction angle, Young's modulus).
static void grad_of(const bm_diff_params& q, const double* g, bm_diff_grad* out) {
    const double E = q.young_modulus, nu = q.poisson_ratio;
    const double mu = E / (2.0 * (1.0 + nu)), lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    // koh = (3 lam + 2 mu) / (2 mu)
    const double gmu = g[GMU] + g[GKOH] * (-3.0 * lam / (2.0 * mu * mu));
    const double glam = g[GLAM] + g[GKOH] * (3.0 / (2.0 * mu));
    const double phi = q.friction_deg * M_PI / 180.0, s = std::sin(phi);
    const double dalpha_dphi_deg = std::sqrt(2.0 / 3.0) * 6.0 * std::cos(phi) / ((3.0 - s) * (3.0 - s)) * M_PI / 180.0;
    for (int a = 0; a < 3; ++a) out->tool_velocity[a] = g[GVT + a];
    out->friction_deg = g[GALPHA] * dalpha_dphi_deg;
    out->young_modulus = gmu / (2.0 * (1.0 + nu)) + glam * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    out->alpha = g[GALPHA];
    out->mu = gmu;
    out->lambda = glam;
}

// A rollout's parameters in the solver's constants (the tool centre at time
// 0; the kernels move it, par_at).
template <class R> static Par<R> par_of(const bm_diff_desc& desc, const bm_diff_params& q) {
    Par<R> P{};
    P.dx = (R)desc.dx;
    P.inv_dx = R(1) / P.dx;
    const double E = q.young_modulus, nu = q.poisson_ratio;
    const double mu = E / (2.0 * (1.0 + nu)), lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    P.mu = (R)mu;
    P.lam = (R)lam;
    const double sphi = std::sin(q.friction_deg * M_PI / 180.0);
    P.alpha = (R)(std::sqrt(2.0 / 3.0) * 2.0 * sphi / (3.0 - sphi));
    P.koh = (R)((3.0 * lam + 2.0 * mu) / (2.0 * mu));
    P.mu_floor = (R)q.floor_friction;
    P.ref_svd = desc.reference_svd ? 1 : 0;
    for (int a = 0; a < 3; ++a) P.g[a] = (R)q.gravity[a];
    P.tool = q.has_tool ? 1 : 0;
    if (P.tool) {
        for (int a = 0; a < 3; ++a) {
            P.tc0[a] = q.tool_center[a];
            P.vt0[a] = q.tool_velocity[a];
            P.tc[a] = (R)q.tool_center[a];
            P.he[a] = (R)q.tool_half_extents[a];
            P.vt[a] = (R)q.tool_velocity[a];
        }
        // rotation from the unit quaternion (w, x, y, z)
        double w = q.tool_quat[0], x = q.tool_quat[1], y = q.tool_quat[2], z = q.tool_quat[3];
        const double nq = std::sqrt(w * w + x * x + y * y + z * z);
        if (nq > 0) {
            w /= nq;
            x /= nq;
            y /= nq;
            z /= nq;
        } else {
            w = 1;
        }
        const double Rm[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z),     2 * (x * z + w * y),
                              2 * (x * y + w * z),     1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                              2 * (x * z - w * y),     2 * (y * z + w * x),     1 - 2 * (x * x + y * y)};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                P.Rl[i * 3 + j] = (R)Rm[i * 3 + j];
                P.Rw[i * 3 + j] = (R)Rm[j * 3 + i];
            }
        P.friction = (R)q.tool_friction;
    }
    return P;
}

// State slot k of a slot array: the states of all the handle's rollouts.
template <class R> static PS<R> slot(const bm_diff_handle h, void* base, int64_t k) {
    return PS<R>{static_cast<R*>(base) + k * h->nr * QCOMP * h->cap, h->cap};
}
// Bytes of one state slot.
static size_t slot_bytes(const bm_diff_handle h) { return h->rs * QCOMP * (size_t)h->cap * h->nr; }
// Time of substep `step` of the last forward (the kernels' tool centre).
static double time_at(const bm_diff_handle h, int step) { return (double)step * (double)h->dt; }
template <class R> static const Par<R>* pars(const bm_diff_handle h) { return static_cast<const Par<R>*>(h->pard); }
// One row of workgroups per rollout over n items, w per workgroup.
static dim3 rows(const bm_diff_handle h, int64_t n, int w) { return dim3((unsigned)((n + w - 1) / w), (unsigned)h->nr); }

// Every rollout's parameters to the device, in stream order.
template <class R> static int upload_pars(bm_diff_handle h) {
    std::vector<Par<R>> staging;
    for (int r = 0; r < h->nr; ++r) staging.push_back(par_of<R>(h->desc, h->prm[(size_t)r]));
    HIP_TRY(hipMemcpyAsync(h->pard, staging.data(), sizeof(Par<R>) * h->nr, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));  // before `staging` goes
    return BM_OK;
}

// The stress buffer the deposits read, or null where each deposit factors F
// itself (the reference's SVD stopping rule).
template <class R> static R* stress_buffer(bm_diff_handle h) {
    return h->desc.reference_svd ? nullptr : static_cast<R*>(h->ks);
}

// The stress of state `st` into the buffer, before its first deposit.
template <class R> static void seed_stress(bm_diff_handle h, PS<R> st) {
    if (R* ks = stress_buffer<R>(h))
        hipLaunchKernelGGL(k_stress<R>, rows(h, h->n, 256), dim3(256), 0, h->stream, st, h->n, pars<R>(h), ks);
}

// Zero (m, q): once per forward or backward call.  Every substep then leaves
// them zero for the next one (k_update or k_update_adj clears each node after
// its last read).
template <class R> static void clear_mass_momentum(bm_diff_handle h) {
    const int64_t v = 4 * h->nn * h->nr;
    hipLaunchKernelGGL(k_zero<R>, dim3(fill_blocks(v)), dim3(256), 0, h->stream, (R*)h->gm, v);
}

// One forward substep from state `in` to state `out` (grid velocity left in
// gu; m and q zero on entry and on return).
template <class R> static void substep(bm_diff_handle h, PS<R> in, PS<R> out, int step) {
    const Par<R>* const P = pars<R>(h);
    const double t = time_at(h, step);
    const R dt = (R)h->dt;
    R* const ks = stress_buffer<R>(h);
    const int64_t nn = h->nn, nr = h->nr;
    // kept grids: this substep's own (m, q, u), left in place for the backward
    R* const gm = h->all_grids ? static_cast<R*>(h->agmq) + 4 * nn * nr * step : (R*)h->gm;
    R* const gq = gm + nn;
    R* const gu = h->all_grids ? static_cast<R*>(h->agu) + 3 * nn * nr * step : (R*)h->gu;
    hipLaunchKernelGGL(k_p2g<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, in, (const R*)h->m,
                       (const R*)h->V0, h->n, h->box, P, t, dt, gm, gq, h->err, (const R*)ks);
    hipLaunchKernelGGL(k_update<R>, rows(h, nn, 256), dim3(256), 0, h->stream, h->box, P, t, dt, gm, gq, gu,
                       h->all_grids ? 0 : 1);
    hipLaunchKernelGGL(k_g2p<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, in, out, h->n, h->box, P, t, dt,
                       (const R*)gu, h->err, ks, (R*)nullptr);
}

// Slot j of the backward's kept factors (every rollout's), allocated on first use.
template <class R> static R* fac_slot(const bm_diff_handle h, int j) {
    return static_cast<R*>(h->fac) + (int64_t)j * FCOMP * h->cap * h->nr;
}
static int ensure(void** p, size_t bytes);
static int ensure_fac(bm_diff_handle h) {
    if (h->fac_slots >= h->K) return BM_OK;
    if (int rc = ensure(&h->fac, h->rs * FCOMP * (size_t)h->cap * h->nr * h->K)) return rc;
    h->fac_slots = h->K;
    return BM_OK;
}

static int ensure(void** p, size_t bytes) {
    if (*p) hipFree(*p);
    *p = nullptr;
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return fail(BM_ERR_OOM, "diff allocation");
    return BM_OK;
}

static void release_kept_grids(bm_diff_handle h) {
    if (h->agmq) hipFree(h->agmq);
    if (h->agu) hipFree(h->agu);
    h->agmq = h->agu = nullptr;
    h->all_slots = 0;
}

static int read_err(bm_diff_handle h) {
    HIP_TRY(hipStreamSynchronize(h->stream));
    int e = 0;
    HIP_TRY(hipMemcpy(&e, h->err, sizeof(int), hipMemcpyDeviceToHost));
    if (e) {
        HIP_TRY(hipMemsetAsync(h->err, 0, sizeof(int), h->stream));
        return fail(BM_ERR_OUT_OF_DOMAIN, "a particle stencil left the node box");
    }
    return BM_OK;
}

// The forward of every rollout; losses[r] = rollout r's loss (may be null).
template <class R> static int forward_t(bm_diff_handle h, int substeps, double* losses) {
    const int K = h->K;
    const int64_t nr = h->nr;
    // checkpoint slot c holds the state after min(c K, substeps) substeps
    const int64_t nck = (substeps + K - 1) / K + 1;
    if (nck > h->ck_slots) {
        if (int rc = ensure(&h->ck, slot_bytes(h) * nck)) return rc;
        h->ck_slots = nck;
    }
    if (h->seg_slots < K) {
        if (int rc = ensure(&h->seg, slot_bytes(h) * K)) return rc;
        h->seg_slots = K;
    }
    if (int rc = upload_pars<R>(h)) return rc;
    HIP_TRY(hipMemcpyAsync(h->ck, h->init, slot_bytes(h), hipMemcpyDeviceToDevice, h->stream));
    PS<R> cur = slot<R>(h, h->ck, 0);
    const int64_t kept_bytes = (int64_t)substeps * 7 * h->nn * nr * (int64_t)h->rs;
    h->all_grids = kept_bytes <= h->grid_budget;
    if (h->all_grids && h->all_slots < substeps) {
        // keep the grids only where they fit the device's free memory with
        // room to spare; an allocation that fails anyway leaves the
        // per-segment grids in use
        size_t free_b = 0, total_b = 0;
        h->all_grids = hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                       kept_bytes + (int64_t(1) << 30) <= (int64_t)free_b;
        if (h->all_slots > 0 || !h->all_grids) release_kept_grids(h);
        if (h->all_grids) {
            h->all_grids = ensure(&h->agmq, h->rs * 4 * h->nn * nr * substeps) == BM_OK &&
                           ensure(&h->agu, h->rs * 3 * h->nn * nr * substeps) == BM_OK;
            if (h->all_grids) {
                h->all_slots = substeps;
            } else {
                (void)hipGetLastError();  // the failed allocation's error is handled here
                release_kept_grids(h);
            }
        }
    }
    if (h->all_grids) {
        hipLaunchKernelGGL(k_zero<R>, dim3(fill_blocks(4 * h->nn * nr * substeps)), dim3(256), 0, h->stream,
                           static_cast<R*>(h->agmq), 4 * h->nn * nr * substeps);
    } else {
        clear_mass_momentum<R>(h);
    }
    seed_stress<R>(h, cur);
    int pp = 0;  // states between checkpoints alternate between two scratch slots
    for (int s = 0; s < substeps; ++s) {
        const int ns = s + 1;
        PS<R> nxt;
        if (ns % K == 0 || ns == substeps) {
            nxt = slot<R>(h, h->ck, (ns + K - 1) / K);
        } else {
            nxt = slot<R>(h, h->seg, pp);
            pp ^= 1;
        }
        substep<R>(h, cur, nxt, s);
        cur = nxt;
    }
    HIP_TRY(hipMemsetAsync(h->grad, 0, sizeof(double) * GN * nr, h->stream));
    hipLaunchKernelGGL(k_loss<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, cur, (const R*)h->x0,
                       (const R*)h->tgt, h->n, slot<R>(h, h->abar, 0), h->grad);
    HIP_TRY(hipGetLastError());
    if (int rc = read_err(h)) return rc;
    std::vector<double> g((size_t)(GN * nr));
    HIP_TRY(hipMemcpy(g.data(), h->grad, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
    if (losses)
        for (int64_t r = 0; r < nr; ++r) losses[r] = g[(size_t)(r * GN + GLOSS)];
    h->substeps = substeps;
    h->have_forward = true;
    return BM_OK;
}

// The adjoint of the loss with respect to the final state, into the first
// adjoint slot.  Every backward starts from it (the reverse walk overwrites
// the slots), so backward can be called again after one forward.
template <class R> static void seed_loss_adjoint(bm_diff_handle h) {
    const int64_t last = (h->substeps + h->K - 1) / h->K;
    hipLaunchKernelGGL(k_loss<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, slot<R>(h, h->ck, last),
                       (const R*)h->x0, (const R*)h->tgt, h->n, slot<R>(h, h->abar, 0), h->grad);
}

// The second segment buffer, stream and events of the overlapped recompute.
static int ensure_side(bm_diff_handle h) {
    if (!h->side) {
        if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess) {
            h->side = nullptr;
            return fail(BM_ERR_HIP, "stream creation failed");
        }
        for (int b = 0; b < 2; ++b)
            if (hipEventCreateWithFlags(&h->ev_rec[b], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&h->ev_walk[b], hipEventDisableTiming) != hipSuccess)
                return fail(BM_ERR_HIP, "event creation failed");
    }
    if (h->seg2_slots < h->K) {
        h->seg2_slots = 0;
        if (int rc = ensure(&h->seg2, slot_bytes(h) * h->K)) return rc;
        if (int rc = ensure(&h->fac2, h->rs * FCOMP * (size_t)h->cap * h->nr * h->K)) return rc;
        h->seg2_slots = h->K;
    }
    return BM_OK;
}

// The backward over the grids the forward kept: each segment's particle
// states (and every substep's SVD factors) are recomputed from its checkpoint
// with the kept velocities, then walked in reverse.  The recompute of the
// segment before runs on the side stream meanwhile, into the other of two
// buffers: at one rollout a launch fills half a wave per SIMD, so the two
// overlap.
template <class R> static int backward_kept(bm_diff_handle h, double* g) {
    const int K = h->K;
    const int T = h->substeps;
    const int nseg = (T + K - 1) / K;
    const int64_t nn = h->nn, nr = h->nr;
    R* const gmq = static_cast<R*>(h->agmq);
    R* const gus = static_cast<R*>(h->agu);
    const Par<R>* const P = pars<R>(h);
    if (int rc = ensure_fac(h)) return rc;
    if (int rc = ensure_side(h)) return rc;
    void* const segb[2] = {h->seg, h->seg2};
    void* const facb[2] = {h->fac, h->fac2};
    auto fac_of = [&](int b, int j) { return static_cast<R*>(facb[b]) + (int64_t)j * FCOMP * h->cap * nr; };
    const R dt = (R)h->dt;
    // the side stream follows everything queued on the handle's stream so far
    HIP_TRY(hipEventRecord(h->ev_walk[0], h->stream));
    HIP_TRY(hipStreamWaitEvent(h->side, h->ev_walk[0], 0));
    auto recompute = [&](int sg, int b) {
        const int s0 = sg * K, len = std::min(K, T - s0);
        hipMemcpyAsync(slot<R>(h, segb[b], 0).p, slot<R>(h, h->ck, sg).p, slot_bytes(h), hipMemcpyDeviceToDevice,
                       h->side);
        // the last substep's new state is the next checkpoint: factors only
        for (int j = 0; j < len; ++j)
            hipLaunchKernelGGL(k_g2p<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->side, slot<R>(h, segb[b], j),
                               j + 1 < len ? slot<R>(h, segb[b], j + 1) : PS<R>{nullptr, h->cap}, h->n, h->box, P,
                               time_at(h, s0 + j), dt, (const R*)(gus + 3 * nn * nr * (s0 + j)), h->err, (R*)nullptr,
                               fac_of(b, j));
        hipEventRecord(h->ev_rec[b], h->side);
    };
    recompute(nseg - 1, 0);
    HIP_TRY(hipMemsetAsync(h->grad, 0, sizeof(double) * GN * nr, h->stream));
    seed_loss_adjoint<R>(h);
    hipLaunchKernelGGL(k_zero<R>, dim3(fill_blocks(3 * nn * nr)), dim3(256), 0, h->stream, (R*)h->gub, 3 * nn * nr);
    PS<R> a = slot<R>(h, h->abar, 0), a2 = slot<R>(h, h->abar2, 0);
    for (int sg = nseg - 1, b = 0; sg >= 0; --sg, b ^= 1) {
        if (sg > 0) {
            // the other buffer holds the segment after this one: its walk (the
            // previous iteration) must be done before the recompute overwrites it
            if (sg < nseg - 1) hipStreamWaitEvent(h->side, h->ev_walk[b ^ 1], 0);
            recompute(sg - 1, b ^ 1);
        }
        hipStreamWaitEvent(h->stream, h->ev_rec[b], 0);
        const int s0 = sg * K, len = std::min(K, T - s0);
        for (int j = len - 1; j >= 0; --j) {
            const PS<R> sj = slot<R>(h, segb[b], j);
            const double t = time_at(h, s0 + j);
            R* const gm = gmq + 4 * nn * nr * (s0 + j);
            R* const gu = gus + 3 * nn * nr * (s0 + j);
            hipLaunchKernelGGL(k_g2p_adj<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, sj, a, a2, h->n, h->box,
                               P, t, dt, (const R*)gu, (R*)h->gub, h->grad, (const R*)fac_of(b, j));
            hipLaunchKernelGGL(k_update_adj<R>, rows(h, nn, 256), dim3(256), 0, h->stream, h->box, P, t, dt, gm,
                               gm + nn, (R*)h->gub, (R*)h->gqb, (R*)h->gmb, h->grad, 0);
            hipLaunchKernelGGL(k_p2g_adj<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, sj, (const R*)h->m,
                               (const R*)h->V0, h->n, h->box, P, t, dt, (const R*)h->gqb, (const R*)h->gmb, a2,
                               h->grad, j > 0 ? (const R*)fac_of(b, j - 1) : (const R*)nullptr);
            std::swap(a, a2);
        }
        hipEventRecord(h->ev_walk[b], h->stream);
    }
    // every recompute was waited for by a walk: the handle's stream covers the side stream
    HIP_TRY(hipGetLastError());
    if (int rc = read_err(h)) return rc;
    HIP_TRY(hipMemcpy(g, h->grad, sizeof(double) * GN * nr, hipMemcpyDeviceToHost));
    return BM_OK;
}

template <class R> static int backward_t(bm_diff_handle h, double* g) {
    const int K = h->K;
    const int T = h->substeps;
    const int nseg = (T + K - 1) / K;
    const int64_t nn = h->nn, nr = h->nr;
    if (h->all_grids) return backward_kept<R>(h, g);
    if (h->grid_slots < K) {
        if (int rc = ensure(&h->sgmq, h->rs * 4 * nn * nr * K)) return rc;
        if (int rc = ensure(&h->sgu, h->rs * 3 * nn * nr * K)) return rc;
        h->grid_slots = K;
    }
    R* const gmq = static_cast<R*>(h->sgmq);
    R* const gus = static_cast<R*>(h->sgu);
    const Par<R>* const P = pars<R>(h);
    if (int rc = ensure_fac(h)) return rc;
    HIP_TRY(hipMemsetAsync(h->grad, 0, sizeof(double) * GN * nr, h->stream));
    seed_loss_adjoint<R>(h);
    // the slots' (m, q) and the velocity adjoint start at zero; the update's
    // adjoint clears each after its last read, so they stay zero between uses
    hipLaunchKernelGGL(k_zero<R>, dim3(fill_blocks(4 * nn * nr * K)), dim3(256), 0, h->stream, gmq, 4 * nn * nr * K);
    hipLaunchKernelGGL(k_zero<R>, dim3(fill_blocks(3 * nn * nr)), dim3(256), 0, h->stream, (R*)h->gub, 3 * nn * nr);
    PS<R> a = slot<R>(h, h->abar, 0), a2 = slot<R>(h, h->abar2, 0);
    const R dt = (R)h->dt;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * K, len = std::min(K, T - s0);
        // recompute the segment from its checkpoint, keeping each substep's
        // grid (m, q, u) and factors for the reverse walk; the last
        // substep's particle update is the next checkpoint, already stored
        hipMemcpyAsync(slot<R>(h, h->seg, 0).p, slot<R>(h, h->ck, sg).p, slot_bytes(h), hipMemcpyDeviceToDevice,
                       h->stream);
        seed_stress<R>(h, slot<R>(h, h->seg, 0));
        R* const ks = stress_buffer<R>(h);
        for (int j = 0; j < len; ++j) {
            const double t = time_at(h, s0 + j);
            R* const gm = gmq + 4 * nn * nr * j;
            R* const gu = gus + 3 * nn * nr * j;
            hipLaunchKernelGGL(k_p2g<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, slot<R>(h, h->seg, j),
                               (const R*)h->m, (const R*)h->V0, h->n, h->box, P, t, dt, gm, gm + nn, h->err,
                               (const R*)ks);
            hipLaunchKernelGGL(k_update<R>, rows(h, nn, 256), dim3(256), 0, h->stream, h->box, P, t, dt, gm, gm + nn,
                               gu, 0);
            hipLaunchKernelGGL(k_g2p<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, slot<R>(h, h->seg, j),
                               j + 1 < len ? slot<R>(h, h->seg, j + 1) : PS<R>{nullptr, h->cap}, h->n, h->box, P, t,
                               dt, (const R*)gu, h->err, ks, fac_slot<R>(h, j));
        }
        for (int j = len - 1; j >= 0; --j) {
            const PS<R> sj = slot<R>(h, h->seg, j);
            const double t = time_at(h, s0 + j);
            R* const gm = gmq + 4 * nn * nr * j;
            R* const gu = gus + 3 * nn * nr * j;
            hipLaunchKernelGGL(k_g2p_adj<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, sj, a, a2, h->n, h->box,
                               P, t, dt, (const R*)gu, (R*)h->gub, h->grad, (const R*)fac_slot<R>(h, j));
            hipLaunchKernelGGL(k_update_adj<R>, rows(h, nn, 256), dim3(256), 0, h->stream, h->box, P, t, dt, gm,
                               gm + nn, (R*)h->gub, (R*)h->gqb, (R*)h->gmb, h->grad, 1);
            hipLaunchKernelGGL(k_p2g_adj<R>, rows(h, h->n, PWG), dim3(PWG), 0, h->stream, sj, (const R*)h->m,
                               (const R*)h->V0, h->n, h->box, P, t, dt, (const R*)h->gqb, (const R*)h->gmb, a2,
                               h->grad, j > 0 ? (const R*)fac_slot<R>(h, j - 1) : (const R*)nullptr);
            std::swap(a, a2);
        }
    }
    HIP_TRY(hipGetLastError());
    if (int rc = read_err(h)) return rc;
    HIP_TRY(hipMemcpy(g, h->grad, sizeof(double) * GN * nr, hipMemcpyDeviceToHost));
    return BM_OK;
}

template <class R> static int download_t(bm_diff_handle h, int r, bm_particles* out) {
    const int64_t n = h->n;
    const int64_t last = (h->substeps + h->K - 1) / h->K;
    std::vector<R> st((size_t)(QCOMP * n));
    const R* src = slot<R>(h, h->ck, last).p + (int64_t)r * QCOMP * h->cap;
    HIP_TRY(hipMemcpy2D(st.data(), sizeof(R) * n, src, sizeof(R) * h->cap, sizeof(R) * n, QCOMP,
                        hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o = h->order[(size_t)i];
        for (int k = 0; k < 3; ++k) {
            if (out->x) out->x[3 * o + k] = (float)st[(size_t)((QX + k) * n + i)];
            if (out->v) out->v[3 * o + k] = (float)st[(size_t)((QV + k) * n + i)];
        }
        for (int k = 0; k < 9; ++k) {
            if (out->F) out->F[9 * o + k] = (float)st[(size_t)((QF + k) * n + i)];
            if (out->C) out->C[9 * o + k] = (float)st[(size_t)((QC + k) * n + i)];
        }
    }
    return BM_OK;
}

// The device order of the particles: grouped by 2 x 2 x 2 stencil bases
// (floor(x / dx - 1/2) per axis), the caller's order kept inside a base.
// A 64-lane wave of a bed with 8 particles per cell then covers a 4^3-node
// tile, so the scatter kernels' LDS tiles stay small and their flush
// atomics few; every result is a sum over particles or restored to the
// caller's order on download.
static std::vector<int64_t> device_order(const bm_diff_handle h, int64_t n, const float* x) {
    const double inv_dx = 1.0 / (double)h->desc.dx;
    const int ob[3] = {h->box.ox, h->box.oy, h->box.oz};
    const int nb[3] = {h->box.nx, h->box.ny, h->box.nz};
    std::vector<int64_t> key((size_t)n), ord((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t g = 0, sub = 0;
        for (int a = 0; a < 3; ++a) {
            const double f = std::floor((double)x[3 * i + a] * inv_dx - 0.5) - ob[a];
            const int64_t c = f < 0 ? 0 : (f > nb[a] ? nb[a] : (int64_t)f);
            g = g * (nb[a] / 2 + 1) + (c >> 1);
            sub = sub * 2 + (c & 1);
        }
        key[(size_t)i] = g * 8 + sub;
        ord[(size_t)i] = i;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return key[(size_t)a] < key[(size_t)b]; });
    return ord;
}

template <class R> static int upload_t(bm_diff_handle h, int64_t n, const bm_particles* p, const float* tgt) {
    std::vector<R> st((size_t)(QCOMP * n)), mm((size_t)n), vv((size_t)n), x0((size_t)(3 * n)), tg((size_t)(3 * n));
    h->order = device_order(h, n, p->x);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o = h->order[(size_t)i];  // slot i holds the caller's particle o
        for (int k = 0; k < 3; ++k) {
            st[(size_t)((QX + k) * n + i)] = (R)p->x[3 * o + k];
            st[(size_t)((QV + k) * n + i)] = (R)p->v[3 * o + k];
            x0[(size_t)(3 * i + k)] = (R)p->x[3 * o + k];
            tg[(size_t)(3 * i + k)] = tgt ? (R)tgt[3 * o + k] : R(0);
        }
        for (int k = 0; k < 9; ++k) {
            st[(size_t)((QC + k) * n + i)] = (R)p->C[9 * o + k];
            st[(size_t)((QF + k) * n + i)] = (R)p->F[9 * o + k];
        }
        mm[(size_t)i] = (R)p->m[o];
        vv[(size_t)i] = (R)p->V0[o];
    }
    // stream-ordered with the handle's (non-blocking) stream; the host
    // vectors live until the synchronisation
    HIP_TRY(hipMemcpy2DAsync(h->init, sizeof(R) * h->cap, st.data(), sizeof(R) * n, sizeof(R) * n, QCOMP,
                             hipMemcpyHostToDevice, h->stream));
    // every rollout starts from the same state
    const size_t one = sizeof(R) * QCOMP * (size_t)h->cap;
    for (int r = 1; r < h->nr; ++r)
        HIP_TRY(hipMemcpyAsync(static_cast<char*>(h->init) + r * one, h->init, one, hipMemcpyDeviceToDevice,
                               h->stream));
    HIP_TRY(hipMemcpyAsync(h->m, mm.data(), sizeof(R) * n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->V0, vv.data(), sizeof(R) * n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->x0, x0.data(), sizeof(R) * 3 * n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->tgt, tg.data(), sizeof(R) * 3 * n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return BM_OK;
}

extern "C" {

int bm_diff_create(const bm_diff_desc* desc, bm_diff_handle* out) {
    if (!desc || !out) return fail(BM_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(desc->dx > 0.0f)) return fail(BM_ERR_INVALID_ARG, "dx must be > 0");
    for (int a = 0; a < 3; ++a)
        if (desc->grid_dims[a] <= 0) return fail(BM_ERR_INVALID_ARG, "grid_dims must be > 0");
    if (desc->max_particles <= 0) return fail(BM_ERR_INVALID_ARG, "max_particles must be > 0");
    if (desc->precision != 32 && desc->precision != 64) return fail(BM_ERR_INVALID_ARG, "precision must be 32 or 64");
    if (desc->checkpoint_every < 0) return fail(BM_ERR_INVALID_ARG, "checkpoint_every must be >= 0");
    if (desc->rollouts < 0 || desc->rollouts > BM_DIFF_MAX_ROLLOUTS)
        return fail(BM_ERR_INVALID_ARG, "rollouts must be in 0..%d", BM_DIFF_MAX_ROLLOUTS);
    auto* h = new bm_diff();
    h->desc = *desc;
    h->nr = desc->rollouts > 0 ? desc->rollouts : 1;
    h->prm.assign((size_t)h->nr, bm_diff_params{});
    h->prm_set.assign((size_t)h->nr, 0);
    h->fp64 = desc->precision == 64;
    h->rs = h->fp64 ? 8 : 4;
    h->K = desc->checkpoint_every > 0 ? desc->checkpoint_every : 16;
    h->box = Box{desc->grid_origin[0], desc->grid_origin[1], desc->grid_origin[2],
                 desc->grid_dims[0],   desc->grid_dims[1],   desc->grid_dims[2]};
    h->nn = (int64_t)desc->grid_dims[0] * desc->grid_dims[1] * desc->grid_dims[2];
    h->cap = desc->max_particles;
    auto bail = [&](int rc) {
        bm_diff_destroy(h);
        return rc;
    };
    if (desc->stream) {
        h->stream = (hipStream_t)desc->stream;
    } else {
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(BM_ERR_HIP, "stream creation failed"));
        h->own_stream = true;
    }
    const size_t r = h->rs, nr = (size_t)h->nr;
    int rc = BM_OK;
    // the particles' m, V0, x(0) and target: one copy for every rollout
    if (!rc) rc = ensure(&h->m, r * h->cap);
    if (!rc) rc = ensure(&h->V0, r * h->cap);
    if (!rc) rc = ensure(&h->x0, r * 3 * h->cap);
    if (!rc) rc = ensure(&h->tgt, r * 3 * h->cap);
    if (!rc) rc = ensure(&h->init, r * QCOMP * h->cap * nr);
    if (!rc) rc = ensure(&h->gm, r * 4 * h->nn * nr);
    if (!rc) rc = ensure(&h->gu, r * 3 * h->nn * nr);
    if (!rc) rc = ensure(&h->gub, r * 3 * h->nn * nr);
    if (!rc) rc = ensure(&h->gqb, r * 3 * h->nn * nr);
    if (!rc) rc = ensure(&h->gmb, r * h->nn * nr);
    if (!rc) rc = ensure(&h->ks, r * 9 * h->cap * nr);
    if (!rc) rc = ensure(&h->abar, r * QCOMP * h->cap * nr);
    if (!rc) rc = ensure(&h->abar2, r * QCOMP * h->cap * nr);
    if (!rc) rc = ensure(&h->pard, sizeof(Par<double>) * nr);
    if (!rc) rc = ensure((void**)&h->grad, sizeof(double) * GN * nr);
    if (!rc) rc = ensure((void**)&h->err, sizeof(int));
    if (rc) return bail(rc);
    if (hipMemset(h->err, 0, sizeof(int)) != hipSuccess) return bail(fail(BM_ERR_HIP, "diff initialisation"));
    *out = h;
    return BM_OK;
}

int bm_diff_destroy(bm_diff_handle h) {
    if (!h) return BM_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void* p : {h->m, h->V0, h->x0, h->tgt, h->init, h->gm, h->gu, h->gub, h->gqb, h->gmb, h->ck, h->seg, h->abar,
                    h->abar2, h->sgmq, h->sgu, h->ks, h->agmq, h->agu, h->pard, h->fac, (void*)h->grad, (void*)h->err})
        if (p) hipFree(p);
    if (h->side) hipStreamSynchronize(h->side);
    for (void* p : {h->seg2, h->fac2})
        if (p) hipFree(p);
    for (hipEvent_t e : {h->ev_rec[0], h->ev_rec[1], h->ev_walk[0], h->ev_walk[1]})
        if (e) hipEventDestroy(e);
    if (h->side) hipStreamDestroy(h->side);
    if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
    delete h;
    return BM_OK;
}

int bm_diff_set_params(bm_diff_handle h, const bm_diff_params* p) {
    if (int rc = check_diff(h)) return rc;
    if (int rc = check_params(p)) return rc;
    h->prm.assign((size_t)h->nr, *p);
    h->prm_set.assign((size_t)h->nr, 1);
    h->have_forward = false;
    return BM_OK;
}

int bm_diff_set_rollout_params(bm_diff_handle h, int32_t r, const bm_diff_params* p) {
    if (int rc = check_diff(h)) return rc;
    if (r < 0 || r >= h->nr) return fail(BM_ERR_INVALID_ARG, "rollout %d of %d", (int)r, h->nr);
    if (int rc = check_params(p)) return rc;
    h->prm[(size_t)r] = *p;
    h->prm_set[(size_t)r] = 1;
    h->have_forward = false;
    return BM_OK;
}

int bm_diff_set_grid_budget(bm_diff_handle h, int64_t bytes) {
    if (int rc = check_diff(h)) return rc;
    if (bytes < 0) return fail(BM_ERR_INVALID_ARG, "grid budget must be >= 0");
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->grid_budget = bytes;
    h->have_forward = false;
    return BM_OK;
}

int bm_diff_upload(bm_diff_handle h, int64_t n, const bm_particles* p, const float* target_displacement) {
    if (int rc = check_diff(h)) return rc;
    if (n <= 0 || !p || !p->x || !p->v || !p->F || !p->C || !p->m || !p->V0)
        return fail(BM_ERR_INVALID_ARG, "bad particle arrays");
    if (n > h->cap) return fail(BM_ERR_OOM, "capacity %lld exceeded", (long long)h->cap);
    for (int64_t i = 0; i < n; ++i) {
        const float sc = 0x1p-100f;  // NaN and Inf survive the sum; finite values cannot overflow it
        float acc = p->m[i] * sc + p->V0[i] * sc;
        for (int k = 0; k < 3; ++k) acc += p->x[3 * i + k] * sc + p->v[3 * i + k] * sc;
        for (int k = 0; k < 9; ++k) acc += p->F[9 * i + k] * sc + p->C[9 * i + k] * sc;
        if (!std::isfinite(acc)) return fail(BM_ERR_NONFINITE, "particle %lld holds a non-finite value", (long long)i);
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->n = n;
    h->have_forward = false;
    return h->fp64 ? upload_t<double>(h, n, p, target_displacement) : upload_t<float>(h, n, p, target_displacement);
}

int bm_diff_forward_batch(bm_diff_handle h, float dt, int32_t substeps, double* losses) {
    BM_TRACE("bm_diff_forward");
    if (int rc = check_diff(h)) return rc;
    if (!all_params_set(h)) return fail(BM_ERR_STATE, "parameters not set");
    if (h->n <= 0) return fail(BM_ERR_STATE, "no particles");
    if (!(dt > 0.0f) || substeps <= 0) return fail(BM_ERR_INVALID_ARG, "dt and substeps must be > 0");
    h->dt = dt;
    return h->fp64 ? forward_t<double>(h, substeps, losses) : forward_t<float>(h, substeps, losses);
}

int bm_diff_forward(bm_diff_handle h, float dt, int32_t substeps, double* loss) {
    if (int rc = check_diff(h)) return rc;
    std::vector<double> losses((size_t)h->nr);
    const int rc = bm_diff_forward_batch(h, dt, substeps, losses.data());
    if (rc == BM_OK && loss) *loss = losses[0];
    return rc;
}

int bm_diff_backward_batch(bm_diff_handle h, bm_diff_grad* out) {
    BM_TRACE("bm_diff_backward");
    if (int rc = check_diff(h)) return rc;
    if (!out) return fail(BM_ERR_INVALID_ARG, "null output");
    if (!h->have_forward) return fail(BM_ERR_STATE, "backward needs a forward pass with the current parameters");
    std::vector<double> g((size_t)(GN * h->nr));
    if (int rc = h->fp64 ? backward_t<double>(h, g.data()) : backward_t<float>(h, g.data())) return rc;
    for (int r = 0; r < h->nr; ++r) grad_of(h->prm[(size_t)r], g.data() + (size_t)r * GN, out + r);
    return BM_OK;
}

int bm_diff_backward(bm_diff_handle h, bm_diff_grad* out) {
    if (int rc = check_diff(h)) return rc;
    if (!out) return fail(BM_ERR_INVALID_ARG, "null output");
    std::vector<bm_diff_grad> all((size_t)h->nr);
    const int rc = bm_diff_backward_batch(h, all.data());
    if (rc == BM_OK) *out = all[0];
    return rc;
}

int bm_diff_map_vjp(int32_t map, int64_t n, const double* F, const double* Gbar, const bm_diff_params* params,
                    int32_t precision, double* G, double* Fbar) {
    if (map != 0 && map != 1) return fail(BM_ERR_INVALID_ARG, "map must be 0 (return map) or 1 (stress)");
    if (precision != 32 && precision != 64) return fail(BM_ERR_INVALID_ARG, "precision must be 32 or 64");
    if (n < 0 || (n > 0 && (!F || !Gbar || !params || !G || !Fbar))) return fail(BM_ERR_INVALID_ARG, "null array");
    if (n == 0) return BM_OK;
    bm_diff_desc unit{};
    unit.dx = 1.0f;
    const size_t bytes = sizeof(double) * 9 * (size_t)n;
    double *dF = nullptr, *dGb = nullptr, *dG = nullptr, *dFb = nullptr;
    int rc = BM_OK;
    if (hipMalloc(&dF, bytes) != hipSuccess || hipMalloc(&dGb, bytes) != hipSuccess ||
        hipMalloc(&dG, bytes) != hipSuccess || hipMalloc(&dFb, bytes) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(BM_ERR_OOM, "map_vjp buffers");
    }
    if (rc == BM_OK && (hipMemcpy(dF, F, bytes, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(dGb, Gbar, bytes, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(BM_ERR_HIP, "map_vjp upload");
    if (rc == BM_OK) {
        const dim3 grid((unsigned)((n + 63) / 64));
        if (precision == 64)
            hipLaunchKernelGGL(k_map_vjp<double>, grid, dim3(64), 0, nullptr, map, n, (const double*)dF,
                               (const double*)dGb, par_of<double>(unit, *params), dG, dFb);
        else
            hipLaunchKernelGGL(k_map_vjp<float>, grid, dim3(64), 0, nullptr, map, n, (const double*)dF,
                               (const double*)dGb, par_of<float>(unit, *params), dG, dFb);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(G, dG, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(Fbar, dFb, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(BM_ERR_HIP, "map_vjp kernel or download");
    }
    hipFree(dF);
    hipFree(dGb);
    hipFree(dG);
    hipFree(dFb);
    return rc;
}

int bm_diff_download_rollout(bm_diff_handle h, int32_t r, bm_particles* out) {
    if (int rc = check_diff(h)) return rc;
    if (!out) return fail(BM_ERR_INVALID_ARG, "null output");
    if (r < 0 || r >= h->nr) return fail(BM_ERR_INVALID_ARG, "rollout %d of %d", (int)r, h->nr);
    if (!h->have_forward) return fail(BM_ERR_STATE, "no forward pass");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return h->fp64 ? download_t<double>(h, r, out) : download_t<float>(h, r, out);
}

int bm_diff_download(bm_diff_handle h, bm_particles* out) { return bm_diff_download_rollout(h, 0, out); }

}  // extern "C"
