// mpm_device.hpp -- per-particle device math for the MLS-MPM soil substep.
//
// Register-resident 3x3 algebra, the quadratic B-spline stencil, the Jacobi
// SVD and the Drucker-Prager return map.  Every routine is a template on the
// scalar type so the FP64 SVD fallback shares the FP32 code.  Operation order
// follows the reference (include/basements/core/math/{vec3,matrix3,svd}.h,
// physics/mpm/mpm_solver.h); whether the compiler may contract a*b+c into an
// FMA is decided per translation unit (reference-order kernels are built with
// -ffp-contract=off, the fused fast-path kernels with contraction on).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bm {

template <class R> struct v3 {
    R x, y, z;
};
template <class R> struct m3 {
    R a[9];  // row-major
};

template <class R> __device__ __forceinline__ v3<R> mk(R x, R y, R z) { return v3<R>{x, y, z}; }
template <class R> __device__ __forceinline__ v3<R> operator+(v3<R> a, v3<R> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> __device__ __forceinline__ v3<R> operator-(v3<R> a, v3<R> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> __device__ __forceinline__ v3<R> operator*(v3<R> a, R s) { return mk(a.x * s, a.y * s, a.z * s); }
// vec3.h:80-101: zero below 1e-6, else a true divide per component.
template <class R> __device__ __forceinline__ v3<R> divg(v3<R> a, R s) {
    if (fabs(s) < R(1e-6f)) return mk<R>(0, 0, 0);
    return mk(a.x / s, a.y / s, a.z / s);
}
template <class R> __device__ __forceinline__ R dot(v3<R> a, v3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> __device__ __forceinline__ v3<R> cross(v3<R> a, v3<R> b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

template <class R> __device__ __forceinline__ m3<R> mmul(const m3<R>& p, const m3<R>& q) {
    m3<R> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = p.a[i * 3] * q.a[j] + p.a[i * 3 + 1] * q.a[3 + j] + p.a[i * 3 + 2] * q.a[6 + j];
    return r;
}
// p^T q without forming the transpose.
template <class R> __device__ __forceinline__ m3<R> mtmul(const m3<R>& p, const m3<R>& q) {
    m3<R> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = p.a[i] * q.a[j] + p.a[3 + i] * q.a[3 + j] + p.a[6 + i] * q.a[6 + j];
    return r;
}
template <class R> __device__ __forceinline__ v3<R> mvec(const m3<R>& p, v3<R> v) {
    return mk(p.a[0] * v.x + p.a[1] * v.y + p.a[2] * v.z, p.a[3] * v.x + p.a[4] * v.y + p.a[5] * v.z,
              p.a[6] * v.x + p.a[7] * v.y + p.a[8] * v.z);
}

// ------------------------------------------------------------------ stencil
// mpm_solver.h:170-184: fx = x/dx - 1/2, base = floor(fx), quadratic B-spline
// weights of the fraction; node i sits at i*dx.
struct Stencil {
    float f[3];
    int b[3];
    float w[3][3];
};
__device__ __forceinline__ void make_stencil(float x, float y, float z, float inv_dx, Stencil& s) {
    // Rounded exactly as the reference rounds it in every build mode: the
    // weights decide the fixed-point mass deposit, which must not depend on
    // whether the surrounding kernel allows FMA contraction.
#pragma clang fp contract(off)
    const float p[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.f[a] = p[a] * inv_dx - 0.5f;
        s.b[a] = (int)floorf(s.f[a]);
        const float fr = s.f[a] - (float)s.b[a];
        s.w[a][0] = 0.5f * (1.0f - fr) * (1.0f - fr);
        s.w[a][1] = 0.75f - (fr - 0.5f) * (fr - 0.5f);
        s.w[a][2] = 0.5f * fr * fr;
    }
}

// ------------------------------------------------------------------ SVD
// svd.h:18-108, restated: cyclic Jacobi on A = F^T F, pairs (0,1),(0,2),(1,2);
// a pair is skipped when |a_ij| < 1e-6; at most MAXS sweeps; a sweep ends the
// iteration when its running max |a_ij| (sampled before each rotation) is
// below tol.  sigma = sqrt(max(0, a_ii)), unsorted; U = F V / sigma with the
// unit axis when sigma <= 1e-6.
template <class R, int MAXS>
__device__ __forceinline__ void jacobi_svd(const m3<R>& F, m3<R>& U, v3<R>& S, m3<R>& V, R skip, R tol) {
    m3<R> A = mtmul(F, F);
#pragma unroll
    for (int i = 0; i < 9; ++i) V.a[i] = (i % 4 == 0) ? R(1) : R(0);
    for (int sweep = 0; sweep < MAXS; ++sweep) {
        R max_off = R(0);
#pragma unroll
        for (int pr = 0; pr < 3; ++pr) {
            const int i = pr < 2 ? 0 : 1;
            const int j = pr == 0 ? 1 : 2;
            const int k = 3 - i - j;
            const R aij = A.a[i * 3 + j], aii = A.a[i * 3 + i], ajj = A.a[j * 3 + j];
            max_off = fmax(max_off, fabs(aij));
            if (fabs(aij) < skip) continue;
            const R tau = (ajj - aii) / (R(2) * aij);
            R t;
            if (tau >= R(0))
                t = R(1) / (tau + sqrt(R(1) + tau * tau));
            else
                t = R(-1) / (-tau + sqrt(R(1) + tau * tau));
            const R c = R(1) / sqrt(R(1) + t * t);
            const R s = c * t;
            const R nii = c * c * aii - R(2) * s * c * aij + s * s * ajj;
            const R njj = s * s * aii + R(2) * s * c * aij + c * c * ajj;
            A.a[i * 3 + i] = nii;
            A.a[j * 3 + j] = njj;
            A.a[i * 3 + j] = R(0);
            A.a[j * 3 + i] = R(0);
            const R aik = A.a[i * 3 + k], ajk = A.a[j * 3 + k];
            A.a[i * 3 + k] = c * aik - s * ajk;
            A.a[k * 3 + i] = A.a[i * 3 + k];
            A.a[j * 3 + k] = s * aik + c * ajk;
            A.a[k * 3 + j] = A.a[j * 3 + k];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const R vi = V.a[r * 3 + i], vj = V.a[r * 3 + j];
                V.a[r * 3 + i] = c * vi - s * vj;
                V.a[r * 3 + j] = s * vi + c * vj;
            }
        }
        if (max_off < tol) break;
    }
    S = mk(sqrt(fmax(R(0), A.a[0])), sqrt(fmax(R(0), A.a[4])), sqrt(fmax(R(0), A.a[8])));
    const m3<R> FV = mmul(F, V);
    const R sv[3] = {S.x, S.y, S.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (sv[c] > R(1e-6f)) {
            U.a[c] = FV.a[c] / sv[c];
            U.a[3 + c] = FV.a[3 + c] / sv[c];
            U.a[6 + c] = FV.a[6 + c] / sv[c];
        } else {
            U.a[c] = (c == 0) ? R(1) : R(0);
            U.a[3 + c] = (c == 1) ? R(1) : R(0);
            U.a[6 + c] = (c == 2) ? R(1) : R(0);
        }
    }
}

// The reference's FP32 SVD exactly (5 sweeps, skip 1e-6, stop 1e-5).
__device__ __forceinline__ void svd_ref(const m3<float>& F, m3<float>& U, v3<float>& S, m3<float>& V) {
    jacobi_svd<float, 5>(F, U, S, V, 1e-6f, 1e-5f);
}

// FP32 Jacobi with an FP64 redo when the FP32 result has lost orthogonality.
// Loss of orthogonality in U = F V / sigma is what near-coincident singular
// values (a tiny eigen-gap of F^T F) or an unconverged sweep produce; the FP64
// pass iterates to 1e-12 relative off-diagonal mass instead.  Returns 1 when
// the fallback ran.
__device__ __forceinline__ int svd_guarded(const m3<float>& F, m3<float>& U, v3<float>& S, m3<float>& V) {
    jacobi_svd<float, 5>(F, U, S, V, 1e-6f, 1e-5f);
    const v3<float> u0 = mk(U.a[0], U.a[3], U.a[6]), u1 = mk(U.a[1], U.a[4], U.a[7]), u2 = mk(U.a[2], U.a[5], U.a[8]);
    const float o = fmaxf(fabsf(dot(u0, u1)), fmaxf(fabsf(dot(u0, u2)), fabsf(dot(u1, u2))));
    if (!(o > 2e-4f)) return 0;
    m3<double> Fd, Ud, Vd;
    v3<double> Sd;
#pragma unroll
    for (int i = 0; i < 9; ++i) Fd.a[i] = (double)F.a[i];
    double fro = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) fro += Fd.a[i] * Fd.a[i];
    jacobi_svd<double, 12>(Fd, Ud, Sd, Vd, 1e-30, 1e-13 * fro);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        U.a[i] = (float)Ud.a[i];
        V.a[i] = (float)Vd.a[i];
    }
    S = mk((float)Sd.x, (float)Sd.y, (float)Sd.z);
    return 1;
}

// ------------------------------------------------------------------ constitutive
struct Material {
    float mu, lambda, alpha, kohesion;
};

// Kirchhoff stress from U and the clamped singular values (mpm_solver.h:195-212).
__device__ __forceinline__ m3<float> hencky_from_svd(const m3<float>& U, v3<float> s, float mu, float lam) {
    const float lx = logf(s.x), ly = logf(s.y), lz = logf(s.z);
    const float tr = lx + ly + lz;
    const float t0 = 2.0f * mu * lx + lam * tr, t1 = 2.0f * mu * ly + lam * tr, t2 = 2.0f * mu * lz + lam * tr;
    // (U diag(t)) U^T
    m3<float> r;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a0 = U.a[i * 3] * t0, a1 = U.a[i * 3 + 1] * t1, a2 = U.a[i * 3 + 2] * t2;
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = a0 * U.a[j * 3] + a1 * U.a[j * 3 + 1] + a2 * U.a[j * 3 + 2];
    }
    return r;
}

// Drucker-Prager projection in log-singular-value space (mpm_solver.h:125-162):
// returns the projected singular values; *lns receives their logarithms
// (the Hencky strain of the projected state, reused for the next stress).
__device__ __forceinline__ v3<float> dp_project(v3<float> s, const Material& M, v3<float>* lns = nullptr) {
    s.x = fmaxf(s.x, 1e-6f);
    s.y = fmaxf(s.y, 1e-6f);
    s.z = fmaxf(s.z, 1e-6f);
    v3<float> e = mk(logf(s.x), logf(s.y), logf(s.z));
    const float tr = e.x + e.y + e.z;
    const float t3 = tr / 3.0f;
    v3<float> eh = mk(e.x - t3, e.y - t3, e.z - t3);
    const float H = sqrtf(dot(eh, eh));
    if (tr > 0.0f) {
        if (lns) *lns = mk(0.0f, 0.0f, 0.0f);
        return mk(1.0f, 1.0f, 1.0f);
    }
    const float limit = -M.alpha * M.kohesion * tr;
    if (H > limit) {
        const float sc = limit / H;
        eh = eh * sc;
        e = mk(eh.x + t3, eh.y + t3, eh.z + t3);
        if (lns) *lns = e;
        return mk(expf(e.x), expf(e.y), expf(e.z));
    }
    if (lns) *lns = e;
    return s;
}

// Kirchhoff stress from U and the Hencky strain ln(sigma) directly.
__device__ __forceinline__ m3<float> hencky_from_log(const m3<float>& U, v3<float> l, float mu, float lam) {
    const float tr = l.x + l.y + l.z;
    const float t0 = 2.0f * mu * l.x + lam * tr, t1 = 2.0f * mu * l.y + lam * tr, t2 = 2.0f * mu * l.z + lam * tr;
    m3<float> r;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a0 = U.a[i * 3] * t0, a1 = U.a[i * 3 + 1] * t1, a2 = U.a[i * 3 + 2] * t2;
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = a0 * U.a[j * 3] + a1 * U.a[j * 3 + 1] + a2 * U.a[j * 3 + 2];
    }
    return r;
}

// (U diag(s)) V^T
__device__ __forceinline__ m3<float> recompose(const m3<float>& U, v3<float> s, const m3<float>& V) {
    m3<float> r;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a0 = U.a[i * 3] * s.x, a1 = U.a[i * 3 + 1] * s.y, a2 = U.a[i * 3 + 2] * s.z;
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = a0 * V.a[j * 3] + a1 * V.a[j * 3 + 1] + a2 * V.a[j * 3 + 2];
    }
    return r;
}

// ------------------------------------------------------------------ fixed-point mass
// Deposit of one particle: each node with a nonzero float weight receives
// max(1, rint(w * q_p)) quanta, the stencil centre (1,1,1) the remainder, so a
// particle always deposits exactly q_p quanta and a node is active iff its
// integer mass is nonzero.
// Quanta of one node: max(1, rint(w q_p)) for w != 0.  q_p < 2^23 and
// w <= 1, so the double product is exact and the result fits 32 bits.
// Fixed-point mass of one stencil node: w * q_p rounded to the nearest
// integer in FP32 -- floor(fma(w, q_p, 1/2)), the sum rounded once (q_p <
// 2^23 is exact in FP32; w * q_p + 1/2 < 2^22) -- and at least one quantum
// on every node with w > 0 (w >= 0, so its bit pattern is nonzero exactly
// then), none where w = 0.  The stencil centre takes the particle's
// remainder, so each particle deposits exactly q_p.  Unbiased: rounding up
// every node would move mass from each stencil's centre to its rim, which
// cancels in the interior but not at a free surface.  Formed as
// median(w q_p + 1/2, w 2^100, 1): for w = 0 that is 1/2 (no quantum); for
// w > 0 (a stencil's nonzero weights are >= 2^-60) w 2^100 exceeds both
// others, so the median is max(w q_p + 1/2, 1).  Four instructions (fma,
// multiply, med3, convert).
__device__ __forceinline__ int mass_quanta(float w, float qp) {
    return (int)__builtin_amdgcn_fmed3f(__builtin_fmaf(w, qp, 0.5f), w * 0x1p100f, 1.0f);
}

}  // namespace bm
