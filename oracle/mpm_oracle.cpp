// mpm_oracle.cpp -- CPU restatement of the reference MLS-MPM soil substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is compiled
// into oracle/liboracle_mpm.so and may be loaded only by tests/, by the
// smoke check and by the CPU-baseline leg of bench.py, always as the checker
// or the timed CPU baseline, never as the product path.  The product path
// (basements_amd/) never links or calls it.
//
// It restates, in scalar-templated form (R = float or double), the algorithm
// of these reference files (paths relative to the reference root):
//   include/basements/physics/mpm/mpm_solver.h      P2G / update / G2P / plasticity
//   include/basements/physics/mpm/spgrid_cpu.h      4^3-block sparse grid, persistence
//   include/basements/core/math/svd.h               one-sided Jacobi on F^T F
//   include/basements/core/math/vec3.h, matrix3.h   operation order, epsilon guards
//   include/basements/physics/coupling/mpm_rigid_coupler.h + src/.../mpm_rigid_coupler.cpp  (M0)
//   include/basements/physics/coupling/mpm_rigid_coupler_pic.h                            (M1)
//   include/basements/physics/dynamics/integrator.h, physics/rigid_body.h                (rigid body)
//
// With R = float and MASS_FLOAT it is meant to be bit-identical to the
// reference compiled with g++ -O2 -mavx2 (no FMA contraction): every sum is
// formed in the reference's order.  With MASS_FIXED the per-node mass is the
// deterministic int64 fixed-point deposit the GPU path uses (each particle
// deposits exactly q_p quanta; the centre node of its 3x3x3 stencil takes the
// remainder), so grid mass is order-independent and bit-comparable.
// With R = double it is the FP64 oracle used for tolerance calibration.
//
// Parity pins (tests/test_oracle_*.py): bit-exact against oracle/_ref (the
// reference headers compiled unchanged by oracle/Makefile) on SVD, P2G, grid
// update, G2P, M0, M1, whole steps and the rigid integrator; closed forms
// (B-spline moments, Hencky stress of diagonal F, DP cone / tension cases,
// SVD reconstruction) for the FP64 instantiation.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- small math
template <class R> struct V3 {
    R x, y, z;
};
template <class R> static inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> static inline V3<R> add(V3<R> a, V3<R> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> static inline V3<R> sub(V3<R> a, V3<R> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> static inline V3<R> mul(V3<R> a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
// vec3.h:80-101 -- zero vector below 1e-6, otherwise a true per-lane divide.
template <class R> static inline V3<R> divg(V3<R> a, R s) {
    if (std::abs(s) < R(1e-6f)) return v3<R>(0, 0, 0);
    return v3(a.x / s, a.y / s, a.z / s);
}
// vec3.h:142-156 -- horizontal add order (x*x'+y*y') + (z*z'+0).
template <class R> static inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> static inline R len(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> static inline V3<R> normalized(V3<R> a) {
    R l = len(a);
    if (l < R(1e-6f)) return v3<R>(0, 0, 0);
    return divg(a, l);
}
template <class R> static inline V3<R> cross(V3<R> a, V3<R> b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

template <class R> struct M3 {
    R m[3][3];
};
template <class R> static inline M3<R> ident() {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = (i == j) ? R(1) : R(0);
    return r;
}
template <class R> static inline M3<R> zero3() {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = R(0);
    return r;
}
template <class R> static inline M3<R> diag3(V3<R> s) {
    M3<R> r = zero3<R>();
    r.m[0][0] = s.x;
    r.m[1][1] = s.y;
    r.m[2][2] = s.z;
    return r;
}
// matrix3.h:219-229 -- c_ij = (a_i0 b_0j + a_i1 b_1j) + a_i2 b_2j.
template <class R> static inline M3<R> mm(const M3<R>& a, const M3<R>& b) {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
template <class R> static inline M3<R> madd(const M3<R>& a, const M3<R>& b) {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
    return r;
}
template <class R> static inline M3<R> mscale(const M3<R>& a, R s) {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s;
    return r;
}
template <class R> static inline M3<R> transp(const M3<R>& a) {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
template <class R> static inline V3<R> mv(const M3<R>& a, V3<R> v) {
    return v3(a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
              a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
              a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z);
}
template <class R> static inline M3<R> outer(V3<R> a, V3<R> b) {
    M3<R> r;
    const R av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = av[i] * bv[j];
    return r;
}

// Quaternion (w, x, y, z), quaternion.h:21-200.
template <class R> struct Q4 {
    R w, x, y, z;
};
// matrix3.h:432-448
template <class R> static inline M3<R> quat_to_mat(Q4<R> q) {
    R xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    R xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    R wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M3<R> r;
    r.m[0][0] = R(1) - R(2) * (yy + zz);
    r.m[0][1] = R(2) * (xy - wz);
    r.m[0][2] = R(2) * (xz + wy);
    r.m[1][0] = R(2) * (xy + wz);
    r.m[1][1] = R(1) - R(2) * (xx + zz);
    r.m[1][2] = R(2) * (yz - wx);
    r.m[2][0] = R(2) * (xz - wy);
    r.m[2][1] = R(2) * (yz + wx);
    r.m[2][2] = R(1) - R(2) * (xx + yy);
    return r;
}

// ---------------------------------------------------------------- SVD
// svd.h:18-108: cyclic Jacobi on A = F^T F over pairs (0,1),(0,2),(1,2),
// pair skipped when |a_ij| < 1e-6, at most 5 sweeps, stop after a sweep whose
// running max |a_ij| (sampled before each rotation) is < 1e-5; sigma =
// sqrt(max(0, a_ii)) unsorted; U = F V / sigma, unit axis when sigma <= 1e-6.
template <class R> static void svd3(const M3<R>& F, M3<R>& U, V3<R>& S, M3<R>& V) {
    M3<R> A = mm(transp(F), F);
    V = ident<R>();
    for (int sweep = 0; sweep < 5; ++sweep) {
        R max_off = R(0);
        for (int i = 0; i < 2; ++i) {
            for (int j = i + 1; j < 3; ++j) {
                const R aij = A.m[i][j], aii = A.m[i][i], ajj = A.m[j][j];
                max_off = std::max(max_off, std::abs(aij));
                if (std::abs(aij) < R(1e-6f)) continue;
                const R tau = (ajj - aii) / (R(2) * aij);
                R t;
                if (tau >= R(0))
                    t = R(1) / (tau + std::sqrt(R(1) + tau * tau));
                else
                    t = R(-1) / (-tau + std::sqrt(R(1) + tau * tau));
                const R c = R(1) / std::sqrt(R(1) + t * t);
                const R s = c * t;
                const R nii = c * c * aii - R(2) * s * c * aij + s * s * ajj;
                const R njj = s * s * aii + R(2) * s * c * aij + c * c * ajj;
                A.m[i][i] = nii;
                A.m[j][j] = njj;
                A.m[i][j] = R(0);
                A.m[j][i] = R(0);
                const int k = 3 - i - j;  // the remaining index
                const R aik = A.m[i][k], ajk = A.m[j][k];
                A.m[i][k] = c * aik - s * ajk;
                A.m[k][i] = A.m[i][k];
                A.m[j][k] = s * aik + c * ajk;
                A.m[k][j] = A.m[j][k];
                for (int r = 0; r < 3; ++r) {
                    const R vi = V.m[r][i], vj = V.m[r][j];
                    V.m[r][i] = c * vi - s * vj;
                    V.m[r][j] = s * vi + c * vj;
                }
            }
        }
        if (max_off < R(1e-5f)) break;
    }
    const R s0 = std::sqrt(std::max(R(0), A.m[0][0]));
    const R s1 = std::sqrt(std::max(R(0), A.m[1][1]));
    const R s2 = std::sqrt(std::max(R(0), A.m[2][2]));
    S = v3(s0, s1, s2);
    const M3<R> FV = mm(F, V);
    const R sv[3] = {s0, s1, s2};
    for (int c = 0; c < 3; ++c) {
        V3<R> col = v3(FV.m[0][c], FV.m[1][c], FV.m[2][c]);
        V3<R> u;
        if (sv[c] > R(1e-6f)) {
            u = divg(col, sv[c]);
        } else {
            u = v3<R>(c == 0 ? 1 : 0, c == 1 ? 1 : 0, c == 2 ? 1 : 0);
        }
        U.m[0][c] = u.x;
        U.m[1][c] = u.y;
        U.m[2][c] = u.z;
    }
}

// ---------------------------------------------------------------- state
template <class R> struct Particle {
    V3<R> x, v;
    M3<R> F, C;
    R m, V0;
};

template <class R> struct Node {
    V3<R> vel;  // momentum before the update, velocity after (spgrid_cpu.h:16-24)
    V3<R> force;
    R mass;
    int64_t mass_q;  // fixed-point mass (MASS_FIXED only)
    bool active;
};

template <class R> struct Block {
    Node<R> n[64];
    int ox, oy, oz;
    void clear() {
        for (auto& q : n) {
            q.vel = v3<R>(0, 0, 0);
            q.force = v3<R>(0, 0, 0);
            q.mass = R(0);
            q.mass_q = 0;
            q.active = false;
        }
    }
};

enum { MASS_FLOAT = 0, MASS_FIXED = 1 };

// spgrid_cpu.h:92-98 -- the block key, kept identical so that iteration over
// the hash map visits blocks in the same order as the reference.
static inline uint64_t block_key(int bx, int by, int bz) {
    uint64_t r = 0;
    r |= ((uint64_t)(bx + 100000) & 0x1FFFFF);
    r |= ((uint64_t)(by + 100000) & 0x1FFFFF) << 21;
    r |= ((uint64_t)(bz + 100000) & 0x1FFFFF) << 42;
    return r;
}

template <class R> struct Sim {
    R dx, inv_dx;
    std::unordered_map<uint64_t, Block<R>> blocks;
    std::vector<Particle<R>> p;
    V3<R> g = v3<R>(R(0), R(-9.81f), R(0));
    R alpha = R(0.40f), E = R(1.0e5f), nu = R(0.3f), rho = R(1500.0f), mu_floor = R(0.4f);
    int mass_mode = MASS_FLOAT;
    int q_exp = 0;  // mass quantum is 2^q_exp kg in MASS_FIXED mode

    explicit Sim(R dx_) : dx(dx_), inv_dx(R(1) / dx_) {}

    R mu_l() const { return E / (R(2) * (R(1) + nu)); }
    R la_l() const { return E * nu / ((R(1) + nu) * (R(1) - R(2) * nu)); }

    Node<R>* node(int i, int j, int k) {
        const int bx = i >> 2, by = j >> 2, bz = k >> 2;
        const uint64_t key = block_key(bx, by, bz);
        auto it = blocks.find(key);
        Block<R>* b;
        if (it != blocks.end()) {
            b = &it->second;
        } else {
            b = &blocks[key];
            b->ox = bx * 4;
            b->oy = by * 4;
            b->oz = bz * 4;
            b->clear();
        }
        return &b->n[((i & 3) * 4 + (j & 3)) * 4 + (k & 3)];
    }

    void clear() {
        for (auto& kv : blocks) kv.second.clear();
    }

    // mpm_solver.h:170-184 -- stencil base, fraction and quadratic B-spline weights.
    void stencil(const V3<R>& x, R f[3], int b[3], R w[3][3]) const {
        f[0] = x.x * inv_dx - R(0.5f);
        f[1] = x.y * inv_dx - R(0.5f);
        f[2] = x.z * inv_dx - R(0.5f);
        for (int a = 0; a < 3; ++a) {
            b[a] = (int)std::floor(f[a]);
            const R fr = f[a] - (R)b[a];
            w[a][0] = R(0.5f) * (R(1) - fr) * (R(1) - fr);
            w[a][1] = R(0.75f) - (fr - R(0.5f)) * (fr - R(0.5f));
            w[a][2] = R(0.5f) * fr * fr;
        }
    }

    // Kirchhoff stress of Hencky strain, mpm_solver.h:191-212.
    M3<R> hencky_tau(const M3<R>& F) const {
        M3<R> U, V;
        V3<R> s;
        svd3(F, U, s, V);
        s.x = std::max(s.x, R(1e-6f));
        s.y = std::max(s.y, R(1e-6f));
        s.z = std::max(s.z, R(1e-6f));
        const R lx = std::log(s.x), ly = std::log(s.y), lz = std::log(s.z);
        const R tr = lx + ly + lz;
        const R mu_ = mu_l(), la_ = la_l();
        const V3<R> td = v3(R(2) * mu_ * lx + la_ * tr, R(2) * mu_ * ly + la_ * tr, R(2) * mu_ * lz + la_ * tr);
        return mm(mm(U, diag3(td)), transp(U));
    }

    // Fixed-point deposit of one stencil (the GPU path does the same).
    static void quantize_weights(const R w[3][3], int64_t qp, int64_t d[27]) {
        int64_t acc = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const int idx = (i * 3 + j) * 3 + k;
                    const float wf = (float)(w[0][i] * w[1][j] * w[2][k]);
                    // w q_p to the nearest integer in FP32: floor(fma(w, q_p, 1/2)), rounded once
                    // (q_p < 2^23 is exact); at least 1 where w > 0, 0 where w = 0
                    int64_t q = (int64_t)std::fma(wf, (float)qp, 0.5f);
                    if (wf > 0.0f && q < 1) q = 1;
                    d[idx] = q;
                    if (idx != 13) acc += d[idx];
                }
        d[13] = qp - acc;
    }

    int64_t particle_quanta(R m) const { return (int64_t)std::nearbyint(std::ldexp((double)m, -q_exp)); }

    // mpm_solver.h:168-248
    void p2g_one(const Particle<R>& P) {
        R f[3], w[3][3];
        int b[3];
        stencil(P.x, f, b, w);
        const M3<R> tau = hencky_tau(P.F);
        const R ff = R(-4.0f) * inv_dx * inv_dx * P.V0;
        const M3<R> K = mscale(tau, ff);
        int64_t dq[27];
        if (mass_mode == MASS_FIXED) quantize_weights(w, particle_quanta(P.m), dq);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const R wt = w[0][i] * w[1][j] * w[2][k];
                    if (wt == R(0)) continue;
                    const int ni = b[0] + i, nj = b[1] + j, nk = b[2] + k;
                    const V3<R> d = v3(((R)ni - f[0] - R(0.5f)) * dx, ((R)nj - f[1] - R(0.5f)) * dx,
                                       ((R)nk - f[2] - R(0.5f)) * dx);
                    const V3<R> aff = mv(P.C, d);
                    Node<R>* N = node(ni, nj, nk);
                    const R mc = wt * P.m;
                    N->mass += mc;
                    N->mass_q += (mass_mode == MASS_FIXED) ? dq[(i * 3 + j) * 3 + k] : 0;
                    N->vel = add(N->vel, mul(add(P.v, aff), mc));
                    N->force = add(N->force, mul(mv(K, d), wt));
                    N->active = true;
                }
    }

    void p2g() {
        clear();
        for (auto& P : p) p2g_one(P);
        if (mass_mode == MASS_FIXED) {
            for (auto& kv : blocks)
                for (auto& N : kv.second.n)
                    N.mass = (R)std::ldexp((double)N.mass_q, q_exp);
        }
    }

    // mpm_solver.h:250-291
    void update(R dt) {
        for (auto& kv : blocks) {
            Block<R>& B = kv.second;
            for (int i = 0; i < 64; ++i) {
                Node<R>& N = B.n[i];
                if (!N.active || N.mass <= R(1e-6f)) continue;
                N.vel = divg(add(N.vel, mul(N.force, dt)), N.mass);
                N.vel = add(N.vel, mul(g, dt));
                const int jy = B.oy + ((i >> 2) & 3);
                const R ny = ((R)jy + R(0.5f)) * dx;
                if (ny < dx && N.vel.y < R(0)) {
                    const R vn = -N.vel.y;
                    N.vel.y = R(0);
                    const R tx = N.vel.x, tz = N.vel.z;
                    const R tm = std::sqrt(tx * tx + tz * tz);
                    const R budget = mu_floor * vn;
                    if (tm > R(1e-12f)) {
                        const R sc = std::max(R(0), (tm - budget) / tm);
                        N.vel.x = tx * sc;
                        N.vel.z = tz * sc;
                    }
                }
            }
        }
    }

    // mpm_solver.h:114-166
    void plasticity(Particle<R>& P, R dt) const {
        const M3<R> Ftr = mm(madd(ident<R>(), mscale(P.C, dt)), P.F);
        M3<R> U, V;
        V3<R> s;
        svd3(Ftr, U, s, V);
        s.x = std::max(s.x, R(1e-6f));
        s.y = std::max(s.y, R(1e-6f));
        s.z = std::max(s.z, R(1e-6f));
        V3<R> e = v3(std::log(s.x), std::log(s.y), std::log(s.z));
        const R tr = e.x + e.y + e.z;
        V3<R> eh = sub(e, v3(tr / R(3), tr / R(3), tr / R(3)));
        const R H = len(eh);
        if (tr > R(0)) {
            s = v3<R>(1, 1, 1);
        } else {
            const R mu_ = mu_l(), la_ = la_l();
            const R koh = (R(3) * la_ + R(2) * mu_) / (R(2) * mu_);
            const R limit = -alpha * koh * tr;
            if (H > limit) {
                const R sc = limit / H;
                eh = mul(eh, sc);
                e = add(eh, v3(tr / R(3), tr / R(3), tr / R(3)));
                s = v3(std::exp(e.x), std::exp(e.y), std::exp(e.z));
            }
        }
        P.F = mm(mm(U, diag3(s)), transp(V));
    }

    // mpm_solver.h:293-348
    void g2p_one(Particle<R>& P, R dt) {
        R f[3], w[3][3];
        int b[3];
        stencil(P.x, f, b, w);
        P.v = v3<R>(0, 0, 0);
        P.C = zero3<R>();
        const R kC = R(4.0f) * inv_dx * inv_dx;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    const R wt = w[0][i] * w[1][j] * w[2][k];
                    if (wt == R(0)) continue;
                    const int ni = b[0] + i, nj = b[1] + j, nk = b[2] + k;
                    const Node<R>* N = node(ni, nj, nk);
                    const V3<R> d = v3(((R)ni - f[0] - R(0.5f)) * dx, ((R)nj - f[1] - R(0.5f)) * dx,
                                       ((R)nk - f[2] - R(0.5f)) * dx);
                    const V3<R> wv = mul(N->vel, wt);
                    P.v = add(P.v, wv);
                    P.C = madd(P.C, mscale(outer(wv, d), kC));
                }
        P.x = add(P.x, mul(P.v, dt));
        plasticity(P, dt);
        if (P.x.y < R(0)) {
            P.x.y = R(0);
            P.v.y = R(0);
        }
    }

    void g2p(R dt) {
        for (auto& P : p) g2p_one(P, dt);
    }

    void step(R dt) {
        p2g();
        update(dt);
        g2p(dt);
    }
};

// ---------------------------------------------------------------- couplers
template <class R> struct Collider {
    V3<R> c;
    Q4<R> q;
    V3<R> he;
    V3<R> vl, wa;
    R friction, mass;
    int shape;  // 0 box, 1 cylinder (local Y axis, he = (R, L, R))
};
template <class R> struct Reaction {
    V3<R> f, t;
    int n;
};

// mpm_rigid_coupler.h:90-103
template <class R> static bool shape_inside(const Collider<R>& b, V3<R> l) {
    if (b.shape == 1) {
        const R Rr = b.he.x, L = b.he.y;
        const R r2 = l.x * l.x + l.z * l.z;
        return std::abs(l.y) <= L && r2 <= Rr * Rr;
    }
    return std::abs(l.x) <= b.he.x && std::abs(l.y) <= b.he.y && std::abs(l.z) <= b.he.z;
}
// mpm_rigid_coupler.h:106-132
template <class R> static V3<R> shape_normal(const Collider<R>& b, V3<R> l) {
    if (b.shape == 1) {
        const R Rr = b.he.x, L = b.he.y;
        const R r = std::sqrt(l.x * l.x + l.z * l.z);
        const R ad = L - std::abs(l.y), rd = Rr - r;
        if (ad < rd) return v3<R>(0, l.y >= R(0) ? R(1) : R(-1), 0);
        if (r > R(1e-9f)) return v3<R>(l.x / r, 0, l.z / r);
        return v3<R>(1, 0, 0);
    }
    const R ddx = b.he.x - std::abs(l.x), ddy = b.he.y - std::abs(l.y), ddz = b.he.z - std::abs(l.z);
    if (ddx <= ddy && ddx <= ddz) return v3<R>(l.x >= R(0) ? R(1) : R(-1), 0, 0);
    if (ddy <= ddz) return v3<R>(0, l.y >= R(0) ? R(1) : R(-1), 0);
    return v3<R>(0, 0, l.z >= R(0) ? R(1) : R(-1));
}

template <class R> static V3<R> aabb_half(const M3<R>& Rl, V3<R> he) {
    return v3(std::abs(Rl.m[0][0]) * he.x + std::abs(Rl.m[0][1]) * he.y + std::abs(Rl.m[0][2]) * he.z,
              std::abs(Rl.m[1][0]) * he.x + std::abs(Rl.m[1][1]) * he.y + std::abs(Rl.m[1][2]) * he.z,
              std::abs(Rl.m[2][0]) * he.x + std::abs(Rl.m[2][1]) * he.y + std::abs(Rl.m[2][2]) * he.z);
}

template <class R> static bool block_culled(const Block<R>& B, R dx, V3<R> lo, V3<R> hi) {
    const R blx = ((R)B.ox + R(0.5f)) * dx, bly = ((R)B.oy + R(0.5f)) * dx, blz = ((R)B.oz + R(0.5f)) * dx;
    const R bhx = ((R)(B.ox + 3) + R(0.5f)) * dx, bhy = ((R)(B.oy + 3) + R(0.5f)) * dx,
            bhz = ((R)(B.oz + 3) + R(0.5f)) * dx;
    if (bhx < lo.x || blx > hi.x) return true;
    if (bhy < lo.y || bly > hi.y) return true;
    if (bhz < lo.z || blz > hi.z) return true;
    return false;
}

// src/physics/coupling/mpm_rigid_coupler.cpp:58-143 (M0, Dirichlet clamp)
template <class R> static Reaction<R> couple_m0(Sim<R>& S, const Collider<R>& b, R dt) {
    Reaction<R> out{v3<R>(0, 0, 0), v3<R>(0, 0, 0), 0};
    if (dt <= R(0)) return out;
    const R dx = S.dx;
    const M3<R> Rl = quat_to_mat(b.q);
    const M3<R> Rw = transp(Rl);
    const V3<R> h = aabb_half(Rl, b.he);
    const V3<R> lo = sub(b.c, h), hi = add(b.c, h);
    for (auto& kv : S.blocks) {
        Block<R>& B = kv.second;
        if (block_culled(B, dx, lo, hi)) continue;
        for (int li = 0; li < 4; ++li)
            for (int lj = 0; lj < 4; ++lj)
                for (int lk = 0; lk < 4; ++lk) {
                    Node<R>& N = B.n[(li * 4 + lj) * 4 + lk];
                    if (!N.active || N.mass <= R(1e-8f)) continue;
                    const V3<R> xn = v3(((R)(B.ox + li) + R(0.5f)) * dx, ((R)(B.oy + lj) + R(0.5f)) * dx,
                                        ((R)(B.oz + lk) + R(0.5f)) * dx);
                    const V3<R> loc = mv(Rw, sub(xn, b.c));
                    if (!shape_inside(b, loc)) continue;
                    const V3<R> nl = shape_normal(b, loc);
                    const V3<R> r = sub(xn, b.c);
                    const V3<R> vb = add(b.vl, cross(b.wa, r));
                    const V3<R> vn0 = N.vel;
                    const V3<R> n = normalized(mv(Rl, nl));
                    const R vnn = dot(vn0, n), vnb = dot(vb, n);
                    const R vnc = std::max(vnn, vnb);
                    const V3<R> vtn = sub(vn0, mul(n, vnn));
                    const V3<R> vtb = sub(vb, mul(n, vnb));
                    const V3<R> vtnew = add(vtn, mul(sub(vtb, vtn), std::min(R(1), b.friction)));
                    const V3<R> va = add(mul(n, vnc), vtnew);
                    const V3<R> dv = sub(va, vn0);
                    const V3<R> dp = mul(dv, N.mass);
                    N.vel = va;
                    out.f = sub(out.f, divg(dp, dt));
                    out.t = sub(out.t, divg(cross(r, dp), dt));
                    ++out.n;
                }
    }
    return out;
}

// mpm_rigid_coupler_pic.h:88-184 (M1, mass-weighted PIC-rigid exchange)
template <class R> static Reaction<R> couple_m1(Sim<R>& S, const Collider<R>& b, R dt) {
    Reaction<R> out{v3<R>(0, 0, 0), v3<R>(0, 0, 0), 0};
    if (dt <= R(0)) return out;
    const R dx = S.dx;
    const M3<R> Rl = quat_to_mat(b.q);
    const M3<R> Rw = transp(Rl);
    const R vol = R(8.0f) * b.he.x * b.he.y * b.he.z;
    const int approx = std::max(1, (int)std::round(vol / (dx * dx * dx)));
    const R mb = b.mass / (R)approx;
    const V3<R> h = aabb_half(Rl, b.he);
    const V3<R> lo = sub(b.c, h), hi = add(b.c, h);
    V3<R> fa = v3<R>(0, 0, 0), ta = v3<R>(0, 0, 0);
    for (auto& kv : S.blocks) {
        Block<R>& B = kv.second;
        if (block_culled(B, dx, lo, hi)) continue;
        for (int li = 0; li < 4; ++li)
            for (int lj = 0; lj < 4; ++lj)
                for (int lk = 0; lk < 4; ++lk) {
                    Node<R>& N = B.n[(li * 4 + lj) * 4 + lk];
                    const V3<R> xn = v3(((R)(B.ox + li) + R(0.5f)) * dx, ((R)(B.oy + lj) + R(0.5f)) * dx,
                                        ((R)(B.oz + lk) + R(0.5f)) * dx);
                    const V3<R> loc = mv(Rw, sub(xn, b.c));
                    if (!shape_inside(b, loc)) continue;
                    const V3<R> r = sub(xn, b.c);
                    const V3<R> vb = add(b.vl, cross(b.wa, r));
                    V3<R> va;
                    if (!N.active || N.mass < R(1e-9f)) {
                        N.active = true;
                        N.mass = mb;
                        N.vel = vb;
                        va = vb;
                    } else {
                        const V3<R> p_old = mul(N.vel, N.mass);
                        const V3<R> p_add = mul(vb, mb);
                        const R mn = N.mass + mb;
                        va = divg(add(p_old, p_add), mn);
                        N.vel = va;
                        N.mass = mn;
                    }
                    const V3<R> dv = sub(va, vb);
                    const V3<R> dF = mul(dv, mb / dt);
                    fa = add(fa, dF);
                    ta = add(ta, cross(r, dF));
                    ++out.n;
                }
    }
    out.f = fa;
    out.t = ta;
    return out;
}

// ---------------------------------------------------------------- rigid body
// integrator.h:19-57 for one dynamic body (force/torque cleared afterwards).
template <class R> struct Body {
    V3<R> x;
    Q4<R> q;
    V3<R> v, w;
    R inv_mass;
    M3<R> inv_I;
    R lin_damp, ang_damp;
    V3<R> f, t;
};
// kinematic: the body moves at its own velocities whatever the forces (a
// dragged tool, as the research harness's tine drag moves its tine).
template <class R> static void integrate_body(Body<R>& b, R dt, bool kinematic = false) {
    if (!kinematic) {
        b.v = add(b.v, mul(mul(b.f, b.inv_mass), dt));
        b.w = add(b.w, mul(mv(b.inv_I, b.t), dt));
        b.v = mul(b.v, R(1) - b.lin_damp * dt);
        b.w = mul(b.w, R(1) - b.ang_damp * dt);
    }
    b.x = add(b.x, mul(b.v, dt));
    const Q4<R> qv{R(0), b.w.x, b.w.y, b.w.z};
    const Q4<R>& o = b.q;
    const Q4<R> pr{qv.w * o.w - qv.x * o.x - qv.y * o.y - qv.z * o.z, qv.w * o.x + qv.x * o.w + qv.y * o.z - qv.z * o.y,
                   qv.w * o.y - qv.x * o.z + qv.y * o.w + qv.z * o.x, qv.w * o.z + qv.x * o.y - qv.y * o.x + qv.z * o.w};
    const R hs = R(0.5f) * dt;
    b.q.w += pr.w * hs;
    b.q.x += pr.x * hs;
    b.q.y += pr.y * hs;
    b.q.z += pr.z * hs;
    const R n = std::sqrt(b.q.w * b.q.w + b.q.x * b.q.x + b.q.y * b.q.y + b.q.z * b.q.z);
    if (n < R(1e-6f)) {
        b.q = Q4<R>{1, 0, 0, 0};
    } else {
        const R inv = R(1) / n;
        b.q = Q4<R>{b.q.w * inv, b.q.x * inv, b.q.y * inv, b.q.z * inv};
    }
    b.f = v3<R>(0, 0, 0);
    b.t = v3<R>(0, 0, 0);
}

}  // namespace orc

// ================================================================= C ABI
// All array I/O is double precision; float instances round on input (values
// that began as float32 round-trip exactly).
using namespace orc;

namespace {
template <class R> M3<R> load_m3(const double* a) {
    M3<R> r;
    for (int i = 0; i < 9; ++i) r.m[i / 3][i % 3] = (R)a[i];
    return r;
}
template <class R> void store_m3(const M3<R>& r, double* a) {
    for (int i = 0; i < 9; ++i) a[i] = (double)r.m[i / 3][i % 3];
}
// Collider packed as 20 doubles: c(3) q(w,x,y,z)(4) he(3) vl(3) wa(3) friction mass shape pad
template <class R> Collider<R> load_col(const double* a) {
    Collider<R> c;
    c.c = v3<R>((R)a[0], (R)a[1], (R)a[2]);
    c.q = Q4<R>{(R)a[3], (R)a[4], (R)a[5], (R)a[6]};
    c.he = v3<R>((R)a[7], (R)a[8], (R)a[9]);
    c.vl = v3<R>((R)a[10], (R)a[11], (R)a[12]);
    c.wa = v3<R>((R)a[13], (R)a[14], (R)a[15]);
    c.friction = (R)a[16];
    c.mass = (R)a[17];
    c.shape = (int)a[18];
    return c;
}
template <class R> void store_reaction(const Reaction<R>& r, double* o) {
    o[0] = r.f.x; o[1] = r.f.y; o[2] = r.f.z;
    o[3] = r.t.x; o[4] = r.t.y; o[5] = r.t.z;
    o[6] = (double)r.n;
}

// The coupled substep in the contract order of mpm_rigid_coupler.h:209-211 and
// the body half of bench_compare_all.cpp:642-663: force += applied load;
// P2G; grid update; reaction = coupler(grid, collider of the body, dt);
// force += reaction.force, torque += reaction.torque; TimeIntegrator::
// integrate; G2P.  col: the collider template (shape, half extents,
// friction, mass; pose and velocities come from the body).  body: 31 values
// as orc_*_integrate_body, in/out.  applied: force(3) torque(3).  rec (may be
// null): per substep the body after integration and the reaction,
// x(3) q(4) v(3) w(3) force(3) torque(3) nodes_inside = 20 values.
// motion: 1 dynamic (TimeIntegrator::integrate), 2 kinematic.
template <class R> void sim_step_coupled(void* h, int method, const double* col, double* bd, const double* applied,
                                         double dt, int nsteps, double* rec, int motion) {
    Sim<R>& S = *(Sim<R>*)h;
    Collider<R> c = load_col<R>(col);
    Body<R> b;
    b.x = v3<R>((R)bd[0], (R)bd[1], (R)bd[2]);
    b.q = Q4<R>{(R)bd[3], (R)bd[4], (R)bd[5], (R)bd[6]};
    b.v = v3<R>((R)bd[7], (R)bd[8], (R)bd[9]);
    b.w = v3<R>((R)bd[10], (R)bd[11], (R)bd[12]);
    b.inv_mass = (R)bd[13];
    b.inv_I = load_m3<R>(bd + 14);
    b.lin_damp = (R)bd[23];
    b.ang_damp = (R)bd[24];
    b.f = v3<R>((R)bd[25], (R)bd[26], (R)bd[27]);
    b.t = v3<R>((R)bd[28], (R)bd[29], (R)bd[30]);
    const V3<R> af = v3<R>((R)applied[0], (R)applied[1], (R)applied[2]);
    const V3<R> at = v3<R>((R)applied[3], (R)applied[4], (R)applied[5]);
    for (int s_ = 0; s_ < nsteps; ++s_) {
        b.f = add(b.f, af);
        b.t = add(b.t, at);
        S.p2g();
        S.update((R)dt);
        c.c = b.x;
        c.q = b.q;
        c.vl = b.v;
        c.wa = b.w;
        const Reaction<R> r = method == 0 ? couple_m0<R>(S, c, (R)dt) : couple_m1<R>(S, c, (R)dt);
        b.f = add(b.f, r.f);
        b.t = add(b.t, r.t);
        integrate_body<R>(b, (R)dt, motion == 2);
        S.g2p((R)dt);
        if (rec) {
            double* o = rec + 20 * (int64_t)s_;
            o[0] = b.x.x; o[1] = b.x.y; o[2] = b.x.z;
            o[3] = b.q.w; o[4] = b.q.x; o[5] = b.q.y; o[6] = b.q.z;
            o[7] = b.v.x; o[8] = b.v.y; o[9] = b.v.z;
            o[10] = b.w.x; o[11] = b.w.y; o[12] = b.w.z;
            store_reaction<R>(r, o + 13);
        }
    }
    bd[0] = b.x.x; bd[1] = b.x.y; bd[2] = b.x.z;
    bd[3] = b.q.w; bd[4] = b.q.x; bd[5] = b.q.y; bd[6] = b.q.z;
    bd[7] = b.v.x; bd[8] = b.v.y; bd[9] = b.v.z;
    bd[10] = b.w.x; bd[11] = b.w.y; bd[12] = b.w.z;
    bd[25] = b.f.x; bd[26] = b.f.y; bd[27] = b.f.z;
    bd[28] = b.t.x; bd[29] = b.t.y; bd[30] = b.t.z;
}

template <class R> void* sim_create(double dx, int mass_mode, int q_exp) {
    auto* s = new Sim<R>((R)dx);
    s->mass_mode = mass_mode;
    s->q_exp = q_exp;
    return s;
}
template <class R> void sim_params(void* h, const double* prm) {
    auto* s = (Sim<R>*)h;
    // E, nu, rho, alpha, mu_floor, gx, gy, gz
    s->E = (R)prm[0];
    s->nu = (R)prm[1];
    s->rho = (R)prm[2];
    s->alpha = (R)prm[3];
    s->mu_floor = (R)prm[4];
    s->g = v3<R>((R)prm[5], (R)prm[6], (R)prm[7]);
}
// Particle state as SoA doubles: x(3n) v(3n) F(9n) C(9n) m(n) V0(n); row-major 3x3 per particle.
template <class R> void sim_set_particles(void* h, int64_t n, const double* x, const double* v, const double* F,
                                          const double* C, const double* m, const double* V0) {
    auto* s = (Sim<R>*)h;
    s->p.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        Particle<R>& P = s->p[(size_t)i];
        P.x = v3<R>((R)x[3 * i], (R)x[3 * i + 1], (R)x[3 * i + 2]);
        P.v = v3<R>((R)v[3 * i], (R)v[3 * i + 1], (R)v[3 * i + 2]);
        P.F = load_m3<R>(F + 9 * i);
        P.C = load_m3<R>(C + 9 * i);
        P.m = (R)m[i];
        P.V0 = (R)V0[i];
    }
}
template <class R> void sim_get_particles(void* h, double* x, double* v, double* F, double* C) {
    auto* s = (Sim<R>*)h;
    for (size_t i = 0; i < s->p.size(); ++i) {
        const Particle<R>& P = s->p[i];
        x[3 * i] = P.x.x; x[3 * i + 1] = P.x.y; x[3 * i + 2] = P.x.z;
        v[3 * i] = P.v.x; v[3 * i + 1] = P.v.y; v[3 * i + 2] = P.v.z;
        store_m3(P.F, F + 9 * i);
        store_m3(P.C, C + 9 * i);
    }
}
// Dense dump of the node box [org, org+dims): mass_q(int64), mass, vel(3), force(3), active, allocated.
template <class R> void sim_grid_dense(void* h, const int* org, const int* dims, int64_t* mq, double* mass,
                                       double* vel, double* force, uint8_t* active, uint8_t* alloc) {
    auto* s = (Sim<R>*)h;
    const int64_t nn = (int64_t)dims[0] * dims[1] * dims[2];
    for (int64_t i = 0; i < nn; ++i) {
        mq[i] = 0; mass[i] = 0; active[i] = 0; alloc[i] = 0;
        for (int c = 0; c < 3; ++c) { vel[3 * i + c] = 0; force[3 * i + c] = 0; }
    }
    for (auto& kv : s->blocks) {
        Block<R>& B = kv.second;
        for (int li = 0; li < 4; ++li)
            for (int lj = 0; lj < 4; ++lj)
                for (int lk = 0; lk < 4; ++lk) {
                    const int gi = B.ox + li - org[0], gj = B.oy + lj - org[1], gk = B.oz + lk - org[2];
                    if (gi < 0 || gj < 0 || gk < 0 || gi >= dims[0] || gj >= dims[1] || gk >= dims[2]) continue;
                    const int64_t o = ((int64_t)gi * dims[1] + gj) * dims[2] + gk;
                    const Node<R>& N = B.n[(li * 4 + lj) * 4 + lk];
                    mq[o] = N.mass_q;
                    mass[o] = N.mass;
                    vel[3 * o] = N.vel.x; vel[3 * o + 1] = N.vel.y; vel[3 * o + 2] = N.vel.z;
                    force[3 * o] = N.force.x; force[3 * o + 1] = N.force.y; force[3 * o + 2] = N.force.z;
                    active[o] = N.active ? 1 : 0;
                    alloc[o] = 1;
                }
    }
}
// Overwrite node state inside the dense box (used to feed identical grids to the GPU and the oracle).
template <class R> void sim_grid_set_dense(void* h, const int* org, const int* dims, const double* mass,
                                           const double* vel, const uint8_t* active) {
    auto* s = (Sim<R>*)h;
    for (int gi = 0; gi < dims[0]; ++gi)
        for (int gj = 0; gj < dims[1]; ++gj)
            for (int gk = 0; gk < dims[2]; ++gk) {
                const int64_t o = ((int64_t)gi * dims[1] + gj) * dims[2] + gk;
                if (!active[o]) continue;
                Node<R>* N = s->node(gi + org[0], gj + org[1], gk + org[2]);
                N->mass = (R)mass[o];
                N->vel = v3<R>((R)vel[3 * o], (R)vel[3 * o + 1], (R)vel[3 * o + 2]);
                N->active = true;
            }
}
}  // namespace

#define ORC_INSTANTIATE(SUF, R)                                                                                    \
    extern "C" void* orc_##SUF##_create(double dx, int mass_mode, int q_exp) {                                     \
        return sim_create<R>(dx, mass_mode, q_exp);                                                                \
    }                                                                                                              \
    extern "C" void orc_##SUF##_destroy(void* h) { delete (Sim<R>*)h; }                                            \
    extern "C" void orc_##SUF##_params(void* h, const double* prm) { sim_params<R>(h, prm); }                     \
    extern "C" void orc_##SUF##_set_particles(void* h, int64_t n, const double* x, const double* v,                \
                                              const double* F, const double* C, const double* m,                  \
                                              const double* V0) {                                                  \
        sim_set_particles<R>(h, n, x, v, F, C, m, V0);                                                             \
    }                                                                                                              \
    extern "C" void orc_##SUF##_get_particles(void* h, double* x, double* v, double* F, double* C) {               \
        sim_get_particles<R>(h, x, v, F, C);                                                                       \
    }                                                                                                              \
    extern "C" void orc_##SUF##_p2g(void* h) { ((Sim<R>*)h)->p2g(); }                                             \
    extern "C" void orc_##SUF##_update(void* h, double dt) { ((Sim<R>*)h)->update((R)dt); }                       \
    extern "C" void orc_##SUF##_g2p(void* h, double dt) { ((Sim<R>*)h)->g2p((R)dt); }                             \
    extern "C" void orc_##SUF##_step(void* h, double dt, int nsteps) {                                             \
        for (int s_ = 0; s_ < nsteps; ++s_) ((Sim<R>*)h)->step((R)dt);                                            \
    }                                                                                                              \
    extern "C" void orc_##SUF##_couple(void* h, int method, const double* col, double dt, double* out) {           \
        Collider<R> c = load_col<R>(col);                                                                          \
        Reaction<R> r = method == 0 ? couple_m0<R>(*(Sim<R>*)h, c, (R)dt) : couple_m1<R>(*(Sim<R>*)h, c, (R)dt);  \
        store_reaction<R>(r, out);                                                                                 \
    }                                                                                                              \
    extern "C" void orc_##SUF##_step_coupled(void* h, int method, const double* col, double* body,                \
                                             const double* applied, double dt, int nsteps, double* rec,            \
                                             int motion) {                                                         \
        sim_step_coupled<R>(h, method, col, body, applied, dt, nsteps, rec, motion);                               \
    }                                                                                                              \
    extern "C" int64_t orc_##SUF##_num_blocks(void* h) { return (int64_t)((Sim<R>*)h)->blocks.size(); }           \
    extern "C" void orc_##SUF##_grid_dense(void* h, const int* org, const int* dims, int64_t* mq, double* mass,    \
                                           double* vel, double* force, uint8_t* active, uint8_t* alloc) {          \
        sim_grid_dense<R>(h, org, dims, mq, mass, vel, force, active, alloc);                                      \
    }                                                                                                              \
    extern "C" void orc_##SUF##_grid_set_dense(void* h, const int* org, const int* dims, const double* mass,       \
                                               const double* vel, const uint8_t* active) {                         \
        sim_grid_set_dense<R>(h, org, dims, mass, vel, active);                                                    \
    }                                                                                                              \
    extern "C" int64_t orc_##SUF##_grid_mass_q_total(void* h) {                                                    \
        int64_t t = 0;                                                                                             \
        for (auto& kv : ((Sim<R>*)h)->blocks)                                                                      \
            for (auto& N : kv.second.n) t += N.mass_q;                                                             \
        return t;                                                                                                  \
    }                                                                                                              \
    extern "C" void orc_##SUF##_svd(const double* F, double* U, double* S, double* V) {                            \
        M3<R> u, v;                                                                                                \
        V3<R> s;                                                                                                   \
        svd3<R>(load_m3<R>(F), u, s, v);                                                                           \
        store_m3<R>(u, U);                                                                                         \
        store_m3<R>(v, V);                                                                                         \
        S[0] = s.x; S[1] = s.y; S[2] = s.z;                                                                        \
    }                                                                                                              \
    extern "C" void orc_##SUF##_hencky_tau(void* h, const double* F, double* tau) {                                \
        store_m3<R>(((Sim<R>*)h)->hencky_tau(load_m3<R>(F)), tau);                                                 \
    }                                                                                                              \
    extern "C" void orc_##SUF##_plasticity(void* h, const double* F, const double* C, double dt, double* Fout) {   \
        Particle<R> P;                                                                                             \
        P.F = load_m3<R>(F);                                                                                       \
        P.C = load_m3<R>(C);                                                                                       \
        ((Sim<R>*)h)->plasticity(P, (R)dt);                                                                        \
        store_m3<R>(P.F, Fout);                                                                                    \
    }                                                                                                              \
    extern "C" void orc_##SUF##_weights(void* h, const double* x, double* fr, int* base, double* w) {              \
        R f[3], ww[3][3];                                                                                          \
        int b[3];                                                                                                  \
        ((Sim<R>*)h)->stencil(v3<R>((R)x[0], (R)x[1], (R)x[2]), f, b, ww);                                         \
        for (int a = 0; a < 3; ++a) {                                                                              \
            fr[a] = f[a];                                                                                          \
            base[a] = b[a];                                                                                        \
            for (int c = 0; c < 3; ++c) w[a * 3 + c] = ww[a][c];                                                   \
        }                                                                                                          \
    }                                                                                                              \
    /* body: x(3) q(4) v(3) w(3) inv_mass inv_I(9) lin_damp ang_damp f(3) t(3) = 31 doubles, in/out */            \
    extern "C" void orc_##SUF##_integrate_body(double* bd, double dt) {                                            \
        Body<R> b;                                                                                                 \
        b.x = v3<R>((R)bd[0], (R)bd[1], (R)bd[2]);                                                                 \
        b.q = Q4<R>{(R)bd[3], (R)bd[4], (R)bd[5], (R)bd[6]};                                                       \
        b.v = v3<R>((R)bd[7], (R)bd[8], (R)bd[9]);                                                                 \
        b.w = v3<R>((R)bd[10], (R)bd[11], (R)bd[12]);                                                              \
        b.inv_mass = (R)bd[13];                                                                                    \
        b.inv_I = load_m3<R>(bd + 14);                                                                             \
        b.lin_damp = (R)bd[23];                                                                                    \
        b.ang_damp = (R)bd[24];                                                                                    \
        b.f = v3<R>((R)bd[25], (R)bd[26], (R)bd[27]);                                                              \
        b.t = v3<R>((R)bd[28], (R)bd[29], (R)bd[30]);                                                              \
        integrate_body<R>(b, (R)dt);                                                                               \
        bd[0] = b.x.x; bd[1] = b.x.y; bd[2] = b.x.z;                                                               \
        bd[3] = b.q.w; bd[4] = b.q.x; bd[5] = b.q.y; bd[6] = b.q.z;                                                \
        bd[7] = b.v.x; bd[8] = b.v.y; bd[9] = b.v.z;                                                               \
        bd[10] = b.w.x; bd[11] = b.w.y; bd[12] = b.w.z;                                                            \
        bd[25] = b.f.x; bd[26] = b.f.y; bd[27] = b.f.z;                                                            \
        bd[28] = b.t.x; bd[29] = b.t.y; bd[30] = b.t.z;                                                            \
    }

ORC_INSTANTIATE(f32, float)
ORC_INSTANTIATE(f64, double)
