// mpm_fused.hip -- the fused substep kernel of BM_MODE_FAST: G2P of substep
// n and P2G of substep n+1 in one pass over the particles.
//
// One wave (64 lanes) runs one unit: up to unit_p particles of one 4^3-node
// grid block (particles are binned by the block of their stencil base
// floor(x/dx - 1/2), mpm_fast_kernels.hip).  The P2G deposit goes into an
// 8^3-node LDS tile (the block's nodes, the two layers the quadratic stencil
// reaches and one drift layer on each side) that is flushed to HBM once.
//
// LDS accumulation without same-address collisions: the wave first ranks its
// particles per stencil-base cell (64 cells per block = 64 lanes), then works
// in rounds -- in round r lane c holds ranks 2r and 2r+1 of cell c.  For a
// fixed stencil offset every lane of a round then addresses a different node,
// so the scatter either uses collision-free LDS atomics (RMW = false) or a
// plain read-modify-write in program order (RMW = true; one wave, so no
// other writer exists).  The two particles of a lane share their 27 nodes
// and are deposited with one read-modify-write per node.  Particles that
// changed cell during their G2P, drifted out of the block since the last
// re-binning, or exceed the rank table go through LDS atomics (or global
// atomics when their stencil left the tile).  GTILE stages the G2P node
// velocities in LDS as well; otherwise the gather reads the grid through the
// caches.

#include "mpm_fast_device.hpp"

namespace bm {

constexpr int RMAX = 16;        // particle ranks per cell held in the round table
constexpr int UNIT_MAX = 1024;  // largest unit (leftover-list capacity)

// LDS slot of tile node (ix, iy, iz): rows of 8 nodes along z are rotated by
// 4 on odd y, so the 8 lanes that share a 128-byte LDS cycle in a round
// (cells (cx, cy..cy+1, 0..3)) address 8 distinct 16-byte bank groups.
__device__ __forceinline__ int tix(int ix, int iy, int iz) { return ((ix * TE + iy) * TE) + ((iz + 4 * (iy & 1)) & 7); }
__device__ __forceinline__ void tile_coords(int k, int& ix, int& iy, int& iz) {
    ix = k >> 6;
    iy = (k >> 3) & 7;
    iz = ((k & 7) - 4 * (iy & 1)) & 7;
}

// The fixed-point node mass is accumulated in LDS as two 32-bit sums: the
// low 22 bits of each deposit (in the w word of the momentum record) and the
// rest.  A unit holds at most UNIT_MAX = 1024 particles and each particle
// deposits once per node, so neither sum can overflow.
constexpr int MLO_BITS = 22;
constexpr unsigned MLO_MASK = (1u << MLO_BITS) - 1u;

// floor(x/dx - 1/2) rounded exactly as make_stencil rounds it.
__device__ __forceinline__ int base_of(float x, float inv_dx) {
#pragma clang fp contract(off)
    return (int)floorf(x * inv_dx - 0.5f);
}

// What a particle deposits: its stencil at the new position and the folded
// momentum coefficients.  rx..rz: stencil base relative to the tile origin.
struct Dep {
    Stencil s;
    v3<float> mvel;
    m3<float> Aff;
    double qp;
    int rx, ry, rz;
    int state;  // 0 nothing to deposit, 1 tile, 2 outside the tile
};

template <bool DO_G2P, bool RMW, bool GTILE>
struct Fused {
    struct Smem {
        float4 vt[(DO_G2P && GTILE) ? TNODES : 1];  // G2P node velocities
        float4 pt[TNODES];                          // deposited momentum; w = low mass bits (uint)
        unsigned int mh[TNODES];                    // high mass bits
        unsigned short tab[RMAX][64];               // round table: tab[r][cell] = particle
        unsigned short left[UNIT_MAX];              // particles outside the table
        int cnt[64];
        int nleft;
        int tb[27];
    };

    // G2P (when DO_G2P), plasticity and write-back of particle i; returns
    // the particle's deposit.
    static __device__ __forceinline__ Dep advance(Smem& S, int64_t i, const PView& P, const GridBuf& gin,
                                                  const Domain& d, const Phys& ph, const FastArgs& fa, int* err,
                                                  int tx0, int ty0, int tz0) {
        Dep D;
        D.state = 0;
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
            if (!stencil_ok(d, lx, ly, lz)) {
                atomicOr(err, 1);
                return D;
            }
            const float kC = 4.0f * inv_dx * inv_dx;
            const int rx = lx - tx0, ry = ly - ty0, rz = lz - tz0;
            if (GTILE && rx >= 0 && ry >= 0 && rz >= 0 && rx <= TE - 3 && ry <= TE - 3 && rz <= TE - 3) {
                gather(s, dx, kC, [&](int a, int b, int c) { return S.vt[tix(rx + a, ry + b, rz + c)]; }, vel, Cm);
            } else {
                gather(s, dx, kC, [&](int a, int b, int c) { return gin.mv[node_of(d, lx + a, ly + b, lz + c)]; }, vel,
                       Cm);
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
                fa.redo[atomicAdd(fa.n_redo, 1)] = (int)i;  // state untouched: redone from scratch in FP64
                return D;
            }
            const Material M{ph.mu, ph.lambda, ph.alpha, ph.kohesion};
            const v3<float> sn = dp_fast(sg, M, lns);
            F = recompose(Us, sn, V);
            if (y < 0.0f) {
                y = 0.0f;
                vel.y = 0.0f;
            }
            check_finite(err, -1, x, y, z, vel.x, vel.y, vel.z);
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
            if (!svd_fast(F, Us, sg, V)) {
                fa.redo[atomicAdd(fa.n_redo, 1)] = (int)i;
                return D;
            }
            lns = mk(flog(fmaxf(sg.x, 1e-6f)), flog(fmaxf(sg.y, 1e-6f)), flog(fmaxf(sg.z, 1e-6f)));
        }
        const float mass = P.p[PM * P.cap + i];
        const float V0 = P.p[PV0 * P.cap + i];
        p2g_coeffs(mass, V0, vel, Cm, Us, lns, ph, inv_dx, fa.dt_fold, D.mvel, D.Aff);
        make_stencil(x, y, z, inv_dx, D.s);
        const int lx = D.s.b[0] - d.ox, ly = D.s.b[1] - d.oy, lz = D.s.b[2] - d.oz;
        if (!stencil_ok(d, lx, ly, lz)) {
            atomicOr(err, 1);
            return D;
        }
        D.qp = rint((double)mass * fa.qscale_inv);
        D.rx = lx - tx0;
        D.ry = ly - ty0;
        D.rz = lz - tz0;
        D.state = (D.rx >= 0 && D.ry >= 0 && D.rz >= 0 && D.rx <= TE - 3 && D.ry <= TE - 3 && D.rz <= TE - 3) ? 1 : 2;
        return D;
    }

    static __device__ __forceinline__ bool in_cell(const Dep& D, int cell) {
        return D.state == 1 && cell >= 0 && D.rx - 1 == (cell >> 4) && D.ry - 1 == ((cell >> 2) & 3) &&
               D.rz - 1 == (cell & 3);
    }

    // Ordered read-modify-write of one node record (no other lane of the
    // round addresses this node for this stencil offset).
    static __device__ __forceinline__ void rmw(Smem& S, int k, v3<float> mom, long long dq) {
        float4 q = S.pt[k];
        q.x += mom.x;
        q.y += mom.y;
        q.z += mom.z;
        q.w = __uint_as_float(__float_as_uint(q.w) + ((unsigned)dq & MLO_MASK));
        S.pt[k] = q;
        S.mh[k] += (unsigned)(dq >> MLO_BITS);
        __asm__ volatile("" ::: "memory");
    }

    // Deposit of one particle: RMW when the lane owns the stencil's cell in
    // this round, LDS atomics when it does not, global atomics off the tile.
    static __device__ __forceinline__ int deposit(Smem& S, const Dep& D, bool own, const GridBuf& gout,
                                                  uint8_t* bever, const Domain& d, int tx0, int ty0, int tz0) {
        if (D.state == 0) return 0;
        const float dx = d.dx;
        if (D.state == 2) {
            const int lx = D.rx + tx0, ly = D.ry + ty0, lz = D.rz + tz0;
            scatter(D.s, dx, D.qp, D.mvel, D.Aff, [&](int a, int b, int c, float w, int dq, v3<float> mom) {
                if (w != 0.0f) deposit_global(gout, bever, d, lx + a, ly + b, lz + c, dq, mom);
            });
            return 1;
        }
        if (RMW && own) {
            scatter(D.s, dx, D.qp, D.mvel, D.Aff, [&](int a, int b, int c, float w, int dq, v3<float> mom) {
                rmw(S, tix(D.rx + a, D.ry + b, D.rz + c), mom, dq);
            });
        } else {
            // zero weights deposit zero: no per-node branch is needed
            scatter(D.s, dx, D.qp, D.mvel, D.Aff, [&](int a, int b, int c, float w, int dq, v3<float> mom) {
                const int k = tix(D.rx + a, D.ry + b, D.rz + c);
                float* q = (float*)&S.pt[k];
                atomicAdd(q + 0, mom.x);
                atomicAdd(q + 1, mom.y);
                atomicAdd(q + 2, mom.z);
                atomicAdd((unsigned*)(q + 3), (unsigned)dq & MLO_MASK);
                atomicAdd(&S.mh[k], (unsigned)(dq >> MLO_BITS));
            });
        }
        return 0;
    }

    // Two particles with the same stencil base owned by this lane: one
    // read-modify-write per node for both.
    static __device__ __forceinline__ void deposit2(Smem& S, const Dep& A, const Dep& B, float dx) {
        long long accA = 0, accB = 0;
        v3<float> mcA, mcB;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float dxA = ((float)(A.s.b[0] + a) - A.s.f[0] - 0.5f) * dx;
            const float dxB = ((float)(B.s.b[0] + a) - B.s.f[0] - 0.5f) * dx;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const float dyA = ((float)(A.s.b[1] + b) - A.s.f[1] - 0.5f) * dx;
                const float dyB = ((float)(B.s.b[1] + b) - B.s.f[1] - 0.5f) * dx;
                const float pxA = A.mvel.x + A.Aff.a[0] * dxA + A.Aff.a[1] * dyA;
                const float pyA = A.mvel.y + A.Aff.a[3] * dxA + A.Aff.a[4] * dyA;
                const float pzA = A.mvel.z + A.Aff.a[6] * dxA + A.Aff.a[7] * dyA;
                const float pxB = B.mvel.x + B.Aff.a[0] * dxB + B.Aff.a[1] * dyB;
                const float pyB = B.mvel.y + B.Aff.a[3] * dxB + B.Aff.a[4] * dyB;
                const float pzB = B.mvel.z + B.Aff.a[6] * dxB + B.Aff.a[7] * dyB;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float dzA = ((float)(A.s.b[2] + c) - A.s.f[2] - 0.5f) * dx;
                    const float dzB = ((float)(B.s.b[2] + c) - B.s.f[2] - 0.5f) * dx;
                    const float wA = A.s.w[0][a] * A.s.w[1][b] * A.s.w[2][c];
                    const float wB = B.s.w[0][a] * B.s.w[1][b] * B.s.w[2][c];
                    const v3<float> mA = mk(wA * (pxA + A.Aff.a[2] * dzA), wA * (pyA + A.Aff.a[5] * dzA),
                                            wA * (pzA + A.Aff.a[8] * dzA));
                    const v3<float> mB = mk(wB * (pxB + B.Aff.a[2] * dzB), wB * (pyB + B.Aff.a[5] * dzB),
                                            wB * (pzB + B.Aff.a[8] * dzB));
                    if (a == 1 && b == 1 && c == 1) {  // centres take the remainders, last
                        mcA = mA;
                        mcB = mB;
                        continue;
                    }
                    const long long qA = mass_quanta(wA, A.qp), qB = mass_quanta(wB, B.qp);
                    accA += qA;
                    accB += qB;
                    rmw(S, tix(A.rx + a, A.ry + b, A.rz + c), mA + mB, qA + qB);
                }
            }
        }
        rmw(S, tix(A.rx + 1, A.ry + 1, A.rz + 1), mcA + mcB, ((long long)A.qp - accA) + ((long long)B.qp - accB));
    }

    static __device__ void run(PView P, const int4* __restrict__ units, GridBuf gin, GridBuf gout, uint8_t* bever,
                               Domain d, Phys ph, FastArgs fa, int* err, unsigned long long* counters) {
        __shared__ Smem S;
        // XCD-aware unit order: consecutive workgroups land on different XCDs,
        // so each XCD gets a contiguous (spatially coherent) run of units.
        const int per = (gridDim.x + 7) / 8;
        const int u = (blockIdx.x % 8) * per + blockIdx.x / 8;
        if (u >= fa.n_units) return;
        const int4 U = units[u];
        const int blk = U.x;
        const int bz = blk % d.nbz, by = (blk / d.nbz) % d.nby, bx = blk / (d.nbz * d.nby);
        const int tx0 = bx * 4 - 1, ty0 = by * 4 - 1, tz0 = bz * 4 - 1;  // domain-local tile origin
        const int lane = threadIdx.x;

        for (int k = lane; k < TNODES; k += 64) {
            if (DO_G2P && GTILE) {
                int ix, iy, iz;
                tile_coords(k, ix, iy, iz);
                const int lx = tx0 + ix, ly = ty0 + iy, lz = tz0 + iz;
                float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
                if (lx >= 0 && ly >= 0 && lz >= 0 && lx < d.nx && ly < d.ny && lz < d.nz)
                    q = gin.mv[node_of(d, lx, ly, lz)];
                S.vt[k] = q;
            }
            S.pt[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            S.mh[k] = 0u;
        }
        S.cnt[lane] = 0;
        if (lane < 27) S.tb[lane] = 0;
        if (lane == 0) S.nleft = 0;
        __syncthreads();

        // ---- rank the unit's particles per stencil-base cell of the block
        const float inv_dx = d.inv_dx;
        for (int pi = lane; pi < U.z; pi += 64) {
            const int64_t i = (int64_t)U.y + pi;
            const int cx = base_of(P.p[(PX + 0) * P.cap + i], inv_dx) - d.ox - tx0 - 1;
            const int cy = base_of(P.p[(PX + 1) * P.cap + i], inv_dx) - d.oy - ty0 - 1;
            const int cz = base_of(P.p[(PX + 2) * P.cap + i], inv_dx) - d.oz - tz0 - 1;
            if (cx >= 0 && cy >= 0 && cz >= 0 && cx < 4 && cy < 4 && cz < 4) {
                const int cell = (cx * 4 + cy) * 4 + cz;
                const int r = atomicAdd(&S.cnt[cell], 1);
                if (r < RMAX) {
                    S.tab[r][cell] = (unsigned short)pi;
                    continue;
                }
            }
            S.left[atomicAdd(&S.nleft, 1)] = (unsigned short)pi;
        }
        __syncthreads();

        const int mine = min(S.cnt[lane], RMAX);
        int rounds = mine;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
        int off = 0;
        for (int r = 0; r < rounds; r += 2) {
            Dep A, B;
            A.state = 0;
            B.state = 0;
            if (r < mine) A = advance(S, (int64_t)U.y + S.tab[r][lane], P, gin, d, ph, fa, err, tx0, ty0, tz0);
            if (r + 1 < mine) B = advance(S, (int64_t)U.y + S.tab[r + 1][lane], P, gin, d, ph, fa, err, tx0, ty0, tz0);
            const bool ownA = in_cell(A, lane), ownB = in_cell(B, lane);
            if (RMW && ownA && ownB) {
                deposit2(S, A, B, d.dx);
            } else {
                off += deposit(S, A, ownA, gout, bever, d, tx0, ty0, tz0);
                off += deposit(S, B, ownB, gout, bever, d, tx0, ty0, tz0);
            }
        }
        const int nl = S.nleft;
        for (int k = lane; k < nl; k += 64) {
            const Dep D = advance(S, (int64_t)U.y + S.left[k], P, gin, d, ph, fa, err, tx0, ty0, tz0);
            off += deposit(S, D, false, gout, bever, d, tx0, ty0, tz0);
        }
        if (off) atomicAdd(&counters[1], (unsigned long long)off);
        __syncthreads();

        // ---- flush the tile into the next grid
        for (int k = lane; k < TNODES; k += 64) {
            const float4 q = S.pt[k];
            const unsigned long long m =
                ((unsigned long long)S.mh[k] << MLO_BITS) + (unsigned long long)__float_as_uint(q.w);
            if (m == 0ull) continue;
            int ix, iy, iz;
            tile_coords(k, ix, iy, iz);
            const int lx = tx0 + ix, ly = ty0 + iy, lz = tz0 + iz;
            const int64_t nid = node_of(d, lx, ly, lz);
            atomicAdd((unsigned long long*)&gout.mq[nid], m);
            float* nm = (float*)&gout.mv[nid];
            atomicAdd(nm + 0, q.x);
            atomicAdd(nm + 1, q.y);
            atomicAdd(nm + 2, q.z);
            // which of the 3x3x3 blocks around the unit's block holds the node
            S.tb[(((lx >> 2) - bx + 1) * 3 + ((ly >> 2) - by + 1)) * 3 + ((lz >> 2) - bz + 1)] = 1;
        }
        __syncthreads();
        if (lane < 27 && S.tb[lane]) {
            const int qx = lane / 9, qy = (lane / 3) % 3, qz = lane % 3;
            touch_block_f(gout, bever, ((bx + qx - 1) * d.nby + (by + qy - 1)) * d.nbz + (bz + qz - 1));
        }
    }
};

template <bool DO_G2P, bool RMW, bool GTILE>
__global__ void __launch_bounds__(64) k_fused(PView P, const int4* __restrict__ units, GridBuf gin, GridBuf gout,
                                              uint8_t* bever, Domain d, Phys ph, FastArgs fa, int* err,
                                              unsigned long long* counters) {
    Fused<DO_G2P, RMW, GTILE>::run(P, units, gin, gout, bever, d, ph, fa, err, counters);
}

void launch_redo(bool do_g2p, const PView& P, const GridBuf& gin, const GridBuf& gout, uint8_t* bever,
                 const Domain& d, const Phys& ph, const FastArgs& fa, int* err, unsigned long long* counters,
                 hipStream_t s);

template <bool DO_G2P, bool RMW, bool GTILE>
static void go(int nwg, hipStream_t s, const PView& P, const int4* units, const GridBuf& gin, const GridBuf& gout,
               uint8_t* bever, const Domain& d, const Phys& ph, const FastArgs& fa, int* err,
               unsigned long long* counters) {
    hipLaunchKernelGGL((k_fused<DO_G2P, RMW, GTILE>), dim3(nwg), dim3(64), 0, s, P, units, gin, gout, bever, d, ph, fa,
                       err, counters);
}

// The fused substep kernel and its FP64 redo pass.  `variant` bit 0 selects
// plain read-modify-write (else collision-free LDS atomics), bit 1 stages
// the G2P nodes in LDS.  The redo kernel's grid is fixed (it strides over
// the device-side count) so no host round trip is needed.
void launch_fused(bool do_g2p, const PView& P, const int4* units, int n_units, const GridBuf& gin,
                  const GridBuf& gout, uint8_t* bever, const Domain& d, const Phys& ph, float dt_g2p, float dt_fold,
                  double qscale_inv, int* err, unsigned long long* counters, int* redo, int* n_redo, int variant,
                  hipStream_t s) {
    if (n_units <= 0) return;
    FastArgs fa{dt_g2p, dt_fold, n_units, qscale_inv, redo, n_redo, nullptr, nullptr};
    const int nwg = ((n_units + 7) / 8) * 8;
    // n_redo was zeroed by the grid clear that precedes every fast pass
    const bool rmw = variant & 1, gtile = variant & 2;
    if (do_g2p) {
        if (rmw && gtile) go<true, true, true>(nwg, s, P, units, gin, gout, bever, d, ph, fa, err, counters);
        else if (rmw) go<true, true, false>(nwg, s, P, units, gin, gout, bever, d, ph, fa, err, counters);
        else if (gtile) go<true, false, true>(nwg, s, P, units, gin, gout, bever, d, ph, fa, err, counters);
        else go<true, false, false>(nwg, s, P, units, gin, gout, bever, d, ph, fa, err, counters);
    } else {
        if (rmw) go<false, true, false>(nwg, s, P, units, gin, gout, bever, d, ph, fa, err, counters);
        else go<false, false, false>(nwg, s, P, units, gin, gout, bever, d, ph, fa, err, counters);
    }
    launch_redo(do_g2p, P, gin, gout, bever, d, ph, fa, err, counters, s);
}

}  // namespace bm
