// mpm_internal.hpp -- handle layout and kernel launchers shared by the
// translation units of libbm_mpm.so (host side only sees these declarations).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bm_mpm.h"

namespace bm {

// One node buffer of the dense block-linear grid.  Node (i,j,k) of the domain
// lives in block ((i>>2)*NBY + (j>>2))*NBZ + (k>>2), slot ((i&3)*4+(j&3))*4+(k&3)
// -- the reference's in-block order (spgrid_cpu.h:205-218).
struct GridBuf {
    long long* mq;    // fixed-point node mass
    float4* mv;       // xyz: momentum (after P2G) / velocity (after update); w: node mass after update
    float4* f;        // xyz: stress force (reference mode)
    uint32_t* bflag;  // 1 while the block is in blist
    int32_t* blist;   // blocks touched since the last clear
    int32_t* bcount;  // length of blist
};

struct Domain {
    int ox, oy, oz;      // node index of the lowest corner
    int nx, ny, nz;      // nodes per axis (multiples of 4)
    int nbx, nby, nbz;   // blocks per axis
    float dx, inv_dx;
};

// Particle SoA view: component c of particle i at p[c * cap + i].
// Components: x 0-2, v 3-5, F 6-14, C 15-23, m 24, V0 25.
enum { PX = 0, PV = 3, PF = 6, PC = 15, PM = 24, PV0 = 25, PCOMP = 26 };

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
    return ((int64_t)(lx >> 2) * d.nby + (ly >> 2)) * d.nbz + (lz >> 2);
}
__host__ __device__ inline int64_t node_of(const Domain& d, int lx, int ly, int lz) {
    return block_of(d, lx, ly, lz) * 64 + ((lx & 3) * 4 + (ly & 3)) * 4 + (lz & 3);
}

// Collider in the form the coupler kernel wants: rotation matrices built on
// the host from the quaternion (matrix3.h:432-448), the |R| he world AABB
// (mpm_rigid_coupler.cpp:70-76) and the block range to scan.
struct ColliderDev {
    float c[3];
    float Rl[9];  // local -> world
    float Rw[9];  // world -> local (transpose)
    float he[3];
    float vl[3], wa[3];
    float friction;
    float mb;            // M1 per-node body mass
    float lo[3], hi[3];  // world AABB of the body
    int shape;
    int b0[3];  // first block (domain-local block coordinates) of the scanned range
    int nb[3];  // blocks in the range per axis
};

// ------------------------------------------------------------- launchers
void launch_clear(const GridBuf& g, int nblocks, hipStream_t s);
void launch_p2g_ref(const PView& P, const GridBuf& g, uint8_t* bever, const Domain& d, const Phys& ph,
                    double qscale_inv, int* err, hipStream_t s);
void launch_update(const GridBuf& g, int nblocks, const Domain& d, const Phys& ph, double qscale, float dt,
                   hipStream_t s);
void launch_g2p_ref(const PView& P, const GridBuf& g, const Domain& d, const Phys& ph, float dt, int* err,
                    hipStream_t s);
int couple_partials_needed(const ColliderDev& b);
void launch_couple(int method, const ColliderDev& b, const GridBuf& g, uint8_t* bever, const Domain& d, float dt,
                   double* partials, double* out7, hipStream_t s);
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
int launch_rebin(const PView& P, float* pdst, const int* pid_src, int* pid_dst, const Domain& d, int nblocks,
                 int unit_p, BinScratch& bs, hipStream_t s, int* n_units_host);
void launch_iota(int* p, int64_t n, int64_t first, hipStream_t s);
void launch_fused(bool do_g2p, const PView& P, const int4* units, int n_units, const GridBuf& gin,
                  const GridBuf& gout, uint8_t* bever, const Domain& d, const Phys& ph, float dt_g2p, float dt_fold,
                  double qscale_inv, int* err, unsigned long long* counters, int* redo, int* n_redo, int variant,
                  hipStream_t s);
// Largest unit the fast kernels accept: 512 particles of at most 2^23 - 1
// mass quanta each keep every LDS node-mass sum below 2^32.
constexpr int FUSED_UNIT_MAX = 512;
// Two-kernel form (variant bit 2): per-particle G2P writing 12 deposit
// coefficients per particle into `coef`, then the ranked per-block P2G.
void launch_split(bool do_g2p, const PView& P, const int4* units, int n_units, const GridBuf& gin,
                  const GridBuf& gout, uint8_t* bever, const Domain& d, const Phys& ph, float dt_g2p, float dt_fold,
                  double qscale_inv, int* err, unsigned long long* counters, int* redo, int* n_redo, float* coef,
                  hipStream_t s, void (*between)(void*) = nullptr, void* ctx = nullptr);

}  // namespace bm
