// mpm_fast_kernels.hip -- particle re-binning and the FP64 redo pass of the
// fast substep (BM_MODE_FAST).
//
// Particles are kept ordered by the 4^3-node grid block that holds their
// stencil base floor(x/dx - 1/2); every few substeps they are re-binned by a
// counting sort (count per block, exclusive scan, scatter) and the work units
// (up to unit_p particles of one block) are rebuilt.  The fused substep
// kernel itself lives in mpm_fused.hip.

#include <hipcub/hipcub.hpp>

#include "mpm_fast_device.hpp"

namespace bm {

// --------------------------------------------------------------- binning
__global__ void k_fill_i32(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Block of each particle and its rank inside the block.  Consecutive lanes
// mostly share a block, so the lanes of a wave agree on one counter
// increment per distinct block (the leader adds the group size) instead of
// one contended atomic per particle.
__global__ void __launch_bounds__(256) k_bin(PView P, Domain d, int* bcnt, int* key, int* slot, SceneView sv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool live = i < P.n;
    if (sv.chunks) {  // one scene per workgroup
        const int4 ch = sv.chunks[blockIdx.x];
        i = ch.y + threadIdx.x;
        live = (int)threadIdx.x < ch.z;
        d = sv.sc[ch.x].d;
    }
    int b = -1;
    if (live) {
        const float x = P.p[(PX + 0) * P.cap + i], y = P.p[(PX + 1) * P.cap + i], z = P.p[(PX + 2) * P.cap + i];
        const int lx = (int)floorf(x * d.inv_dx - 0.5f) - d.ox;
        const int ly = (int)floorf(y * d.inv_dx - 0.5f) - d.oy;
        const int lz = (int)floorf(z * d.inv_dx - 0.5f) - d.oz;
        // out-of-domain particles go to the nearest block; the substep
        // kernels detect and report them
        const int bx = clampi(lx >> 2, 0, d.nbx - 1), by = clampi(ly >> 2, 0, d.nby - 1),
                  bz = clampi(lz >> 2, 0, d.nbz - 1);
        b = (int)(d.b0 + ((int64_t)bx * d.nby + by) * d.nbz + bz);
    }
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long pending = __ballot(live);
    int my_slot = 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lb = __shfl(b, leader);
        const unsigned long long grp = __ballot(live && b == lb);
        int base = 0;
        if (lane == leader) base = atomicAdd(&bcnt[lb], __popcll(grp));
        base = __shfl(base, leader);
        if (live && b == lb) my_slot = base + __popcll(grp & below);
        pending &= ~grp;
    }
    if (live) {
        key[i] = b;
        slot[i] = my_slot;
    }
}

// VC = false leaves the velocity and C components behind (their values are
// dead: the next particle pass is a G2P that rewrites them).
template <bool VC>
__global__ void k_reorder(PView src, float* dst, const int* key, const int* slot, const int* bstart,
                          const int* pid_src, int* pid_dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= src.n) return;
    const int64_t j = (int64_t)bstart[key[i]] + slot[i];
#pragma unroll
    for (int c = 0; c < PCOMP; ++c) {
        if (!VC && ((c >= PV && c < PV + 3) || (c >= PC && c < PC + 9))) continue;
        dst[c * src.cap + j] = src.p[c * src.cap + i];
    }
    pid_dst[j] = pid_src[i];
}

// Back to insertion order for a download: slot base + i holds particle
// pid[base + i] (an index in [base, base + n)); every component moves to that
// particle's slot of dst.  Entries out of range are left out (the host checks
// the ids it reads back and reports a corrupted order).
__global__ void k_unpermute(PView src, float* dst, const int* pid, int64_t base, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t j = (int64_t)pid[base + i];
    if (j < base || j >= base + n) return;
#pragma unroll
    for (int c = 0; c < PCOMP; ++c) dst[c * src.cap + j] = src.p[c * src.cap + base + i];
}

__global__ void k_units_count(const int* bcnt, int nblocks, int unit_p, int* ucnt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblocks) ucnt[b] = (bcnt[b] + unit_p - 1) / unit_p;
}

__global__ void k_units_fill(const int* bcnt, const int* bstart, const int* ucnt, const int* ustart, int nblocks,
                             int unit_p, int4* units, int* n_units, const int* block_scene) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const int c = bcnt[b], nu = ucnt[b], u0 = ustart[b], s0 = bstart[b];
    const int scene = block_scene ? block_scene[b] : 0;
    for (int u = 0; u < nu; ++u) {
        const int cnt = min(unit_p, c - u * unit_p);
        units[u0 + u] = make_int4(b, s0 + u * unit_p, cnt, scene);
    }
    if (b == nblocks - 1) *n_units = u0 + nu;
}

__global__ void k_iota(int* p, int64_t n, int64_t first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)(first + i);
}

// ================================================================ launchers
static inline int grid_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

size_t scan_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (int*)nullptr, (int*)nullptr, n);
    return bytes;
}

// Re-bin the particles by grid block and rebuild the work units.  Returns the
// unit count on the host (one stream synchronisation per re-binning).
int launch_rebin(const PView& P, float* pdst, const int* pid_src, int* pid_dst, const Domain& d, int nblocks,
                 int unit_p, BinScratch& bs, hipStream_t s, int* n_units_host, SceneView sv, int n_chunks,
                 bool with_vc) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nblocks, 256)), dim3(256), 0, s, bs.bcnt, nblocks, 0);
    if (sv.chunks) {
        if (n_chunks > 0)
            hipLaunchKernelGGL(k_bin, dim3(n_chunks), dim3(256), 0, s, P, d, bs.bcnt, bs.key, bs.slot, sv);
    } else if (P.n > 0) {
        hipLaunchKernelGGL(k_bin, dim3(grid_for(P.n, 256)), dim3(256), 0, s, P, d, bs.bcnt, bs.key, bs.slot, sv);
    }
    size_t tb = bs.temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(bs.temp, tb, bs.bcnt, bs.bstart, nblocks, s);
    if (P.n > 0)
        hipLaunchKernelGGL(with_vc ? k_reorder<true> : k_reorder<false>, dim3(grid_for(P.n, 256)), dim3(256), 0, s,
                           P, pdst, bs.key, bs.slot, bs.bstart, pid_src, pid_dst);
    hipLaunchKernelGGL(k_units_count, dim3(grid_for(nblocks, 256)), dim3(256), 0, s, bs.bcnt, nblocks, unit_p,
                       bs.ucnt);
    tb = bs.temp_bytes;
    (void)hipcub::DeviceScan::ExclusiveSum(bs.temp, tb, bs.ucnt, bs.ustart, nblocks, s);
    hipLaunchKernelGGL(k_units_fill, dim3(grid_for(nblocks, 256)), dim3(256), 0, s, bs.bcnt, bs.bstart, bs.ucnt,
                       bs.ustart, nblocks, unit_p, bs.units, bs.n_units, sv.block_scene);
    if (!n_units_host) return 0;  // the deposit kernel reads the count on the device
    if (hipMemcpyAsync(n_units_host, bs.n_units, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    return 0;
}

void launch_unpermute(const PView& P, float* dst, const int* pid, int64_t base, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_unpermute, dim3(grid_for(n, 256)), dim3(256), 0, s, P, dst, pid, base, n);
}

void launch_iota(int* p, int64_t n, int64_t first, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, first);
}


}  // namespace bm
