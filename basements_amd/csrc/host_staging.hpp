// host_staging.hpp -- moving particle state between the caller's host layout
// (bm_particles: x, v as 3 floats per particle, F, C as 9, m, V0 as 1) and
// the device's component-major arrays, through a pinned staging buffer.
//
// The transposes run on several host threads; the copies are single pinned
// DMA transfers (a pitched 2-D copy covers all 26 components at once).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "mpm_internal.hpp"

namespace bm {

// f(lo, hi) over [0, n) split across up to 16 host threads (serial below
// 64k items).
template <class Fn>
void parallel_ranges(int64_t n, Fn f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t want = std::min<int64_t>(std::min<int64_t>(hw, 16), (n + 65535) / 65536);
    if (want <= 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve((size_t)want - 1);
    const int64_t per = (n + want - 1) / want;
    for (int64_t t = 1; t < want; ++t) {
        const int64_t lo = t * per, hi = std::min(n, lo + per);
        if (lo < hi) pool.emplace_back([=] { f(lo, hi); });
    }
    f((int64_t)0, std::min(n, per));
    for (auto& th : pool) th.join();
}

// A growable pinned host buffer (hipHostMalloc), freed with its owner.
struct PinnedStage {
    void* p = nullptr;
    size_t bytes = 0;
    bool ensure(size_t need) {
        if (need <= bytes) return true;
        const size_t sz = std::max(need, bytes * 3 / 2);
        if (p) hipHostFree(p);
        p = nullptr;
        bytes = 0;
        if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = sz;
        return true;
    }
    void release() {
        if (p) hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~PinnedStage() { release(); }
    float* f() const { return static_cast<float*>(p); }
};

// Caller layout -> component-major st[c * n + i].  Null fields are skipped.
// Returns false when any value is not finite.
inline bool soa_from_particles(const float* x, const float* v, const float* F, const float* Cm, const float* m,
                               const float* V0, int64_t n, float* st) {
    std::atomic<bool> bad{false};
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        float acc = 0.0f;  // NaN and Inf survive the sum (scaled so finite values cannot overflow it)
        for (int64_t i = lo; i < hi; ++i) {
            for (int k = 0; k < 3; ++k) {
                if (x) st[(PX + k) * n + i] = x[3 * i + k];
                if (v) st[(PV + k) * n + i] = v[3 * i + k];
            }
            for (int k = 0; k < 9; ++k) {
                if (F) st[(PF + k) * n + i] = F[9 * i + k];
                if (Cm) st[(PC + k) * n + i] = Cm[9 * i + k];
            }
            if (m) st[PM * n + i] = m[i];
            if (V0) st[PV0 * n + i] = V0[i];
        }
        for (int64_t i = lo; i < hi; ++i)
            for (int c = 0; c < PCOMP; ++c) {
                const bool present = (c < PV ? x : c < PF ? v : c < PC ? F : c < PM ? Cm : c == PM ? m : V0) != nullptr;
                if (present) acc += st[c * n + i] * 0x1p-100f;
            }
        if (!std::isfinite(acc)) bad.store(true, std::memory_order_relaxed);
    });
    return !bad.load();
}

// Component-major st -> caller layout; slot i holds particle order[i] - base
// (order == nullptr: slot i is particle i).  Returns false when an order
// entry is out of range.  Null fields are skipped.
inline bool particles_from_soa(const float* st, int64_t n, const int* order, int64_t base, bm_particles* p) {
    std::atomic<bool> bad{false};
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            const int64_t o = order ? (int64_t)order[i] - base : i;
            if (o < 0 || o >= n) {
                bad.store(true, std::memory_order_relaxed);
                continue;
            }
            for (int k = 0; k < 3; ++k) {
                if (p->x) p->x[3 * o + k] = st[(PX + k) * n + i];
                if (p->v) p->v[3 * o + k] = st[(PV + k) * n + i];
            }
            for (int k = 0; k < 9; ++k) {
                if (p->F) p->F[9 * o + k] = st[(PF + k) * n + i];
                if (p->C) p->C[9 * o + k] = st[(PC + k) * n + i];
            }
            if (p->m) p->m[o] = st[PM * n + i];
            if (p->V0) p->V0[o] = st[PV0 * n + i];
        }
    });
    return !bad.load();
}

}  // namespace bm
