// host_transpose.hpp -- the host-side layout changes of particle state
// between the caller's layout (bm_particles: x, v as 3 floats per particle,
// F, C as 9, m, V0 as 1) and component-major arrays st[c * n + i], on
// several host threads, with the finiteness test folded into the upload
// pass.  Plain C++ (no HIP): tests/cpp/staging_sanitize.cpp drives it under
// AddressSanitizer, UndefinedBehaviorSanitizer and ThreadSanitizer.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/bm_mpm.h"
#include "particle_layout.hpp"

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

// Caller layout -> component-major st[c * n + i].  Null fields are skipped.
// Returns false when any value is not finite.  Blocks of 1024 particles, one
// component row at a time (each output row is one sequential stream); the
// finiteness test is an exponent check folded into the same pass.
inline bool soa_from_particles(const float* x, const float* v, const float* F, const float* Cm, const float* m,
                               const float* V0, int64_t n, float* st) {
    std::atomic<bool> bad{false};
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        uint32_t nonfinite = 0;
        auto rows = [&](const float* src, int c0, int w, int64_t b0, int64_t b1) {
            if (!src) return;
            for (int k = 0; k < w; ++k) {
                float* o = st + (int64_t)(c0 + k) * n;
                for (int64_t i = b0; i < b1; ++i) {
                    const float a = src[(int64_t)w * i + k];
                    uint32_t u;
                    std::memcpy(&u, &a, sizeof u);
                    nonfinite |= (uint32_t)((u & 0x7f800000u) == 0x7f800000u);
                    o[i] = a;
                }
            }
        };
        for (int64_t b0 = lo; b0 < hi; b0 += 1024) {
            const int64_t b1 = std::min(hi, b0 + 1024);
            rows(x, PX, 3, b0, b1);
            rows(v, PV, 3, b0, b1);
            rows(F, PF, 9, b0, b1);
            rows(Cm, PC, 9, b0, b1);
            rows(m, PM, 1, b0, b1);
            rows(V0, PV0, 1, b0, b1);
        }
        if (nonfinite) bad.store(true, std::memory_order_relaxed);
    });
    return !bad.load();
}

// Component-major st -> caller layout; slot i holds particle order[i] - base
// (order == nullptr: slot i is particle i).  Returns false when an order
// entry is out of range.  Null fields are skipped.
inline bool particles_from_soa(const float* st, int64_t n, const int* order, int64_t base, bm_particles* p) {
    if (!order) {  // insertion order already: sequential streams both ways
        parallel_ranges(n, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                for (int k = 0; k < 3; ++k) {
                    if (p->x) p->x[3 * i + k] = st[(PX + k) * n + i];
                    if (p->v) p->v[3 * i + k] = st[(PV + k) * n + i];
                }
                for (int k = 0; k < 9; ++k) {
                    if (p->F) p->F[9 * i + k] = st[(PF + k) * n + i];
                    if (p->C) p->C[9 * i + k] = st[(PC + k) * n + i];
                }
                if (p->m) p->m[i] = st[PM * n + i];
                if (p->V0) p->V0[i] = st[PV0 * n + i];
            }
        });
        return true;
    }
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

// True when every order entry lies in [base, base + n).
inline bool order_in_range(const int* order, int64_t n, int64_t base) {
    std::atomic<bool> bad{false};
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        bool b = false;
        for (int64_t i = lo; i < hi; ++i) b |= (int64_t)order[i] < base || (int64_t)order[i] >= base + n;
        if (b) bad.store(true, std::memory_order_relaxed);
    });
    return !bad.load();
}

}  // namespace bm
