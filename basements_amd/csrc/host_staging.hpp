// host_staging.hpp -- moving particle state between the caller's host layout
// (bm_particles: x, v as 3 floats per particle, F, C as 9, m, V0 as 1) and
// the device's component-major arrays, through a pinned staging buffer.
//
// The transposes (host_transpose.hpp) run on several host threads; the
// copies are single pinned DMA transfers (a pitched 2-D copy covers all 26
// components at once).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

#include "host_transpose.hpp"

namespace bm {

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

}  // namespace bm
