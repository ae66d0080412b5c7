// particle_layout.hpp -- component indices of the device particle state
// (component-major SoA: component c of particle i at p[c * capacity + i]),
// shared by the kernels and the host-side transposes.
#pragma once

namespace bm {

// x 0-2, v 3-5, F 6-14 (row-major), C 15-23 (row-major), m 24, V0 25.
enum { PX = 0, PV = 3, PF = 6, PC = 15, PM = 24, PV0 = 25, PCOMP = 26 };

}  // namespace bm
