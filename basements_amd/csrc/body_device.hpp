// body_device.hpp -- the device body's end of a coupled substep: the
// coupler's partial reactions summed in a fixed order, then the body
// integrated.  Shared by the coupler's own body-end kernel and the fast
// form's FP64 redo kernel, which runs it in its first workgroup so that no
// launch of its own sits between the coupler and the particle pass.  Every
// a * b + c is rounded separately (FMA contraction off), whatever the
// including translation unit's flags: the integration matches the FP32
// TimeIntegrator bit for bit.
#pragma once

#include "mpm_internal.hpp"

namespace bm {

// TimeIntegrator::integrate for one body (integrator.h:19-57): semi-implicit
// Euler, per-body damping, q += 1/2 (0, w) q dt, normalised, forces cleared.
__device__ inline void integrate_body_d(bm_body& b, float dt) {
#pragma clang fp contract(off)
    if (!b.integrate) return;
    float v[3], w[3];
    if (b.integrate == BM_BODY_KINEMATIC) {  // moves at its own velocities; the reaction is reported only
        for (int a = 0; a < 3; ++a) {
            v[a] = b.linear_velocity[a];
            w[a] = b.angular_velocity[a];
            b.position[a] = b.position[a] + v[a] * dt;
        }
    } else {
        for (int a = 0; a < 3; ++a) v[a] = b.linear_velocity[a] + (b.force[a] * b.inv_mass) * dt;
        const float* I = b.inv_inertia;
        const float it[3] = {I[0] * b.torque[0] + I[1] * b.torque[1] + I[2] * b.torque[2],
                             I[3] * b.torque[0] + I[4] * b.torque[1] + I[5] * b.torque[2],
                             I[6] * b.torque[0] + I[7] * b.torque[1] + I[8] * b.torque[2]};
        for (int a = 0; a < 3; ++a) w[a] = b.angular_velocity[a] + it[a] * dt;
        const float ld = 1.0f - b.linear_damping * dt, ad = 1.0f - b.angular_damping * dt;
        for (int a = 0; a < 3; ++a) {
            v[a] = v[a] * ld;
            w[a] = w[a] * ad;
            b.position[a] = b.position[a] + v[a] * dt;
            b.linear_velocity[a] = v[a];
            b.angular_velocity[a] = w[a];
        }
    }
    // Quaternion product (0, w) * q, quaternion.h
    const float ow = b.orientation[0], ox = b.orientation[1], oy = b.orientation[2], oz = b.orientation[3];
    const float pw = 0.0f * ow - w[0] * ox - w[1] * oy - w[2] * oz;
    const float px = 0.0f * ox + w[0] * ow + w[1] * oz - w[2] * oy;
    const float py = 0.0f * oy - w[0] * oz + w[1] * ow + w[2] * ox;
    const float pz = 0.0f * oz + w[0] * oy - w[1] * ox + w[2] * ow;
    const float hs = 0.5f * dt;
    float q[4] = {ow + pw * hs, ox + px * hs, oy + py * hs, oz + pz * hs};
    const float n = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (n < 1e-6f) {
        q[0] = 1.0f;
        q[1] = q[2] = q[3] = 0.0f;
    } else {
        const float inv = 1.0f / n;
        for (int a = 0; a < 4; ++a) q[a] = q[a] * inv;
    }
    for (int a = 0; a < 4; ++a) b.orientation[a] = q[a];
    for (int a = 0; a < 3; ++a) {
        b.force[a] = 0.0f;
        b.torque[a] = 0.0f;
    }
}

// End of the body's coupled substep (bench_compare_all.cpp:659-663): the
// partials are summed as by k_reduce_partials (the first 64 lanes; every lane
// of the workgroup reaches the barriers), the reaction (rounded to FP32, as
// CouplingReaction holds it) is added to the accumulators, then the body is
// integrated.
__device__ inline void body_end(bm_body* body, const double* partials, int nparts, double* out7, float dt) {
#pragma clang fp contract(off)
    __shared__ double red[7][64];
    const int t = threadIdx.x;
    if (t < 64) {
        double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = t; i < nparts; i += 64)
#pragma unroll
            for (int c = 0; c < 7; ++c) s[c] += partials[(int64_t)i * 7 + c];
#pragma unroll
        for (int c = 0; c < 7; ++c) red[c][t] = s[c];
    }
    __syncthreads();
    for (int w = 32; w >= 1; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int c = 0; c < 7; ++c) red[c][t] += red[c][t + w];
        __syncthreads();
    }
    if (t != 0) return;
    double r[7];
    for (int c = 0; c < 7; ++c) r[c] = red[c][0];
    for (int c = 0; c < 7; ++c) out7[c] = r[c];
    bm_body b = *body;
    for (int a = 0; a < 3; ++a) {
        b.force[a] = b.force[a] + (float)r[a];
        b.torque[a] = b.torque[a] + (float)r[3 + a];
    }
    integrate_body_d(b, dt);
    *body = b;
}

}  // namespace bm
