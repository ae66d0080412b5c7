// ref_driver.cpp -- C entry points around the reference's own MPM headers.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference sources where they lie (include/basements/physics/mpm/*.h,
// physics/coupling/*.h, src/physics/coupling/mpm_rigid_coupler.cpp,
// physics/dynamics/integrator.h) into oracle/_ref/libref_mpm.so.  Nothing of
// the reference is copied: the driver only calls it.  It is used to pin the
// restatement in oracle/mpm_oracle.cpp and as the reference CPU baseline.
//
// MPMSolver keeps its phases (p2g_single, update_grid, g2p_single) private
// (mpm_solver.h:97-349); the reference exposes no split-phase API, so the
// access specifier is widened for this translation unit only.  Every standard
// header the reference pulls in is included first so the widening reaches
// only the reference's own classes.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <immintrin.h>
#include <string>
#include <unordered_map>
#include <vector>
#include <array>
#include <memory>
#include <functional>

#define private public
#include "basements/physics/mpm/mpm_solver.h"
#undef private
#include "basements/physics/coupling/mpm_rigid_coupler.h"
#include "basements/physics/coupling/mpm_rigid_coupler_pic.h"
#include "basements/physics/dynamics/integrator.h"

using basements::math::Matrix3;
using basements::math::Quaternion;
using basements::math::Vec3;
using basements::mpm::MPMSolver;
namespace cpl = basements::physics::coupling;

namespace {
Matrix3 load_m3(const double* a) {
    Matrix3 r;
    for (int i = 0; i < 9; ++i) r.m[i / 3][i % 3] = (float)a[i];
    return r;
}
void store_m3(const Matrix3& r, double* a) {
    for (int i = 0; i < 9; ++i) a[i] = r.m[i / 3][i % 3];
}
}  // namespace

extern "C" {

void* ref_create(double dx) {
    auto* s = new MPMSolver();
    s->initialize((float)dx);
    return s;
}
void ref_destroy(void* h) { delete (MPMSolver*)h; }

// prm: E, nu, rho, alpha, mu_floor, gx, gy, gz (same packing as the oracle)
void ref_params(void* h, const double* prm) {
    auto* s = (MPMSolver*)h;
    s->set_material((float)prm[0], (float)prm[1], (float)prm[2]);
    s->set_plastic_alpha((float)prm[3]);
    s->set_floor_friction((float)prm[4]);
    s->set_gravity(Vec3((float)prm[5], (float)prm[6], (float)prm[7]));
}

void ref_add_particle(void* h, const double* x, const double* v, double m) {
    ((MPMSolver*)h)->add_particle(Vec3((float)x[0], (float)x[1], (float)x[2]), Vec3((float)v[0], (float)v[1], (float)v[2]),
                                  (float)m);
}

// Overwrite the full state of every particle (external mutation between steps, mpm_solver.h:56-59).
void ref_set_particles(void* h, int64_t n, const double* x, const double* v, const double* F, const double* C,
                       const double* m, const double* V0) {
    auto& P = ((MPMSolver*)h)->get_particles_mutable();
    P.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        auto& p = P[(size_t)i];
        p.position = Vec3((float)x[3 * i], (float)x[3 * i + 1], (float)x[3 * i + 2]);
        p.velocity = Vec3((float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2]);
        p.F = load_m3(F + 9 * i);
        p.C = load_m3(C + 9 * i);
        p.mass = (float)m[i];
        p.volume = (float)V0[i];
    }
}

int64_t ref_num_particles(void* h) { return (int64_t)((MPMSolver*)h)->get_particles().size(); }

void ref_get_particles(void* h, double* x, double* v, double* F, double* C, double* m, double* V0) {
    const auto& P = ((MPMSolver*)h)->get_particles();
    for (size_t i = 0; i < P.size(); ++i) {
        const auto& p = P[i];
        x[3 * i] = p.position.x; x[3 * i + 1] = p.position.y; x[3 * i + 2] = p.position.z;
        v[3 * i] = p.velocity.x; v[3 * i + 1] = p.velocity.y; v[3 * i + 2] = p.velocity.z;
        store_m3(p.F, F + 9 * i);
        store_m3(p.C, C + 9 * i);
        m[i] = p.mass;
        V0[i] = p.volume;
    }
}

void ref_step(void* h, double dt, int n) {
    for (int i = 0; i < n; ++i) ((MPMSolver*)h)->step((float)dt);
}
// The three phases of MPMSolver::step (mpm_solver.h:71-87), callable one by one.
void ref_p2g(void* h, double dt) {
    auto* s = (MPMSolver*)h;
    s->grid.clear();
    for (auto& p : s->particles) s->p2g_single(p, (float)dt);
}
void ref_update(void* h, double dt) { ((MPMSolver*)h)->update_grid((float)dt); }
void ref_g2p(void* h, double dt) {
    auto* s = (MPMSolver*)h;
    for (auto& p : s->particles) s->g2p_single(p, (float)dt);
}

// col: c(3) q(w,x,y,z)(4) he(3) vl(3) wa(3) friction mass shape pad ; out: f(3) t(3) n
void ref_couple(void* h, int method, const double* col, double dt, double* out) {
    cpl::RigidColliderState c;
    c.center = Vec3((float)col[0], (float)col[1], (float)col[2]);
    c.orientation = Quaternion((float)col[3], (float)col[4], (float)col[5], (float)col[6]);
    c.half_extents = Vec3((float)col[7], (float)col[8], (float)col[9]);
    c.linear_velocity = Vec3((float)col[10], (float)col[11], (float)col[12]);
    c.angular_velocity = Vec3((float)col[13], (float)col[14], (float)col[15]);
    c.friction = (float)col[16];
    c.mass = (float)col[17];
    c.shape = ((int)col[18] == 1) ? cpl::ShapeKind::CYLINDER : cpl::ShapeKind::BOX;
    auto& g = ((MPMSolver*)h)->get_grid_mutable();
    cpl::CouplingReaction r =
        method == 0 ? cpl::MPMRigidCoupler::apply(g, c, (float)dt) : cpl::MPMRigidCouplerPIC::apply(g, c, (float)dt);
    out[0] = r.force.x; out[1] = r.force.y; out[2] = r.force.z;
    out[3] = r.torque.x; out[4] = r.torque.y; out[5] = r.torque.z;
    out[6] = r.nodes_inside;
}

int64_t ref_num_blocks(void* h) { return (int64_t)((MPMSolver*)h)->get_grid().get_blocks().size(); }

void ref_grid_dense(void* h, const int* org, const int* dims, double* mass, double* vel, double* force,
                    uint8_t* active, uint8_t* alloc) {
    const int64_t nn = (int64_t)dims[0] * dims[1] * dims[2];
    for (int64_t i = 0; i < nn; ++i) {
        mass[i] = 0; active[i] = 0; alloc[i] = 0;
        for (int c = 0; c < 3; ++c) { vel[3 * i + c] = 0; force[3 * i + c] = 0; }
    }
    for (const auto& kv : ((MPMSolver*)h)->get_grid().get_blocks()) {
        const auto& B = kv.second;
        for (int li = 0; li < 4; ++li)
            for (int lj = 0; lj < 4; ++lj)
                for (int lk = 0; lk < 4; ++lk) {
                    const int gi = B.origin_x + li - org[0], gj = B.origin_y + lj - org[1], gk = B.origin_z + lk - org[2];
                    if (gi < 0 || gj < 0 || gk < 0 || gi >= dims[0] || gj >= dims[1] || gk >= dims[2]) continue;
                    const int64_t o = ((int64_t)gi * dims[1] + gj) * dims[2] + gk;
                    const auto& N = B.nodes[(li * 4 + lj) * 4 + lk];
                    mass[o] = N.mass;
                    vel[3 * o] = N.velocity.x; vel[3 * o + 1] = N.velocity.y; vel[3 * o + 2] = N.velocity.z;
                    force[3 * o] = N.force.x; force[3 * o + 1] = N.force.y; force[3 * o + 2] = N.force.z;
                    active[o] = N.active ? 1 : 0;
                    alloc[o] = 1;
                }
    }
}

void ref_grid_set_dense(void* h, const int* org, const int* dims, const double* mass, const double* vel,
                        const uint8_t* active) {
    auto& g = ((MPMSolver*)h)->get_grid_mutable();
    for (int gi = 0; gi < dims[0]; ++gi)
        for (int gj = 0; gj < dims[1]; ++gj)
            for (int gk = 0; gk < dims[2]; ++gk) {
                const int64_t o = ((int64_t)gi * dims[1] + gj) * dims[2] + gk;
                if (!active[o]) continue;
                auto* N = g.get_node(gi + org[0], gj + org[1], gk + org[2]);
                N->mass = (float)mass[o];
                N->velocity = Vec3((float)vel[3 * o], (float)vel[3 * o + 1], (float)vel[3 * o + 2]);
                N->active = true;
            }
}

void ref_svd(const double* F, double* U, double* S, double* V) {
    Matrix3 u, v;
    Vec3 s;
    basements::math::SVD::compute(load_m3(F), u, s, v);
    store_m3(u, U);
    store_m3(v, V);
    S[0] = s.x; S[1] = s.y; S[2] = s.z;
}

// bd: x(3) q(4) v(3) w(3) inv_mass inv_I(9) lin_damp ang_damp f(3) t(3) = 31 doubles, in/out
void ref_integrate_body(double* bd, double dt) {
    basements::physics::RigidBody b;
    b.type = basements::physics::BodyType::DYNAMIC;
    b.position = Vec3((float)bd[0], (float)bd[1], (float)bd[2]);
    b.orientation = Quaternion((float)bd[3], (float)bd[4], (float)bd[5], (float)bd[6]);
    b.linear_velocity = Vec3((float)bd[7], (float)bd[8], (float)bd[9]);
    b.angular_velocity = Vec3((float)bd[10], (float)bd[11], (float)bd[12]);
    b.inv_mass = (float)bd[13];
    b.inv_inertia_tensor = load_m3(bd + 14);
    b.linear_damping = (float)bd[23];
    b.angular_damping = (float)bd[24];
    b.force = Vec3((float)bd[25], (float)bd[26], (float)bd[27]);
    b.torque = Vec3((float)bd[28], (float)bd[29], (float)bd[30]);
    basements::dynamics::TimeIntegrator::integrate(&b, 1, (float)dt);
    bd[0] = b.position.x; bd[1] = b.position.y; bd[2] = b.position.z;
    bd[3] = b.orientation.w; bd[4] = b.orientation.x; bd[5] = b.orientation.y; bd[6] = b.orientation.z;
    bd[7] = b.linear_velocity.x; bd[8] = b.linear_velocity.y; bd[9] = b.linear_velocity.z;
    bd[10] = b.angular_velocity.x; bd[11] = b.angular_velocity.y; bd[12] = b.angular_velocity.z;
    bd[25] = b.force.x; bd[26] = b.force.y; bd[27] = b.force.z;
    bd[28] = b.torque.x; bd[29] = b.torque.y; bd[30] = b.torque.z;
}

}  // extern "C"
