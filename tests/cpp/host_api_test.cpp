// host_api_test.cpp -- the reference's MPM unit tests, restated against the
// C++ host mirror (basements_amd/host/mpm_solver.hpp) so they read as the
// reference's own: tests/test_mpm_constitutive.cpp:58-159,
// tests/test_mpm_stress_isolation.cpp:24-118, tests/test_mpm_cpu.cpp:8-64,
// tests/test_mpm_autodiff_tape_neutral.cpp:78-130.
// Needs a GPU.  Exit status 0 when every check passes; one line per test.
#include <mpm_solver.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

using namespace basements_amd::math;
using namespace basements_amd::mpm;

namespace {

int g_fail = 0;
std::string g_test;

#define CHECK(cond)                                                                    \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::printf("  FAIL %s: %s (line %d)\n", g_test.c_str(), #cond, __LINE__); \
            ++g_fail;                                                                  \
        }                                                                              \
    } while (0)

float max_grid_speed(const SPGridView& g) {
    float top = 0.0f;
    for (const auto& pair : g.get_blocks())
        for (const auto& n : pair.second.nodes)
            if (n.active) top = std::max(top, n.velocity.length());
    return top;
}

// singular values of F from the eigenvalues of F^T F (cyclic Jacobi, FP64)
void singular_values(const Matrix3& F, double s[3]) {
    double a[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            a[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) a[i][j] += (double)F.m[k][i] * F.m[k][j];
        }
    for (int sweep = 0; sweep < 50; ++sweep)
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (std::fabs(a[p][q]) < 1e-300) continue;
                const double th = 0.5 * std::atan2(2 * a[p][q], a[q][q] - a[p][p]);
                const double c = std::cos(th), sn = std::sin(th);
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - sn * akq;
                    a[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - sn * aqk;
                    a[q][k] = sn * apk + c * aqk;
                }
            }
    for (int i = 0; i < 3; ++i) s[i] = std::sqrt(std::max(0.0, a[i][i]));
}

void ZeroStrainProducesZeroGridForce() {
    MPMSolver solver;
    solver.initialize(0.1f);
    solver.set_material(1.0e5f, 0.3f, 1500.0f);
    solver.set_gravity(Vec3::zero());
    solver.add_particle(Vec3(0.20f, 0.20f, 0.20f), Vec3::zero(), 0.01f);
    solver.step(1.0f / 240.0f);
    CHECK(max_grid_speed(solver.get_grid()) < 1e-6f);
}

void SmallStrainProducesLinearResponse() {
    auto peak = [](float eps) {
        MPMSolver solver;
        solver.initialize(0.1f);
        solver.set_material(1.0e5f, 0.3f, 1500.0f);
        solver.set_gravity(Vec3::zero());
        solver.add_particle(Vec3(0.20f, 0.20f, 0.20f), Vec3::zero(), 0.01f);
        auto& particles = solver.get_particles_mutable();
        particles[0].F = Matrix3::scale(Vec3(1.0f, 1.0f + eps, 1.0f));
        solver.step(1.0f / 240.0f);
        return max_grid_speed(solver.get_grid());
    };
    const float v_small = peak(0.005f), v_large = peak(0.010f);
    CHECK(v_small > 1e-8f);
    CHECK(std::fabs(v_large / v_small - 2.0f) <= 2.0f * 0.07f);
}

void DruckerPragerProjectionRespectsCone() {
    MPMSolver solver;
    solver.initialize(0.1f);
    solver.set_material(1.0e5f, 0.3f, 1500.0f);
    solver.set_plastic_alpha(0.6f);
    solver.set_gravity(Vec3::zero());
    solver.add_particle(Vec3(0.20f, 0.20f, 0.20f), Vec3::zero(), 0.01f);
    auto& particles = solver.get_particles_mutable();
    particles[0].F = Matrix3::scale(Vec3(1.5f, 0.5f, 1.5f));
    solver.step(1.0f / 240.0f);
    double s[3];
    singular_values(particles[0].F, s);
    double eps[3], tr = 0.0;
    for (int i = 0; i < 3; ++i) {
        eps[i] = std::log(std::max(s[i], 1e-6));
        tr += eps[i];
    }
    double hat = 0.0;
    for (int i = 0; i < 3; ++i) hat += (eps[i] - tr / 3.0) * (eps[i] - tr / 3.0);
    hat = std::sqrt(hat);
    if (tr > 0.0) {
        for (int i = 0; i < 3; ++i) CHECK(std::fabs(eps[i]) < 1e-4);
    } else {
        const double E = 1.0e5, nu = 0.3, mu = E / (2 * (1 + nu)), lam = E * nu / ((1 + nu) * (1 - 2 * nu));
        const double kohesion = (3 * lam + 2 * mu) / (2 * mu);
        CHECK(hat <= -0.6 * kohesion * tr * 1.05 + 1e-4);
    }
}

void CompressedColumnExpandsVertically() {
    MPMSolver solver;
    solver.initialize(0.1f);
    solver.set_material(1.0e5f, 0.3f, 1500.0f);
    solver.set_gravity(Vec3(0, 0, 0));
    solver.add_particle(Vec3(0.20f, 0.18f, 0.20f), Vec3::zero(), 0.01f);
    solver.add_particle(Vec3(0.20f, 0.22f, 0.20f), Vec3::zero(), 0.01f);
    auto& particles = solver.get_particles_mutable();
    particles[0].F = Matrix3::scale(Vec3(1.0f, 0.9f, 1.0f));
    particles[1].F = Matrix3::scale(Vec3(1.0f, 0.9f, 1.0f));
    solver.step(1.0f / 240.0f);
    CHECK(particles[0].velocity.y < -1e-4f);
    CHECK(particles[1].velocity.y > +1e-4f);
    for (int i = 0; i < 2; ++i) {
        CHECK(std::fabs(particles[i].velocity.x) <= 1e-4f);
        CHECK(std::fabs(particles[i].velocity.z) <= 1e-4f);
    }
}

void HydrostaticCompressionIsIsotropic() {
    MPMSolver solver;
    solver.initialize(0.1f);
    solver.set_material(1.0e5f, 0.3f, 1500.0f);
    solver.set_gravity(Vec3(0, 0, 0));
    const Vec3 c(0.30f, 0.30f, 0.30f);
    const float o = 0.02f;
    const Vec3 offs[6] = {Vec3(-o, 0, 0), Vec3(o, 0, 0), Vec3(0, -o, 0), Vec3(0, o, 0), Vec3(0, 0, -o), Vec3(0, 0, o)};
    for (const Vec3& d : offs) solver.add_particle(c + d, Vec3::zero(), 0.01f);
    auto& particles = solver.get_particles_mutable();
    for (auto& p : particles) p.F = Matrix3::scale(Vec3(0.95f, 0.95f, 0.95f));
    solver.step(1.0f / 240.0f);
    const float vx0 = particles[0].velocity.x, vx1 = particles[1].velocity.x;
    const float vy0 = particles[2].velocity.y, vy1 = particles[3].velocity.y;
    const float vz0 = particles[4].velocity.z, vz1 = particles[5].velocity.z;
    CHECK(vx0 < -1e-4f && vx1 > 1e-4f && vy0 < -1e-4f && vy1 > 1e-4f && vz0 < -1e-4f && vz1 > 1e-4f);
    const float mx = 0.5f * (std::fabs(vx0) + std::fabs(vx1)), my = 0.5f * (std::fabs(vy0) + std::fabs(vy1)),
                mz = 0.5f * (std::fabs(vz0) + std::fabs(vz1));
    const float avg = (mx + my + mz) / 3.0f;
    CHECK(std::fabs(mx - avg) <= 0.2f * avg && std::fabs(my - avg) <= 0.2f * avg && std::fabs(mz - avg) <= 0.2f * avg);
}

void SandCollapseInit() {
    MPMSolver solver;
    solver.initialize(0.1f);
    int n = 0;
    for (float x = 0.5f; x < 1.0f; x += 0.05f)
        for (float y = 0.5f; y < 1.0f; y += 0.05f)
            for (float z = 0.5f; z < 1.0f; z += 0.05f) {
                solver.add_particle(Vec3(x, y, z), Vec3(0, 0, 0));
                ++n;
            }
    CHECK(n > 100);
    for (int i = 0; i < 10; ++i) solver.step(0.01f);
    const auto& particles = solver.get_particles();
    float avg_y = 0.0f;
    for (const auto& p : particles) avg_y += p.position.y;
    avg_y /= (float)particles.size();
    CHECK(avg_y < 0.74f);
}

void BoundaryCollision() {
    MPMSolver solver;
    solver.initialize(0.1f);
    solver.add_particle(Vec3(0.5f, 0.05f, 0.5f), Vec3(0, -10.0f, 0));
    for (int i = 0; i < 10; ++i) solver.step(0.01f);
    CHECK(solver.get_particles()[0].position.y >= 0.0f);
}

// Host-mirror behaviour the reference's tests rely on: total node mass equals
// the particles' mass, the grid snapshot has the reference's block shape, and
// a density change rescales V0 (mpm_solver.h:42-51).
void MirrorBookkeeping() {
    MPMSolver solver;
    solver.initialize(0.05f);
    solver.set_gravity(Vec3::zero());
    double m_total = 0.0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            solver.add_particle(Vec3(0.3f + 0.02f * i, 0.3f + 0.02f * j, 0.31f), Vec3::zero(), 0.01f + 0.001f * i);
            m_total += 0.01f + 0.001f * i;
        }
    solver.set_material(1.0e5f, 0.3f, 3000.0f);  // rho 1500 -> 3000 halves V0
    CHECK(std::fabs(solver.get_particles()[0].volume - 0.01f / 1500.0f * 0.5f) < 1e-12f);
    solver.p2g(1e-4f);
    const SPGridView g = solver.get_grid();
    double m_grid = 0.0;
    int blocks = 0;
    for (const auto& pair : g.get_blocks()) {
        ++blocks;
        const GridBlock& b = pair.second;
        CHECK(pair.first == SPGridView::get_block_key(b.origin_x / 4, b.origin_y / 4, b.origin_z / 4));
        for (const auto& n : b.nodes) m_grid += n.mass;
    }
    CHECK(blocks > 0);
    CHECK(std::fabs(m_grid - m_total) <= 1e-6 * m_total);
    solver.grid_update(1e-4f);
    solver.g2p(1e-4f);
    CHECK(solver.get_particles().size() == 25);
}

// tests/test_mpm_autodiff_tape_neutral.cpp:78-130: step(dt), step(dt, nullptr)
// and step(dt, &tape) give the same particles, and the tape stays empty.
void TapeNeutral() {
    auto build = [](MPMSolver& s) {
        s.initialize(0.1f);
        s.set_material(1.0e5f, 0.3f, 1500.0f);
        s.set_plastic_alpha(0.4f);
        s.set_gravity(Vec3(0, -9.81f, 0));
        s.add_particle(Vec3(0.20f, 0.50f, 0.20f), Vec3(0, 0, 0), 0.01f);
        s.add_particle(Vec3(0.30f, 0.55f, 0.20f), Vec3(0, 0, 0), 0.01f);
        s.add_particle(Vec3(0.20f, 0.50f, 0.30f), Vec3(0, 0, 0), 0.01f);
        s.add_particle(Vec3(0.30f, 0.55f, 0.30f), Vec3(0.01f, 0, 0), 0.01f);
    };
    auto same = [](const std::vector<Particle>& a, const std::vector<Particle>& b) {
        if (a.size() != b.size()) return false;
        auto eq = [](float x, float y) { return std::memcmp(&x, &y, sizeof(float)) == 0; };
        for (size_t i = 0; i < a.size(); ++i) {
            if (!eq(a[i].position.x, b[i].position.x) || !eq(a[i].position.y, b[i].position.y) ||
                !eq(a[i].position.z, b[i].position.z) || !eq(a[i].velocity.x, b[i].velocity.x) ||
                !eq(a[i].velocity.y, b[i].velocity.y) || !eq(a[i].velocity.z, b[i].velocity.z))
                return false;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    if (!eq(a[i].F.m[r][c], b[i].F.m[r][c]) || !eq(a[i].C.m[r][c], b[i].C.m[r][c])) return false;
        }
        return true;
    };
    MPMSolver a, b, c;
    build(a);
    build(b);
    build(c);
    AutodiffTape tape;
    const float dt = 1.0f / 480.0f;
    bool ok = true;
    for (int i = 0; i < 50 && ok; ++i) {
        a.step(dt);
        b.step(dt, nullptr);
        c.step(dt, &tape);
        ok = same(a.get_particles(), b.get_particles()) && same(a.get_particles(), c.get_particles());
    }
    CHECK(ok);
    CHECK(tape.size() == 0);
}

}  // namespace

int main() {
    const std::vector<std::pair<const char*, std::function<void()>>> tests = {
        {"ZeroStrainProducesZeroGridForce", ZeroStrainProducesZeroGridForce},
        {"SmallStrainProducesLinearResponse", SmallStrainProducesLinearResponse},
        {"DruckerPragerProjectionRespectsCone", DruckerPragerProjectionRespectsCone},
        {"CompressedColumnExpandsVertically", CompressedColumnExpandsVertically},
        {"HydrostaticCompressionIsIsotropic", HydrostaticCompressionIsIsotropic},
        {"SandCollapseInit", SandCollapseInit},
        {"BoundaryCollision", BoundaryCollision},
        {"MirrorBookkeeping", MirrorBookkeeping},
        {"TapeNeutral", TapeNeutral},
    };
    for (const auto& t : tests) {
        g_test = t.first;
        const int before = g_fail;
        try {
            t.second();
        } catch (const std::exception& e) {
            std::printf("  FAIL %s: exception %s\n", t.first, e.what());
            ++g_fail;
        }
        std::printf("%s %s\n", g_fail == before ? "PASS" : "FAIL", t.first);
        std::fflush(stdout);
    }
    return g_fail ? 1 : 0;
}
