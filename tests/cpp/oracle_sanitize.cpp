// oracle_sanitize.cpp -- drives the oracle (oracle/mpm_oracle.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer: both precisions, both
// mass modes, both couplers and shapes, the SVD on degenerate and random
// matrices, plasticity, the stencil weights and the rigid integrator.  Built
// and run by tests/test_oracle_sanitize.py on the host CPU; exits non-zero
// on any sanitizer report (-fno-sanitize-recover=all).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#define ORC_DECL(SUF)                                                                                              \
    extern "C" void* orc_##SUF##_create(double dx, int mass_mode, int q_exp);                                      \
    extern "C" void orc_##SUF##_destroy(void* h);                                                                  \
    extern "C" void orc_##SUF##_params(void* h, const double* prm);                                                \
    extern "C" void orc_##SUF##_set_particles(void* h, int64_t n, const double* x, const double* v,                \
                                              const double* F, const double* C, const double* m, const double* V0); \
    extern "C" void orc_##SUF##_get_particles(void* h, double* x, double* v, double* F, double* C);               \
    extern "C" void orc_##SUF##_step(void* h, double dt, int nsteps);                                              \
    extern "C" void orc_##SUF##_p2g(void* h);                                                                      \
    extern "C" void orc_##SUF##_update(void* h, double dt);                                                        \
    extern "C" void orc_##SUF##_g2p(void* h, double dt);                                                           \
    extern "C" void orc_##SUF##_couple(void* h, int method, const double* col, double dt, double* out);           \
    extern "C" int64_t orc_##SUF##_grid_mass_q_total(void* h);                                                     \
    extern "C" void orc_##SUF##_svd(const double* F, double* U, double* S, double* V);                             \
    extern "C" void orc_##SUF##_plasticity(void* h, const double* F, const double* C, double dt, double* Fout);    \
    extern "C" void orc_##SUF##_weights(void* h, const double* x, double* fr, int* base, double* w);               \
    extern "C" void orc_##SUF##_integrate_body(double* bd, double dt);
ORC_DECL(f32)
ORC_DECL(f64)

#define ORC_RUN(SUF)                                                                                               \
    static int run_##SUF(int mass_mode) {                                                                          \
        const int n = 1500;                                                                                        \
        std::mt19937_64 rng(1234 + mass_mode);                                                                     \
        std::uniform_real_distribution<double> U01(0.0, 1.0);                                                      \
        std::normal_distribution<double> N01(0.0, 1.0);                                                            \
        std::vector<double> x(3 * n), v(3 * n), F(9 * n), C(9 * n), m(n), V0(n);                                   \
        for (int i = 0; i < n; ++i) {                                                                              \
            for (int k = 0; k < 3; ++k) {                                                                          \
                x[3 * i + k] = 0.2 + 0.3 * U01(rng);                                                               \
                v[3 * i + k] = 0.2 * N01(rng);                                                                     \
            }                                                                                                      \
            for (int k = 0; k < 9; ++k) {                                                                          \
                F[9 * i + k] = (k % 4 == 0 ? 1.0 : 0.0) + 0.05 * N01(rng);                                         \
                C[9 * i + k] = N01(rng);                                                                           \
            }                                                                                                      \
            m[i] = 2e-4;                                                                                           \
            V0[i] = m[i] / 1600.0;                                                                                 \
        }                                                                                                          \
        const double prm[8] = {3.5e5, 0.3, 1600.0, 0.386, 0.4, 0.0, -9.81, 0.0};                                   \
        void* h = orc_##SUF##_create(1.0 / 32.0, mass_mode, -30);                                                  \
        orc_##SUF##_params(h, prm);                                                                                \
        orc_##SUF##_set_particles(h, n, x.data(), v.data(), F.data(), C.data(), m.data(), V0.data());              \
        orc_##SUF##_step(h, 1e-4, 5);                                                                              \
        for (int method = 0; method < 2; ++method)                                                                 \
            for (int shape = 0; shape < 2; ++shape) {                                                              \
                double col[20] = {0.35, 0.35, 0.35, 0.9238795, 0.3826834, 0.0, 0.0, 0.06, 0.05, 0.06,              \
                                  0.1, -0.2, 0.0, 0.5, 0.0, -0.3, 0.5, 2.0, (double)shape, 0.0};                   \
                double out[7];                                                                                     \
                orc_##SUF##_p2g(h);                                                                                \
                orc_##SUF##_update(h, 1e-4);                                                                       \
                orc_##SUF##_couple(h, method, col, 1e-4, out);                                                     \
                orc_##SUF##_g2p(h, 1e-4);                                                                          \
                for (double o : out)                                                                               \
                    if (!std::isfinite(o)) return 1;                                                               \
            }                                                                                                      \
        orc_##SUF##_get_particles(h, x.data(), v.data(), F.data(), C.data());                                      \
        for (double a : x)                                                                                         \
            if (!std::isfinite(a)) return 2;                                                                       \
        if (mass_mode == 1 && orc_##SUF##_grid_mass_q_total(h) <= 0) return 3;                                     \
        /* SVD: identity, repeated singular values, rank deficiency, a reflection, random */                       \
        const double mats[5][9] = {{1, 0, 0, 0, 1, 0, 0, 0, 1},                                                    \
                                   {2, 0, 0, 0, 2, 0, 0, 0, 1},                                                    \
                                   {1, 2, 3, 2, 4, 6, 1, 1, 1},                                                    \
                                   {0, 1, 0, 1, 0, 0, 0, 0, 1},                                                    \
                                   {0.3, -1.2, 0.7, 2.1, 0.05, -0.4, 0.9, 0.8, 1.6}};                              \
        for (const auto& M : mats) {                                                                               \
            double Um[9], S[3], Vm[9], Fo[9], w[9], fr[3];                                                         \
            int base[3];                                                                                           \
            orc_##SUF##_svd(M, Um, S, Vm);                                                                         \
            orc_##SUF##_plasticity(h, M, M, 1e-4, Fo);                                                             \
            const double xp[3] = {M[0] * 0.1 + 0.3, 0.3, 0.3};                                                     \
            orc_##SUF##_weights(h, xp, fr, base, w);                                                               \
            for (double s : S)                                                                                     \
                if (!std::isfinite(s)) return 4;                                                                   \
        }                                                                                                          \
        double bd[31] = {0, 0.5, 0, 1, 0, 0, 0, 0.1, 0, 0, 0, 0.2, 0, 0.5, 1, 0, 0, 0, 1, 0, 0, 0, 1,              \
                         0.01, 0.01, 0, -9.81, 0, 0, 0, 0};                                                        \
        for (int s = 0; s < 10; ++s) orc_##SUF##_integrate_body(bd, 1e-4);                                         \
        orc_##SUF##_destroy(h);                                                                                    \
        return 0;                                                                                                  \
    }
ORC_RUN(f32)
ORC_RUN(f64)

int main() {
    for (int mm = 0; mm < 2; ++mm) {
        if (int rc = run_f32(mm)) {
            std::printf("f32 mass_mode %d failed: %d\n", mm, rc);
            return rc;
        }
        if (int rc = run_f64(mm)) {
            std::printf("f64 mass_mode %d failed: %d\n", mm, rc);
            return rc;
        }
    }
    std::printf("oracle sanitizer run clean\n");
    return 0;
}
