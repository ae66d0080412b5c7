// Drives the host-side particle transposes (basements_amd/csrc/host_transpose.hpp)
// under AddressSanitizer + UndefinedBehaviorSanitizer, and separately under
// ThreadSanitizer: sizes that run serially and on several threads, null
// fields, non-finite values at range boundaries, permuted downloads with a
// base offset, and out-of-range order entries.  Every buffer is sized
// exactly, so any read or write past its end is a sanitizer report.
// Prints "clean" and exits 0 when every check held.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

#include "host_transpose.hpp"

namespace {

int failures = 0;

void check(bool ok, const char* what, long long n) {
    if (!ok) {
        std::printf("FAIL: %s (n=%lld)\n", what, n);
        ++failures;
    }
}

struct Host {
    std::vector<float> x, v, F, C, m, V0;
    explicit Host(int64_t n) : x(3 * n), v(3 * n), F(9 * n), C(9 * n), m(n), V0(n) {}
    bm_particles view() { return bm_particles{x.data(), v.data(), F.data(), C.data(), m.data(), V0.data()}; }
};

void fill(Host& h, std::mt19937& rng) {
    std::uniform_real_distribution<float> u(-4.0f, 4.0f);
    for (auto* a : {&h.x, &h.v, &h.F, &h.C, &h.m, &h.V0})
        for (auto& e : *a) e = u(rng);
}

// st[c * n + i] must equal the caller-layout field for every non-null field.
bool soa_matches(const Host& h, const std::vector<float>& st, int64_t n, bool with_v, bool with_C) {
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            if (st[(bm::PX + k) * n + i] != h.x[3 * i + k]) return false;
            if (with_v && st[(bm::PV + k) * n + i] != h.v[3 * i + k]) return false;
        }
        for (int k = 0; k < 9; ++k) {
            if (st[(bm::PF + k) * n + i] != h.F[9 * i + k]) return false;
            if (with_C && st[(bm::PC + k) * n + i] != h.C[9 * i + k]) return false;
        }
        if (st[bm::PM * n + i] != h.m[i] || st[bm::PV0 * n + i] != h.V0[i]) return false;
    }
    return true;
}

void run_size(int64_t n, std::mt19937& rng) {
    Host h(n);
    fill(h, rng);
    std::vector<float> st((size_t)bm::PCOMP * n, 0.0f);

    // all fields
    check(bm::soa_from_particles(h.x.data(), h.v.data(), h.F.data(), h.C.data(), h.m.data(), h.V0.data(), n,
                                 st.data()),
          "finite upload accepted", n);
    check(soa_matches(h, st, n, true, true), "upload transpose", n);

    // null fields leave their rows untouched
    std::vector<float> st2((size_t)bm::PCOMP * n, 7.5f);
    check(bm::soa_from_particles(h.x.data(), nullptr, h.F.data(), nullptr, h.m.data(), h.V0.data(), n, st2.data()),
          "upload with null fields", n);
    check(soa_matches(h, st2, n, false, false), "upload transpose, null fields", n);
    bool untouched = true;
    for (int c : {(int)bm::PV, bm::PV + 1, bm::PV + 2, (int)bm::PC, bm::PC + 4, bm::PC + 8})
        for (int64_t i = 0; i < n; ++i) untouched &= st2[(size_t)c * n + i] == 7.5f;
    check(untouched, "null-field rows untouched", n);

    // in-order download: exact round trip
    Host back(n);
    bm_particles bv = back.view();
    check(bm::particles_from_soa(st.data(), n, nullptr, 0, &bv), "in-order download", n);
    check(back.x == h.x && back.v == h.v && back.F == h.F && back.C == h.C && back.m == h.m && back.V0 == h.V0,
          "in-order round trip", n);

    // permuted download with a base offset: slot i is particle order[i] - base
    const int64_t base = 1000;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), (int)base);
    std::shuffle(order.begin(), order.end(), rng);
    check(bm::order_in_range(order.data(), n, base), "order in range", n);
    Host perm(n);
    bm_particles pv = perm.view();
    check(bm::particles_from_soa(st.data(), n, order.data(), base, &pv), "permuted download", n);
    bool placed = true;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t o = order[i] - base;
        for (int k = 0; k < 3; ++k) placed &= perm.x[3 * o + k] == st[(bm::PX + k) * n + i];
        for (int k = 0; k < 9; ++k) placed &= perm.C[9 * o + k] == st[(bm::PC + k) * n + i];
        placed &= perm.V0[o] == st[bm::PV0 * n + i];
    }
    check(placed, "permuted placement", n);

    // permuted download with null fields writes only the others
    Host part(n);
    bm_particles partial = part.view();
    partial.v = nullptr;
    partial.F = nullptr;
    check(bm::particles_from_soa(st.data(), n, order.data(), base, &partial), "permuted download, null fields", n);

    if (n == 0) return;

    // out-of-range entries at the ends of the range and of the thread slices
    for (int64_t at : {(int64_t)0, n / 2, n - 1}) {
        for (int bad : {(int)base - 1, (int)(base + n)}) {
            std::vector<int> o2 = order;
            o2[at] = bad;
            check(!bm::order_in_range(o2.data(), n, base), "out-of-range order refused", n);
            Host dst(n);
            bm_particles dv = dst.view();
            check(!bm::particles_from_soa(st.data(), n, o2.data(), base, &dv), "out-of-range download refused", n);
        }
    }

    // a non-finite value anywhere in any field is caught
    const float nonfinite[] = {std::numeric_limits<float>::quiet_NaN(), std::numeric_limits<float>::infinity(),
                               -std::numeric_limits<float>::infinity()};
    int which = 0;
    for (int64_t at : {(int64_t)0, n / 3, n - 1}) {
        for (auto* field : {&h.x, &h.v, &h.F, &h.C, &h.m, &h.V0}) {
            const int64_t w = (int64_t)field->size() / n;
            const int64_t idx = w * at + (w - 1);
            const float keep = (*field)[idx];
            (*field)[idx] = nonfinite[which++ % 3];
            check(!bm::soa_from_particles(h.x.data(), h.v.data(), h.F.data(), h.C.data(), h.m.data(), h.V0.data(), n,
                                          st.data()),
                  "non-finite value refused", n);
            (*field)[idx] = keep;
        }
    }
    // largest finite and subnormal values pass
    h.F[0] = std::numeric_limits<float>::max();
    h.C[0] = std::numeric_limits<float>::denorm_min();
    h.m[n - 1] = -std::numeric_limits<float>::max();
    check(bm::soa_from_particles(h.x.data(), h.v.data(), h.F.data(), h.C.data(), h.m.data(), h.V0.data(), n,
                                 st.data()),
          "extreme finite values accepted", n);
    // a NaN in a null field's source is never read
    h.v[0] = std::numeric_limits<float>::quiet_NaN();
    check(bm::soa_from_particles(h.x.data(), nullptr, h.F.data(), h.C.data(), h.m.data(), h.V0.data(), n, st.data()),
          "null field not read", n);
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    // serial (< 64k), exactly one slice, several threads with ragged slices
    for (int64_t n : {(int64_t)0, (int64_t)1, (int64_t)7, (int64_t)65536, (int64_t)65537, (int64_t)300001})
        run_size(n, rng);
    if (failures) {
        std::printf("%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("clean\n");
    return 0;
}
