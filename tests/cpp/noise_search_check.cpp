// Exactness check of the device noise-site and channel searches
// (hazard_upper_bound, pick_channel in clifft_amd/csrc/svm_device.h) against
// std::upper_bound and the reference's linear running-sum scan
// (src/clifft/svm/svm.h:229-243, svm_kernels.inl:1894-1909), on random
// hazard tables (uniform, clustered, zero and tiny hazards) and channel
// lists (equiprobable and skewed, with zero-probability channels).
#include <cstdint>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <random>
#include <cmath>
#define CSVM_INLINE static inline
#include "noise_search.inc"  // written by tests/test_noise_search.py from svm_device.h
int main() {
    std::mt19937_64 g(7);
    long bad = 0, tests = 0;
    for (int t = 0; t < 1500; ++t) {
        uint32_t n = 1 + g() % (t % 3 == 0 ? 5 : 300);
        std::vector<double> p(n);
        int mode = t % 5;
        for (auto& x : p) {
            double u = std::uniform_real_distribution<double>(0, 1)(g);
            if (mode == 0) x = 1e-3;
            else if (mode == 1) x = (g() % 4 == 0) ? 0.0 : u * 1e-2;
            else if (mode == 2) x = (g() % 50 == 0) ? 0.5 : 1e-12;
            else if (mode == 3) x = u;
            else x = (g() % 2) ? 0.0 : 1e-300 * u;
        }
        std::vector<double> hz(n);
        double acc = 0;
        for (uint32_t i = 0; i < n; ++i) { acc += -std::log1p(-std::min(p[i], 0.999)); hz[i] = acc; }
        double H = hz[n - 1];
        uint32_t G = H > 0 ? std::min<uint32_t>(n, 1u << 16) : 0;
        double scale = G ? (double)G / H : 0;
        std::vector<uint32_t> guide(G + 1, n);
        for (uint32_t b = 0; b < G; ++b) guide[b] = std::upper_bound(hz.begin(), hz.end(), (double)b / scale) - hz.begin();
        for (int q = 0; q < 300; ++q) {
            double target;
            int k = q % 4;
            if (k == 0) target = std::uniform_real_distribution<double>(0, H * 1.2 + 1e-300)(g);
            else if (k == 1) target = hz[g() % n];
            else if (k == 2) target = std::nextafter(hz[g() % n], 0.0);
            else target = G ? (double)(g() % (G + 1)) / scale : 0.0;
            if (!(target >= 0)) target = 0;
            uint32_t want = std::upper_bound(hz.begin(), hz.end(), target) - hz.begin();
            uint32_t got = hazard_upper_bound(hz.data(), n, guide.data(), G, scale, target);
            ++tests;
            if (got != want) { if (bad < 5) printf("hz mismatch n=%u t=%g want %u got %u\n", n, target, want, got); ++bad; }
        }
        // channels
        uint32_t nc = 1 + g() % 16;
        std::vector<double> cum(nc);
        double a = 0;
        const bool equi = (t % 2) == 0;
        const double pe = std::uniform_real_distribution<double>(1e-6, 0.1)(g);
        for (uint32_t c = 0; c < nc; ++c) { a += equi ? pe : ((g() % 3 == 0) ? 0.0 : std::uniform_real_distribution<double>(0, 0.1)(g)); cum[c] = a; }
        for (int q = 0; q < 50; ++q) {
            double u = (q == 0) ? std::nextafter(1.0, 0.0) : (q == 1 ? 0.0 : std::uniform_real_distribution<double>(0, 1)(g));
            double r = u * cum[nc - 1];
            uint32_t want = nc;
            for (uint32_t c = 0; c < nc; ++c) if (r < cum[c]) { want = c; break; }
            uint32_t got = pick_channel(cum.data(), 0, nc, r, u);
            ++tests;
            if (got != want) { if (bad < 10) printf("ch mismatch\n"); ++bad; }
        }
    }
    printf("tests %ld bad %ld\n", tests, bad);
    return bad != 0;
}
