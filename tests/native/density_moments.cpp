// Host check of the density moment sums (deorbit_amd/csrc/moments.h) against
// direct sums of the reference's kernel (sph.hpp:19-24, kW) over random
// neighbour sets: the first iteration at h, and second iterations at h1 across
// the band the moments cover, including candidates placed on the band edges.
// Prints one line per case: "case max_rel_err band_count"; exits 1 when a
// relative error exceeds 1e-13.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "moments.h"

using dsph::DensMom;

static double kw_shape(double r, double h) {       // sph.hpp kW without 1/(pi h^3)
    const double q = r / h;
    if (q < 1.0) return 1.0 - 1.5 * q * q + 0.75 * q * q * q;
    if (q < 2.0) { const double t = 2.0 - q; return 0.25 * t * t * t; }
    return 0.0;
}

int main() {
    std::mt19937_64 rng(20261204);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double d = dsph::kDensBandRel;
    double worst = 0.0;
    int overflow_seen = 0, cases = 0;
    for (int trial = 0; trial < 3000; trial++) {
        const double h = 5.0 + 25.0 * U(rng);
        const int n = 40 + (int)(300 * U(rng));
        std::vector<double> r(n), m(n);
        for (int j = 0; j < n; j++) {
            const double x = (2 * U(rng) - 1) * 2.3 * h, y = (2 * U(rng) - 1) * 2.3 * h, z = (2 * U(rng) - 1) * 2.3 * h;
            r[j] = std::sqrt(x * x + y * y + z * z);
            m[j] = 1e5 + 1e7 * U(rng);
        }
        // a few candidates exactly on the band edges and at h, 2h
        const double edges[] = {1.0 - d, 1.0 + d, 2.0 - 2 * d, 2.0 + 2 * d, 1.0, 2.0, 0.0};
        for (int e = 0; e < 7 && e < n; e++) r[e] = edges[e] * h;
        DensMom M;
        M.clear();
        std::vector<int> band;
        for (int j = 0; j < n; j++) {
            const double r2 = r[j] * r[j];
            if (M.add(m[j], r[j], r2, r[j] / h, 2.0 * h)) {
                band.push_back(j);
                M.nb++;
                M.sb += r[j] < 2.0 * h ? m[j] * kw_shape(r[j], h) : 0.0;
            }
        }
        if (M.nb > dsph::kDensBand) overflow_seen++;
        const double probes[] = {h, h * (1 - 0.5 * d), h * (1 + 0.5 * d), h * (1 + (U(rng) - 0.5) * d)};
        for (double h1 : probes) {
            if (!M.covers(h, h1) && M.nb <= dsph::kDensBand) {
                std::printf("covers() refused h1 = h (1 %+.3e)\n", h1 / h - 1);
                return 1;
            }
            double s = (h1 == h) ? M.at(h, h) + M.sb : M.at(h, h1);
            if (h1 != h)
                for (int j : band) s += r[j] < 2.0 * h1 ? m[j] * kw_shape(r[j], h1) : 0.0;
            double ref = 0.0;
            for (int j = 0; j < n; j++) ref += r[j] < 2.0 * h1 ? m[j] * kw_shape(r[j], h1) : 0.0;
            const double e = std::fabs(s - ref) / ref;
            if (e > worst) worst = e;
            cases++;
        }
        // outside the band the moments must not be used
        if (M.covers(h, h * (1 + 0.51 * d)) || M.covers(h, h * (1 - 0.51 * d))) {
            std::printf("covers() accepted h1 outside the band\n");
            return 1;
        }
    }
    std::printf("moments %d cases, max rel err %.3e, band overflow in %d sets\n", cases, worst, overflow_seen);
    return worst < 1e-13 ? 0 : 1;
}
