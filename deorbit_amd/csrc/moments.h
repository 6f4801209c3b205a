#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace dsph {

// Moment sums of a density walk (host and device).
//
// A walk that lets the second iteration of a density round (sph.hpp:189-198)
// skip its own walk.  With q = r/h the M4 shape is 1 - 1.5 q^2 + 0.75 q^3 for
// q < 1 and (2 - q)^3 / 4 for 1 <= q < 2, so for every h1 with
// |h1/h - 1| <= d/2
//   sum_j m_j f(r_j / h1) = A0 - 1.5 A2 / h1^2 + 0.75 A3 / h1^3
//                           + (B3 + 3 D B2 + 3 D^2 B1 + D^3 B0) / (4 h1^3)
//                           + sum over the band candidates of m_j f(r_j / h1)
// with An = sum m r^n over q < 1 - d, Bn = sum m u^n with u = 2h - r over
// 1 + d <= q < 2 - 2d, and D = 2 (h1 - h): these candidates stay on the same
// side of h1 as of h and within 2 h1 (u + D = 2 h1 - r > 0).  The band
// candidates (1 - d <= q < 1 + d and 2 - 2d <= q < 2 + 2d) are kept by index
// and summed directly at h1 with the reference's own test r < 2 h1; about
// 3 N (2d) (2^3 + 1) / 2^3 of a particle's N neighbours fall in them.  The B
// sum has no cancellation (|D| < u); the A sum cancels at most fourfold, as
// the shape itself does at q = 1.  A particle whose h1 leaves the band, whose
// bands held more than kDensBand candidates, or whose 2 h1 may exceed the
// reference's cell width (emulation) walks again.
constexpr int kDensBand = 16;
constexpr double kDensBandRel = 1e-2;

struct DensMom {
    double a0, a2, a3, b0, b1, b2, b3;
    double sb;      // the band candidates' terms at the walk's h
    int nb;         // band candidates met (the first kDensBand are kept)
    __host__ __device__ __forceinline__ void clear() {
        a0 = a2 = a3 = b0 = b1 = b2 = b3 = sb = 0.0;
        nb = 0;
    }
    // The moments' share of a candidate of mass m at distance r (r2 as formed,
    // q = r/h, two_h = 2h); true for a band candidate, which it leaves out.
    __host__ __device__ __forceinline__ bool add(double m, double r, double r2, double q, double two_h) {
        const bool inner = q < 1.0 - kDensBandRel;
        const bool outer = (q >= 1.0 + kDensBandRel) & (q < 2.0 - 2.0 * kDensBandRel);
        const double a = inner ? m : 0.0, ar2 = a * r2;
        a0 += a;
        a2 += ar2;
        a3 = fma(ar2, r, a3);
        const double b = outer ? m : 0.0, u = two_h - r, bu = b * u, bu2 = bu * u;
        b0 += b;
        b1 += bu;
        b2 += bu2;
        b3 = fma(bu2, u, b3);
        return !(inner | outer) & (q < 2.0 + 2.0 * kDensBandRel);
    }
    // sum_j m_j f(r_j / h1) over the moments of a walk at h (band excluded)
    __host__ __device__ __forceinline__ double at(double h, double h1) const {
        const double ih = 1.0 / h1, ih2 = ih * ih, ih3 = ih2 * ih;
        const double D = 2.0 * (h1 - h);
        const double bs = fma(D, fma(D, fma(D, b0, 3.0 * b1), 3.0 * b2), b3);
        return a0 - 1.5 * a2 * ih2 + 0.75 * a3 * ih3 + 0.25 * bs * ih3;
    }
    // may the iteration at h1 use the moments of a walk at h?
    __host__ __device__ __forceinline__ bool covers(double h, double h1) const {
        return nb <= kDensBand && h1 >= h * (1.0 - 0.5 * kDensBandRel) && h1 <= h * (1.0 + 0.5 * kDensBandRel);
    }
};

}  // namespace dsph
