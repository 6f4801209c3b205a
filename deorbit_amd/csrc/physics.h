// Point physics of the SPH step, templated on the arithmetic type R (float for
// the mixed-precision path, double for the all-FP64 path).
//
// Formulas follow hydrocodes/common/eos.hpp (Tillotson P and its analytic
// partials) and hydrocodes/sph/sph.hpp (cubic-spline M4 kernel, Smith's
// principal-stress formula, von Mises radial return).  In the FP64
// instantiation the expression order matches the reference so per-particle
// values agree with it to the last few ulps.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsph {

constexpr double kPi = 3.14159265358979323846;

// Material constants in the arithmetic type of the instantiation.
template <class R>
struct MatT {
    R rho0, A, B, a, b, alpha, beta, u0, uiv, ucv;
    R G, Y, Emod, gamma;
    R wm, wk;
    int32_t type;      // 0 ideal gas, 1 Tillotson
    int32_t pad;
};

// ---- precision-dispatched math -------------------------------------------
__device__ __forceinline__ float d_exp(float x) { return expf(x); }
__device__ __forceinline__ double d_exp(double x) { return exp(x); }
__device__ __forceinline__ float d_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float d_cbrt(float x) { return cbrtf(x); }
__device__ __forceinline__ double d_cbrt(double x) { return cbrt(x); }
__device__ __forceinline__ float d_acos(float x) { return acosf(x); }
__device__ __forceinline__ double d_acos(double x) { return acos(x); }
__device__ __forceinline__ float d_cos(float x) { return cosf(x); }
__device__ __forceinline__ double d_cos(double x) { return cos(x); }
__device__ __forceinline__ float d_fabs(float x) { return fabsf(x); }
__device__ __forceinline__ double d_fabs(double x) { return fabs(x); }

// max/min with the <algorithm> argument semantics: first argument wins ties.
template <class R> __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }
template <class R> __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }

// Smallest positive density floor, representable in both types (sph.hpp:195).
template <class R> __device__ __forceinline__ R rho_floor();
template <> __device__ __forceinline__ float rho_floor<float>() { return 1e-30f; }
template <> __device__ __forceinline__ double rho_floor<double>() { return 1e-30; }

// ---- Tillotson EOS (eos.hpp:52-113) -------------------------------------
template <class R>
__device__ __forceinline__ R till_p(const MatT<R>& m, R rho, R u) {
    if (rho <= R(0)) return R(0);
    R eta = rho / m.rho0, mu = eta - R(1);
    R w0 = u / (m.u0 * eta * eta) + R(1);
    R pc = (m.a + m.b / w0) * rho * u + m.A * mu + m.B * mu * mu;
    if (rho >= m.rho0 || u <= m.uiv) return pc;
    R z = m.rho0 / rho - R(1);
    R pe = m.a * rho * u + (m.b * rho * u / w0 + m.A * mu * d_exp(-m.beta * z)) * d_exp(-m.alpha * z * z);
    if (u >= m.ucv) return pe;
    return ((u - m.uiv) * pe + (m.ucv - u) * pc) / (m.ucv - m.uiv);
}

template <class R>
__device__ __forceinline__ R eos_pressure(const MatT<R>& m, R rho, R u) {
    if (m.type == 0) return (m.gamma - R(1)) * rho * u;
    return till_p(m, rho, u);
}

template <class R>
__device__ __forceinline__ void till_cond(const MatT<R>& m, R rho, R u, R& P, R& dPdr, R& dPdu) {
    R eta = rho / m.rho0, mu = eta - R(1);
    R w = R(1) + u / (m.u0 * eta * eta), aob = m.a + m.b / w;
    P = aob * rho * u + m.A * mu + m.B * mu * mu;
    dPdu = aob * rho - m.b * rho * u / (w * w * m.u0 * eta * eta);
    R dwdr = R(-2) * u / (m.u0 * eta * eta * eta * m.rho0);
    dPdr = aob * u + rho * u * (-m.b / (w * w)) * dwdr + m.A / m.rho0 + R(2) * m.B * mu / m.rho0;
}

template <class R>
__device__ __forceinline__ void till_expd(const MatT<R>& m, R rho, R u, R& P, R& dPdr, R& dPdu) {
    R eta = rho / m.rho0, mu = eta - R(1);
    R w = R(1) + u / (m.u0 * eta * eta);
    R z = m.rho0 / rho - R(1), dzdr = -m.rho0 / (rho * rho);
    R e1 = d_exp(-m.alpha * z * z), e2 = d_exp(-m.beta * z);
    R g = m.b * rho * u / w + m.A * mu * e2;
    P = m.a * rho * u + g * e1;
    R dwdr = R(-2) * u / (m.u0 * eta * eta * eta * m.rho0);
    R dgu = m.b * rho / w - m.b * rho * u / (w * w * m.u0 * eta * eta);
    dPdu = m.a * rho + dgu * e1;
    R dbruw = m.b * u / w + m.b * rho * u * (R(-1) / (w * w)) * dwdr;
    R dgr = dbruw + m.A * e2 / m.rho0 + m.A * mu * (e2 * (-m.beta) * dzdr);
    R de1 = e1 * (-m.alpha * R(2) * z) * dzdr;
    dPdr = m.a * u + dgr * e1 + g * de1;
}

template <class R>
__device__ __forceinline__ R eos_sound_speed(const MatT<R>& m, R rho, R u) {
    if (m.type == 0) {
        R P = (m.gamma - R(1)) * rho * u;
        return d_sqrt(smax(m.gamma * P / smax(rho, rho_floor<R>()), R(1e-12)));
    }
    if (rho <= R(0)) return R(1e-3);
    R P, dPdr, dPdu;
    if (rho >= m.rho0 || u <= m.uiv) {
        till_cond(m, rho, u, P, dPdr, dPdu);
    } else if (u >= m.ucv) {
        till_expd(m, rho, u, P, dPdr, dPdu);
    } else {
        R pc, pcr, pcu, pe, per, peu;
        till_cond(m, rho, u, pc, pcr, pcu);
        till_expd(m, rho, u, pe, per, peu);
        R inv = R(1) / (m.ucv - m.uiv);
        P = ((u - m.uiv) * pe + (m.ucv - u) * pc) * inv;
        dPdu = ((pe + (u - m.uiv) * peu) + (-pc + (m.ucv - u) * pcu)) * inv;
        dPdr = ((u - m.uiv) * per + (m.ucv - u) * pcr) * inv;
    }
    return d_sqrt(smax(dPdr + P / (rho * rho) * dPdu, R(1e-6)));
}

// ---- largest eigenvalue of a symmetric 3x3 (sph.hpp:41-54) --------------
// Components in the order {xx, yy, zz, xy, xz, yz}.
template <class R>
__device__ __forceinline__ R eig_max(R xx, R yy, R zz, R xy, R xz, R yz) {
    R p1 = xy * xy + xz * xz + yz * yz;
    R q = (xx + yy + zz) / R(3);
    if (p1 < R(1e-30)) {
        R m = xx;
        if (m < yy) m = yy;
        if (m < zz) m = zz;
        return m;
    }
    R p2 = (xx - q) * (xx - q) + (yy - q) * (yy - q) + (zz - q) * (zz - q) + R(2) * p1;
    R p = d_sqrt(p2 / R(6));
    R a = (xx - q) / p, b = (yy - q) / p, c = (zz - q) / p;
    R d = xy / p, e = xz / p, f = yz / p;
    R r = (a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e)) / R(2);
    r = smax(R(-1), smin(R(1), r));
    R phi = d_acos(r) / R(3);
    return q + R(2) * p * d_cos(phi);
}

// ---- von Mises radial return (sph.hpp:320-331) ---------------------------
template <class R>
__device__ __forceinline__ void yield_limit(R Y, R s[6]) {
    if (Y <= R(0)) {
        for (int k = 0; k < 6; k++) s[k] = R(0);
        return;
    }
    R J2 = R(0.5) * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    R vm = d_sqrt(R(3) * J2);
    if (vm > Y) {
        R f = Y / vm;
        for (int k = 0; k < 6; k++) s[k] *= f;
    }
}

// Cubic-spline M4 shape without normalisation: f(q) with W = f(q)/(pi h^3).
template <class R>
__host__ __device__ __forceinline__ R kern_shape(R q) {
    if (q < R(1)) return R(1) - R(1.5) * q * q + R(0.75) * q * q * q;
    if (q < R(2)) {
        R t = R(2) - q;
        return R(0.25) * t * t * t;
    }
    return R(0);
}

}  // namespace dsph
