// Per-particle kernels: damage growth + second half kick + trapezoidal u/S +
// yield (the end of System::step), diagnostics reductions, and the IC-order
// import/export of fields.
#pragma once
#include "kernels_pair.h"

namespace dsph {

// Benz-Asphaug damage growth (sph.hpp:336-348) on every particle, then for
// non-fixed particles the closing half kick, the trapezoidal update of u and S
// with the u floor, and the von Mises return (sph.hpp:364-376).  Gravity and
// the material-0 damping factor follow the gpu_ic finish kernel
// (sph_force.metal:283-303) and are inactive at their defaults.
template <class R>
__global__ __launch_bounds__(256) void k_finish(Carried<R> c, Work<R> w, int do_damage, int do_kick) {
    const int64_t n = w.n;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t inf = c.info[i];
    const int seg = info_seg(inf), mat = info_mat(inf);
    const SegPar& p = w.par[seg];
    const double dt = w.st[seg].dt_cur;
    if (do_damage && p.damage_on) {
        const MatT<R>& m = w.mats[mat];
        const R D = c.D[i];
        if (m.Emod > R(0) && D < R(1)) {
            const R sg = w.sigmax[i];
            const R el = sg / m.Emod;
            if (sg > R(0) && el > c.eps_act[i]) {
                const R cg = R(0.4) * d_sqrt(m.Emod / smax(w.rho[i], rho_floor<R>()));
                const R rs = R(0.5) * (R)p.h_init / (R)p.eta;
                const R d13 = d_cbrt(D) + (cg / rs) * (R)dt;
                c.D[i] = smin(R(1), d13 * d13 * d13);
            }
        }
    }
    if (!do_kick || info_fixed(inf)) return;
    const double half = 0.5 * dt;
    double vx = c.vx[i] + (double)w.ax[i] * half;
    double vy = c.vy[i] + (double)w.ay[i] * half;
    double vz = c.vz[i] + (double)w.az[i] * half;
    if (p.g != 0.0) vz -= p.g * half;
    if (mat == 0 && p.damp != 1.0) { vx *= p.damp; vy *= p.damp; vz *= p.damp; }
    c.vx[i] = vx; c.vy[i] = vy; c.vz[i] = vz;
    double u = c.u[i] + 0.5 * dt * ((double)w.dudt_old[i] + (double)w.dudt[i]);
    if (u < 1e-12) u = 1e-12;
    c.u[i] = u;
    R s[6];
    const R hdt = (R)(0.5 * dt);
#pragma unroll
    for (int k = 0; k < 6; k++)
        s[k] = c.S[(size_t)k * n + i] + hdt * (w.dSdt_old[(size_t)k * n + i] + w.dSdt[(size_t)k * n + i]);
    yield_limit((R)w.mats[mat].Y, s);
#pragma unroll
    for (int k = 0; k < 6; k++) c.S[(size_t)k * n + i] = s[k];
}

// ---------------------------------------------------------------- diagnostics
// Compensated FP64 sums: each partial is a (sum, correction) pair; pairs are
// combined with TwoSum in a fixed order, so results are run-to-run identical.
struct KSum {
    double s, c;
};
__device__ __forceinline__ void ksum_add(KSum& a, double v) {
    const double t = a.s + v;
    const double bp = t - a.s;
    a.c += (a.s - (t - bp)) + (v - bp);
    a.s = t;
}
__device__ __forceinline__ KSum ksum_merge(KSum a, KSum b) {
    KSum r = a;
    ksum_add(r, b.s);
    r.c += b.c;
    return r;
}

constexpr int kDiagSums = 6;      // mass, e_kin, e_int, mom x, y, z
constexpr int kDiagExt = 8;       // vmax, umax, -hmin, hmax, -rho_min, rho_max, dmax, nonfinite
struct DiagPart {
    KSum sum[kDiagSums];
    double ext[kDiagExt];
};

struct DiagChunk {
    int32_t seg;
    int32_t pad;
    int64_t beg, end;
};

template <class R>
__global__ __launch_bounds__(256) void k_diag_partial(Carried<R> c, Work<R> w, const DiagChunk* chunks,
                                                      DiagPart* parts) {
    __shared__ DiagPart sh[256];
    const DiagChunk ch = chunks[blockIdx.x];
    DiagPart acc;
    for (int k = 0; k < kDiagSums; k++) acc.sum[k] = KSum{0.0, 0.0};
    for (int k = 0; k < kDiagExt; k++) acc.ext[k] = -__builtin_huge_val();
    acc.ext[7] = 0.0;
    for (int64_t i = ch.beg + threadIdx.x; i < ch.end; i += blockDim.x) {
        const double m = c.mass[i], vx = c.vx[i], vy = c.vy[i], vz = c.vz[i], u = c.u[i];
        const double v2 = vx * vx + vy * vy + vz * vz;
        const bool fin = isfinite(c.x[i]) && isfinite(c.y[i]) && isfinite(c.z[i]) && isfinite(v2) && isfinite(u);
        if (!fin) { acc.ext[7] += 1.0; continue; }
        ksum_add(acc.sum[0], m);
        ksum_add(acc.sum[1], 0.5 * m * v2);
        ksum_add(acc.sum[2], m * u);
        ksum_add(acc.sum[3], m * vx);
        ksum_add(acc.sum[4], m * vy);
        ksum_add(acc.sum[5], m * vz);
        acc.ext[0] = fmax(acc.ext[0], sqrt(v2));
        acc.ext[1] = fmax(acc.ext[1], u);
        acc.ext[2] = fmax(acc.ext[2], -(double)c.h[i]);
        acc.ext[3] = fmax(acc.ext[3], (double)c.h[i]);
        acc.ext[4] = fmax(acc.ext[4], -(double)w.rho[i]);
        acc.ext[5] = fmax(acc.ext[5], (double)w.rho[i]);
        acc.ext[6] = fmax(acc.ext[6], (double)c.D[i]);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            DiagPart a = sh[threadIdx.x];
            const DiagPart& b = sh[threadIdx.x + off];
            for (int k = 0; k < kDiagSums; k++) a.sum[k] = ksum_merge(a.sum[k], b.sum[k]);
            for (int k = 0; k < 7; k++) a.ext[k] = fmax(a.ext[k], b.ext[k]);
            a.ext[7] += b.ext[7];
            sh[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}

// One thread per segment folds that segment's chunk partials in chunk order.
__global__ void k_diag_final(const DiagPart* parts, const int32_t* seg_chunk0, int nseg, DiagPart* out) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
        DiagPart acc;
        for (int k = 0; k < kDiagSums; k++) acc.sum[k] = KSum{0.0, 0.0};
        for (int k = 0; k < kDiagExt; k++) acc.ext[k] = -__builtin_huge_val();
        acc.ext[7] = 0.0;
        for (int q = seg_chunk0[s]; q < seg_chunk0[s + 1]; q++) {
            const DiagPart& b = parts[q];
            for (int k = 0; k < kDiagSums; k++) acc.sum[k] = ksum_merge(acc.sum[k], b.sum[k]);
            for (int k = 0; k < 7; k++) acc.ext[k] = fmax(acc.ext[k], b.ext[k]);
            acc.ext[7] += b.ext[7];
        }
        out[s] = acc;
    }
}

// dst[k] = src[perm[k]]: carries a per-particle field through a re-sort.
template <class T>
__global__ void k_gather(const T* src, T* dst, const uint32_t* perm, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

// ---------------------------------------------------------------- IC-order I/O
template <class T>
__global__ void k_export(const T* src, double* dst, int width, int comp, const uint32_t* orig, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[(size_t)orig[k] * width + comp] = (double)src[k];
}
template <class T>
__global__ void k_import(const double* src, T* dst, int width, int comp, const uint32_t* orig, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = (T)src[(size_t)orig[k] * width + comp];
}
// kind 0: material (int32), 1: fixed (uint8), 2: segment (int32)
__global__ void k_export_info(const uint32_t* info, void* dst, int kind, const uint32_t* orig, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = info[k];
    const size_t o = orig[k];
    if (kind == 0) ((int32_t*)dst)[o] = info_mat(v);
    else if (kind == 1) ((uint8_t*)dst)[o] = (uint8_t)info_fixed(v);
    else ((int32_t*)dst)[o] = info_seg(v);
}
__global__ void k_import_info(const void* src, uint32_t* info, int kind, const uint32_t* orig, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = info[k];
    const size_t o = orig[k];
    if (kind == 0) info[k] = info_pack(((const int32_t*)src)[o], info_fixed(v), info_seg(v));
    else if (kind == 1) info[k] = info_pack(info_mat(v), ((const uint8_t*)src)[o] != 0, info_seg(v));
}

}  // namespace dsph
