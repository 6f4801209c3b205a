// Neighbour-pair kernels: adaptive-h density (with the EOS / principal-stress
// epilogue fused into the last iteration of a round) and the fused
// strain-rate + momentum + energy pass.
//
// One lane owns one particle i of the cell-sorted order and walks the sort
// cells within its own search radius (culling cells whose nearest point is out
// of range), reading neighbour records from the cell table's contiguous
// ranges.  Consecutive lanes hold spatially adjacent particles, so the
// neighbour records a wave touches overlap heavily and are served from L1/L2.
// Sums are gather-only (no atomics), so results do not depend on scheduling.
#pragma once
#include "kernels_build.h"

namespace dsph {

// Own-cell offset of particle i (FP64 path: derived from absolute position).
template <class R>
__device__ __forceinline__ void own_offset(const Carried<R>& c, const SegGrid& g, int64_t i, uint32_t cc,
                                           const typename Vec4<R>::T& rec, R& ox, R& oy, R& oz) {
    if (sizeof(R) == 4) {
        ox = rec.x; oy = rec.y; oz = rec.z;
    } else {
        const int cx = cc & 1023u, cy = (cc >> 10) & 1023u, cz = cc >> 20;
        ox = (R)(c.x[i] - (g.lo[0] + (double)cx * g.cw));
        oy = (R)(c.y[i] - (g.lo[1] + (double)cy * g.cw));
        oz = (R)(c.z[i] - (g.lo[2] + (double)cz * g.cw));
    }
}

// Distance along one axis from offset o (in [0, cw)) to the stencil cell d.
template <class R>
__device__ __forceinline__ R cell_gap(int d, R o, R cw) {
    const R lo = (R)d * cw - o, hi = o - (R)(d + 1) * cw;
    const R g = lo > hi ? lo : hi;
    return g > R(0) ? g : R(0);
}

// Fast reciprocal / square root: hardware approximations in FP32 (~1 ulp),
// IEEE operations in FP64.
__device__ __forceinline__ float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double f_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float f_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double f_sqrt(double x) { return sqrt(x); }

// Branch-free M4 shape and slope on q = r/h (q >= 0):
//   f(q)  = t^3/4 - s^3          (1 - 1.5q^2 + 0.75q^3 for q < 1, (2-q)^3/4 for 1 <= q < 2)
//   f'(q) = 3 s^2 - 0.75 t^2     (-3q + 2.25q^2 for q < 1, -0.75(2-q)^2 for 1 <= q < 2)
// with t = max(2 - q, 0), s = max(1 - q, 0); W = f/(pi h^3), dW/dr = f'/(pi h^4).
template <class R>
__device__ __forceinline__ R m4_shape(R q) {
    const R t = smax(R(2) - q, R(0)), s = smax(R(1) - q, R(0));
    return R(0.25) * t * t * t - s * s * s;
}
template <class R>
__device__ __forceinline__ R m4_slope(R q) {
    const R t = smax(R(2) - q, R(0)), s = smax(R(1) - q, R(0));
    return R(3) * s * s - R(0.75) * t * t;
}

// Would the reference's own 27-cell stencil (sph.hpp:163-182) reach particle
// j from particle i in this density round?  Needed only when 2h_i can exceed
// the reference's cell width; reproduces its neighbour set exactly.
__device__ __forceinline__ int oracle_bin(const SegGrid& g, const SegPar& p, int a, double v) {
    int b = (int)((v - p.dlo[a]) / g.ocw[a]);
    if (b < 0) b = 0;
    if (b >= g.onc[a]) b = g.onc[a] - 1;
    return b;
}
__device__ __forceinline__ bool oracle_reaches(const SegGrid& g, const SegPar& p, const int ob[3], double xj,
                                               double yj, double zj) {
    const double v[3] = {xj, yj, zj};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (g.onc[a] <= 1) continue;
        const int d = oracle_bin(g, p, a, v[a]) - ob[a];
        if (p.per[a]) {
            const int m = g.onc[a];
            const int dd = ((d % m) + m) % m;
            if (!(dd == 0 || dd == 1 || dd == m - 1)) return false;
        } else if (d < -1 || d > 1) {
            return false;
        }
    }
    return true;
}

// Per-particle EOS, damage scaling and artificial-stress coefficient
// (sph.hpp:261-275), then the force-pass neighbour records of particle i.
template <class R>
__device__ __forceinline__ void eos_epilogue(const Carried<R>& c, const Work<R>& w, int64_t i, const SegPar& p,
                                             int mat, R rho, R h, R ox, R oy, R oz) {
    using V4 = typename Vec4<R>::T;
    const int64_t n = w.n;
    const MatT<R> m = w.mats[mat];
    const R u = (R)c.u[i];
    const R P = eos_pressure(m, rho, u);
    const R cs = eos_sound_speed(m, rho, u);
    const R csig = d_sqrt(cs * cs + (R(4) / R(3)) * m.G / smax(rho, rho_floor<R>()));
    R s6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) s6[k] = c.S[(size_t)k * n + i];
    R pf = P, dsc = R(1);
    if (p.damage_on) {
        w.sigmax[i] = -P + eig_max(s6[0], s6[1], s6[2], s6[3], s6[4], s6[5]);
        dsc = R(1) - c.D[i];
        if (P < R(0)) pf = dsc * P;
    }
    const R eps_as = (R)p.eps_as;
    const R rart = (eps_as > R(0) && pf < R(0)) ? eps_as * (-pf) / (rho * rho) : R(0);
    w.P[i] = P;
    w.cs[i] = cs;
    w.csig[i] = csig;
    w.Pf[i] = pf;
    w.dscale[i] = dsc;
    w.Rart[i] = rart;
    const R ir2 = R(1) / (rho * rho);
    const R mi = (R)c.mass[i];
    V4 a, b, t, sa, sb;
    if (sizeof(R) == 4) {
        a.x = ox; a.y = oy; a.z = oz;
    } else {
        a.x = (R)c.x[i]; a.y = (R)c.y[i]; a.z = (R)c.z[i];
    }
    a.w = h;
    b.x = (R)c.vx[i]; b.y = (R)c.vy[i]; b.z = (R)c.vz[i]; b.w = mi;
    t.x = rho; t.y = pf * ir2; t.z = csig; t.w = rart;
    const R sc = dsc * ir2;
    sa.x = s6[0] * sc; sa.y = s6[1] * sc; sa.z = s6[2] * sc; sa.w = s6[3] * sc;
    sb.x = s6[4] * sc; sb.y = s6[5] * sc; sb.z = mi / rho; sb.w = R(0);
    // per-cell upper bound of h for the force stencil (positive floats order as integers)
    const uint32_t cc = w.cellc[i];
    const SegGrid& g = w.grid[info_seg(c.info[i])];
    if (isfinite((double)ox)) {
        float hf = (float)h;
        if ((double)hf < (double)h) hf = __uint_as_float(__float_as_uint(hf) + 1u);   // round up
        atomicMax((unsigned int*)&w.cell_hmax[cell_key(g, cc & 1023u, (cc >> 10) & 1023u, cc >> 20)],
                  __float_as_uint(hf));
    }
    w.frec_pos[i] = a;
    w.frec_vel[i] = b;
    w.frec_th[i] = t;
    w.frec_sa[i] = sa;
    w.frec_sb[i] = sb;
}

// One density iteration of round r (sph.hpp:189-201):
//   rho_i = max(sum_j m_j W(r_ij, h_i) over r_ij < 2 h_i, 1e-30)
//   h_i   = clamp(eta (m_i / rho_i)^(1/3), h_init / 2, 2 h_init)
// LAST marks the second iteration of the round: it also reduces max h for the
// round test and runs the EOS epilogue.
template <class R, bool LAST>
__global__ __launch_bounds__(256) void k_density(Carried<R> c, Work<R> w, int r, int remap, SegTargets<1> tg) {
    using V4 = typename Vec4<R>::T;
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    int seg = 0;
    bool act = false;
    if (i < w.n) {
        seg = info_seg(c.info[i]);
        act = w.st[seg].active[r] != 0;
    }
    double hnew_d = 0.0;
    if (act) {
        const SegPar& p = w.par[seg];
        const SegGrid& g = w.grid[seg];
        const R h = c.h[i];
        const uint32_t cc = w.cellc[i];
        const V4 me = w.dens_rec[i];
        R ox, oy, oz;
        own_offset<R>(c, g, i, cc, me, ox, oy, oz);
        const bool finite = isfinite((double)ox) && isfinite((double)oy) && isfinite((double)oz);
        R rho;
        if (finite) {
            const int cx = cc & 1023u, cy = (cc >> 10) & 1023u, cz = cc >> 20;
            const R cw = (R)g.cw;
            const R two_h = R(2) * h, th2 = two_h * two_h;
            int rad = (int)ceil((double)two_h * g.inv_cw);
            rad = rad < 1 ? 1 : (rad > kMaxStencil ? kMaxStencil : rad);
            const R inv_h = R(1) / h;
            const R sig = inv_h * inv_h * inv_h * R(1.0 / kPi);
            const bool emu = (double)two_h > 0.999 * g.ocw_min;
            int ob[3] = {0, 0, 0};
            if (emu) {
                ob[0] = oracle_bin(g, p, 0, c.x[i]);
                ob[1] = oracle_bin(g, p, 1, c.y[i]);
                ob[2] = oracle_bin(g, p, 2, c.z[i]);
            }
            // contribution of one candidate; FP32 tests r^2 (W is continuous at 2h),
            // FP64 tests r < 2h exactly as the reference does
            auto term = [&](const V4& q, R bx, R by, R bz) -> R {
                const R rx = bx - q.x, ry = by - q.y, rz = bz - q.z;
                const R r2 = rx * rx + ry * ry + rz * rz;
                const R rr = f_sqrt(r2);
                const bool in = (sizeof(R) == 4) ? (r2 < th2) : (rr < two_h);
                return in ? q.w * (sig * m4_shape(rr * inv_h)) : R(0);
            };
            R s = R(0);
            for (int dx = -rad; dx <= rad; dx++) {
                const int nx = cx + dx;
                if ((unsigned)nx >= (unsigned)g.dim[0]) continue;
                const R gx = cell_gap(dx, ox, cw);
                const R gx2 = gx * gx;
                if (gx2 >= th2) continue;
                for (int dy = -rad; dy <= rad; dy++) {
                    const int ny = cy + dy;
                    if ((unsigned)ny >= (unsigned)g.dim[1]) continue;
                    const R gy = cell_gap(dy, oy, cw);
                    const R gxy = gx2 + gy * gy;
                    if (gxy >= th2) continue;
                    for (int dz = -rad; dz <= rad; dz++) {
                        const int nz = cz + dz;
                        if ((unsigned)nz >= (unsigned)g.dim[2]) continue;
                        const R gz = cell_gap(dz, oz, cw);
                        if (gxy + gz * gz >= th2) continue;
                        const uint2 rg = w.table[cell_key(g, nx, ny, nz)];
                        R bx, by, bz;
                        if (sizeof(R) == 4) {
                            bx = ox - (R)dx * cw; by = oy - (R)dy * cw; bz = oz - (R)dz * cw;
                        } else {
                            bx = me.x; by = me.y; bz = me.z;
                        }
                        if (!emu) {
                            uint32_t k = rg.x;
                            for (; k + 1 < rg.y; k += 2) {       // two loads in flight
                                const V4 q0 = w.dens_rec[k], q1 = w.dens_rec[k + 1];
                                s += term(q0, bx, by, bz);
                                s += term(q1, bx, by, bz);
                            }
                            if (k < rg.y) s += term(w.dens_rec[k], bx, by, bz);
                        } else {
                            for (uint32_t k = rg.x; k < rg.y; k++) {
                                if (!oracle_reaches(g, p, ob, c.x[k], c.y[k], c.z[k])) continue;
                                s += term(w.dens_rec[k], bx, by, bz);
                            }
                        }
                    }
                }
            }
            rho = smax(s, rho_floor<R>());
        } else {
            rho = rho_floor<R>();
        }
        const R h0 = (R)p.h_init;
        const R mi = (R)c.mass[i];
        const R hn = smin(R(2) * h0, smax(R(0.5) * h0, (R)p.eta * d_cbrt(mi / rho)));
        w.rho[i] = rho;
        c.h[i] = hn;
        hnew_d = (double)hn;
        if (LAST) eos_epilogue<R>(c, w, i, p, info_mat(c.info[i]), rho, hn, ox, oy, oz);
    }
    if (LAST) {
        const double v[1] = {hnew_d};
        block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, act);
    }
}

// EOS epilogue on its own (System::compute_forces entry without a density pass).
template <class R>
__global__ __launch_bounds__(256) void k_eos_records(Carried<R> c, Work<R> w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.n) return;
    const uint32_t inf = c.info[i];
    const int seg = info_seg(inf);
    const SegGrid& g = w.grid[seg];
    R ox, oy, oz;
    own_offset<R>(c, g, i, w.cellc[i], w.dens_rec[i], ox, oy, oz);
    eos_epilogue<R>(c, w, i, w.par[seg], info_mat(inf), w.rho[i], c.h[i], ox, oy, oz);
}

// Fused strain rate + forces (sph.hpp:212-258 and 277-317; the Metal
// strain_forces kernel, sph_force.metal:308-365, fuses the same two loops).
// For each pair j != i with 0 < r < 2 hbar, hbar = (h_i + h_j)/2:
//   L_ab  += (m_j/rho_j) (v_j - v_i)_a (grad_i W)_b
//   term   = Pf_i/rho_i^2 + Pf_j/rho_j^2 + Pi_ij (Monaghan viscosity if approaching)
//   du    += m_j/2 term (v_ij . grad W)   [+ Price conductivity when alpha_u > 0]
//   a     -= m_j (term + (R_i + R_j)(W/W(hbar/eta))^4) grad W
//   a     += m_j (S_i d_i/rho_i^2 + S_j d_j/rho_j^2) . grad W
// then the Jaumann stress rate, the deviatoric work and the CFL step.
template <class R>
__global__ __launch_bounds__(256) void k_force(Carried<R> c, Work<R> w, int remap, SegTargets<1> tg) {
    using V4 = typename Vec4<R>::T;
    const int64_t n = w.n;
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    const bool in = i < n;
    int seg = 0;
    double dt_i = 0.0;
    bool dt_valid = false;
    if (in) {
        const uint32_t inf = c.info[i];
        seg = info_seg(inf);
        const int mat = info_mat(inf);
        const SegPar& p = w.par[seg];
        const SegGrid& g = w.grid[seg];
        const V4 pi = w.frec_pos[i];
        const V4 vi = w.frec_vel[i];
        const V4 ti = w.frec_th[i];
        const V4 sai = w.frec_sa[i];
        const V4 sbi = w.frec_sb[i];
        const uint32_t cc = w.cellc[i];
        R ox, oy, oz;
        own_offset<R>(c, g, i, cc, w.dens_rec[i], ox, oy, oz);
        const bool finite = isfinite((double)ox) && isfinite((double)oy) && isfinite((double)oz);
        const R hi = pi.w, rho_i = ti.x, pio = ti.y, csig_i = ti.z, rart_i = ti.w;
        const R alpha = (R)p.alpha, beta = (R)p.beta, eps = (R)p.eps, eps_as = (R)p.eps_as;
        const R alpha_u = (R)p.alpha_u;
        const R inv_fdp = (eps_as > R(0)) ? R(1) / kern_shape(R(1) / (R)p.eta) : R(0);
        R L00 = 0, L01 = 0, L02 = 0, L10 = 0, L11 = 0, L12 = 0, L20 = 0, L21 = 0, L22 = 0;
        R ax = 0, ay = 0, az = 0, du = 0;
        if (finite) {
            const int cx = cc & 1023u, cy = (cc >> 10) & 1023u, cz = cc >> 20;
            const R cw = (R)g.cw;
            const R rad_r = hi + (R)w.st[seg].hmax_final;
            const R rad2 = rad_r * rad_r;
            int rad = (int)ceil((double)rad_r * g.inv_cw);
            rad = rad < 1 ? 1 : (rad > kMaxStencil ? kMaxStencil : rad);
            const R ui = (R)c.u[i];
            const R Pi = w.P[i];
            for (int dx = -rad; dx <= rad; dx++) {
                const int nx = cx + dx;
                if ((unsigned)nx >= (unsigned)g.dim[0]) continue;
                const R gx0 = cell_gap(dx, ox, cw);
                const R gx2 = gx0 * gx0;
                if (gx2 >= rad2) continue;
                for (int dy = -rad; dy <= rad; dy++) {
                    const int ny = cy + dy;
                    if ((unsigned)ny >= (unsigned)g.dim[1]) continue;
                    const R gy0 = cell_gap(dy, oy, cw);
                    const R gxy = gx2 + gy0 * gy0;
                    if (gxy >= rad2) continue;
                    for (int dz = -rad; dz <= rad; dz++) {
                        const int nz = cz + dz;
                        if ((unsigned)nz >= (unsigned)g.dim[2]) continue;
                        const R gz0 = cell_gap(dz, oz, cw);
                        const R gg = gxy + gz0 * gz0;
                        if (gg >= rad2) continue;
                        const uint32_t key = cell_key(g, nx, ny, nz);
                        const R reach = hi + (R)w.cell_hmax[key];
                        if (gg >= reach * reach) continue;
                        const uint2 rg = w.table[key];
                        R bx, by, bz;
                        if (sizeof(R) == 4) {
                            bx = ox - (R)dx * cw; by = oy - (R)dy * cw; bz = oz - (R)dz * cw;
                        } else {
                            bx = pi.x; by = pi.y; bz = pi.z;
                        }
                        for (uint32_t k = rg.x; k < rg.y; k++) {
                            const V4 q = w.frec_pos[k];
                            const R rx = bx - q.x, ry = by - q.y, rz = bz - q.z;
                            const R r2 = rx * rx + ry * ry + rz * rz;
                            const R th = hi + q.w;                  // 2 hbar
                            if (!(r2 < th * th) || r2 == R(0)) continue;
                            const R r = f_sqrt(r2);
                            if (sizeof(R) == 8 && !(r < th)) continue;
                            const R hbar = R(0.5) * th;
                            const R ihb = f_rcp(hbar);
                            const R qq = r * ihb;
                            const R sgm = ihb * ihb * ihb * R(1.0 / kPi);
                            const R shp = m4_slope(qq);
                            const R fr = sgm * ihb * shp * f_rcp(r);    // dW/dr / r
                            const R gxw = rx * fr, gyw = ry * fr, gzw = rz * fr;
                            const V4 vj = w.frec_vel[k];
                            const V4 tj = w.frec_th[k];
                            const V4 saj = w.frec_sa[k];
                            const V4 sbj = w.frec_sb[k];
                            const R dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
                            const R Vj = sbj.z;
                            const R t0 = Vj * dvx, t1 = Vj * dvy, t2v = Vj * dvz;
                            L00 += t0 * gxw; L01 += t0 * gyw; L02 += t0 * gzw;
                            L10 += t1 * gxw; L11 += t1 * gyw; L12 += t1 * gzw;
                            L20 += t2v * gxw; L21 += t2v * gyw; L22 += t2v * gzw;
                            const R dvr = -(dvx * rx + dvy * ry + dvz * rz);   // v_ij . r_ij
                            R pij = R(0);
                            if (dvr < R(0)) {
                                const R mu = hbar * dvr * f_rcp(r2 + eps * hbar * hbar);
                                const R cbar = R(0.5) * (csig_i + tj.z);
                                const R rbar = R(0.5) * (rho_i + tj.x);
                                pij = (-alpha * cbar * mu + beta * mu * mu) * f_rcp(rbar);
                            }
                            const R term = pio + tj.y + pij;
                            const R vg = -(dvx * gxw + dvy * gyw + dvz * gzw);  // v_ij . grad W
                            const R mj = vj.w;
                            du += R(0.5) * mj * term * vg;
                            R mterm = term;
                            if (eps_as > R(0)) {
                                R f = m4_shape(qq) * inv_fdp;
                                f = f * f;
                                f = f * f;
                                mterm += (rart_i + tj.w) * f;
                            }
                            const R mm = mj * mterm;
                            const R sxx = sai.x + saj.x, syy = sai.y + saj.y, szz = sai.z + saj.z;
                            const R sxy = sai.w + saj.w, sxz = sbi.x + sbj.x, syz = sbi.y + sbj.y;
                            ax += (sxx * gxw + sxy * gyw + sxz * gzw) * mj - gxw * mm;
                            ay += (sxy * gxw + syy * gyw + syz * gzw) * mj - gyw * mm;
                            az += (sxz * gxw + syz * gyw + szz * gzw) * mj - gzw * mm;
                            if (alpha_u > R(0)) {
                                const R Pj = w.P[k];
                                const R uj = (R)c.u[k];
                                const R rbar = R(0.5) * (rho_i + tj.x);
                                const R vsig = d_sqrt(d_fabs(Pi - Pj) / rbar);
                                du += alpha_u * mj * vsig * (ui - uj) * (sgm * ihb * shp) / rbar;
                            }
                        }
                    }
                }
            }
        }
        // strain rate, spin, Jaumann rate and deviatoric power (sph.hpp:230-256)
        const R G = w.mats[mat].G;
        R epsw = R(0);
        R d6[6] = {0, 0, 0, 0, 0, 0};
        if (G > R(0)) {
            R s[6];
#pragma unroll
            for (int k = 0; k < 6; k++) s[k] = c.S[(size_t)k * n + i];
            const R exx = L00, eyy = L11, ezz = L22;
            const R exy = R(0.5) * (L01 + L10), exz = R(0.5) * (L02 + L20), eyz = R(0.5) * (L12 + L21);
            const R wxy = R(0.5) * (L01 - L10), wxz = R(0.5) * (L02 - L20), wyz = R(0.5) * (L12 - L21);
            const R tr = (exx + eyy + ezz) / R(3);
            const R W[3][3] = {{0, wxy, wxz}, {-wxy, 0, wyz}, {-wxz, -wyz, 0}};
            const R M[3][3] = {{s[0], s[3], s[4]}, {s[3], s[1], s[5]}, {s[4], s[5], s[2]}};
            R J[3][3];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int bb = 0; bb < 3; bb++) {
                    R ws = 0, sw = 0;
#pragma unroll
                    for (int k = 0; k < 3; k++) { ws += W[a][k] * M[k][bb]; sw += M[a][k] * W[k][bb]; }
                    J[a][bb] = ws - sw;
                }
            d6[0] = R(2) * G * (exx - tr) + J[0][0];
            d6[1] = R(2) * G * (eyy - tr) + J[1][1];
            d6[2] = R(2) * G * (ezz - tr) + J[2][2];
            d6[3] = R(2) * G * exy + J[0][1];
            d6[4] = R(2) * G * exz + J[0][2];
            d6[5] = R(2) * G * eyz + J[1][2];
            epsw = (s[0] * exx + s[1] * eyy + s[2] * ezz + R(2) * (s[3] * exy + s[4] * exz + s[5] * eyz)) / rho_i;
        }
#pragma unroll
        for (int k = 0; k < 6; k++) w.dSdt[(size_t)k * n + i] = d6[k];
        w.eps_work[i] = epsw;
        w.ax[i] = ax; w.ay[i] = ay; w.az[i] = az;
        w.dudt[i] = du + epsw * w.dscale[i];
        // CFL step (sph.hpp:350-355), evaluated in FP64
        dt_i = p.cfl * (double)hi / ((double)csig_i * (1.0 + p.alpha) + 1e-30);
        dt_valid = true;
    }
    const double v[1] = {dt_i};
    block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, dt_valid);
}

}  // namespace dsph
