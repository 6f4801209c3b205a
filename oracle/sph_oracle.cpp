// CPU FP64 oracle: a restatement of the reference SPH step for parity checks.
// TEST INFRASTRUCTURE ONLY -- loaded by tests/, the smoke check and bench.py's
// cpu_baseline leg; never by the product library.
//
// The arithmetic follows hydrocodes/sph/sph.hpp and hydrocodes/common/eos.hpp
// operation for operation (same expression order, same branch tests, same
// neighbour visiting order), because the point of this file is to be
// bit-identical to the reference when both are built with g++ -O2
// -ffp-contract=off.  The storage is this project's own: structure-of-arrays
// vectors, a CSR cell list instead of vector<vector<int>>, and free functions.
#include "sph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- kernel
// Cubic-spline M4, sigma = 1/(pi h^3), support 2h (sph.hpp:17-28).
inline double kern_w(double r, double h) {
    double q = r / h;
    double s = 1.0 / (M_PI * h * h * h);
    if (q < 1.0) return s * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
    if (q < 2.0) {
        double t = 2.0 - q;
        return s * 0.25 * t * t * t;
    }
    return 0.0;
}
inline double kern_dwdr(double r, double h) {
    double q = r / h;
    double s = 1.0 / (M_PI * h * h * h);
    if (q < 1.0) return s * (-3.0 * q + 2.25 * q * q) / h;
    if (q < 2.0) {
        double t = 2.0 - q;
        return s * (-0.75 * t * t) / h;
    }
    return 0.0;
}

// std::max / std::min semantics spelled out (first argument wins ties, NaN
// handling identical to <algorithm>).
inline double mx(double a, double b) { return (a < b) ? b : a; }
inline double mn(double a, double b) { return (b < a) ? b : a; }

// ---------------------------------------------------------------- EOS
struct Mat {
    int type = 0;
    double gamma = 1.4;
    double rho0 = 0, a = 0, b = 0, A = 0, B = 0, alpha = 0, beta = 0;
    double u0 = 0, uiv = 0, ucv = 0;
    double G = 0, Y = 0, Emod = 0, wm = 0, wk = 0;
};

Mat make_mat(const orc_material& m) {
    Mat r;
    r.type = m.type;
    r.gamma = m.gamma;
    r.rho0 = m.rho0; r.A = m.A; r.B = m.B; r.a = m.a; r.b = m.b;
    r.alpha = m.alpha; r.beta = m.beta; r.u0 = m.u0; r.uiv = m.uiv; r.ucv = m.ucv;
    r.G = m.G; r.Y = m.Y; r.wm = m.wm; r.wk = m.wk;
    // Young's modulus with K ~ A (eos.hpp:37); only the Tillotson factory sets it.
    r.Emod = (m.type == 1 && m.G > 0) ? 9.0 * m.A * m.G / (3.0 * m.A + m.G) : 0.0;
    return r;
}

// Tillotson pressure (eos.hpp:52-63).
double till_p(const Mat& m, double rho, double u) {
    if (rho <= 0) return 0.0;
    double eta = rho / m.rho0, mu = eta - 1.0;
    double w0 = u / (m.u0 * eta * eta) + 1.0;
    double pc = (m.a + m.b / w0) * rho * u + m.A * mu + m.B * mu * mu;
    if (rho >= m.rho0 || u <= m.uiv) return pc;
    double z = m.rho0 / rho - 1.0;
    double pe = m.a * rho * u + (m.b * rho * u / w0 + m.A * mu * std::exp(-m.beta * z))
                                    * std::exp(-m.alpha * z * z);
    if (u >= m.ucv) return pe;
    return ((u - m.uiv) * pe + (m.ucv - u) * pc) / (m.ucv - m.uiv);
}

double eos_p(const Mat& m, double rho, double u) {            // eos.hpp:65-68
    if (m.type == 0) return (m.gamma - 1.0) * rho * u;
    return till_p(m, rho, u);
}

// Condensed-branch P and analytic partials (eos.hpp:72-79).
void till_cond(const Mat& m, double rho, double u, double& P, double& dPdr, double& dPdu) {
    double eta = rho / m.rho0, mu = eta - 1.0;
    double w = 1.0 + u / (m.u0 * eta * eta), aob = m.a + m.b / w;
    P = aob * rho * u + m.A * mu + m.B * mu * mu;
    dPdu = aob * rho - m.b * rho * u / (w * w * m.u0 * eta * eta);
    double dwdr = -2.0 * u / (m.u0 * eta * eta * eta * m.rho0);
    dPdr = aob * u + rho * u * (-m.b / (w * w)) * dwdr + m.A / m.rho0 + 2.0 * m.B * mu / m.rho0;
}
// Expanded-branch P and analytic partials (eos.hpp:80-94).
void till_expd(const Mat& m, double rho, double u, double& P, double& dPdr, double& dPdu) {
    double eta = rho / m.rho0, mu = eta - 1.0;
    double w = 1.0 + u / (m.u0 * eta * eta);
    double z = m.rho0 / rho - 1.0, dzdr = -m.rho0 / (rho * rho);
    double e1 = std::exp(-m.alpha * z * z), e2 = std::exp(-m.beta * z);
    double g = m.b * rho * u / w + m.A * mu * e2;
    P = m.a * rho * u + g * e1;
    double dwdr = -2.0 * u / (m.u0 * eta * eta * eta * m.rho0);
    double dgu = m.b * rho / w - m.b * rho * u / (w * w * m.u0 * eta * eta);
    dPdu = m.a * rho + dgu * e1;
    double dbruw = m.b * u / w + m.b * rho * u * (-1.0 / (w * w)) * dwdr;
    double dgr = dbruw + m.A * e2 / m.rho0 + m.A * mu * (e2 * (-m.beta) * dzdr);
    double de1 = e1 * (-m.alpha * 2.0 * z) * dzdr;
    dPdr = m.a * u + dgr * e1 + g * de1;
}
// c_s = sqrt(max(dP/drho + P/rho^2 dP/du, 1e-6)) (eos.hpp:95-113).
double eos_cs(const Mat& m, double rho, double u) {
    if (m.type == 0) {
        double P = (m.gamma - 1.0) * rho * u;
        return std::sqrt(mx(m.gamma * P / mx(rho, 1e-30), 1e-12));
    }
    if (rho <= 0) return 1e-3;
    double P, dPdr, dPdu;
    if (rho >= m.rho0 || u <= m.uiv) {
        till_cond(m, rho, u, P, dPdr, dPdu);
    } else if (u >= m.ucv) {
        till_expd(m, rho, u, P, dPdr, dPdu);
    } else {
        double pc, pcr, pcu, pe, per, peu;
        till_cond(m, rho, u, pc, pcr, pcu);
        till_expd(m, rho, u, pe, per, peu);
        double inv = 1.0 / (m.ucv - m.uiv);
        P = ((u - m.uiv) * pe + (m.ucv - u) * pc) * inv;
        dPdu = ((pe + (u - m.uiv) * peu) + (-pc + (m.ucv - u) * pcu)) * inv;
        dPdr = ((u - m.uiv) * per + (m.ucv - u) * pcr) * inv;
    }
    return std::sqrt(mx(dPdr + P / (rho * rho) * dPdu, 1e-6));
}

// ---------------------------------------------------------------- Sym3
// Components kept as six scalars {xx, yy, zz, xy, xz, yz}.
struct T6 { double xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0; };

// Largest eigenvalue, Smith (1961) trigonometric form (sph.hpp:41-54).
double top_eig(const T6& s) {
    double p1 = s.xy * s.xy + s.xz * s.xz + s.yz * s.yz;
    double q = (s.xx + s.yy + s.zz) / 3.0;
    if (p1 < 1e-30) {
        double m = s.xx;                 // std::max({xx, yy, zz}) scans left to right
        if (m < s.yy) m = s.yy;
        if (m < s.zz) m = s.zz;
        return m;
    }
    double p2 = (s.xx - q) * (s.xx - q) + (s.yy - q) * (s.yy - q) + (s.zz - q) * (s.zz - q) + 2 * p1;
    double p = std::sqrt(p2 / 6.0);
    double a = (s.xx - q) / p, b = (s.yy - q) / p, c = (s.zz - q) / p;
    double d = s.xy / p, e = s.xz / p, f = s.yz / p;
    double r = (a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e)) / 2.0;
    r = mx(-1.0, mn(1.0, r));
    double phi = std::acos(r) / 3.0;
    return q + 2.0 * p * std::cos(phi);
}

// ---------------------------------------------------------------- threads
// Static-chunk fork-join (sph.hpp:56-70): serial below 2000 items.
template <class F>
void par_for(int n, unsigned nt, F f) {
    if (n < 2000 || nt <= 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> pool;
    int chunk = (n + (int)nt - 1) / (int)nt;
    for (unsigned t = 0; t < nt; t++) {
        int lo = (int)t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([&f, lo, hi] { for (int i = lo; i < hi; i++) f(i); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

// ================================================================ system
struct orc_sys {
    orc_params prm;
    std::vector<Mat> mats;
    unsigned nthreads = 1;
    int n = 0;
    // particle state, structure of arrays
    std::vector<double> x, y, z, vx, vy, vz, ax, ay, az;
    std::vector<double> mass, rho, u, dudt, P, cs, csig, hh, eps_work;
    std::vector<double> D, eps_act, sigmax, Pf, Rart, dscale;
    std::vector<T6> S, dSdt;
    std::vector<int> mat;
    std::vector<char> fixed;
    // cell list (CSR): cell c holds items[start[c] .. start[c+1]) in ascending index
    int nc[3] = {1, 1, 1};
    double cw[3] = {1, 1, 1};
    double h_max = 0.01;
    std::vector<int> cstart, items, cell_of;

    // minimum-image separation a - b (sph.hpp:130-136)
    inline void sep(int i, int j, double& dx, double& dy, double& dz) const {
        dx = x[i] - x[j];
        dy = y[i] - y[j];
        dz = z[i] - z[j];
        if (prm.periodic[0]) { double L = prm.hi[0] - prm.lo[0]; dx -= L * std::round(dx / L); }
        if (prm.periodic[1]) { double L = prm.hi[1] - prm.lo[1]; dy -= L * std::round(dy / L); }
        if (prm.periodic[2]) { double L = prm.hi[2] - prm.lo[2]; dz -= L * std::round(dz / L); }
    }
    inline int bin(int ax, double c) const {                     // sph.hpp:142-146
        int i = (int)((c - prm.lo[ax]) / cw[ax]);
        if (i < 0) i = 0;
        if (i >= nc[ax]) i = nc[ax] - 1;
        return i;
    }
    void build_cells() {                                          // sph.hpp:147-162
        h_max = 1e-30;
        for (int i = 0; i < n; i++) h_max = mx(h_max, hh[i]);
        double rc = 2.0 * h_max;
        for (int a = 0; a < 3; a++) {
            double ext = prm.hi[a] - prm.lo[a];
            if (ext <= 0) ext = rc;
            int m = std::max(1, (int)(ext / rc));
            if (prm.periodic[a] && m < 3) m = 1;
            nc[a] = m;
            cw[a] = ext / m;
        }
        size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
        cstart.assign(ncell + 1, 0);
        cell_of.resize(n);
        for (int i = 0; i < n; i++) {
            int c = (bin(0, x[i]) * nc[1] + bin(1, y[i])) * nc[2] + bin(2, z[i]);
            cell_of[i] = c;
            cstart[c + 1]++;
        }
        for (size_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
        items.resize(n);
        std::vector<int> fill(cstart.begin(), cstart.end() - 1);
        for (int i = 0; i < n; i++) items[fill[cell_of[i]]++] = i;   // ascending i per cell
    }
    // 27-cell stencil in (d0, d1, d2) lexicographic order (sph.hpp:163-182).
    template <class F>
    void each_neighbor(int i, F fn) const {
        int b[3] = {bin(0, x[i]), bin(1, y[i]), bin(2, z[i])};
        int lo[3], hi[3];
        for (int a = 0; a < 3; a++) { lo[a] = nc[a] > 1 ? -1 : 0; hi[a] = nc[a] > 1 ? 1 : 0; }
        for (int d0 = lo[0]; d0 <= hi[0]; d0++) {
            int j0 = b[0] + d0;
            if (prm.periodic[0]) j0 = ((j0 % nc[0]) + nc[0]) % nc[0];
            else if (j0 < 0 || j0 >= nc[0]) continue;
            for (int d1 = lo[1]; d1 <= hi[1]; d1++) {
                int j1 = b[1] + d1;
                if (prm.periodic[1]) j1 = ((j1 % nc[1]) + nc[1]) % nc[1];
                else if (j1 < 0 || j1 >= nc[1]) continue;
                for (int d2 = lo[2]; d2 <= hi[2]; d2++) {
                    int j2 = b[2] + d2;
                    if (prm.periodic[2]) j2 = ((j2 % nc[2]) + nc[2]) % nc[2];
                    else if (j2 < 0 || j2 >= nc[2]) continue;
                    int c = (j0 * nc[1] + j1) * nc[2] + j2;
                    for (int k = cstart[c]; k < cstart[c + 1]; k++) fn(items[k]);
                }
            }
        }
    }

    // ------------------------------------------------------------ phases
    void density_h(int iters) {                                    // sph.hpp:184-208
        const double h0 = prm.h_init, eta = prm.eta;
        for (int round = 0; round < 4; round++) {
            build_cells();
            double hmax0 = h_max;
            for (int it = 0; it < iters; it++) {
                par_for(n, nthreads, [&](int i) {
                    double s = 0, hi = hh[i];
                    each_neighbor(i, [&](int j) {
                        double dx, dy, dz;
                        sep(i, j, dx, dy, dz);
                        double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (r < 2.0 * hi) s += mass[j] * kern_w(r, hi);
                    });
                    rho[i] = mx(s, 1e-30);
                    hh[i] = mn(2.0 * h0, mx(0.5 * h0, eta * std::cbrt(mass[i] / rho[i])));
                });
            }
            double hm = 1e-30;
            for (int i = 0; i < n; i++) hm = mx(hm, hh[i]);
            if (hm <= hmax0 * 1.02) break;
        }
        build_cells();
    }

    void strain_stress() {                                         // sph.hpp:212-258
        par_for(n, nthreads, [&](int i) {
            double G = mats[mat[i]].G;
            if (G <= 0) { dSdt[i] = T6{}; eps_work[i] = 0; return; }
            double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            each_neighbor(i, [&](int j) {
                if (j == i) return;
                double rx, ry, rz;
                sep(i, j, rx, ry, rz);
                double r = std::sqrt(rx * rx + ry * ry + rz * rz);
                double hbar = 0.5 * (hh[i] + hh[j]);
                if (r >= 2.0 * hbar || r == 0) return;
                double f = kern_dwdr(r, hbar) / r;
                double g[3] = {rx * f, ry * f, rz * f};
                double vol = mass[j] / rho[j];
                double dv[3] = {vx[j] - vx[i], vy[j] - vy[i], vz[j] - vz[i]};
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++) L[a][c] += vol * dv[a] * g[c];
            });
            double exx = L[0][0], eyy = L[1][1], ezz = L[2][2];
            double exy = 0.5 * (L[0][1] + L[1][0]), exz = 0.5 * (L[0][2] + L[2][0]),
                   eyz = 0.5 * (L[1][2] + L[2][1]);
            double wxy = 0.5 * (L[0][1] - L[1][0]), wxz = 0.5 * (L[0][2] - L[2][0]),
                   wyz = 0.5 * (L[1][2] - L[2][1]);
            double tr = (exx + eyy + ezz) / 3.0;
            const T6& s = S[i];
            // Jaumann rotation term J = W S - S W, W antisymmetric (W01 = wxy ...).
            double W[3][3] = {{0, wxy, wxz}, {-wxy, 0, wyz}, {-wxz, -wyz, 0}};
            double M[3][3] = {{s.xx, s.xy, s.xz}, {s.xy, s.yy, s.yz}, {s.xz, s.yz, s.zz}};
            double J[3][3];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) {
                    double ws = 0, sw = 0;
                    for (int k = 0; k < 3; k++) { ws += W[a][k] * M[k][c]; sw += M[a][k] * W[k][c]; }
                    J[a][c] = ws - sw;
                }
            T6& d = dSdt[i];
            d.xx = 2 * G * (exx - tr) + J[0][0];
            d.yy = 2 * G * (eyy - tr) + J[1][1];
            d.zz = 2 * G * (ezz - tr) + J[2][2];
            d.xy = 2 * G * exy + J[0][1];
            d.xz = 2 * G * exz + J[0][2];
            d.yz = 2 * G * eyz + J[1][2];
            eps_work[i] = (s.xx * exx + s.yy * eyy + s.zz * ezz
                           + 2 * (s.xy * exy + s.xz * exz + s.yz * eyz)) / rho[i];
        });
    }

    void forces() {                                                // sph.hpp:260-318
        par_for(n, nthreads, [&](int i) {
            const Mat& m = mats[mat[i]];
            P[i] = eos_p(m, rho[i], u[i]);
            cs[i] = eos_cs(m, rho[i], u[i]);
            csig[i] = std::sqrt(cs[i] * cs[i] + (4.0 / 3.0) * m.G / mx(rho[i], 1e-30));
            double pf = P[i], dsc = 1.0;
            if (prm.damage_on) {
                sigmax[i] = -P[i] + top_eig(S[i]);
                dsc = 1.0 - D[i];
                if (P[i] < 0) pf = dsc * P[i];
            }
            Pf[i] = pf;
            dscale[i] = dsc;
            Rart[i] = (prm.eps_as > 0 && pf < 0) ? prm.eps_as * (-pf) / (rho[i] * rho[i]) : 0.0;
        });
        strain_stress();
        const double alpha = prm.alpha, beta = prm.beta, eps = prm.eps, eta = prm.eta;
        const double eps_as = prm.eps_as, alpha_u = prm.alpha_u;
        par_for(n, nthreads, [&](int i) {
            double acx = 0, acy = 0, acz = 0, du = 0;
            double iri2 = 1.0 / (rho[i] * rho[i]);
            double pio = Pf[i] * iri2;
            double di = dscale[i];
            const T6& si = S[i];
            each_neighbor(i, [&](int j) {
                if (j == i) return;
                double rx, ry, rz;
                sep(i, j, rx, ry, rz);
                double r = std::sqrt(rx * rx + ry * ry + rz * rz);
                double hbar = 0.5 * (hh[i] + hh[j]);
                if (r >= 2.0 * hbar || r == 0) return;
                double f = kern_dwdr(r, hbar) / r;
                double gx = rx * f, gy = ry * f, gz = rz * f;
                double wx = vx[i] - vx[j], wy = vy[i] - vy[j], wz = vz[i] - vz[j];
                double rbar = 0.5 * (rho[i] + rho[j]);
                double pij = 0, dvr = wx * rx + wy * ry + wz * rz;
                if (dvr < 0) {                                      // Monaghan AV, approaching
                    double mu = hbar * dvr / (r * r + eps * hbar * hbar);
                    double cbar = 0.5 * (csig[i] + csig[j]);
                    pij = (-alpha * cbar * mu + beta * mu * mu) / rbar;
                }
                double jrj2 = 1.0 / (rho[j] * rho[j]);
                double term = pio + Pf[j] * jrj2 + pij;
                du += 0.5 * mass[j] * term * (wx * gx + wy * gy + wz * gz);
                double mterm = term;
                if (eps_as > 0) {                                   // artificial stress, momentum only
                    double wdp = kern_w(hbar / eta, hbar);
                    double fr = (wdp > 0 ? kern_w(r, hbar) / wdp : 0.0);
                    double f4 = fr * fr;
                    f4 *= f4;
                    mterm += (Rart[i] + Rart[j]) * f4;
                }
                double mm = mass[j] * mterm;
                acx -= gx * mm;
                acy -= gy * mm;
                acz -= gz * mm;
                // deviatoric stress, damage-scaled on each side
                const T6& sj = S[j];
                double ci = iri2 * di, cj = jrj2 * dscale[j];
                double tx = (si.xx * gx + si.xy * gy + si.xz * gz) * ci
                          + (sj.xx * gx + sj.xy * gy + sj.xz * gz) * cj;
                double ty = (si.xy * gx + si.yy * gy + si.yz * gz) * ci
                          + (sj.xy * gx + sj.yy * gy + sj.yz * gz) * cj;
                double tz = (si.xz * gx + si.yz * gy + si.zz * gz) * ci
                          + (sj.xz * gx + sj.yz * gy + sj.zz * gz) * cj;
                acx += tx * mass[j];
                acy += ty * mass[j];
                acz += tz * mass[j];
                if (alpha_u > 0) {                                  // Price conductivity
                    double vsig = std::sqrt(std::abs(P[i] - P[j]) / rbar);
                    du += alpha_u * mass[j] * vsig * (u[i] - u[j]) * kern_dwdr(r, hbar) / rbar;
                }
            });
            ax[i] = acx; ay[i] = acy; az[i] = acz;
            dudt[i] = du + eps_work[i] * di;
        });
    }

    void yield_one(int i) {                                        // sph.hpp:320-331
        double Y = mats[mat[i]].Y;
        T6& s = S[i];
        if (Y <= 0) { s = T6{}; return; }
        double J2 = 0.5 * (s.xx * s.xx + s.yy * s.yy + s.zz * s.zz)
                    + (s.xy * s.xy + s.xz * s.xz + s.yz * s.yz);
        double vm = std::sqrt(3.0 * J2);
        if (vm > Y) {
            double f = Y / vm;
            s.xx *= f; s.yy *= f; s.zz *= f; s.xy *= f; s.xz *= f; s.yz *= f;
        }
    }

    void damage(double dt) {                                       // sph.hpp:336-348
        const double rs = 0.5 * prm.h_init / prm.eta;
        par_for(n, nthreads, [&](int i) {
            const Mat& m = mats[mat[i]];
            if (m.Emod <= 0 || D[i] >= 1.0) return;
            double e_loc = sigmax[i] / m.Emod;
            if (sigmax[i] > 0 && e_loc > eps_act[i]) {
                double cg = 0.4 * std::sqrt(m.Emod / mx(rho[i], 1e-30));
                double d13 = std::cbrt(D[i]) + (cg / rs) * dt;
                D[i] = mn(1.0, d13 * d13 * d13);
            }
        });
    }

    double cfl_dt() const {                                        // sph.hpp:350-355
        double dt = 1e30;
        for (int i = 0; i < n; i++) dt = mn(dt, prm.cfl * hh[i] / (csig[i] * (1.0 + prm.alpha) + 1e-30));
        return dt;
    }

    void advance(double dt) {                                      // sph.hpp:357-377
        std::vector<double> du_old = dudt;
        std::vector<T6> dS_old = dSdt;
        double half = 0.5 * dt;
        for (int i = 0; i < n; i++)
            if (!fixed[i]) { vx[i] += ax[i] * half; vy[i] += ay[i] * half; vz[i] += az[i] * half; }
        for (int i = 0; i < n; i++)
            if (!fixed[i]) { x[i] += vx[i] * dt; y[i] += vy[i] * dt; z[i] += vz[i] * dt; }
        density_h(2);
        forces();
        if (prm.damage_on) damage(dt);
        for (int i = 0; i < n; i++) {
            if (fixed[i]) continue;
            vx[i] += ax[i] * half; vy[i] += ay[i] * half; vz[i] += az[i] * half;
            u[i] += 0.5 * dt * (du_old[i] + dudt[i]);
            if (u[i] < 1e-12) u[i] = 1e-12;
            T6& s = S[i];
            const T6& o = dS_old[i];
            const T6& d = dSdt[i];
            s.xx += 0.5 * dt * (o.xx + d.xx);
            s.yy += 0.5 * dt * (o.yy + d.yy);
            s.zz += 0.5 * dt * (o.zz + d.zz);
            s.xy += 0.5 * dt * (o.xy + d.xy);
            s.xz += 0.5 * dt * (o.xz + d.xz);
            s.yz += 0.5 * dt * (o.yz + d.yz);
            yield_one(i);
        }
    }
};

// ================================================================ C API
extern "C" {

orc_sys* orc_create(const orc_params* p, const orc_material* mats, int32_t nmat) {
    orc_sys* s = new orc_sys();
    s->prm = *p;
    for (int k = 0; k < nmat; k++) s->mats.push_back(make_mat(mats[k]));
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (p->nthreads > 0) nt = (unsigned)p->nthreads;
    else if (nt > 16) nt = 16;
    s->nthreads = nt;
    s->h_max = p->h_init;
    return s;
}

void orc_destroy(orc_sys* s) { delete s; }

void orc_add(orc_sys* s, int64_t n, const double* pos3, const double* vel3, const double* mass,
             const double* u, const int32_t* mat, const uint8_t* fixed) {
    for (int64_t k = 0; k < n; k++) {
        s->x.push_back(pos3[3 * k]); s->y.push_back(pos3[3 * k + 1]); s->z.push_back(pos3[3 * k + 2]);
        s->vx.push_back(vel3[3 * k]); s->vy.push_back(vel3[3 * k + 1]); s->vz.push_back(vel3[3 * k + 2]);
        s->ax.push_back(0); s->ay.push_back(0); s->az.push_back(0);
        s->mass.push_back(mass[k]); s->rho.push_back(0); s->u.push_back(u[k]); s->dudt.push_back(0);
        s->P.push_back(0); s->cs.push_back(0); s->csig.push_back(0); s->hh.push_back(s->prm.h_init);
        s->eps_work.push_back(0); s->S.push_back(T6{}); s->dSdt.push_back(T6{});
        s->D.push_back(0); s->eps_act.push_back(1e30); s->sigmax.push_back(0);
        s->Pf.push_back(0); s->Rart.push_back(0); s->dscale.push_back(1.0);
        s->mat.push_back(mat[k]); s->fixed.push_back(fixed ? (fixed[k] ? 1 : 0) : 0);
        s->n++;
    }
}

int64_t orc_n(const orc_sys* s) { return s->n; }

void orc_seed_damage(orc_sys* s, double dx, uint32_t seed) {        // sph.hpp:117-128
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(1e-9, 1.0);
    double V = dx * dx * dx;
    for (int i = 0; i < s->n; i++) {
        const Mat& m = s->mats[s->mat[i]];
        if (m.wk > 0 && m.wm > 0) s->eps_act[i] = std::pow(U(rng) / (m.wk * V), 1.0 / m.wm);
        else s->eps_act[i] = 1e30;
    }
}

void orc_init(orc_sys* s) { s->density_h(2); s->forces(); }
double orc_timestep(orc_sys* s) { return s->cfl_dt(); }
void orc_step(orc_sys* s, double dt) { s->advance(dt); }
void orc_build_cells(orc_sys* s) { s->build_cells(); }
void orc_compute_density_h(orc_sys* s, int32_t iters) { s->density_h(iters); }
void orc_compute_strain_stress(orc_sys* s) { s->strain_stress(); }
void orc_compute_forces(orc_sys* s) { s->forces(); }
void orc_grow_damage(orc_sys* s, double dt) { s->damage(dt); }
void orc_yield_limit(orc_sys* s, int64_t i) { s->yield_one((int)i); }

static int field_width(int f) {
    switch (f) {
        case ORC_POS: case ORC_VEL: case ORC_ACC: return 3;
        case ORC_S: case ORC_DSDT: return 6;
        default: return (f >= 0 && f < ORC_NFIELDS) ? 1 : 0;
    }
}

static std::vector<double>* scalar_field(orc_sys* s, int f) {
    switch (f) {
        case ORC_MASS: return &s->mass;     case ORC_RHO: return &s->rho;
        case ORC_U: return &s->u;           case ORC_DUDT: return &s->dudt;
        case ORC_P: return &s->P;           case ORC_CS: return &s->cs;
        case ORC_CSIG: return &s->csig;     case ORC_HH: return &s->hh;
        case ORC_EPS_WORK: return &s->eps_work;
        case ORC_D: return &s->D;           case ORC_EPS_ACT: return &s->eps_act;
        case ORC_SIGMAX: return &s->sigmax; case ORC_PF: return &s->Pf;
        case ORC_RART: return &s->Rart;     case ORC_DSCALE: return &s->dscale;
        default: return nullptr;
    }
}

int32_t orc_get(const orc_sys* cs, int32_t f, double* out) {
    orc_sys* s = const_cast<orc_sys*>(cs);
    int n = s->n;
    if (field_width(f) == 0) return -1;
    if (f == ORC_POS || f == ORC_VEL || f == ORC_ACC) {
        const std::vector<double>* v[3];
        if (f == ORC_POS) { v[0] = &s->x; v[1] = &s->y; v[2] = &s->z; }
        else if (f == ORC_VEL) { v[0] = &s->vx; v[1] = &s->vy; v[2] = &s->vz; }
        else { v[0] = &s->ax; v[1] = &s->ay; v[2] = &s->az; }
        for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) out[3 * i + k] = (*v[k])[i];
    } else if (f == ORC_S || f == ORC_DSDT) {
        const std::vector<T6>& t = (f == ORC_S) ? s->S : s->dSdt;
        for (int i = 0; i < n; i++) {
            const T6& e = t[i];
            double* o = out + 6 * i;
            o[0] = e.xx; o[1] = e.yy; o[2] = e.zz; o[3] = e.xy; o[4] = e.xz; o[5] = e.yz;
        }
    } else if (f == ORC_MAT) {
        for (int i = 0; i < n; i++) out[i] = (double)s->mat[i];
    } else if (f == ORC_FIXED) {
        for (int i = 0; i < n; i++) out[i] = (double)s->fixed[i];
    } else {
        std::memcpy(out, scalar_field(s, f)->data(), sizeof(double) * n);
    }
    return 0;
}

int32_t orc_set(orc_sys* s, int32_t f, const double* in) {
    int n = s->n;
    if (field_width(f) == 0) return -1;
    if (f == ORC_POS || f == ORC_VEL || f == ORC_ACC) {
        std::vector<double>* v[3];
        if (f == ORC_POS) { v[0] = &s->x; v[1] = &s->y; v[2] = &s->z; }
        else if (f == ORC_VEL) { v[0] = &s->vx; v[1] = &s->vy; v[2] = &s->vz; }
        else { v[0] = &s->ax; v[1] = &s->ay; v[2] = &s->az; }
        for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) (*v[k])[i] = in[3 * i + k];
    } else if (f == ORC_S || f == ORC_DSDT) {
        std::vector<T6>& t = (f == ORC_S) ? s->S : s->dSdt;
        for (int i = 0; i < n; i++) {
            const double* o = in + 6 * i;
            T6& e = t[i];
            e.xx = o[0]; e.yy = o[1]; e.zz = o[2]; e.xy = o[3]; e.xz = o[4]; e.yz = o[5];
        }
    } else if (f == ORC_MAT) {
        for (int i = 0; i < n; i++) s->mat[i] = (int)in[i];
    } else if (f == ORC_FIXED) {
        for (int i = 0; i < n; i++) s->fixed[i] = in[i] != 0.0 ? 1 : 0;
    } else {
        std::memcpy(scalar_field(s, f)->data(), in, sizeof(double) * n);
    }
    return 0;
}

void orc_cell_info(const orc_sys* s, int32_t nc[3], double cw[3], double* h_max) {
    for (int a = 0; a < 3; a++) { nc[a] = s->nc[a]; cw[a] = s->cw[a]; }
    *h_max = s->h_max;
}

double orc_kW(double r, double h) { return kern_w(r, h); }
double orc_kdWdr(double r, double h) { return kern_dwdr(r, h); }
double orc_eig_max(const double s6[6]) {
    T6 t;
    t.xx = s6[0]; t.yy = s6[1]; t.zz = s6[2]; t.xy = s6[3]; t.xz = s6[4]; t.yz = s6[5];
    return top_eig(t);
}
double orc_material_emod(const orc_material* m) { return make_mat(*m).Emod; }
double orc_pressure(const orc_material* m, double rho, double u) { return eos_p(make_mat(*m), rho, u); }
double orc_sound_speed(const orc_material* m, double rho, double u) { return eos_cs(make_mat(*m), rho, u); }

}  // extern "C"
