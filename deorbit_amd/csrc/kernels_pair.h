// Neighbour-pair kernels: adaptive-h density (with the EOS / principal-stress
// epilogue fused into the last iteration of a round) and the fused
// strain-rate + momentum + energy pass.
//
// One lane owns one particle i of the cell-sorted order and walks the sort
// columns within its own search radius, reading in each one contiguous run of
// neighbour records cut to its reach along z (see for_pieces).  Consecutive
// lanes hold neighbouring particles of one column, so the neighbour records a
// wave touches overlap heavily and are served from L1/L2.  Sums are
// gather-only (no atomics), so results do not depend on scheduling.
#pragma once
#include "kernels_build.h"
#include "moments.h"

namespace dsph {

// Offset of particle i from its own frame: the corner of its column (x, y)
// and the middle of its slice block (z).  FP32 path: the density record; FP64 path:
// derived from the absolute position.
template <class R>
__device__ __forceinline__ void own_offset(const Carried<R>& c, const SegGrid& g, int64_t i, uint32_t cc,
                                           const typename Vec4<R>::T& rec, R& ox, R& oy, R& oz) {
    if (sizeof(R) == 4) {
        ox = rec.x; oy = rec.y; oz = rec.z;
    } else {
        const int cx = cc & 1023u, cy = (cc >> 10) & 1023u, cz = cc >> 20;
        ox = (R)(c.x[i] - (g.lo[0] + (double)cx * g.cw[0]));
        oy = (R)(c.y[i] - (g.lo[1] + (double)cy * g.cw[1]));
        oz = (R)(c.z[i] - (g.lo[2] + (double)block_origin(cz) * g.cw[2]));
    }
}
// Offset from the corner of the own slice, given the offset oz from the own
// block's frame origin (what the distance culling of z slices measures from).
template <class R>
__device__ __forceinline__ R slice_offset(const SegGrid& g, int cz, R oz) {
    return oz - (R)(cz - block_origin(cz)) * (R)g.cw[2];
}

// Distance along one axis from offset o (in [0, cw)) to the stencil cell d.
template <class R>
__device__ __forceinline__ R cell_gap(int d, R o, R cw) {
    const R lo = (R)d * cw - o, hi = o - (R)(d + 1) * cw;
    const R g = lo > hi ? lo : hi;
    return g > R(0) ? g : R(0);
}

// Minimum-image separation on periodic axes (sph.hpp:130-136):
// d -= L round(d / L), round half away from zero.
__device__ __forceinline__ float d_round(float x) { return roundf(x); }
__device__ __forceinline__ double d_round(double x) { return round(x); }
template <class R>
__device__ __forceinline__ void min_image(const SegGrid& g, R& rx, R& ry, R& rz) {
    if (g.per[0]) { const R L = (R)g.L[0]; rx -= L * d_round(rx / L); }
    if (g.per[1]) { const R L = (R)g.L[1]; ry -= L * d_round(ry / L); }
    if (g.per[2]) { const R L = (R)g.L[2]; rz -= L * d_round(rz / L); }
}

// The cells a stencil walk visits along one axis.  Offset d names the cell
// c + d, wrapped on a periodic axis (whose cells tile the period exactly, so a
// wrapped cell's image sits d cell widths away).  The frame shift is the
// corner of the visited image's frame relative to the own frame's corner;
// frames are cells along x and y and blocks of kSliceBlock slices along z
// (blk).  A periodic axis with fewer cells than the stencil is walked once
// over all its cells (offsets -(dim-1)/2 .. dim-1-(dim-1)/2) without culling.
// o is the offset from the own cell's corner (for the distance culling).
template <class R, bool PER>
struct AxisWalk {
    int c, dim, lo, hi, per, blk;
    bool cull;
    R o, cw;
    __device__ __forceinline__ void init(const SegGrid& g, int a, int cc, R oo, double reach) {
        c = cc; dim = g.dim[a]; per = PER && g.per[a]; o = oo; cw = (R)g.cw[a];
        blk = (a == 2) ? kSliceBlock : 1;
        int rad = (int)ceil(reach * g.inv_cw[a]);
        rad = rad < 1 ? 1 : (rad > kMaxStencil ? kMaxStencil : rad);
        cull = true;
        lo = -rad; hi = rad;
        if (per && 2 * rad + 1 > dim) {
            lo = -((dim - 1) / 2);
            hi = dim - 1 - (dim - 1) / 2;
            cull = false;
        }
    }
    // false when offset d leaves a non-periodic axis; ishift is the shift in cells
    __device__ __forceinline__ bool cell(int d, int& n, R& shift, int& ishift, R& gap) const {
        n = c + d;
        if (per) {
            if (n < 0) n += dim;
            else if (n >= dim) n -= dim;
        } else if ((unsigned)n >= (unsigned)dim) {
            return false;
        }
        // image frame corner: frame start of n, moved by the wrap (c + d - n)
        const int fs = n - n % blk, fc = c - c % blk;
        ishift = fs + (c + d - n) - fc;
        shift = (R)ishift * cw;
        gap = cull ? cell_gap(d, o, cw) : R(0);
        return true;
    }
};

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
// (q is finite wherever the result is used, so the clamps are plain max
// instructions; the NaN-propagating smax stays in the EOS and h updates)
__device__ __forceinline__ float ramp(float x) { return fmaxf(x, 0.0f); }
__device__ __forceinline__ double ramp(double x) { return fmax(x, 0.0); }
template <class R>
__device__ __forceinline__ R m4_shape(R q) {
    const R t = ramp(R(2) - q), s = ramp(R(1) - q);
    return R(0.25) * t * t * t - s * s * s;
}
template <class R>
__device__ __forceinline__ R m4_slope(R q) {
    const R t = ramp(R(2) - q), s = ramp(R(1) - q);
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
// With P64 the pressure is evaluated in FP64 from rho64, the FP64 density
// (mixed path): P ~ A (rho/rho0 - 1) magnifies a relative density error by
// 1/|rho/rho0 - 1|, which is ~1e4 for iron at 10 MPa.
template <class R, bool P64 = false>
__device__ __forceinline__ void eos_epilogue(const Carried<R>& c, const Work<R>& w, int64_t i, const SegPar& p,
                                             int mat, R rho, R h, R ox, R oy, R oz, double rho64 = 0.0) {
    using V4 = typename Vec4<R>::T;
    const int64_t n = w.n;
    const MatT<R> m = w.mats[mat];
    const R u = (R)c.u[i];
    const R P = P64 ? (R)eos_pressure<double>(w.mats64[mat], rho64, c.u[i]) : eos_pressure(m, rho, u);
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
    w.frec_pos[i] = a;
    w.frec_vel[i] = b;
    w.frec_th[i] = t;
    w.frec_sa[i] = sa;
    w.frec_sb[i] = sb;
}

// Raise the upper bound of h of particle i's slice block, which the force
// stencil culls with, to its final h (rounded up to float; positive floats
// order as integers), for the valid lanes.  Called after the EOS epilogue by
// every lane of the wave: one atomic per block the wave touches.
template <class R>
__device__ __forceinline__ void raise_h_bound(const Work<R>& w, const SegGrid& g, uint32_t cc, double h,
                                              bool valid) {
    const int cx = cc & 1023u, cy = (cc >> 10) & 1023u, cz = cc >> 20;
    float hf = (float)h;
    if ((double)hf < h) hf = __uint_as_float(__float_as_uint(hf) + 1u);
    wave_max_at(w.cell_hmax, valid ? block_key(g, cx, cy, cz) : 0u, hf, valid);
}

// ---------------------------------------------------------------- neighbour lists
// Slot of quad q (entries 4q..4q+3) of particle i's list (see Work::nbr).
__device__ __forceinline__ size_t nbr_quad(int64_t i, uint32_t q, int cap) {
    return ((size_t)(i >> 6) * (size_t)(cap >> 2) + (size_t)q) * 64u + (size_t)(i & 63);
}
__device__ __forceinline__ uint32_t quad_get(const uint4& v, int s) {
    return s == 0 ? v.x : (s == 1 ? v.y : (s == 2 ? v.z : v.w));
}

// Position of particle i relative to the frame of neighbour k: FP32 path, the
// own frame offset shifted by the frame delta (cells along x and y, slice
// blocks along z), taking on a periodic axis the image the stencil walk uses;
// FP64 path, absolute position.
template <class R, bool PER>
__device__ __forceinline__ void rel_base(const Work<R>& w, const SegGrid& g, uint32_t k, int cx, int cy, int cz,
                                         R ox, R oy, R oz, R ax, R ay, R az, R& bx, R& by, R& bz) {
    if (sizeof(R) == 4) {
        const uint32_t ck = w.cellc[k];
        const int cj[3] = {(int)(ck & 1023u), (int)((ck >> 10) & 1023u), (int)(ck >> 20)};
        const int ci[3] = {cx, cy, cz};
        R sh[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int d = cj[a] - ci[a];
            int wrap = 0;                  // cells moved to reach the image
            if (PER && g.per[a]) {
                const int m = g.dim[a], lo = -((m - 1) / 2), hi = m - 1 - (m - 1) / 2;
                if (d < lo) wrap = m;
                else if (d > hi) wrap = -m;
            }
            const int fj = (a == 2) ? block_start(cj[a]) : cj[a], fi = (a == 2) ? block_start(ci[a]) : ci[a];
            sh[a] = (R)(fj - fi + wrap) * (R)g.cw[a];
        }
        bx = ox - sh[0]; by = oy - sh[1]; bz = oz - sh[2];
    } else {
        bx = ax; by = ay; bz = az;
    }
}

// ---------------------------------------------------------------- density
// Per-particle constants of one density iteration.
template <class R, bool PER>
struct DensLane {
    using V4 = typename Vec4<R>::T;
    V4 me;                 // own density record
    R h, ox, oy, oz, two_h, th2, inv_h, sig;
    R ozs;                 // offset from the own slice's corner (distance culling)
    int cx, cy, cz;
    bool emu;              // 2h may exceed the reference's cell width this round
    bool per;              // some axis of the segment is periodic
    int ob[3];             // the reference's cell of i (emulation only)
    const SegGrid* gp;

    __device__ void setup(const Carried<R>& c, const SegPar& p, const SegGrid& g, int64_t i, uint32_t cc) {
        cx = cc & 1023u; cy = (cc >> 10) & 1023u; cz = cc >> 20;
        ozs = slice_offset<R>(g, cz, oz);
        gp = &g;
        per = PER && (g.per[0] || g.per[1] || g.per[2]);
        two_h = R(2) * h;
        th2 = two_h * two_h;
        inv_h = R(1) / h;
        sig = inv_h * inv_h * inv_h * R(1.0 / kPi);
        emu = (double)two_h > 0.999 * g.ocw_min;
        ob[0] = ob[1] = ob[2] = 0;
        if (emu) {
            ob[0] = oracle_bin(g, p, 0, c.x[i]);
            ob[1] = oracle_bin(g, p, 1, c.y[i]);
            ob[2] = oracle_bin(g, p, 2, c.z[i]);
        }
    }
    // contribution of one candidate; FP32 tests r^2 (W is continuous at 2h),
    // FP64 tests r < 2h exactly as the reference does
    // (a slot that is not `live` adds exactly zero)
    __device__ __forceinline__ R term(const V4& q, R bx, R by, R bz, bool live = true) const {
        R rx = bx - q.x, ry = by - q.y, rz = bz - q.z;
        if (PER && per) min_image(*gp, rx, ry, rz);
        const R r2 = rx * rx + ry * ry + rz * rz;
        const R rr = f_sqrt(r2);
        // `&`, not `&&`: a short-circuit lets the compiler branch around (and sink
        // the record load into) every slot's test
        const bool in = live & ((sizeof(R) == 4) ? (r2 < th2) : (rr < two_h));
        // the mass multiplies unconditionally, so the record stays one 16-byte load
        return q.w * (in ? sig * m4_shape(rr * inv_h) : R(0));
    }
    // frame of a visited piece (shifts from for_pieces)
    __device__ __forceinline__ void frame(R sx, R sy, R sz, R& bx, R& by, R& bz) const {
        if (sizeof(R) == 4) {
            bx = ox - sx; by = oy - sy; bz = oz - sz;
        } else {
            bx = me.x; by = me.y; bz = me.z;
        }
    }
};

// Frame shift of a visited piece in whole cells (x, y) and slices (z, for the
// slots below and from the piece's split), for walkers that form the shift in a
// precision of their own: shift = multiple * cell width.
struct PieceFrame {
    int ix, iy, iz0, iz1;
};

// Floor to int in the working precision.
__device__ __forceinline__ int ifloor(float x) { return (int)floorf(x); }
__device__ __forceinline__ int ifloor(double x) { return (int)floor(x); }
// Slices added at both ends of a column's z range so that rounding never
// drops a slice within reach (the extra candidates fail the range test).
constexpr float kSliceMargin = 1e-3f;

// Visit the candidates of particle i as pieces: runs [k0, k1) of the sorted
// order with one frame shift (sx, sy) across x and y and, along z, sz0 for
// slots below ksplit and sz1 from ksplit on (see AxisWalk);
// fn(k0, k1, ksplit, sx, sy, sz0, sz1, reach, frame) per piece (frame: the
// same shifts as whole multiples of the cell widths).  reach_max bounds
// the reach; with HB the reach of a piece is reach_of(its bound of h) and the
// piece is culled with it (else reach is reach_max).  (ox, oy) are offsets
// from the own column's corner, ozs from the own slice's corner.
//
// Without periodic axes the walk goes column by column: every column whose
// nearest point is within reach, and in it one contiguous run of the slices
// within reach along z (at most two slice blocks within the tuning limits,
// hence one frame split).  A wave's
// lanes sit in one column and walk the same column sequence with runs of
// similar length (MIR mirrors the sequence along y for half of them, below), and the
// z range is cut to the lane's own reach rather than to whole cubic cells.
// Periodic segments walk cell by cell.
template <class R, bool PER, bool HB, bool MIR = false, class RF, class F>
__device__ __forceinline__ void for_pieces(const SegGrid& g, const Work<R>& w, int cx, int cy, int cz, R ox, R oy,
                                           R ozs, R reach_max, RF&& reach_of, F&& fn) {
    const R rmax2 = reach_max * reach_max;
    if constexpr (!PER) {
        // grid fields in registers; the reach trouble flag is raised once after
        // the walk (an atomic inside the loops would make every column reload them)
        const WalkGrid G(g);
        bool far = false;
        const R wx = (R)g.cw[0], wy = (R)g.cw[1], wz = (R)g.cw[2], iwz = (R)g.inv_cw[2];
        int rx = (int)ceil((double)reach_max * g.inv_cw[0]), ry = (int)ceil((double)reach_max * g.inv_cw[1]);
        rx = rx < 1 ? 1 : (rx > kMaxStencil ? kMaxStencil : rx);
        ry = ry < 1 ? 1 : (ry > kMaxStencil ? kMaxStencil : ry);
        const int fi = block_start(cz), zmax = G.dim[2] - 1;
        // With MIR, lanes in the upper half of their column along y take the
        // columns in mirrored order along y: at each step of the walk the wave's
        // lanes then sit in (at most two) columns where they have similar numbers
        // of pairs in range, which evens out the turns of a pair body that runs
        // over the pairs in range only.  Along y only: the two columns a step
        // meets are then close in memory (y runs fastest inside a column tile),
        // and mirroring along x as well spreads a wave's loads over four.
        const int mx = 1, my = (MIR && oy >= R(0.5) * wy) ? -1 : 1;
        for (int ex = -rx; ex <= rx; ex++) {
            const int dx = mx * ex;
            const int nx = cx + dx;
            if ((unsigned)nx >= (unsigned)G.dim[0]) continue;
            const R gx = cell_gap(dx, ox, wx);
            const R gx2 = gx * gx;
            if (gx2 >= rmax2) continue;
            for (int ey = -ry; ey <= ry; ey++) {
                const int dy = my * ey;
                const int ny = cy + dy;
                if ((unsigned)ny >= (unsigned)G.dim[1]) continue;
                const R gy = cell_gap(dy, oy, wy);
                const R gxy = gx2 + gy * gy;
                if (gxy >= rmax2) continue;
                const uint32_t col = cell_key(G, nx, ny, 0);
                R rz = f_sqrt(rmax2 - gxy);
                int z0 = cz + ifloor((ozs - rz) * iwz - (R)kSliceMargin);
                int z1 = cz + ifloor((ozs + rz) * iwz + (R)kSliceMargin);
                z0 = z0 < 0 ? 0 : z0;
                z1 = z1 > zmax ? zmax : z1;
                if (z0 > z1) continue;
                const int f0 = block_start(z0), f1 = block_start(z1);
                if (f1 - f0 > kSliceBlock) {
                    // more than two blocks: only a smoothing length far beyond 2 h_init
                    // set from outside reaches this far; report it instead of
                    // summing in the wrong frame
                    far = true;
                    continue;
                }
                R rb = reach_max;
                if constexpr (HB) {
                    // the larger bound of h of the (at most two) blocks met
                    const uint32_t cb = block_key(G, nx, ny, 0);
                    const R hb0 = (R)w.cell_hmax[cb + f0 / kSliceBlock];
                    const R hb1 = (f1 != f0) ? (R)w.cell_hmax[cb + f1 / kSliceBlock] : hb0;
                    rb = reach_of(hb0 > hb1 ? hb0 : hb1);
                    const R rb2 = rb * rb;
                    if (gxy >= rb2) continue;
                    rz = f_sqrt(rb2 - gxy);
                    const int qa = cz + ifloor((ozs - rz) * iwz - (R)kSliceMargin);
                    const int qe = cz + ifloor((ozs + rz) * iwz + (R)kSliceMargin);
                    z0 = qa > z0 ? qa : z0;
                    z1 = qe < z1 ? qe : z1;
                    if (z0 > z1) continue;
                }
                const uint32_t k0 = w.cstart[col + z0], k1 = w.cstart[col + z1 + 1];
                if (k0 >= k1) continue;
                const int g0 = block_start(z0), g1 = block_start(z1);
                const uint32_t ks = (g1 != g0) ? w.cstart[col + g1] : k1;
                fn(k0, k1, ks, (R)dx * wx, (R)dy * wy, (R)(g0 - fi) * wz, (R)(g1 - fi) * wz, rb,
                   PieceFrame{dx, dy, g0 - fi, g1 - fi});
            }
        }
        if (far) atomicOr(&w.st[&g - w.grid].bad, 8);
        return;
    }
    AxisWalk<R, PER> X, Y, Z;
    X.init(g, 0, cx, ox, (double)reach_max);
    Y.init(g, 1, cy, oy, (double)reach_max);
    Z.init(g, 2, cz, ozs, (double)reach_max);
    for (int dx = X.lo; dx <= X.hi; dx++) {
        int nx, mx;
        R sx, gx;
        if (!X.cell(dx, nx, sx, mx, gx)) continue;
        const R gx2 = gx * gx;
        if (gx2 >= rmax2) continue;
        for (int dy = Y.lo; dy <= Y.hi; dy++) {
            int ny, my;
            R sy, gy;
            if (!Y.cell(dy, ny, sy, my, gy)) continue;
            const R gxy = gx2 + gy * gy;
            if (gxy >= rmax2) continue;
            for (int dz = Z.lo; dz <= Z.hi; dz++) {
                int nz, mz;
                R sz, gz;
                if (!Z.cell(dz, nz, sz, mz, gz)) continue;
                const R gg = gxy + gz * gz;
                if (gg >= rmax2) continue;
                R rb = reach_max;
                if constexpr (HB) {
                    rb = reach_of((R)w.cell_hmax[block_key(g, nx, ny, nz)]);
                    if (gg >= rb * rb) continue;
                }
                const uint32_t key = cell_key(g, nx, ny, nz);
                const uint32_t k0 = w.cstart[key], k1 = w.cstart[key + 1];
                if (k0 < k1) fn(k0, k1, k1, sx, sy, sz, sz, rb, PieceFrame{mx, my, mz, mz});
            }
        }
    }
}

// Sum over the candidates within 2h of particle i (piece order, sorted order
// inside a piece).
template <class R, bool PER>
__device__ R density_walk(const Carried<R>& c, const Work<R>& w, const SegPar& p, const SegGrid& g,
                          const DensLane<R, PER>& L) {
    using V4 = typename Vec4<R>::T;
    R s = R(0);
    for_pieces<R, PER, false>(g, w, L.cx, L.cy, L.cz, L.ox, L.oy, L.ozs, L.two_h, [](R) { return R(0); },
                              [&](uint32_t k0, uint32_t k1, uint32_t ks, R sx, R sy, R sz0, R sz1, R,
                                  const PieceFrame&) {
        R bx, by, bz0, bz1;
        L.frame(sx, sy, sz0, bx, by, bz0);
        L.frame(sx, sy, sz1, bx, by, bz1);
        if (!L.emu) {
            // four loads in flight, also for a piece's last partial group: its
            // spare slots read the records that follow (the record arrays hold
            // kRecordPad zeroed records past the last particle) and add zero
            for (uint32_t k = k0; k < k1; k += 4) {
                const V4* rec = w.dens_rec + k;
                const V4 q0 = rec[0], q1 = rec[1], q2 = rec[2], q3 = rec[3];
                s += L.term(q0, bx, by, k < ks ? bz0 : bz1);
                s += L.term(q1, bx, by, k + 1 < ks ? bz0 : bz1, k + 1 < k1);
                s += L.term(q2, bx, by, k + 2 < ks ? bz0 : bz1, k + 2 < k1);
                s += L.term(q3, bx, by, k + 3 < ks ? bz0 : bz1, k + 3 < k1);
            }
        } else {
            for (uint32_t k = k0; k < k1; k++) {
                if (!oracle_reaches(g, p, L.ob, c.x[k], c.y[k], c.z[k])) continue;
                s += L.term(w.dens_rec[k], bx, by, k < ks ? bz0 : bz1);
            }
        }
    });
    return s;
}

// ---------------------------------------------------------------- FP64 density terms (mixed path)
// sqrt(r2) from the FP32 reciprocal square root and one Newton step in FP64:
// relative error ~1.5 e^2 for a seed error e ~1e-7 (r2 = 0 gives 0).
__device__ __forceinline__ double sqrt_seeded(double r2) {
    const float y = __builtin_amdgcn_rsqf(fmaxf((float)r2, 1e-30f));
    const double yd = (double)y;
    const double r0 = r2 * yd;
    return fma(fma(-r0, r0, r2), 0.5 * yd, r0);
}

// FP64 quantities of particle i for the FP64 density walk: h and the absolute
// position.
struct DensHP {
    double h, two_h, inv_h, x, y, z;
    __device__ __forceinline__ void setup(double hh) {
        h = hh;
        two_h = 2.0 * hh;
        inv_h = 1.0 / hh;
    }
};

// sum_j m_j f(r_ij / h_i) over r_ij < 2 h_i with FP64 separations x_i - x_j
// (as the reference forms them), kernel shapes, masses and sum; rho_i = sum /
// (pi h_i^3).  Same candidates and order as density_walk: the walk and an FP32
// screen of its records (range widened by 2e-5 in r^2) pick the candidates, the
// FP64 test r < 2h (sph.hpp:191) decides.  Terms of the candidates screened out
// are exact zeros, so they are skipped.
template <bool PER>
__device__ double density_walk_hp(const Carried<float>& c, const Work<float>& w, const SegPar& p,
                                  const SegGrid& g, const DensLane<float, PER>& L, const DensHP& H) {
    double s = 0.0;
    auto term = [&](const double4& q) -> double {
        double rx = H.x - q.x, ry = H.y - q.y, rz = H.z - q.z;
        if (PER && L.per) min_image(g, rx, ry, rz);
        const double r = sqrt_seeded(rx * rx + ry * ry + rz * rz);
        const double qq = r * H.inv_h;
        const double t = fmax(2.0 - qq, 0.0), sv = fmax(1.0 - qq, 0.0);
        const double fv = 0.25 * t * t * t - sv * sv * sv;
        return r < H.two_h ? q.w * fv : 0.0;
    };
    for_pieces<float, PER, false>(g, w, L.cx, L.cy, L.cz, L.ox, L.oy, L.ozs, L.two_h * 1.00001f,
                                  [](float) { return 0.0f; },
                                  [&](uint32_t k0, uint32_t k1, uint32_t ks, float sx, float sy, float sz0, float sz1,
                                      float, const PieceFrame&) {
        float bx, by, bz0, bz1;
        L.frame(sx, sy, sz0, bx, by, bz0);
        L.frame(sx, sy, sz1, bx, by, bz1);
        if (!L.emu) {
            // In chunks of up to 64 candidates: the screen (four records in flight;
            // the record arrays hold kRecordPad records past the last particle, so a
            // group's loads never leave them, and slots past the chunk never count)
            // sets a bit per candidate that may be in range, then the FP64 terms run
            // over the set bits two at a time, in increasing k, both records loaded
            // before either term (the second term adds exactly zero when the chunk
            // has one bit left, and the sum is never -0).
            const float th2 = L.th2 * 1.00002f;
            for (uint32_t c0 = k0; c0 < k1; c0 += 64u) {
                const uint32_t c1 = min(c0 + 64u, k1);
                uint64_t mask = 0;
                for (uint32_t k = c0; k < c1; k += 4) {
                    float4 q[4];
#pragma unroll
                    for (int b = 0; b < 4; b++) q[b] = rec_at(w.dens_rec + b, k);
                    uint32_t nib = 0;
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        float rx = bx - q[b].x, ry = by - q[b].y, rz = (k + b < ks ? bz0 : bz1) - q[b].z;
                        if (PER && L.per) min_image(g, rx, ry, rz);
                        nib |= (uint32_t)((k + b < c1) & (rx * rx + ry * ry + rz * rz < th2)) << b;
                    }
                    mask |= (uint64_t)nib << (k - c0);
                }
                while (mask) {
                    const uint32_t ba = (uint32_t)__builtin_ctzll(mask);
                    mask &= mask - 1;
                    const bool two = mask != 0;
                    const uint32_t bb = two ? (uint32_t)__builtin_ctzll(mask) : ba;
                    mask &= mask - 1;
                    const double4 qa = rec_at(w.dens_rec64, c0 + ba), qb = rec_at(w.dens_rec64, c0 + bb);
                    const double ta = term(qa), tb = term(qb);
                    s += ta;
                    s += two ? tb : 0.0;
                }
            }
        } else {
            for (uint32_t k = k0; k < k1; k++) {
                if (!oracle_reaches(g, p, L.ob, c.x[k], c.y[k], c.z[k])) continue;
                s += term(w.dens_rec64[k]);
            }
        }
    });
    return s;
}

// ---------------------------------------------------------------- moment walks
// (DensMom, kDensBand and kDensBandRel: moments.h)

// Counters of the moment walks, kept only when Work::stats is set: lanes of
// density rounds, lanes whose second iteration walked again, and waves with
// such a lane, added per wave to the segment of the wave's first lane.  Every
// lane of the wave must call this.
template <class R>
__device__ __forceinline__ void count_rewalks(const Work<R>& w, int seg, bool act, bool rewalk) {
    if (!w.stats) return;
    const unsigned long long na = __ballot(act), nr = __ballot(rewalk);
    const int s0 = __shfl(seg, 0);
    if (lane_id() == 0) {
        SegState& S = w.st[s0];
        atomicAdd(&S.dens_lanes, (unsigned long long)__popcll(na));
        if (nr) {
            atomicAdd(&S.dens_rewalks, (unsigned long long)__popcll(nr));
            atomicAdd(&S.dens_rewalk_waves, 1ull);
        }
        if (na) atomicAdd(&S.dens_waves, 1ull);
    }
}

// Lane slots of the band candidates: slot s of the lane at threadIdx.x t sits
// at [s * blockDim.x + t] of the block's dynamic LDS.
__device__ __forceinline__ uint32_t* dens_band_slots() {
    extern __shared__ uint32_t dsph_dyn_lds[];
    return dsph_dyn_lds + threadIdx.x;
}
constexpr size_t dens_band_lds(int block) { return (size_t)kDensBand * (size_t)block * sizeof(uint32_t); }

// density_walk_hp with the moments of the walk (reach and screen widened to
// 2h (1 + d) for the outer band), in the force walk's column order (mirrored
// along y for the lanes in the upper half of their column); fills M and the
// band slots.  Lanes that emulate the reference's cells do not come here.
template <bool PER>
__device__ void density_walk_hp_mom(const Work<float>& w, const SegGrid& g, const DensLane<float, PER>& L,
                                    const DensHP& H, DensMom& M, uint32_t* band) {
    M.clear();
    const uint32_t bstride = blockDim.x;
    auto add = [&](const double4& q, uint32_t k, bool live) {
        double rx = H.x - q.x, ry = H.y - q.y, rz = H.z - q.z;
        if (PER && L.per) min_image(g, rx, ry, rz);
        const double r2 = rx * rx + ry * ry + rz * rz;
        const double r = sqrt_seeded(r2);
        const double qq = r * H.inv_h;
        if (M.add(live ? q.w : 0.0, r, r2, qq, H.two_h) & live) {
            if (M.nb < kDensBand) band[M.nb * bstride] = k;
            M.nb++;
            M.sb += r < H.two_h ? q.w * m4_shape(qq) : 0.0;
        }
    };
    const float grow = (float)(1.0 + kDensBandRel);
    for_pieces<float, PER, false, true>(g, w, L.cx, L.cy, L.cz, L.ox, L.oy, L.ozs, L.two_h * grow * 1.00001f,
                                  [](float) { return 0.0f; },
                                  [&](uint32_t k0, uint32_t k1, uint32_t ks, float sx, float sy, float sz0, float sz1,
                                      float, const PieceFrame&) {
        float bx, by, bz0, bz1;
        L.frame(sx, sy, sz0, bx, by, bz0);
        L.frame(sx, sy, sz1, bx, by, bz1);
        const float th2 = L.th2 * (grow * grow * 1.00002f);
        // Chunks of 32 candidates (32-bit masks): a run longer than a chunk costs
        // the wave the turns of both chunks' busiest lanes, but the mask
        // bookkeeping per pair turn is half that of 64-bit masks.  The terms are
        // added in increasing k either way (the spare second term of a chunk's
        // last turn adds exact zeros), so the sums do not depend on the chunking.
        for (uint32_t c0 = k0; c0 < k1; c0 += 32u) {
            const uint32_t c1 = min(c0 + 32u, k1);
            uint32_t mask = 0;
            for (uint32_t k = c0; k < c1; k += 4) {
                float4 q[4];
#pragma unroll
                for (int b = 0; b < 4; b++) q[b] = rec_at(w.dens_rec + b, k);
                uint32_t nib = 0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    float rx = bx - q[b].x, ry = by - q[b].y, rz = (k + b < ks ? bz0 : bz1) - q[b].z;
                    if (PER && L.per) min_image(g, rx, ry, rz);
                    nib |= (uint32_t)((k + b < c1) & (rx * rx + ry * ry + rz * rz < th2)) << b;
                }
                mask |= nib << (k - c0);
            }
            while (mask) {
                const uint32_t ba = (uint32_t)__builtin_ctz(mask);
                mask &= mask - 1;
                const bool two = mask != 0;
                const uint32_t bb = two ? (uint32_t)__builtin_ctz(mask) : ba;
                mask &= mask - 1;
                const double4 qa = rec_at(w.dens_rec64, c0 + ba), qb = rec_at(w.dens_rec64, c0 + bb);
                add(qa, c0 + ba, true);
                add(qb, c0 + bb, two);
            }
        }
    });
}

// The band candidates' terms at the h of H (FP64 records, as the walk formed r).
template <bool PER>
__device__ double dens_band_hp(const Work<float>& w, const SegGrid& g, const DensLane<float, PER>& L,
                               const DensHP& H, const DensMom& M, const uint32_t* band) {
    double s = 0.0;
    const uint32_t bstride = blockDim.x;
    for (int b = 0; b < M.nb; b++) {
        const double4 q = rec_at(w.dens_rec64, band[b * bstride]);
        double rx = H.x - q.x, ry = H.y - q.y, rz = H.z - q.z;
        if (PER && L.per) min_image(g, rx, ry, rz);
        const double r = sqrt_seeded(rx * rx + ry * ry + rz * rz);
        s += r < H.two_h ? q.w * m4_shape(r * H.inv_h) : 0.0;
    }
    return s;
}

// The FP64 path's walk (density_walk<double>, every slot, four records in
// flight) with the moments of the walk, reach widened to 2h (1 + d); fills M
// and the band slots.  Lanes that emulate the reference's cells do not come
// here.
template <bool PER>
__device__ void density_walk_mom(const Work<double>& w, const SegGrid& g, const DensLane<double, PER>& L,
                                 DensMom& M, uint32_t* band) {
    M.clear();
    const uint32_t bstride = blockDim.x;
    auto add = [&](const double4& q, uint32_t k, bool live) {
        double rx = L.me.x - q.x, ry = L.me.y - q.y, rz = L.me.z - q.z;
        if (PER && L.per) min_image(g, rx, ry, rz);
        const double r2 = rx * rx + ry * ry + rz * rz;
        const double r = sqrt(r2);
        const double qq = r * L.inv_h;
        if (M.add(live ? q.w : 0.0, r, r2, qq, L.two_h) & live) {
            if (M.nb < kDensBand) band[M.nb * bstride] = k;
            M.nb++;
            M.sb += r < L.two_h ? q.w * m4_shape(qq) : 0.0;
        }
    };
    for_pieces<double, PER, false>(g, w, L.cx, L.cy, L.cz, L.ox, L.oy, L.ozs, L.two_h * (1.0 + kDensBandRel),
                                   [](double) { return 0.0; },
                                   [&](uint32_t k0, uint32_t k1, uint32_t, double, double, double, double, double,
                                       const PieceFrame&) {
        // slots past the run read the records that follow (kRecordPad zeroed
        // records past the last particle) and add nothing
        for (uint32_t k = k0; k < k1; k += 4) {
            const double4 q0 = rec_at(w.dens_rec, k), q1 = rec_at(w.dens_rec + 1, k), q2 = rec_at(w.dens_rec + 2, k),
                          q3 = rec_at(w.dens_rec + 3, k);
            add(q0, k, true);
            add(q1, k + 1, k + 1 < k1);
            add(q2, k + 2, k + 2 < k1);
            add(q3, k + 3, k + 3 < k1);
        }
    });
}

// The band candidates' terms at the lane's current h (FP64 path).
template <bool PER>
__device__ double dens_band64(const Work<double>& w, const SegGrid& g, const DensLane<double, PER>& L,
                              const DensMom& M, const uint32_t* band) {
    double s = 0.0;
    const uint32_t bstride = blockDim.x;
    for (int b = 0; b < M.nb; b++) {
        const double4 q = rec_at(w.dens_rec, band[b * bstride]);
        double rx = L.me.x - q.x, ry = L.me.y - q.y, rz = L.me.z - q.z;
        if (PER && L.per) min_image(g, rx, ry, rz);
        const double r = sqrt(rx * rx + ry * ry + rz * rz);
        s += r < L.two_h ? q.w * m4_shape(r * L.inv_h) : 0.0;
    }
    return s;
}

// The new h from an FP64 density (sph.hpp:195-198), in FP64.
__device__ __forceinline__ double density_h64(const SegPar& p, double m, double rho) {
    const double h0 = p.h_init;
    return smin(2.0 * h0, smax(0.5 * h0, p.eta * cbrt(m / rho)));
}

// The new h of particle i from its density (sph.hpp:195-198).
template <class R>
__device__ __forceinline__ R density_h(const Carried<R>& c, const SegPar& p, int64_t i, R rho) {
    const R h0 = (R)p.h_init;
    const R mi = (R)c.mass[i];
    return smin(R(2) * h0, smax(R(0.5) * h0, (R)p.eta * d_cbrt(mi / rho)));
}

// Does particle i read its neighbour list in the second density iteration of
// round 0?  Only while its h has not outgrown the list skin.
template <class R>
__device__ __forceinline__ bool dens_list_usable(const Work<R>& w, int64_t i, R h) {
    return w.nbr_cnt[i] != kNbrOverflow && h <= w.skin_check * w.h0[i];
}

// Store the new density and h; the last iteration of a round also flags h
// growth beyond the list skin and runs the EOS epilogue.
template <class R, bool LAST>
__device__ __forceinline__ R density_store(const Carried<R>& c, const Work<R>& w, int64_t i, int seg,
                                           const SegPar& p, R rho, bool finite, int lists, R ox, R oy, R oz) {
    const R hn = density_h<R>(c, p, i, rho);
    w.rho[i] = rho;
    c.h[i] = hn;
    // lists built this step stay usable by the force pass only while every
    // final h of the segment is within the skin of its build-time value
    if (LAST && lists && finite && !(hn <= w.skin_check * w.h0[i])) atomicOr(&w.st[seg].list_bad, 1u);
    if (LAST) eos_epilogue<R>(c, w, i, p, info_mat(c.info[i]), rho, hn, ox, oy, oz);
    return hn;
}

// One density iteration of round r by walking the cells (sph.hpp:189-201):
//   rho_i = max(sum_j m_j W(r_ij, h_i) over r_ij < 2 h_i, 1e-30)
//   h_i   = clamp(eta (m_i / rho_i)^(1/3), h_init / 2, 2 h_init)
// LAST marks the second iteration of the round: it also reduces max h for the
// round test and runs the EOS epilogue.  In the second iteration of round 0
// with lists built, lanes whose list is usable are left to k_density_list;
// the others walk and mark their list unusable for the force pass too.
template <class R, bool LAST, bool PER>
__global__ __launch_bounds__(256) void k_density(Carried<R> c, Work<R> w, int r, int remap, int lists,
                                                 SegTargets<1> tg) {
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    int seg = 0;
    bool act = false;
    if (i < w.n) {
        seg = info_seg(c.info[i]);
        act = w.st[seg].active[r] != 0;
    }
    const bool by_list = LAST && lists && r == 0;
    double hnew_d = 0.0;
    uint32_t cc = 0;
    bool fin = false;
    if (act) {
        const SegPar& p = w.par[seg];
        const SegGrid& g = w.grid[seg];
        DensLane<R, PER> L;
        L.h = c.h[i];
        if (by_list && dens_list_usable<R>(w, i, L.h)) {
            act = false;                           // k_density_list owns this particle
        } else {
            if (by_list) w.nbr_cnt[i] = kNbrOverflow;
            cc = w.cellc[i];
            L.me = w.dens_rec[i];
            own_offset<R>(c, g, i, cc, L.me, L.ox, L.oy, L.oz);
            const bool finite = isfinite((double)L.ox) && isfinite((double)L.oy) && isfinite((double)L.oz);
            R rho = rho_floor<R>();
            if (finite) {
                L.setup(c, p, g, i, cc);
                rho = smax(density_walk<R, PER>(c, w, p, g, L), rho_floor<R>());
            }
            hnew_d = (double)density_store<R, LAST>(c, w, i, seg, p, rho, finite, lists, L.ox, L.oy, L.oz);
            fin = finite;
        }
    }
    if (LAST) {
        raise_h_bound<R>(w, w.grid[seg], cc, hnew_d, act && fin);
        const double v[1] = {hnew_d};
        block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, act);
    }
}

// A density round that no segment runs: the whole workgroup leaves at once,
// after marking its partial-reduction record empty.  Every thread of the
// workgroup reads the same flag.
template <class R>
__device__ __forceinline__ bool round_idle(const Work<R>& w, int r) {
    if (w.st[0].any_active[r] != 0) return false;
    if (threadIdx.x == 0) w.parts[blockIdx.x].seg = -1;
    return true;
}

// Both iterations of density round r in one pass (sph.hpp:189-201): the sum of
// particle i uses only its own h, and both iterations of a round share the
// round's reference grid, so the second iteration needs nothing but the h the
// first one gives i.  The first iteration's rho and h stay in registers.  In
// FP64 the first iteration is a moment walk and the second walks again only
// where h left the moments' band (dynamic LDS: dens_band_lds(blockDim.x)
// bytes); in FP32 both iterations walk.  Used for every round except round 0
// with lists, whose iterations build and read the lists.  The FP32 variant is
// held at five waves per SIMD (96 VGPRs, a few epilogue values spilled): at its
// natural four waves the walk is slower.  SEG1 (here, in k_density_round_hp and
// in k_force): the context holds one segment, so its parameters and grid are
// read at index 0, which the compiler keeps in scalar registers instead of one
// copy per lane (4 % faster density, 1 % faster force pass at 4M particles).
template <class R, bool PER, bool SEG1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 5 : 1)))
void k_density_round(Carried<R> c, Work<R> w, int r, int remap, int lists, SegTargets<1> tg) {
    if (round_idle(w, r)) return;
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    uint32_t* band = dens_band_slots();
    (void)band;
    bool rewalk = false;
    int seg = 0;
    bool act = false;
    if (i < w.n) {
        seg = info_seg(c.info[i]);
        act = w.st[seg].active[r] != 0;
    }
    double hnew_d = 0.0;
    uint32_t cc = 0;
    bool fin = false;
    if (act) {
        const SegPar& p = w.par[SEG1 ? 0 : seg];
        const SegGrid& g = w.grid[SEG1 ? 0 : seg];
        DensLane<R, PER> L;
        cc = w.cellc[i];
        L.me = w.dens_rec[i];
        own_offset<R>(c, g, i, cc, L.me, L.ox, L.oy, L.oz);
        const bool finite = isfinite((double)L.ox) && isfinite((double)L.oy) && isfinite((double)L.oz);
        fin = finite;
        L.h = c.h[i];
        R rho = rho_floor<R>();
        // FP64: the first iteration is a moment walk (not for lanes that
        // emulate the reference's cells), and the second walks again only
        // where h left the moments' band
        DensMom M;
        bool mom = false;
        if (finite) {
            L.setup(c, p, g, i, cc);
            if constexpr (sizeof(R) == 8) {
                if (!L.emu) {
                    density_walk_mom<PER>(w, g, L, M, band);
                    rho = smax((M.at(L.h, L.h) + M.sb) * L.sig, rho_floor<R>());
                    mom = true;
                }
            }
            if (!mom) rho = smax(density_walk<R, PER>(c, w, p, g, L), rho_floor<R>());
        }
        const R h0 = L.h;
        L.h = density_h<R>(c, p, i, rho);
        rho = rho_floor<R>();
        if (finite) {
            L.setup(c, p, g, i, cc);
            bool done = false;
            if constexpr (sizeof(R) == 8) {
                if (mom && !L.emu && M.covers(h0, L.h)) {
                    rho = smax((M.at(h0, L.h) + dens_band64<PER>(w, g, L, M, band)) * L.sig, rho_floor<R>());
                    done = true;
                }
            }
            if (!done) {
                rewalk = sizeof(R) == 8;
                rho = smax(density_walk<R, PER>(c, w, p, g, L), rho_floor<R>());
            }
        }
        hnew_d = (double)density_store<R, true>(c, w, i, seg, p, rho, finite, lists, L.ox, L.oy, L.oz);
    }
    if constexpr (sizeof(R) == 8) count_rewalks(w, seg, act, rewalk);
    raise_h_bound<R>(w, w.grid[SEG1 ? 0 : seg], cc, hnew_d, act && fin);
    const double v[1] = {hnew_d};
    block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, act);
}

// Both iterations of density round r for the mixed path with FP64 density
// terms (k_density_round with density_walk_hp): h of the first iteration stays
// in FP64 for the second, the density and the new h are FP64 until stored, and
// the pressure of the EOS epilogue is evaluated in FP64 from that density.
// Both iterations need the FP64 terms: an FP32 first iteration moves h by
// ~3e-7, which the second iteration's density passes on to P at up to ~1e-4
// of |P| near 10 MPa.  The first iteration is a moment walk, so the second
// walks again only where h moved out of the moments' band.  Neighbour lists
// are not used.  Dynamic LDS: dens_band_lds(blockDim.x) bytes.
template <bool PER, bool SEG1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_density_round_hp(Carried<float> c, Work<float> w, int r, int remap,
                                                         SegTargets<1> tg) {
    if (round_idle(w, r)) return;
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    uint32_t* band = dens_band_slots();
    int seg = 0;
    bool act = false;
    if (i < w.n) {
        seg = info_seg(c.info[i]);
        act = w.st[seg].active[r] != 0;
    }
    double hnew = 0.0;
    uint32_t cc = 0;
    bool fin = false, rewalk = false;
    if (act) {
        const SegPar& p = w.par[SEG1 ? 0 : seg];
        const SegGrid& g = w.grid[SEG1 ? 0 : seg];
        DensLane<float, PER> L;
        DensHP H;
        cc = w.cellc[i];
        L.me = w.dens_rec[i];
        own_offset<float>(c, g, i, cc, L.me, L.ox, L.oy, L.oz);
        const bool finite = isfinite(L.ox) && isfinite(L.oy) && isfinite(L.oz);
        fin = finite;
        H.x = c.x[i];
        H.y = c.y[i];
        H.z = c.z[i];
        const double m = c.mass[i];
        const double h0 = (double)c.h[i];
        double rho = 1e-30;
        // first iteration: a moment walk, unless the lane emulates the
        // reference's cells
        DensMom M;
        bool mom = false;
        if (finite) {
            L.h = (float)h0;
            L.setup(c, p, g, i, cc);
            H.setup(h0);
            double s;
            if (!L.emu) {
                density_walk_hp_mom<PER>(w, g, L, H, M, band);
                s = M.at(h0, h0) + M.sb;
                mom = true;
            } else {
                s = density_walk_hp<PER>(c, w, p, g, L, H);
            }
            rho = smax(s * (H.inv_h * H.inv_h * H.inv_h / kPi), 1e-30);
        }
        const double h1 = density_h64(p, m, rho);
        // second iteration: from the moments while h1 stays in their band
        rho = 1e-30;
        if (finite) {
            L.h = (float)h1;
            L.setup(c, p, g, i, cc);
            H.setup(h1);
            double s;
            if (mom && !L.emu && M.covers(h0, h1)) {
                s = M.at(h0, h1) + dens_band_hp<PER>(w, g, L, H, M, band);
            } else {
                rewalk = true;
                s = density_walk_hp<PER>(c, w, p, g, L, H);
            }
            rho = smax(s * (H.inv_h * H.inv_h * H.inv_h / kPi), 1e-30);
        }
        const double h = density_h64(p, m, rho);
        hnew = h;
        w.rho[i] = (float)rho;
        c.h[i] = (float)h;
        eos_epilogue<float, true>(c, w, i, p, info_mat(c.info[i]), (float)rho, (float)h, L.ox, L.oy, L.oz, rho);
    }
    count_rewalks(w, seg, act, rewalk);
    raise_h_bound<float>(w, w.grid[SEG1 ? 0 : seg], cc, hnew, act && fin);
    const double v[1] = {hnew};
    block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, act);
}

// Second density iteration of round 0 for the particles whose list is usable:
// the list holds the walk's pairs in the walk's order, so the sum is the one
// the walk would give.  Four entries are read ahead so their loads overlap.
template <class R, bool PER>
__global__ __launch_bounds__(256) void k_density_list(Carried<R> c, Work<R> w, int remap, SegTargets<1> tg) {
    using V4 = typename Vec4<R>::T;
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    int seg = 0;
    bool act = false;
    if (i < w.n) {
        seg = info_seg(c.info[i]);
        act = w.st[seg].active[0] != 0 && dens_list_usable<R>(w, i, c.h[i]);
    }
    double hnew_d = 0.0;
    uint32_t cc = 0;
    bool fin = false;
    if (act) {
        const SegPar& p = w.par[seg];
        const SegGrid& g = w.grid[seg];
        DensLane<R, PER> L;
        L.h = c.h[i];
        cc = w.cellc[i];
        L.me = w.dens_rec[i];
        own_offset<R>(c, g, i, cc, L.me, L.ox, L.oy, L.oz);
        fin = isfinite((double)L.ox);
        L.setup(c, p, g, i, cc);
        const uint32_t nc = w.nbr_cnt[i];
        const int cap = w.nbr_cap;
        R s = R(0);
        auto add = [&](uint32_t k, const V4& q) {
            if (L.emu && !oracle_reaches(g, p, L.ob, c.x[k], c.y[k], c.z[k])) return;
            R bx, by, bz;
            rel_base<R, PER>(w, g, k, L.cx, L.cy, L.cz, L.ox, L.oy, L.oz, L.me.x, L.me.y, L.me.z, bx, by, bz);
            s += L.term(q, bx, by, bz);
        };
        for (uint32_t e = 0; e < nc; e += 4) {
            const uint4 kk = w.nbr[nbr_quad(i, e >> 2, cap)];
            const V4 q0 = w.dens_rec[kk.x], q1 = w.dens_rec[kk.y], q2 = w.dens_rec[kk.z], q3 = w.dens_rec[kk.w];
            add(kk.x, q0);
            if (e + 1 < nc) add(kk.y, q1);
            if (e + 2 < nc) add(kk.z, q2);
            if (e + 3 < nc) add(kk.w, q3);
        }
        const R rho = smax(s, rho_floor<R>());
        hnew_d = (double)density_store<R, true>(c, w, i, seg, p, rho, true, 1, L.ox, L.oy, L.oz);
    }
    raise_h_bound<R>(w, w.grid[seg], cc, hnew_d, act && fin);
    const double v[1] = {hnew_d};
    block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, act);
}

// First density iteration of round 0, which also writes the neighbour lists:
// while walking, every candidate j with
//   r_ij < (1+skin)(h_i + max(h_i, hblock_j))
// enters i's list in walk order (hblock_j: bound of h over j's slice block at
// the start of the step).  That covers r < 2h_i and r < h_i + h_j for as long
// as no h grows by more than the skin.  Lists that do not fit are marked.
template <class R, bool PER>
__global__ __launch_bounds__(256) void k_density_build(Carried<R> c, Work<R> w, int remap) {
    using V4 = typename Vec4<R>::T;
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    if (i >= w.n) return;
    const int seg = info_seg(c.info[i]);
    uint32_t cnt = kNbrOverflow;
    if (w.st[seg].active[0]) {
        const SegPar& p = w.par[seg];
        const SegGrid& g = w.grid[seg];
        DensLane<R, PER> L;
        L.h = c.h[i];
        const uint32_t cc = w.cellc[i];
        L.me = w.dens_rec[i];
        own_offset<R>(c, g, i, cc, L.me, L.ox, L.oy, L.oz);
        const bool finite = isfinite((double)L.ox) && isfinite((double)L.oy) && isfinite((double)L.oz);
        R rho = rho_floor<R>();
        if (finite) {
            L.setup(c, p, g, i, cc);
            const R sk = w.skin_build;
            const R hm = (R)bits_to_double(w.st[seg].hmax[0]);
            const int cap = w.nbr_cap;
            R s = R(0);
            cnt = 0u;
            // entries are gathered four at a time and stored as one 16-byte quad;
            // unused lanes of a last partial quad keep valid indices
            uint4 quad = make_uint4((uint32_t)i, (uint32_t)i, (uint32_t)i, (uint32_t)i);
            for_pieces<R, PER, true>(g, w, L.cx, L.cy, L.cz, L.ox, L.oy, L.ozs, sk * (L.h + smax(L.h, hm)),
                                     [&](R hb) { return sk * (L.h + smax(L.h, hb)); },
                                     [&](uint32_t k0, uint32_t k1, uint32_t ks, R sx, R sy, R sz0, R sz1,
                                         R rb, const PieceFrame&) {
                const R rb2 = rb * rb;
                R bx, by, bz0, bz1;
                L.frame(sx, sy, sz0, bx, by, bz0);
                L.frame(sx, sy, sz1, bx, by, bz1);
                auto visit = [&](uint32_t k, const V4& q) {
                    const R bz = k < ks ? bz0 : bz1;
                    R rx = bx - q.x, ry = by - q.y, rz = bz - q.z;
                    if (PER && L.per) min_image(g, rx, ry, rz);
                    const R r2 = rx * rx + ry * ry + rz * rz;
                    if (r2 < rb2) {
                        if (cnt < (uint32_t)cap) {
                            const uint32_t sl = cnt & 3u;
                            quad.x = sl == 0u ? k : quad.x;
                            quad.y = sl == 1u ? k : quad.y;
                            quad.z = sl == 2u ? k : quad.z;
                            quad.w = sl == 3u ? k : quad.w;
                            if (sl == 3u) w.nbr[nbr_quad(i, cnt >> 2, cap)] = quad;
                        }
                        cnt++;
                    }
                    if (L.emu && !oracle_reaches(g, p, L.ob, c.x[k], c.y[k], c.z[k])) return;
                    s += L.term(q, bx, by, bz);
                };
                uint32_t k = k0;
                for (; k + 1 < k1; k += 2) {         // two loads in flight
                    const V4 q0 = w.dens_rec[k], q1 = w.dens_rec[k + 1];
                    visit(k, q0);
                    visit(k + 1, q1);
                }
                if (k < k1) visit(k, w.dens_rec[k]);
            });
            if (cnt > (uint32_t)cap) {
                cnt = kNbrOverflow;
                atomicAdd(&w.st[seg].list_overflow, 1ull);
            } else if (cnt & 3u) {
                w.nbr[nbr_quad(i, cnt >> 2, cap)] = quad;
            }
            rho = smax(s, rho_floor<R>());
        }
        density_store<R, false>(c, w, i, seg, p, rho, finite, 0, L.ox, L.oy, L.oz);
    }
    w.nbr_cnt[i] = cnt;
}

// EOS epilogue on its own (System::compute_forces entry without a density pass).
template <class R, bool P64>
__global__ __launch_bounds__(256) void k_eos_records(Carried<R> c, Work<R> w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < w.n;
    int seg = 0;
    uint32_t cc = 0;
    double h = 0.0;
    bool fin = false;
    if (in) {
        const uint32_t inf = c.info[i];
        seg = info_seg(inf);
        const SegGrid& g = w.grid[seg];
        cc = w.cellc[i];
        R ox, oy, oz;
        own_offset<R>(c, g, i, cc, w.dens_rec[i], ox, oy, oz);
        eos_epilogue<R, P64>(c, w, i, w.par[seg], info_mat(inf), w.rho[i], c.h[i], ox, oy, oz, (double)w.rho[i]);
        h = (double)c.h[i];
        fin = isfinite((double)ox);
    }
    raise_h_bound<R>(w, w.grid[seg], cc, h, fin);
}

// ---------------------------------------------------------------- forces
constexpr int kForceBatch = 4;   // position records in flight per lane (force cell walk)
static_assert(kForceBatch <= kRecordPad + 1, "a screen group reads at most kRecordPad records past a run");
// Fused strain rate + forces (sph.hpp:212-258 and 277-317; the Metal
// strain_forces kernel, sph_force.metal:308-365, fuses the same two loops).
// For each pair j != i with 0 < r < 2 hbar, hbar = (h_i + h_j)/2:
//   L_ab  += (m_j/rho_j) (v_j - v_i)_a (grad_i W)_b
//   term   = Pf_i/rho_i^2 + Pf_j/rho_j^2 + Pi_ij (Monaghan viscosity if approaching)
//   du    += m_j/2 term (v_ij . grad W)   [+ Price conductivity when alpha_u > 0]
//   a     -= m_j (term + (R_i + R_j)(W/W(hbar/eta))^4) grad W
//   a     += m_j (S_i d_i/rho_i^2 + S_j d_j/rho_j^2) . grad W
// then the Jaumann stress rate, the deviatoric work and the CFL step.
template <class R, bool PER>
struct ForceLane {
    using V4 = typename Vec4<R>::T;
    V4 pi, vi, sai, sbi;
    R hi, rho_i, pio, csig_i, rart_i;
    R alpha, beta, eps, eps_as, alpha_u, inv_fdp, ui, Pi;
    R ox, oy, oz;
    R ozs;                 // offset from the own slice's corner (distance culling)
    int cx, cy, cz;
    bool per;              // some axis of the segment is periodic
    const SegGrid* gp;
    R L00 = 0, L01 = 0, L02 = 0, L10 = 0, L11 = 0, L12 = 0, L20 = 0, L21 = 0, L22 = 0;
    R ax = 0, ay = 0, az = 0, du = 0;

    __device__ void setup(const Carried<R>& c, const Work<R>& w, const SegPar& p, const SegGrid& g, int64_t i,
                          uint32_t cc) {
        pi = w.frec_pos[i];
        vi = w.frec_vel[i];
        const V4 ti = w.frec_th[i];
        sai = w.frec_sa[i];
        sbi = w.frec_sb[i];
        hi = pi.w; rho_i = ti.x; pio = ti.y; csig_i = ti.z; rart_i = ti.w;
        alpha = (R)p.alpha; beta = (R)p.beta; eps = (R)p.eps; eps_as = (R)p.eps_as; alpha_u = (R)p.alpha_u;
        inv_fdp = (eps_as > R(0)) ? R(1) / kern_shape(R(1) / (R)p.eta) : R(0);
        cx = cc & 1023u; cy = (cc >> 10) & 1023u; cz = cc >> 20;
        gp = &g;
        per = PER && (g.per[0] || g.per[1] || g.per[2]);
        ui = alpha_u > R(0) ? (R)c.u[i] : R(0);
        Pi = alpha_u > R(0) ? w.P[i] : R(0);
    }

    // neighbour k at separation (rx, ry, rz), r2 = |r|^2 < th^2, th = 2 hbar
    __device__ __forceinline__ void body(const Carried<R>& c, const Work<R>& w, uint32_t k, R rx, R ry, R rz, R r2,
                                         R th, const V4& vj, const V4& tj, const V4& saj, const V4& sbj) {
        // FP32: r and 1/r from one reciprocal square root; FP64: IEEE sqrt and divide
        R r, ir;
        if constexpr (sizeof(R) == 4) {
            ir = __builtin_amdgcn_rsqf(r2);
            r = r2 * ir;
        } else {
            r = f_sqrt(r2);
            if (!(r < th)) return;
            ir = f_rcp(r);
        }
        const R hbar = R(0.5) * th;
        const R ihb = f_rcp(hbar);
        const R qq = r * ihb;
        const R sgm = ihb * ihb * ihb * R(1.0 / kPi);
        const R shp = m4_slope(qq);
        const R fr = sgm * ihb * shp * ir;          // dW/dr / r
        const R gxw = rx * fr, gyw = ry * fr, gzw = rz * fr;
        const R dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
        const R Vj = sbj.z;
        const R t0 = Vj * dvx, t1 = Vj * dvy, t2v = Vj * dvz;
        L00 += t0 * gxw; L01 += t0 * gyw; L02 += t0 * gzw;
        L10 += t1 * gxw; L11 += t1 * gyw; L12 += t1 * gzw;
        L20 += t2v * gxw; L21 += t2v * gyw; L22 += t2v * gzw;
        // Written without branches so that every record component is used in
        // one basic block and each record is fetched by a single 16-byte load.
        const R dvr = -(dvx * rx + dvy * ry + dvz * rz);   // v_ij . r_ij
        // Monaghan viscosity on approaching pairs (mu = 0 otherwise gives Pi = 0)
        const R cbar = R(0.5) * (csig_i + tj.z);
        const R rbar = R(0.5) * (rho_i + tj.x);
        const R den = r2 + eps * hbar * hbar;
        R mu, pij;
        if constexpr (sizeof(R) == 4) {
            // both quotients from one reciprocal: 1/den = rbar/(den rbar), 1/rbar = den/(den rbar)
            const R inv = f_rcp(den * rbar);
            mu = hbar * fminf(dvr, 0.0f) * (rbar * inv);
            pij = (-alpha * cbar * mu + beta * mu * mu) * (den * inv);
        } else {
            mu = hbar * smin(dvr, R(0)) * f_rcp(den);
            pij = (-alpha * cbar * mu + beta * mu * mu) * f_rcp(rbar);
        }
        const R term = pio + tj.y + pij;
        const R vg = -(dvx * gxw + dvy * gyw + dvz * gzw);  // v_ij . grad W
        const R mj = vj.w;
        du += R(0.5) * mj * term * vg;
        // artificial stress; inv_fdp = 0 when it is off
        R f = m4_shape(qq) * inv_fdp;
        f = f * f;
        f = f * f;
        const R mterm = term + (rart_i + tj.w) * f;
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

    // separation of candidate k seen from (bx, by, bz); true when in range
    // (a slot that is not `live` is never in range)
    __device__ __forceinline__ bool in_range(const V4& q, R bx, R by, R bz, R& rx, R& ry, R& rz, R& r2,
                                             R& th, bool live = true) const {
        rx = bx - q.x; ry = by - q.y; rz = bz - q.z;
        if (PER && per) min_image(*gp, rx, ry, rz);
        r2 = rx * rx + ry * ry + rz * rz;
        th = hi + q.w;
        return live & (r2 < th * th) & (r2 != R(0));   // no short-circuit: one basic block per slot
    }

    // strain rate, spin, Jaumann rate and deviatoric power (sph.hpp:230-256),
    // the outputs of particle i, and its CFL step (sph.hpp:350-355, in FP64)
    __device__ double store(const Carried<R>& c, const Work<R>& w, int64_t i, int mat, const SegPar& p) {
        const int64_t n = w.n;
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
        return p.cfl * (double)hi / ((double)csig_i * (1.0 + p.alpha) + 1e-30);
    }
};

// Does particle i read its neighbour list in the force pass?  Only when the
// lists of this step were built, its own fit, and no h of its segment grew
// beyond the skin (then the list holds every pair the walk would find, in the
// walk's order).
template <class R>
__device__ __forceinline__ bool force_list_usable(const Work<R>& w, int64_t i, int seg, int lists) {
    return lists && w.st[seg].list_bad == 0u && w.nbr_cnt[i] != kNbrOverflow;
}

// Force pass by walking the columns (radius h_i + max h, culled with the
// per-block bound of h), in the mirrored column order (for_pieces, MIR).  With
// lists built, particles whose list is usable are left to k_force_list, which
// sums in the list's order (the density walk's).
template <class R, bool PER, bool SEG1>
__global__ __launch_bounds__(256) void k_force(Carried<R> c, Work<R> w, int remap, int lists, SegTargets<1> tg) {
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    int seg = 0;
    double dt_i = 0.0;
    bool mine = false;
    if (i < w.n) {
        const uint32_t inf = c.info[i];
        seg = info_seg(inf);
        mine = !force_list_usable<R>(w, i, seg, lists);
        if (mine) {
            const SegPar& p = w.par[SEG1 ? 0 : seg];
            const SegGrid& g = w.grid[SEG1 ? 0 : seg];
            const uint32_t cc = w.cellc[i];
            ForceLane<R, PER> F;
            F.setup(c, w, p, g, i, cc);
            own_offset<R>(c, g, i, cc, w.dens_rec[i], F.ox, F.oy, F.oz);
            F.ozs = slice_offset<R>(g, F.cz, F.oz);
            if (isfinite((double)F.ox) && isfinite((double)F.oy) && isfinite((double)F.oz)) {
                for_pieces<R, PER, true, true>(g, w, F.cx, F.cy, F.cz, F.ox, F.oy, F.ozs,
                                               F.hi + (R)w.st[SEG1 ? 0 : seg].hmax_final, [&](R hb) { return F.hi + hb; },
                                         [&](uint32_t k0, uint32_t k1, uint32_t ks, R sx, R sy, R sz0, R sz1,
                                             R, const PieceFrame&) {
                    R bx, by, bz0, bz1;
                    if (sizeof(R) == 4) {
                        bx = F.ox - sx; by = F.oy - sy; bz0 = F.oz - sz0; bz1 = F.oz - sz1;
                    } else {
                        bx = F.pi.x; by = F.pi.y; bz0 = bz1 = F.pi.z;
                    }
                    // In chunks of up to 64 candidates: screen the position records
                    // (kForceBatch in flight; the record arrays hold kRecordPad records
                    // past the last particle, so a group's loads never leave them, and
                    // slots past the chunk are never in range) into a bit mask, then run
                    // the pair body over the set bits only, in increasing k.  A wave's
                    // body loop then takes as many turns as its lane with the most pairs
                    // in range, not as its lane with the longest run.
                    for (uint32_t c0 = k0; c0 < k1; c0 += 64u) {
                        const uint32_t c1 = min(c0 + 64u, k1);
                        uint64_t mask = 0;
                        for (uint32_t k = c0; k < c1; k += kForceBatch) {
                            typename Vec4<R>::T q[kForceBatch];
#pragma unroll
                            for (int b = 0; b < kForceBatch; b++) q[b] = rec_at(w.frec_pos + b, k);
                            uint32_t nib = 0;
#pragma unroll
                            for (int b = 0; b < kForceBatch; b++) {
                                R rx, ry, rz, r2, th;
                                const bool in = F.in_range(q[b], bx, by, k + b < ks ? bz0 : bz1, rx, ry, rz, r2, th,
                                                           k + b < c1);
                                nib |= (uint32_t)in << b;
                            }
                            mask |= (uint64_t)nib << (k - c0);
                        }
                        while (mask) {
                            const uint32_t k = c0 + (uint32_t)__builtin_ctzll(mask);
                            mask &= mask - 1;
                            R rx, ry, rz, r2, th;
                            F.in_range(rec_at(w.frec_pos, k), bx, by, k < ks ? bz0 : bz1, rx, ry, rz, r2, th);
                            F.body(c, w, k, rx, ry, rz, r2, th, rec_at(w.frec_vel, k), rec_at(w.frec_th, k),
                                   rec_at(w.frec_sa, k), rec_at(w.frec_sb, k));
                        }
                    }
                });
            }
            dt_i = F.store(c, w, i, info_mat(inf), p);
        }
    }
    const double v[1] = {dt_i};
    block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, mine);
}

// Force pass over the neighbour lists.  List entries are mostly pairs in
// range, so every record of two entries is read up front and their loads
// overlap.
template <class R, bool PER>
__global__ __launch_bounds__(256) void k_force_list(Carried<R> c, Work<R> w, int remap, SegTargets<1> tg) {
    using V4 = typename Vec4<R>::T;
    const int b = xcd_block(blockIdx.x, gridDim.x, remap);
    const int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    int seg = 0;
    double dt_i = 0.0;
    bool mine = false;
    if (i < w.n) {
        const uint32_t inf = c.info[i];
        seg = info_seg(inf);
        mine = force_list_usable<R>(w, i, seg, 1);
        if (mine) {
            const SegPar& p = w.par[seg];
            const SegGrid& g = w.grid[seg];
            const uint32_t cc = w.cellc[i];
            ForceLane<R, PER> F;
            F.setup(c, w, p, g, i, cc);
            own_offset<R>(c, g, i, cc, w.dens_rec[i], F.ox, F.oy, F.oz);
            const uint32_t nc = w.nbr_cnt[i];
            const int cap = w.nbr_cap;
            // two entries at a time, with all their records in flight together
            auto two = [&](uint32_t k0, uint32_t k1, bool has1) {
                const V4 q0 = w.frec_pos[k0], q1 = w.frec_pos[k1];
                const V4 v0 = w.frec_vel[k0], v1 = w.frec_vel[k1];
                const V4 t0 = w.frec_th[k0], t1 = w.frec_th[k1];
                const V4 a0 = w.frec_sa[k0], a1 = w.frec_sa[k1];
                const V4 b0 = w.frec_sb[k0], b1 = w.frec_sb[k1];
                R bx, by, bz, rx, ry, rz, r2, th;
                rel_base<R, PER>(w, *F.gp, k0, F.cx, F.cy, F.cz, F.ox, F.oy, F.oz, F.pi.x, F.pi.y, F.pi.z, bx, by, bz);
                if (F.in_range(q0, bx, by, bz, rx, ry, rz, r2, th)) F.body(c, w, k0, rx, ry, rz, r2, th, v0, t0, a0, b0);
                rel_base<R, PER>(w, *F.gp, k1, F.cx, F.cy, F.cz, F.ox, F.oy, F.oz, F.pi.x, F.pi.y, F.pi.z, bx, by, bz);
                if (has1 && F.in_range(q1, bx, by, bz, rx, ry, rz, r2, th))
                    F.body(c, w, k1, rx, ry, rz, r2, th, v1, t1, a1, b1);
            };
            for (uint32_t e = 0; e < nc; e += 4) {
                const uint4 kk = w.nbr[nbr_quad(i, e >> 2, cap)];
                two(kk.x, kk.y, e + 1 < nc);
                if (e + 2 < nc) two(kk.z, kk.w, e + 3 < nc);
            }
            dt_i = F.store(c, w, i, info_mat(inf), p);
        }
    }
    const double v[1] = {dt_i};
    block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, mine);
}

}  // namespace dsph
