// Cell-list build: step control, kick + drift with a per-segment bounding box,
// per-segment sort grid, cell keys, a stable LSD radix sort, and the
// permutation of the carried state into cell order.
//
// Replaces the Metal counting sort (sph_force.metal:27-46 with the host scan of
// gpu_ic.cpp:97-100) and the reference's per-round std::vector cell lists
// (sph.hpp:147-162).  Every reduction is an exact min/max, and the sort is
// stable, so a run is bitwise reproducible.
#pragma once
#include "common.h"

namespace dsph {

constexpr int kMaxStencil = 16;               // cap on the per-particle stencil radius in cells
constexpr uint32_t kNbrOverflow = 0xffffffffu; // neighbour count of a list that did not fit

// Positive float rounded up to the next float at or above the double d.
__device__ __forceinline__ float float_at_least(double d) {
    float f = (float)d;
    if ((double)f < d) f = __uint_as_float(__float_as_uint(f) + 1u);
    return f;
}

// ---------------------------------------------------------------- step control
// mode 0: each segment takes its device CFL step; 1: host scalar dt; 2: per-segment host dt.
__global__ void k_step_begin(SegState* st, const SegPar* par, int nseg, int mode, double host_dt,
                             const double* seg_dt) {
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) {
        SegState& S = st[s];
        double dt = (mode == 0) ? bits_to_double(S.dt_next) : (mode == 1 ? host_dt : seg_dt[s]);
        // a CFL step must be positive: zero (an infinite signal speed) would
        // stall the segment's clock
        if (mode == 0 && !(dt > 0.0)) S.bad |= 2;
        // adaptive runs stop exactly at t_end (run_ic.cpp:327-330); a finished
        // segment takes steps of zero
        if (mode == 0 && par[s].t_end > 0.0 && S.t + dt > par[s].t_end) dt = fmax(par[s].t_end - S.t, 0.0);
        if (!(dt > 0.0)) {
            if (!(dt == 0.0)) S.bad |= 2;
            dt = 0.0;
        }
        S.dt_cur = dt;
        S.t += dt;
        S.steps += 1;
    }
}

// Clear the CFL minimum before a force pass recomputes it.
__global__ void k_reset_dt(SegState* st, int nseg) {
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) st[s].dt_next = 0x7ff0000000000000ull;
}

// Reset the per-build reductions (bbox, round maxima, round flags).
__global__ void k_build_begin(SegState* st, int nseg) {
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) {
        SegState& S = st[s];
        for (int a = 0; a < 3; a++) { S.bmin[a] = ~0ull; S.bmax[a] = 0ull; }
        for (int r = 0; r < 5; r++) S.hmax[r] = 0ull;
        for (int r = 0; r < 4; r++) S.active[r] = 0;
        S.rounds = 0;
        S.hmax_final = 0.0;
        S.list_bad = 0u;
    }
}

// Half kick and drift of non-fixed particles (sph.hpp:360-361, with the
// gpu_ic gravity term of sph_force.metal:232-237), then the segment bbox over
// finite positions.  move == 0 only measures the bbox.
template <class R>
__global__ __launch_bounds__(256) void k_kick_drift(Carried<R> c, Work<R> w, int move, SegTargets<6> tg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < w.n;
    int seg = 0;
    double x = 0, y = 0, z = 0;
    bool fin = false;
    if (in) {
        const uint32_t inf = c.info[i];
        seg = info_seg(inf);
        x = c.x[i]; y = c.y[i]; z = c.z[i];
        if (move && !info_fixed(inf)) {
            const double dt = w.st[seg].dt_cur, half = 0.5 * dt;
            const double g = w.par[seg].g;
            double vx = c.vx[i] + (double)w.ax[i] * half;
            double vy = c.vy[i] + (double)w.ay[i] * half;
            double vz = c.vz[i] + (double)w.az[i] * half;
            if (g != 0.0) vz -= g * half;
            c.vx[i] = vx; c.vy[i] = vy; c.vz[i] = vz;
            x += vx * dt; y += vy * dt; z += vz * dt;
            c.x[i] = x; c.y[i] = y; c.z[i] = z;
        }
        fin = isfinite(x) && isfinite(y) && isfinite(z);
        if (!fin) atomicOr(&w.st[seg].bad, 4);
    }
    const double v[6] = {x, y, z, x, y, z};
    block_reduce<6>(tg, sizeof(SegState) / 8, w.parts, seg, v, fin);
}

// Sort grid per segment: columns of width w = cell_factor * h_init across x
// and y cut into slices of w * kSliceRatio along z, over the segment bbox (at
// most 1001 cells per axis), segments laid end to end in one key space.  If the total exceeds the table capacity every width grows by 25 %
// until it fits (coarser cells stay correct: stencils are sized in cells).
// total_out[0] is the total laid out, total_out[1] the total the nominal
// widths would need (saturated at 2^32 - 1), from which the host grows the
// table between calls.
// One workgroup of kGridThreads threads: segments are handled a chunk at a
// time, one per thread, and a block scan lays their key ranges end to end.
constexpr int kGridThreads = 1024;
__global__ __launch_bounds__(kGridThreads) void k_grid(SegGrid* grid, SegState* st, const SegPar* par, int nseg,
                                                       double cell_factor, uint32_t capacity, uint32_t* total_out) {
    __shared__ unsigned long long s_scan[kGridThreads], s_scanb[kGridThreads];
    __shared__ unsigned long long s_carry, s_carryb;
    const int t = threadIdx.x;
    double scale = 1.0;
    unsigned long long total = 0, total_blk = 0, wanted = 0;
    for (int attempt = 0; attempt < 200; attempt++) {
        if (t == 0) { s_carry = 0; s_carryb = 0; }
        __syncthreads();
        for (int s0 = 0; s0 < nseg; s0 += kGridThreads) {
            const int s = s0 + t;
            unsigned long long nc = 0, nbk = 0;
            if (s < nseg) {
                SegGrid& g = grid[s];
                double lo[3], ext[3];
                for (int a = 0; a < 3; a++) {
                    double l = ord_dec(st[s].bmin[a]), h = ord_dec(st[s].bmax[a]);
                    if (!(l <= h)) { l = 0.0; h = 0.0; }          // segment without finite particles
                    lo[a] = l;
                    ext[a] = h - l;
                }
                // periodic axes span the domain instead of the particle bbox
                // (positions are not wrapped; cell_coord clamps strays into the
                // edge cells as the reference's binof does)
                int per[3];
                for (int a = 0; a < 3; a++) {
                    const double L = par[s].dhi[a] - par[s].dlo[a];
                    per[a] = par[s].per[a] && L > 0.0;
                    if (per[a]) {
                        lo[a] = par[s].dlo[a];
                        ext[a] = L;
                    }
                }
                const double w = cell_factor * par[s].h_init * scale;
                // one trailing "park" cell collects particles with a non-finite
                // position (never visited by a stencil)
                nc = 1;
                for (int a = 0; a < 3; a++) {
                    double cw = (a == 2) ? w * kSliceRatio : w;
                    if (ext[a] / cw > 1000.0) cw = ext[a] / 1000.0;   // at most 1001 cells per axis
                    g.lo[a] = lo[a];
                    g.per[a] = per[a];
                    if (per[a]) {
                        // whole cells of at least the nominal width tile the period
                        int m = (int)floor(ext[a] / cw);
                        m = m < 1 ? 1 : m;
                        g.dim[a] = m;
                        g.L[a] = ext[a];
                        g.cw[a] = ext[a] / m;
                    } else {
                        g.dim[a] = (int)floor(ext[a] / cw) + 1;
                        g.L[a] = 0.0;
                        g.cw[a] = cw;
                    }
                    g.inv_cw[a] = 1.0 / g.cw[a];
                }
                // columns padded to whole tiles (col_index)
                const int dmin = g.dim[0] < g.dim[1] ? g.dim[0] : g.dim[1];
                int sh = 0;
                while ((2 << sh) <= kColTileMax && 10 * (2 << sh) <= dmin) sh++;
                const int T = 1 << sh;
                g.tile_sh = sh;
                const int tx = (g.dim[0] + T - 1) >> sh;
                g.tiles_y = (g.dim[1] + T - 1) >> sh;
                const unsigned long long ncol = (unsigned long long)tx * g.tiles_y * T * T;
                nc += ncol * (unsigned long long)g.dim[2];
                g.ncells = (uint32_t)nc;
                g.nblk = (g.dim[2] + kSliceBlock - 1) / kSliceBlock;
                nbk = ncol * (unsigned long long)g.nblk;
            }
            // inclusive scans of the chunk's cell and block counts
            s_scan[t] = nc;
            s_scanb[t] = nbk;
            __syncthreads();
            for (int o = 1; o < kGridThreads; o <<= 1) {
                const unsigned long long add = t >= o ? s_scan[t - o] : 0ull;
                const unsigned long long addb = t >= o ? s_scanb[t - o] : 0ull;
                __syncthreads();
                s_scan[t] += add;
                s_scanb[t] += addb;
                __syncthreads();
            }
            const unsigned long long carry = s_carry, carryb = s_carryb;
            if (s < nseg) {
                grid[s].cell_base = (uint32_t)(carry + s_scan[t] - nc);
                grid[s].blk_base = (uint32_t)(carryb + s_scanb[t] - nbk);
            }
            __syncthreads();
            if (t == kGridThreads - 1) {
                s_carry = carry + s_scan[t];
                s_carryb = carryb + s_scanb[t];
            }
            __syncthreads();
        }
        total = s_carry;
        total_blk = s_carryb;
        __syncthreads();
        if (attempt == 0) wanted = total;
        if (total <= capacity) break;
        scale *= 1.25;
    }
    if (t == 0) {
        // slice blocks of the grid laid out (never more than its cells)
        total_out[2] = (uint32_t)(total_blk < capacity ? total_blk : capacity);
        if (total > capacity) { st[0].bad |= 1; total = capacity; }
        total_out[0] = (uint32_t)total;
        total_out[1] = wanted > 0xffffffffull ? 0xffffffffu : (uint32_t)wanted;
    }
}

__device__ __forceinline__ int cell_coord(const SegGrid& g, int a, double v) {
    double t = floor((v - g.lo[a]) * g.inv_cw[a]);
    int c = (t < 0.0) ? 0 : (t > (double)(g.dim[a] - 1) ? g.dim[a] - 1 : (int)t);
    return c;
}

// Cell key of every particle, and the clearing of the bounds of h (grid-stride).
template <class R>
__global__ __launch_bounds__(256) void k_keys(Carried<R> c, Work<R> w, uint32_t* keys, uint32_t* vals,
                                              const uint32_t* total_cells) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.n) {
        const int seg = info_seg(c.info[i]);
        const double x = c.x[i], y = c.y[i], z = c.z[i];
        const SegGrid& g = w.grid[seg];
        uint32_t key = g.cell_base + g.ncells - 1u;              // park cell
        if (isfinite(x) && isfinite(y) && isfinite(z))
            key = cell_key(g, cell_coord(g, 0, x), cell_coord(g, 1, y), cell_coord(g, 2, z));
        // the sort runs over IC order, so particles of one cell end up in IC
        // order whatever order the previous step left them in
        const uint32_t o = c.orig[i];
        keys[o] = key;
        vals[o] = o;
        w.dev_of[o] = (uint32_t)i;
    }
    // clear the bounds of h of every slice block (total_cells[2] of them)
    const uint32_t total = total_cells[2];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = (uint32_t)(blockIdx.x * blockDim.x + threadIdx.x); k < total; k += stride)
        w.cell_hmax[k] = 0.0f;
}

// First slot of every cell in the sorted keys (a lower bound per cell, so empty
// cells get the slot where they would start), for cells 0 .. total inclusive,
// in two passes.  k_chunk_start records the first slot of every chunk of
// kStartChunk cells: slot i, between keys k[i-1] and k[i] (k[-1] = -inf,
// k[n] = +inf), writes the entries of the chunks that start in (k[i-1], k[i]],
// so every entry is written once.  k_cell_start then gives each wave one chunk
// and each lane kCellChunk consecutive cells of it: a binary search between the
// chunk's first slot and the next chunk's for the lane's first cell, then a
// merge walk along the keys for the rest.
constexpr int kCellChunk = 8;
static_assert(kCellChunk == 8, "k_cell_start stores a lane's cells as two 16-byte vectors");
constexpr int kStartChunk = 64 * kCellChunk;
__device__ __forceinline__ int64_t lower_bound_key(const uint32_t* skeys, int64_t lo, int64_t hi, uint64_t c) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (skeys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
__global__ __launch_bounds__(256) void k_chunk_start(const uint32_t* skeys, int64_t n, uint32_t* lb,
                                                    const uint32_t* total_cells) {
    const int64_t last_chunk = (int64_t)(*total_cells / kStartChunk) + 1;    // entries 0 .. last_chunk
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        const int64_t j0 = i == 0 ? 0 : (int64_t)(skeys[i - 1] / kStartChunk) + 1;
        const int64_t j1 = i == n ? last_chunk : (int64_t)(skeys[i] / kStartChunk);
        for (int64_t j = j0; j <= j1; j++) lb[j] = (uint32_t)i;
    }
}
__global__ __launch_bounds__(256) void k_cell_start(const uint32_t* skeys, int64_t n, uint32_t* cstart,
                                                   const uint32_t* lb, const uint32_t* total_cells) {
    const uint64_t total = *total_cells;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwave = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    for (uint64_t j = wave; j * kStartChunk <= total; j += nwave) {
        const uint64_t c0 = j * kStartChunk + (uint64_t)lane * kCellChunk;
        if (c0 > total) continue;
        int64_t a = lower_bound_key(skeys, lb[j], lb[j + 1], c0);
        if (c0 + kCellChunk <= total + 1) {
            // a lane's cells are 32-byte aligned: two 16-byte stores
            uint32_t v[kCellChunk];
#pragma unroll
            for (int k = 0; k < kCellChunk; k++) {
                while (a < n && skeys[a] < c0 + k) a++;
                v[k] = (uint32_t)a;
            }
            uint4* dst = (uint4*)(cstart + c0);
            dst[0] = make_uint4(v[0], v[1], v[2], v[3]);
            dst[1] = make_uint4(v[4], v[5], v[6], v[7]);
        } else {
            for (uint64_t c = c0; c < total + 1; c++) {
                while (a < n && skeys[a] < c) a++;
                cstart[c] = (uint32_t)a;
            }
        }
    }
}

// ---------------------------------------------------------------- radix sort
// LSD radix sort of (key, value) pairs, 8-bit digits, tiles of ITEMS keys per
// thread (rs_items picks 8 below 2M keys, so that a pass still has two tiles
// per CU, else 16).
constexpr int kRsThreads = 256;
constexpr int rs_items(int64_t n) { return n < (int64_t)2000000 ? 8 : 16; }

// Lanes of the wave holding the same 8-bit digit as this lane.
__device__ __forceinline__ unsigned long long match_digit(uint32_t d, bool valid) {
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

template <int ITEMS>
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* keys, int64_t n, int shift,
                                                       uint32_t* hist, int nb) {
    __shared__ uint32_t h[256];
    const int t = threadIdx.x, lane = lane_id();
    h[t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * (kRsThreads * ITEMS);
    // the whole tile's keys are loaded first, so their loads overlap
    uint32_t kk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int64_t idx = base + (int64_t)k * kRsThreads + t;
        kk[k] = idx < n ? keys[idx] : 0u;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const bool valid = base + (int64_t)k * kRsThreads + t < n;
        const uint32_t d = (kk[k] >> shift) & 255u;
        const unsigned long long m = match_digit(d, valid);
        if (valid && __builtin_ctzll(m) == lane) atomicAdd(&h[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    hist[(size_t)t * nb + blockIdx.x] = h[t];
}

// Exclusive scan of one digit's row of block counts (one workgroup per digit);
// the row total goes to totals[digit].  The scatter adds the digit base.
__global__ __launch_bounds__(256) void k_rs_scan(uint32_t* hist, int nb, uint32_t* totals) {
    __shared__ uint32_t part[256];
    const int t = threadIdx.x;
    uint32_t* row = hist + (size_t)blockIdx.x * nb;
    const int per = (nb + 255) / 256;
    const int beg = t * per, end = min(nb, beg + per);
    uint32_t s = 0;
    for (int i = beg; i < end; i++) s += row[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = beg; i < end; i++) {
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (t == 255) totals[blockIdx.x] = part[255];
}

// Stable scatter of one tile: keys are ranked by digit (wave ballots, then
// waves in order, then rounds in order), placed in LDS in digit order, and
// written out so that consecutive threads write consecutive slots of a
// digit's run.
template <int ITEMS>
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t* kin, const uint32_t* vin,
                                                          uint32_t* kout, uint32_t* vout, int64_t n,
                                                          int shift, const uint32_t* hist, int nb,
                                                          const uint32_t* totals) {
    constexpr int kTile = kRsThreads * ITEMS;
    __shared__ uint32_t gofs[256];    // global slot of digit d's run in this tile minus its local slot
    __shared__ uint32_t lbase[256];   // next local slot of digit d
    __shared__ uint32_t wc[4][256];
    __shared__ uint32_t pre[4][256];
    __shared__ uint32_t sk[kTile], sv[kTile];
    const int t = threadIdx.x, lane = lane_id(), wave = t >> 6;
    // digit base = exclusive scan of the digit totals (each block redoes this 256-entry scan)
    const uint32_t tot = totals[t];
    gofs[t] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = (t >= off) ? gofs[t - off] : 0u;
        __syncthreads();
        gofs[t] += v;
        __syncthreads();
    }
    const uint32_t dbase = gofs[t] - tot;
    // this tile's keys of digit t (the scanned histogram row holds the earlier tiles')
    const uint32_t h0 = hist[(size_t)t * nb + blockIdx.x];
    const uint32_t cnt = ((int)blockIdx.x + 1 < nb ? hist[(size_t)t * nb + blockIdx.x + 1] : tot) - h0;
    lbase[t] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = (t >= off) ? lbase[t - off] : 0u;
        __syncthreads();
        lbase[t] += v;
        __syncthreads();
    }
    const uint32_t lstart = lbase[t] - cnt;
    __syncthreads();
    lbase[t] = lstart;
    gofs[t] = dbase + h0 - lstart;    // modulo 2^32; gofs + local slot is the global slot
    for (int q = 0; q < 4; q++) wc[q][t] = 0;
    __syncthreads();
    const int64_t tile = (int64_t)blockIdx.x * kTile;
    // the whole tile is loaded first, so its loads overlap
    uint32_t kk[ITEMS], vv[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int64_t idx = tile + (int64_t)k * kRsThreads + t;
        kk[k] = idx < n ? kin[idx] : 0u;
        vv[k] = idx < n ? vin[idx] : 0u;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const bool valid = tile + (int64_t)k * kRsThreads + t < n;
        const uint32_t key = kk[k];
        const uint32_t d = (key >> shift) & 255u;
        const unsigned long long m = match_digit(d, valid);
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (valid && __builtin_ctzll(m) == lane) wc[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        const uint32_t c0 = wc[0][t], c1 = wc[1][t], c2 = wc[2][t], c3 = wc[3][t];
        pre[0][t] = 0;
        pre[1][t] = c0;
        pre[2][t] = c0 + c1;
        pre[3][t] = c0 + c1 + c2;
        __syncthreads();
        if (valid) {
            const uint32_t lp = lbase[d] + pre[wave][d] + rank;
            sk[lp] = key;
            sv[lp] = vv[k];
        }
        __syncthreads();
        lbase[t] += c0 + c1 + c2 + c3;
        for (int q = 0; q < 4; q++) wc[q][t] = 0;
        __syncthreads();
    }
    const int64_t left = n - tile;
    const int count = left < kTile ? (int)left : kTile;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const int idx = k * kRsThreads + t;
        if (idx < count) {
            const uint32_t key = sk[idx];
            const uint32_t pos = gofs[(key >> shift) & 255u] + (uint32_t)idx;
            kout[pos] = key;
            vout[pos] = sv[idx];
        }
    }
}

// ---------------------------------------------------------------- permutation
// Gather the carried state into sorted order, carry the last force pass's
// dudt/dSdt along as the "old" derivatives (sph.hpp:358-359), and write the
// density-pass neighbour records.  Positions in the mixed path are stored as
// FP32 offsets from the corner of the particle's own column (x, y) and the
// middle of its slice block (z), so pair separations are formed from numbers
// below a few h; the FP64 path keeps absolute positions.
template <class R>
__global__ __launch_bounds__(256) void k_permute(Carried<R> src, Carried<R> dst, Work<R> w,
                                                 const uint32_t* keys, const uint32_t* vals, SegTargets<1> tg,
                                                 int list_bounds) {
    using V4 = typename Vec4<R>::T;
    const int64_t n = w.n;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < n;
    int seg = 0;
    double hv = 0.0;
    uint32_t bkey = 0;
    float bh = 0.0f;
    bool bvalid = false;
    if (in) {
        const uint32_t s = w.dev_of[vals[i]];   // current slot of the particle sorted to i
        w.perm[i] = s;
        const uint32_t inf = src.info[s];
        seg = info_seg(inf);
        const double x = src.x[s], y = src.y[s], z = src.z[s], m = src.mass[s];
        dst.x[i] = x; dst.y[i] = y; dst.z[i] = z;
        dst.vx[i] = src.vx[s]; dst.vy[i] = src.vy[s]; dst.vz[i] = src.vz[s];
        dst.u[i] = src.u[s]; dst.mass[i] = m;
        const R h = src.h[s];
        dst.h[i] = h;
        dst.D[i] = src.D[s];
        dst.eps_act[i] = src.eps_act[s];
#pragma unroll
        for (int k = 0; k < 6; k++) dst.S[(size_t)k * n + i] = src.S[(size_t)k * n + s];
        dst.info[i] = inf;
        dst.orig[i] = src.orig[s];
        w.dudt_old[i] = w.dudt[s];
#pragma unroll
        for (int k = 0; k < 6; k++) w.dSdt_old[(size_t)k * n + i] = w.dSdt[(size_t)k * n + s];

        const uint32_t key = keys[i];
        const SegGrid& g = w.grid[seg];
        w.h0[i] = h;
        if (key != g.cell_base + g.ncells - 1u) {
            const int cx = cell_coord(g, 0, x), cy = cell_coord(g, 1, y), cz = cell_coord(g, 2, z);
            bkey = block_key(g, cx, cy, cz);
            bh = float_at_least((double)h);
            bvalid = true;
            w.cellc[i] = cellc_pack(cx, cy, cz);
            V4 rec;
            if (sizeof(R) == 4) {
                const double ox = x - (g.lo[0] + (double)cx * g.cw[0]);
                const double oy = y - (g.lo[1] + (double)cy * g.cw[1]);
                const double oz = z - (g.lo[2] + (double)block_origin(cz) * g.cw[2]);
                rec.x = (R)ox; rec.y = (R)oy; rec.z = (R)oz;
                rec.w = (R)m;
                if (w.dens_rec64) w.dens_rec64[i] = make_double4(x, y, z, m);   // the FP64 density terms
            } else {
                rec.x = (R)x; rec.y = (R)y; rec.z = (R)z;
                rec.w = (R)m;
            }
            w.dens_rec[i] = rec;
        } else {
            w.cellc[i] = 0u;
            V4 rec;
            rec.x = rec.y = rec.z = (R)__builtin_nan("");
            rec.w = (R)0;
            w.dens_rec[i] = rec;
            if (sizeof(R) == 4 && w.dens_rec64) w.dens_rec64[i] = make_double4(x, y, z, 0.0);
        }
        hv = (double)h;
    }
    // per-block bound of h at the start of the step, read only by the
    // neighbour-list build (the density epilogue raises it to the final h that
    // the force pass culls with)
    if (list_bounds) wave_max_at(w.cell_hmax, bkey, bh, bvalid);
    const double v[1] = {hv};
    block_reduce<1>(tg, sizeof(SegState) / 8, w.parts, seg, v, in);
}

// Density round bookkeeping, one thread per segment (sph.hpp:185-206).
// r in 0..3: decide whether round r runs and derive the cell grid the
// reference would build for it; r == 4: record the rounds run and the final
// h_max for the force stencils.
__global__ void k_round(SegState* st, SegGrid* grid, const SegPar* par, int nseg, int r) {
    __shared__ int s_any;
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) {
        SegState& S = st[s];
        if (r < 4) {
            int act = 1;
            if (r > 0) {
                const double hm = bits_to_double(S.hmax[r]), h0 = bits_to_double(S.hmax[r - 1]);
                act = S.active[r - 1] && !(hm <= h0 * 1.02);
            }
            S.active[r] = act;
            if (act) atomicOr(&s_any, 1);
            if (act) {
                SegGrid& g = grid[s];
                const double hmax = fmax(bits_to_double(S.hmax[r]), 1e-30);
                const double rc = 2.0 * hmax;
                double cmin = __builtin_huge_val();
                for (int a = 0; a < 3; a++) {
                    double ext = par[s].dhi[a] - par[s].dlo[a];
                    if (ext <= 0) ext = rc;
                    int m = (int)(ext / rc);
                    m = m < 1 ? 1 : m;
                    if (par[s].per[a] && m < 3) m = 1;
                    g.onc[a] = m;
                    g.ocw[a] = ext / m;
                    if (m > 1 && g.ocw[a] < cmin) cmin = g.ocw[a];
                }
                g.ocw_min = cmin;
            }
        } else {
            if (S.list_bad) S.list_walk_steps++;
            if (!S.active[0]) {
                S.rounds = 0;
                S.hmax_final = bits_to_double(S.hmax[0]);
            } else {
                int last = 3;
                while (last > 0 && !S.active[last]) last--;
                S.rounds = last + 1;
                S.hmax_final = bits_to_double(S.hmax[last + 1]);
            }
        }
    }
    __syncthreads();
    if (r < 4 && threadIdx.x == 0) st[0].any_active[r] = s_any;
}

}  // namespace dsph
