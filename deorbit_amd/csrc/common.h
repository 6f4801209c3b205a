// Device-side data layout of the SPH step.
//
// Particles live in structure-of-arrays form in HBM and are re-sorted every
// step into cell order (sort columns across x and y, z slices within a
// column, see below).  Per-particle state that must survive a step ("carried"
// state) is double-buffered so the permutation is a pure gather; everything
// recomputed each step lives in single buffers in the current order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "physics.h"

namespace dsph {

// Packed per-particle descriptor: material (8 bits) | fixed (1 bit) << 8 | segment << 16.
__device__ __host__ __forceinline__ uint32_t info_pack(int mat, int fixed, int seg) {
    return (uint32_t)(mat & 0xff) | ((uint32_t)(fixed & 1) << 8) | ((uint32_t)seg << 16);
}
__device__ __host__ __forceinline__ int info_mat(uint32_t v) { return (int)(v & 0xffu); }
__device__ __host__ __forceinline__ int info_fixed(uint32_t v) { return (int)((v >> 8) & 1u); }
__device__ __host__ __forceinline__ int info_seg(uint32_t v) { return (int)(v >> 16); }

// Per-segment parameters (host-provided, read-only on device).
struct SegPar {
    double eta, h_init, alpha, beta, eps, alpha_u, cfl, eps_as, g, damp, t_end;
    double dlo[3], dhi[3];
    int32_t damage_on;
    int32_t per[3];
};

// Sort cells are columns of width w across x and y, cut along z into slices
// of width w * kSliceRatio; keys run z fastest, so the slices of one column are
// consecutive and any z range of a column is one contiguous index range.
// kSliceBlock consecutive slices of a column (a block, starting at a slice
// index that is a multiple of kSliceBlock) share one FP32 position frame and
// one upper bound of h.  The z range a particle reaches in a column, at most
// 2 * reach / (slice width) + 3 slices with reach <= (1 + list skin) * 4 h_init
// (h <= 2 h_init), meets at most two blocks as long as the tuning keeps the
// column width >= h_init and the skin <= 0.25: 2 * 5 / 0.25 + 3 <= 48 + 1.
constexpr int kSliceBlock = 48;
// Zeroed records past the last particle in the density and force position
// record arrays: the screens read candidates in groups of four without
// clamping the group to its run (slots past the run never count).
constexpr int kRecordPad = 4;
// Particles per context: record k of any record array (at most 32 bytes)
// then sits below 2^32 bytes from the array's base (rec_at).
constexpr int64_t kMaxParticles = ((int64_t)1 << 27) - 64;
constexpr double kSliceRatio = 1.0 / 4.0;
__device__ __host__ __forceinline__ int block_start(int cz) { return cz - cz % kSliceBlock; }
// z of a block's frame origin in slices: its middle, so FP32 offsets from it
// stay within half a block (frame shifts are differences and do not change)
__device__ __host__ __forceinline__ int block_origin(int cz) { return block_start(cz) + kSliceBlock / 2; }

// Per-segment sort grid and oracle-emulation grid (written on device).
struct SegGrid {
    double lo[3];
    double cw[3], inv_cw[3];     // cell width per axis (periodic axes: L / dim)
    int32_t dim[3];
    uint32_t cell_base, ncells;
    // bounds of h, one per slice block: the segment's blocks start at
    // blk_base, column by column, nblk blocks per column
    uint32_t blk_base;
    int32_t nblk;
    int32_t tiles_y;     // column tiles along y (col_index)
    int32_t tile_sh;     // column tiles are 2^tile_sh columns wide
    // periodic axes: the grid spans the domain [dlo, dhi) of length L with
    // dim whole cells, so the image of a cell is exactly L away
    int32_t per[3];
    double L[3];
    // cell grid the reference would use this density round (sph.hpp:147-157)
    double ocw[3];
    int32_t onc[3];
    double ocw_min;     // smallest cw over axes with more than one cell (+inf if none)
};

// Per-segment scalars updated by kernels.
struct SegState {
    unsigned long long bmin[3], bmax[3];   // ordered encodings of the particle bbox
    unsigned long long dt_next;            // bits of the positive CFL step (atomicMin)
    unsigned long long hmax[5];            // bits of max h at the start of rounds 0..4
    double dt_cur, t;
    long long steps;
    int32_t active[4];
    int32_t any_active[4];                 // segment 0 only: does any segment run round r
    int32_t rounds;
    int32_t bad;                           // non-zero: capacity, numeric or reach trouble
    double hmax_final;
    uint32_t list_bad;                     // non-zero: some h outgrew the neighbour-list skin
    uint32_t pad_;
    unsigned long long list_walk_steps;    // force passes that walked cells for that reason
    unsigned long long list_overflow;      // particle lists that did not fit (cumulative)
    // density moment walks (counted only with Work::stats): lanes and waves of
    // density rounds, and those whose second iteration walked again
    unsigned long long dens_lanes, dens_rewalks, dens_waves, dens_rewalk_waves;
};

// Device-side carried particle state, one buffer set.
template <class R>
struct Carried {
    double *x, *y, *z, *vx, *vy, *vz, *u, *mass;
    R *h, *D, *eps_act;
    R* S;              // 6 x n, component-major {xx,yy,zz,xy,xz,yz}
    uint32_t* info;
    uint32_t* orig;    // IC index of the particle
};

// Vector record types for neighbour reads.
template <class R> struct Vec4;
template <> struct Vec4<float> { using T = float4; };
template <> struct Vec4<double> { using T = double4; };

struct BlockPart;

// Everything the kernels read or write besides the carried state.
template <class R>
struct Work {
    using V4 = typename Vec4<R>::T;
    R *ax, *ay, *az;                 // acceleration from the last force pass
    R *rho, *P, *cs, *csig, *Pf, *Rart, *dscale, *sigmax, *eps_work;
    R *dudt, *dSdt;                  // from the last force pass (dSdt: 6 x n)
    R *dudt_old, *dSdt_old;          // previous force pass, in the current order
    V4* dens_rec;                    // (position key, mass)           density pass
    double4* dens_rec64;             // mixed path: (x, y, z, m) in FP64, for the FP64 density terms
    V4* frec_pos;                    // (position key, h)              force pass
    V4* frec_vel;                    // (velocity, mass)
    V4* frec_th;                     // (rho, Pf/rho^2, csig, Rart)
    V4* frec_sa;                     // S*dscale/rho^2 : xx yy zz xy
    V4* frec_sb;                     // S*dscale/rho^2 : xz yz, m/rho, 0
    uint32_t* cellc;                 // packed (cx | cy << 10 | cz << 20) of the sort cell
    uint32_t* dev_of;                // IC index -> slot in the order before this step's sort
    uint32_t* perm;                  // new slot -> slot in the order before this step's sort
    uint32_t* cstart;                // per sort cell c: first slot of c in the current order, with one
                                     // entry past the last cell, so cell c is [cstart[c], cstart[c+1])
    float* cell_hmax;                // per slice block (block_key): upper bound of h in the block
    // Neighbour lists, built by the first density iteration of a step and read
    // by the second iteration and the force pass.  Entries come in quads of
    // four indices; quad q of particle i sits at
    // nbr[((i >> 6) * (nbr_cap / 4) + q) * 64 + (i & 63)], so one wave reading
    // its q-th quads touches 1 KiB of contiguous memory.  nbr_cnt[i] =
    // kNbrOverflow marks a list that did not fit (that particle walks the cells).
    uint4* nbr;
    uint32_t* nbr_cnt;
    R* h0;                           // h at the start of the build (list radius reference)
    int32_t nbr_cap;                 // entries per particle (a multiple of 4); 0 = lists off
    R skin_build;                    // list radius factor (1 + skin)
    R skin_check;                    // lists stay valid while h <= skin_check * h0
    BlockPart* parts;                // per-block partial reductions
    const MatT<R>* mats;
    const MatT<double>* mats64;      // the material table in FP64 (pressure from an FP64 density)
    const SegPar* par;
    SegGrid* grid;
    SegState* st;
    int64_t n;
    int32_t nseg;
    int32_t stats;                   // non-zero: count the density moment walks (SegState::dens_*)
};

// ---- small device helpers ---------------------------------------------------
__device__ __forceinline__ unsigned long long ord_enc(double d) {
    unsigned long long b = (unsigned long long)__double_as_longlong(d);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __host__ __forceinline__ double ord_dec(unsigned long long e) {
    unsigned long long b = (e & 0x8000000000000000ull) ? (e & 0x7fffffffffffffffull) : ~e;
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}
__device__ __host__ __forceinline__ double bits_to_double(unsigned long long b) {
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}
__device__ __forceinline__ unsigned long long double_to_bits(double d) {
    return (unsigned long long)__double_as_longlong(d);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Record k of a record array, addressed by a 32-bit byte offset from the
// array's base (the base stays in scalar registers and one 32-bit offset
// serves every array read at the same k, instead of a 64-bit address sum per
// load); kMaxParticles keeps k * sizeof(T) below 2^32.
template <class T>
__device__ __forceinline__ T rec_at(const T* base, uint32_t k) {
    return *(const T*)((const char*)base + k * (uint32_t)sizeof(T));
}

__device__ __forceinline__ double wave_max_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_min_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

// atomicMax of the non-negative float v at p[key] over the valid lanes of the
// wave, one atomic per distinct key (neighbouring lanes mostly share a key, and
// same-address atomics from one wave serialise).  Every lane of the wave must
// call this.
__device__ __forceinline__ void wave_max_at(float* p, uint32_t key, float v, bool valid) {
    unsigned long long pending = __ballot(valid);
    const int lane = lane_id();
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const uint32_t lkey = (uint32_t)__shfl((int)key, leader);
        const bool mine = valid && key == lkey;
        float x = mine ? v : 0.0f;
        for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
        if (lane == leader) atomicMax((unsigned int*)&p[lkey], __float_as_uint(x));
        pending &= ~__ballot(mine);
    }
}

// Reduce `v` over the lanes of each segment present in the wave and apply one
// atomic per (wave, segment).  Every lane of the wave must call this.
// kind: 0 max, 1 min.  Values are encoded with `enc` (monotone uint64 map).
template <int KIND, bool ORDERED>
__device__ __forceinline__ void seg_atomic(unsigned long long* base, size_t stride_u64, int seg,
                                           double v, bool valid) {
    unsigned long long pending = __ballot(valid);
    const int lane = lane_id();
    while (pending) {
        int leader = __builtin_ctzll(pending);
        int lseg = __shfl(seg, leader);
        bool mine = valid && seg == lseg && ((pending >> lane) & 1ull);
        double x = mine ? v : (KIND == 0 ? -__builtin_huge_val() : __builtin_huge_val());
        x = (KIND == 0) ? wave_max_d(x) : wave_min_d(x);
        if (lane == leader) {
            unsigned long long e = ORDERED ? ord_enc(x) : double_to_bits(x);
            unsigned long long* p = base + (size_t)lseg * stride_u64;
            if (KIND == 0) atomicMax(p, e);
            else atomicMin(p, e);
        }
        pending &= ~__ballot(mine);
    }
}

// ---- block-level segmented reductions ---------------------------------------
// A reduction value k of segment s lands in *(base[k] + s * stride) through
// atomicMin/atomicMax on a monotone 64-bit encoding (ordered encoding for
// signed values, raw IEEE bits for non-negative ones).  Blocks whose valid
// lanes all belong to one segment write a partial record instead of issuing
// atomics; k_parts_reduce folds the records (one atomic per run of equal
// segments per thread), which keeps same-address atomics to a few hundred per
// launch.
constexpr int kMaxRed = 6;
struct BlockPart {
    int32_t seg;          // -1: no record (empty block or handled by atomics)
    int32_t pad;
    double v[kMaxRed];
};
template <int K>
struct SegTargets {
    unsigned long long* base[K];
    int32_t is_max[K];
    int32_t ordered[K];
};

__device__ __forceinline__ unsigned long long red_enc(double x, int ordered) {
    return ordered ? ord_enc(x) : double_to_bits(x);
}
__device__ __forceinline__ void red_apply(unsigned long long* p, double x, int is_max, int ordered) {
    const unsigned long long e = red_enc(x, ordered);
    if (is_max) atomicMax(p, e);
    else atomicMin(p, e);
}

// Every thread of the block must call this (it synchronises the block).
template <int K>
__device__ __forceinline__ void block_reduce(const SegTargets<K>& t, size_t stride, BlockPart* parts, int seg,
                                             const double (&v)[K], bool valid) {
    __shared__ int s_lo, s_hi;
    __shared__ double s_part[K][16];
    const int lane = lane_id(), wave = threadIdx.x >> 6, nwave = (blockDim.x + 63) >> 6;
    if (threadIdx.x == 0) { s_lo = 0x7fffffff; s_hi = -1; }
    __syncthreads();
    {
        int lo = valid ? seg : 0x7fffffff, hi = valid ? seg : -1;
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o));
            hi = max(hi, __shfl_xor(hi, o));
        }
        if (lane == 0) { atomicMin(&s_lo, lo); atomicMax(&s_hi, hi); }
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;
    if (hi < 0) {                                   // no valid lane in the block
        if (threadIdx.x == 0) parts[blockIdx.x].seg = -1;
        __syncthreads();
        return;
    }
    if (lo == hi) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double x = valid ? v[k] : (t.is_max[k] ? -__builtin_huge_val() : __builtin_huge_val());
            x = t.is_max[k] ? wave_max_d(x) : wave_min_d(x);
            if (lane == 0) s_part[k][wave] = x;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            BlockPart bp;
            bp.seg = lo;
            bp.pad = 0;
            for (int k = 0; k < K; k++) {
                double x = s_part[k][0];
                for (int q = 1; q < nwave; q++) x = t.is_max[k] ? fmax(x, s_part[k][q]) : fmin(x, s_part[k][q]);
                bp.v[k] = x;
            }
            parts[blockIdx.x] = bp;
        }
        __syncthreads();
        return;
    }
    // several segments in this block: per-wave, per-segment atomics
    unsigned long long pending = __ballot(valid);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int lseg = __shfl(seg, leader);
        const bool mine = valid && seg == lseg && ((pending >> lane) & 1ull);
#pragma unroll
        for (int k = 0; k < K; k++) {
            double x = mine ? v[k] : (t.is_max[k] ? -__builtin_huge_val() : __builtin_huge_val());
            x = t.is_max[k] ? wave_max_d(x) : wave_min_d(x);
            if (lane == leader) red_apply(t.base[k] + (size_t)lseg * stride, x, t.is_max[k], t.ordered[k]);
        }
        pending &= ~__ballot(mine);
    }
    if (threadIdx.x == 0) parts[blockIdx.x].seg = -1;
    __syncthreads();
}

// Fold the partial records: each thread folds a contiguous chunk into runs of
// equal segment; a thread left with a single run merges it with the other
// single-run lanes of its wave (shuffles, one atomic per segment per wave),
// a thread with several runs applies its runs directly.
template <int K>
__global__ __launch_bounds__(1024) void k_parts_reduce(SegTargets<K> t, size_t stride, const BlockPart* parts,
                                                       int nparts) {
    const int nt = blockDim.x * gridDim.x;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = (nparts + nt - 1) / nt;
    const int beg = min(nparts, tid * per), end = min(nparts, beg + per);
    int cur = -1;
    double acc[K];
    for (int k = 0; k < K; k++) acc[k] = 0.0;
    for (int q = beg; q < end; q++) {
        const int s = parts[q].seg;
        if (s < 0) continue;
        if (s != cur) {
            if (cur >= 0)
                for (int k = 0; k < K; k++) red_apply(t.base[k] + (size_t)cur * stride, acc[k], t.is_max[k], t.ordered[k]);
            cur = s;
            for (int k = 0; k < K; k++) acc[k] = parts[q].v[k];
        } else {
            for (int k = 0; k < K; k++) acc[k] = t.is_max[k] ? fmax(acc[k], parts[q].v[k]) : fmin(acc[k], parts[q].v[k]);
        }
    }
    // the last run of every thread goes through the wave-level merge
    const bool valid = cur >= 0;
    const int lane = lane_id();
    unsigned long long pending = __ballot(valid);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int lseg = __shfl(cur, leader);
        const bool mine = valid && cur == lseg && ((pending >> lane) & 1ull);
#pragma unroll
        for (int k = 0; k < K; k++) {
            double x = mine ? acc[k] : (t.is_max[k] ? -__builtin_huge_val() : __builtin_huge_val());
            x = t.is_max[k] ? wave_max_d(x) : wave_min_d(x);
            if (lane == leader) red_apply(t.base[k] + (size_t)lseg * stride, x, t.is_max[k], t.ordered[k]);
        }
        pending &= ~__ballot(mine);
    }
}

// Block index remap so consecutive tiles of the sorted particle range are
// processed by workgroups of the same XCD (hardware dispatches workgroup b to
// XCD b % 8); keeps a spatial neighbourhood inside one L2.
__device__ __forceinline__ int xcd_block(int b, int nb, int remap) {
    if (!remap) return b;
    const int nx = 8;
    int q = nb / nx, r = nb % nx;
    int x = b % nx, k = b / nx;
    return x * q + (x < r ? x : r) + k;
}

// Columns are numbered in square tiles of 2^tile_sh x 2^tile_sh columns
// (tiles x-major over the grid, columns x-major inside a tile; the grid is
// padded to whole tiles with empty columns), so the columns around a particle,
// which a stencil walk visits, lie close together in memory and the neighbour
// records of the waves in flight on one XCD share its L2.  k_grid picks the
// largest tile up to kColTileMax columns wide whose padding stays within a
// tenth of the segment's smaller column count.
constexpr int kColTileMax = 16;
// (G: SegGrid, or WalkGrid below: any type with these fields)
template <class G>
__device__ __host__ __forceinline__ uint32_t col_index(const G& g, int cx, int cy) {
    const uint32_t ux = (uint32_t)cx, uy = (uint32_t)cy, sh = (uint32_t)g.tile_sh, m = (1u << sh) - 1u;
    return (((((ux >> sh) * (uint32_t)g.tiles_y + (uy >> sh)) << sh) + (ux & m)) << sh) + (uy & m);
}
// Column-major key: columns in tile order, slices z fastest within a column.
template <class G>
__device__ __host__ __forceinline__ uint32_t cell_key(const G& g, int cx, int cy, int cz) {
    return g.cell_base + col_index(g, cx, cy) * (uint32_t)g.dim[2] + (uint32_t)cz;
}
// Index of the bound of h of the block holding slice cz of column (cx, cy)
// (Work::cell_hmax).
template <class G>
__device__ __host__ __forceinline__ uint32_t block_key(const G& g, int cx, int cy, int cz) {
    return g.blk_base + col_index(g, cx, cy) * (uint32_t)g.nblk + (uint32_t)(cz / kSliceBlock);
}
// The fields of a segment's sort grid that a column walk reads, copied into
// registers once per walk (read through a SegGrid reference inside the walk's
// loops, they would be loaded again at every column).
struct WalkGrid {
    uint32_t cell_base, blk_base;
    int32_t nblk, tiles_y, tile_sh;
    int32_t dim[3];
    __device__ __host__ __forceinline__ explicit WalkGrid(const SegGrid& g)
        : cell_base(g.cell_base), blk_base(g.blk_base), nblk(g.nblk), tiles_y(g.tiles_y), tile_sh(g.tile_sh),
          dim{g.dim[0], g.dim[1], g.dim[2]} {}
};
__device__ __host__ __forceinline__ uint32_t cellc_pack(int cx, int cy, int cz) {
    return (uint32_t)cx | ((uint32_t)cy << 10) | ((uint32_t)cz << 20);
}

}  // namespace dsph
