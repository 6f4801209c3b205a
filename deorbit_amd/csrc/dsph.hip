// Host side of the SPH step: device memory, kernel orchestration, hipGraph
// replay, profiling and the extern "C" boundary declared in include/dsph.h.
//
// One Engine<R> per context; R = float is the mixed-precision path (FP32 pair
// sums, EOS and stress; FP64 positions, velocities, internal energy, time step
// and reductions), R = double the all-FP64 comparator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <tuple>
#include <vector>

#include "dsph.h"
#include "kernels_point.h"

namespace dsph {

struct Err {
    int code;
    std::string msg;
};

static thread_local std::string g_last_error;

#define HIPC(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess)                                                                     \
            throw Err{DSPH_E_HIP, std::string(#x) + " failed: " + hipGetErrorString(e_)};        \
    } while (0)

// Largest cell table the engine grows to on its own, in cells per particle
// (the bound set_tuning accepts too).
constexpr int kMaxCellsPerParticle = 64;
constexpr int kDefaultCellsPerParticle = 4;

static void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Err{DSPH_E_HIP, std::string("launch of ") + what + " failed: " + hipGetErrorString(e)};
}

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

enum Phase {
    PH_STEP_BEGIN, PH_KICK_DRIFT, PH_GRID, PH_KEYS, PH_SORT, PH_PERMUTE, PH_ROUND, PH_DENSITY,
    PH_EOS, PH_FORCE, PH_FINISH, PH_DIAG, PH_IO, PH_REDUCE, PH_DENSITY_BUILD, PH_DENSITY_LIST, PH_FORCE_LIST,
    PH_COUNT
};
static const char* kPhaseName[PH_COUNT] = {"step_begin", "kick_drift", "grid", "keys", "sort", "permute",
                                           "round", "density", "eos", "force", "finish", "diag", "io",
                                           "reduce", "density_build", "density_list", "force_list"};

struct HostParticles {
    std::vector<double> pos, vel, mass, u, h, D, eps_act, S;   // S: n x 6 IC-order rows
    std::vector<int32_t> mat, seg;
    std::vector<uint8_t> fixed;
};

// dt_cur := dt[s] without advancing time or the step counter.
__global__ void k_set_dt(SegState* st, int nseg, const double* dt) {
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) st[s].dt_cur = dt[s];
}

class EngineBase {
  public:
    virtual ~EngineBase() {}
};

template <class R>
class Engine : public EngineBase {
  public:
    using V4 = typename Vec4<R>::T;

    // density64: the mixed path's density with FP64 pair terms and the pressure
    // in FP64 (DSPH_MIXED); false keeps FP32 density sums (DSPH_MIXED_FAST)
    Engine(const std::vector<dsph_params>& segp, const std::vector<dsph_material>& mats, hipStream_t s,
           bool density64)
        : segp_(segp), mats_(mats), stream_(s), density64_(sizeof(R) == 4 && density64) {
        nseg_ = (int)segp.size();
        tune_.block_pair = 256;
        tune_.block_point = 256;
        tune_.cell_factor = 1.5;
        tune_.xcd_remap = 1;
        tune_.cell_capacity_per_particle = kDefaultCellsPerParticle;
        tune_.neighbor_capacity = 0;
        tune_.list_skin = 0.06;
        for (const auto& p : segp_)
            if ((p.periodic[0] && p.hi[0] > p.lo[0]) || (p.periodic[1] && p.hi[1] > p.lo[1]) ||
                (p.periodic[2] && p.hi[2] > p.lo[2]))
                periodic_ = true;
        for (auto& e : prof_total_) e = 0.0;
        for (auto& e : prof_count_) e = 0;
    }
    ~Engine() override {
        release_graphs();
        drain_profile(false);
        for (void* p : allocs_) (void)hipFree(p);
        for (auto& ev : ev_pool_) (void)hipEventDestroy(ev);
    }

    int64_t n() const { return n_; }
    int nseg() const { return nseg_; }

    // ---------------------------------------------------------------- memory
    template <class T>
    T* dalloc(size_t count) {
        void* p = nullptr;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) throw Err{DSPH_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
        allocs_.push_back(p);
        return (T*)p;
    }
    void free_all() {
        release_graphs();
        for (void* p : allocs_) (void)hipFree(p);
        allocs_.clear();
    }
    void alloc_carried(Carried<R>& c) {
        c.x = dalloc<double>(n_); c.y = dalloc<double>(n_); c.z = dalloc<double>(n_);
        c.vx = dalloc<double>(n_); c.vy = dalloc<double>(n_); c.vz = dalloc<double>(n_);
        c.u = dalloc<double>(n_); c.mass = dalloc<double>(n_);
        c.h = dalloc<R>(n_); c.D = dalloc<R>(n_); c.eps_act = dalloc<R>(n_);
        c.S = dalloc<R>(6 * n_);
        c.info = dalloc<uint32_t>(n_); c.orig = dalloc<uint32_t>(n_);
    }

    // Allocate device state for the host particle set and upload it (IC order).
    void commit(const HostParticles& hp) {
        free_all();
        n_ = (int64_t)hp.mass.size();
        if (n_ <= 0) throw Err{DSPH_E_ARG, "no particles"};
        if (n_ > kMaxParticles) throw Err{DSPH_E_RANGE, "more than 134,217,664 particles in one context"};
        alloc_carried(bufs_[0]);
        alloc_carried(bufs_[1]);
        cur_ = 0;
        Work<R>& w = w_;
        w.n = n_;
        w.nseg = nseg_;
        {   // DSPH_COUNT_REWALKS=1: count the density moment walks (dsph_diag.density_*)
            const char* e = std::getenv("DSPH_COUNT_REWALKS");
            w.stats = (e && e[0] == '1') ? 1 : 0;
        }
        w.ax = dalloc<R>(n_); w.ay = dalloc<R>(n_); w.az = dalloc<R>(n_);
        w.rho = dalloc<R>(n_); w.P = dalloc<R>(n_); w.cs = dalloc<R>(n_); w.csig = dalloc<R>(n_);
        w.Pf = dalloc<R>(n_); w.Rart = dalloc<R>(n_); w.dscale = dalloc<R>(n_); w.sigmax = dalloc<R>(n_);
        w.eps_work = dalloc<R>(n_);
        w.dudt = dalloc<R>(n_); w.dSdt = dalloc<R>(6 * n_);
        w.dudt_old = dalloc<R>(n_); w.dSdt_old = dalloc<R>(6 * n_);
        w.dens_rec = dalloc<V4>(n_ + kRecordPad);
        w.dens_rec64 = density64_ ? dalloc<double4>(n_) : nullptr;
        w.frec_pos = dalloc<V4>(n_ + kRecordPad); w.frec_vel = dalloc<V4>(n_); w.frec_th = dalloc<V4>(n_);
        HIPC(hipMemsetAsync(w.dens_rec + n_, 0, kRecordPad * sizeof(V4), stream_));
        HIPC(hipMemsetAsync(w.frec_pos + n_, 0, kRecordPad * sizeof(V4), stream_));
        w.frec_sa = dalloc<V4>(n_); w.frec_sb = dalloc<V4>(n_);
        w.cellc = dalloc<uint32_t>(n_);
        w.dev_of = dalloc<uint32_t>(n_);
        w.perm = dalloc<uint32_t>(n_);
        nparts_max_ = (int)blocks_for(n_, 64) + 1;
        w.parts = dalloc<BlockPart>(nparts_max_);
        alloc_table();
        w.h0 = dalloc<R>(n_);
        alloc_lists();
        // sort buffers
        nb_sort_ = (int)blocks_for(n_, kRsThreads * rs_items(n_));
        for (int k = 0; k < 2; k++) { keys_[k] = dalloc<uint32_t>(n_); vals_[k] = dalloc<uint32_t>(n_); }
        hist_ = dalloc<uint32_t>((size_t)256 * nb_sort_);
        digit_totals_ = dalloc<uint32_t>(256);
        d_total_cells_ = dalloc<uint32_t>(3);   // cells laid out, cells wanted, slice blocks laid out
        zero(d_total_cells_, 3);
        // per-segment tables
        std::vector<MatT<R>> hm(mats_.size());
        for (size_t k = 0; k < mats_.size(); k++) hm[k] = to_mat(mats_[k]);
        d_mats_ = dalloc<MatT<R>>(hm.size());
        HIPC(hipMemcpyAsync((void*)d_mats_, hm.data(), hm.size() * sizeof(MatT<R>), hipMemcpyHostToDevice, stream_));
        std::vector<MatT<double>> hm64(mats_.size());
        for (size_t k = 0; k < mats_.size(); k++) hm64[k] = Engine<double>::to_mat(mats_[k]);
        d_mats64_ = dalloc<MatT<double>>(hm64.size());
        HIPC(hipMemcpyAsync((void*)d_mats64_, hm64.data(), hm64.size() * sizeof(MatT<double>), hipMemcpyHostToDevice,
                            stream_));
        std::vector<SegPar> hpar(nseg_);
        for (int s = 0; s < nseg_; s++) hpar[s] = to_par(segp_[s]);
        d_par_ = dalloc<SegPar>(nseg_);
        HIPC(hipMemcpyAsync(d_par_, hpar.data(), nseg_ * sizeof(SegPar), hipMemcpyHostToDevice, stream_));
        d_grid_ = dalloc<SegGrid>(nseg_);
        d_st_ = dalloc<SegState>(nseg_);
        HIPC(hipMemsetAsync(d_grid_, 0, nseg_ * sizeof(SegGrid), stream_));
        std::vector<SegState> hst(nseg_);
        for (auto& s : hst) {
            std::memset(&s, 0, sizeof(SegState));
            s.dt_next = 0x7ff0000000000000ull;
        }
        HIPC(hipMemcpyAsync(d_st_, hst.data(), nseg_ * sizeof(SegState), hipMemcpyHostToDevice, stream_));
        d_segdt_ = dalloc<double>(nseg_);
        w.mats = d_mats_;
        w.mats64 = d_mats64_;
        w.par = d_par_;
        w.grid = d_grid_;
        w.st = d_st_;
        // segment ranges (IC order == any sorted order: segments never interleave)
        seg_off_.assign(nseg_ + 1, 0);
        for (int64_t i = 0; i < n_; i++) seg_off_[hp.seg[i] + 1]++;
        for (int s = 0; s < nseg_; s++) seg_off_[s + 1] += seg_off_[s];
        setup_diag_chunks();
        stage_ = dalloc<double>(6 * n_);
        // zero work arrays
        zero(w.ax, n_); zero(w.ay, n_); zero(w.az, n_); zero(w.rho, n_); zero(w.P, n_); zero(w.cs, n_);
        zero(w.csig, n_); zero(w.Pf, n_); zero(w.Rart, n_); zero(w.sigmax, n_); zero(w.eps_work, n_);
        zero(w.dudt, n_); zero(w.dSdt, 6 * n_); zero(w.dudt_old, n_); zero(w.dSdt_old, 6 * n_);
        fill_value(w.dscale, n_, R(1));
        // carried state, IC order
        Carried<R>& c = bufs_[0];
        std::vector<double> tmp(n_);
        for (int a = 0; a < 3; a++) {
            for (int64_t i = 0; i < n_; i++) tmp[i] = hp.pos[3 * i + a];
            upload(a == 0 ? c.x : (a == 1 ? c.y : c.z), tmp);
            for (int64_t i = 0; i < n_; i++) tmp[i] = hp.vel[3 * i + a];
            upload(a == 0 ? c.vx : (a == 1 ? c.vy : c.vz), tmp);
        }
        upload(c.u, hp.u);
        upload(c.mass, hp.mass);
        upload_conv(c.h, hp.h);
        upload_conv(c.D, hp.D);
        upload_conv(c.eps_act, hp.eps_act);
        std::vector<R> sc(6 * n_);
        for (int k = 0; k < 6; k++)
            for (int64_t i = 0; i < n_; i++) sc[(size_t)k * n_ + i] = (R)hp.S[6 * i + k];
        HIPC(hipMemcpyAsync(c.S, sc.data(), sc.size() * sizeof(R), hipMemcpyHostToDevice, stream_));
        std::vector<uint32_t> info(n_), orig(n_);
        for (int64_t i = 0; i < n_; i++) {
            info[i] = info_pack(hp.mat[i], hp.fixed[i] != 0, hp.seg[i]);
            orig[i] = (uint32_t)i;
        }
        HIPC(hipMemcpyAsync(c.info, info.data(), n_ * 4, hipMemcpyHostToDevice, stream_));
        HIPC(hipMemcpyAsync(c.orig, orig.data(), n_ * 4, hipMemcpyHostToDevice, stream_));
        HIPC(hipStreamSynchronize(stream_));
        built_ = false;
        records_ = false;
        have_dt_ = false;
        h_host_set_ = true;
    }

    // Neighbour lists: nbr_cap entries per particle, in 64-particle blocks.
    // Automatic choice (neighbor_capacity 0): column walks in both paths, which
    // cost less than building the lists and gathering through them; the
    // autotuner also tries lists of 128 entries (about 1.5x the mean list
    // length of a uniform lattice at the default skin).
    int list_capacity() const {
        if (density64_ || tune_.neighbor_capacity < 0) return 0;
        if (tune_.neighbor_capacity > 0) return tune_.neighbor_capacity;
        return 0;
    }
    void alloc_lists() {
        const int cap = (list_capacity() + 3) / 4 * 4;
        const int64_t padded = (n_ + 63) / 64 * 64;
        w_.nbr_cap = cap;
        w_.nbr = cap ? dalloc<uint4>((size_t)padded * (size_t)(cap / 4)) : nullptr;
        w_.nbr_cnt = dalloc<uint32_t>(n_);
        w_.skin_build = (R)(1.0 + tune_.list_skin);
        w_.skin_check = (R)(1.0 + 0.5 * tune_.list_skin);
        lists_built_ = false;
    }
    void free_ptr(void* old) {
        auto it = std::find(allocs_.begin(), allocs_.end(), old);
        if (it != allocs_.end()) { (void)hipFree(*it); allocs_.erase(it); }
    }

    void alloc_table() {
        int64_t cap = std::max<int64_t>(1 << 16, (int64_t)tune_.cell_capacity_per_particle * n_ + 65 * (int64_t)nseg_);
        cap = std::min<int64_t>(cap, (int64_t)0xfffffff0);
        capacity_ = (uint32_t)cap;
        w_.cstart = dalloc<uint32_t>((size_t)capacity_ + 1);
        w_.cell_hmax = dalloc<float>(capacity_);
        chunk_lb_ = dalloc<uint32_t>((size_t)capacity_ / kStartChunk + 3);
        int bits = 1;
        while (bits < 32 && ((1ull << bits) <= (unsigned long long)capacity_)) bits++;
        passes_ = (bits + 7) / 8;
    }

    template <class T>
    void zero(T* p, int64_t count) { HIPC(hipMemsetAsync(p, 0, count * sizeof(T), stream_)); }
    void fill_value(R* p, int64_t count, R v) {
        std::vector<R> h(count, v);
        HIPC(hipMemcpyAsync(p, h.data(), count * sizeof(R), hipMemcpyHostToDevice, stream_));
        HIPC(hipStreamSynchronize(stream_));
    }
    void upload(double* d, const std::vector<double>& h) {
        HIPC(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, stream_));
        HIPC(hipStreamSynchronize(stream_));
    }
    void upload_conv(R* d, const std::vector<double>& h) {
        std::vector<R> t(h.size());
        for (size_t i = 0; i < h.size(); i++) t[i] = (R)h[i];
        HIPC(hipMemcpyAsync(d, t.data(), t.size() * sizeof(R), hipMemcpyHostToDevice, stream_));
        HIPC(hipStreamSynchronize(stream_));
    }

    static MatT<R> to_mat(const dsph_material& m) {
        MatT<R> r;
        std::memset(&r, 0, sizeof(r));
        r.type = m.type;
        r.gamma = (R)m.gamma;
        r.rho0 = (R)m.rho0; r.A = (R)m.A; r.B = (R)m.B; r.a = (R)m.a; r.b = (R)m.b;
        r.alpha = (R)m.alpha; r.beta = (R)m.beta; r.u0 = (R)m.u0; r.uiv = (R)m.uiv; r.ucv = (R)m.ucv;
        r.G = (R)m.G; r.Y = (R)m.Y; r.wm = (R)m.wm; r.wk = (R)m.wk;
        double E = m.Emod;
        if (E == 0.0 && m.type == DSPH_EOS_TILLOTSON && m.G > 0) E = 9.0 * m.A * m.G / (3.0 * m.A + m.G);
        r.Emod = (R)E;
        return r;
    }
    static SegPar to_par(const dsph_params& p) {
        SegPar s;
        std::memset(&s, 0, sizeof(s));
        s.eta = p.eta; s.h_init = p.h_init; s.alpha = p.alpha; s.beta = p.beta; s.eps = p.eps;
        s.alpha_u = p.alpha_u; s.cfl = p.cfl; s.eps_as = p.eps_as; s.g = p.g; s.damp = p.damp;
        s.t_end = p.t_end;
        for (int a = 0; a < 3; a++) { s.dlo[a] = p.lo[a]; s.dhi[a] = p.hi[a]; s.per[a] = p.periodic[a]; }
        s.damage_on = p.damage_on;
        return s;
    }

    // ---------------------------------------------------------------- profiling
    struct Pending {
        int phase;
        hipEvent_t a, b;
    };
    hipEvent_t get_event() {
        if (!ev_pool_.empty()) {
            hipEvent_t e = ev_pool_.back();
            ev_pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPC(hipEventCreate(&e));
        return e;
    }
    void drain_profile(bool accumulate) {
        if (pending_.empty()) return;
        if (accumulate) (void)hipEventSynchronize(pending_.back().b);
        for (auto& p : pending_) {
            if (accumulate) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                    prof_total_[p.phase] += ms;
                    prof_count_[p.phase] += 1;
                }
            }
            ev_pool_.push_back(p.a);
            ev_pool_.push_back(p.b);
        }
        pending_.clear();
    }
    template <class K, class... Args>
    void launch(int phase, const char* name, K kernel, dim3 grid, dim3 block, Args... args) {
        launch_lds(phase, name, kernel, grid, block, 0, args...);
    }
    // launch with `lds` bytes of dynamic LDS per workgroup
    template <class K, class... Args>
    void launch_lds(int phase, const char* name, K kernel, dim3 grid, dim3 block, size_t lds, Args... args) {
        hipEvent_t ea = nullptr, eb = nullptr;
        const bool prof = profiling_ && !capturing_;
        if (prof) {
            ea = get_event();
            eb = get_event();
            HIPC(hipEventRecord(ea, stream_));
        }
        hipLaunchKernelGGL(kernel, grid, block, lds, stream_, args...);
        check_launch(name);
        if (prof) {
            HIPC(hipEventRecord(eb, stream_));
            pending_.push_back(Pending{phase, ea, eb});
        }
    }

    // ---------------------------------------------------------------- reductions
    unsigned long long* st_field(size_t offset) { return (unsigned long long*)((char*)d_st_ + offset); }
    SegTargets<6> bbox_targets() {
        SegTargets<6> t;
        for (int a = 0; a < 3; a++) {
            t.base[a] = st_field(offsetof(SegState, bmin) + 8 * a);
            t.base[3 + a] = st_field(offsetof(SegState, bmax) + 8 * a);
            t.is_max[a] = 0; t.is_max[3 + a] = 1;
            t.ordered[a] = 1; t.ordered[3 + a] = 1;
        }
        return t;
    }
    SegTargets<1> hmax_target(int slot) {
        SegTargets<1> t;
        t.base[0] = st_field(offsetof(SegState, hmax) + 8 * slot);
        t.is_max[0] = 1;
        t.ordered[0] = 0;
        return t;
    }
    SegTargets<1> dt_target() {
        SegTargets<1> t;
        t.base[0] = st_field(offsetof(SegState, dt_next));
        t.is_max[0] = 0;
        t.ordered[0] = 0;
        return t;
    }
    // Fold the per-block partials of the kernel just launched (its own phase
    // keeps only that kernel's time; the fold is accounted as "reduce").
    // About four records per thread, so the fold stays a few microseconds at
    // any particle count (each thread's records are read one after another).
    template <int K>
    void fold_parts(int, const SegTargets<K>& t, unsigned nblocks) {
        const unsigned grid = std::max(1u, (nblocks + 4095u) / 4096u);
        launch(PH_REDUCE, "parts_reduce", k_parts_reduce<K>, dim3(grid), dim3(1024), t, sizeof(SegState) / 8,
               (const BlockPart*)w_.parts, (int)nblocks);
    }

    // ---------------------------------------------------------------- phases
    Carried<R>& cur() { return bufs_[cur_]; }

    void build(int move, bool carry_all) {
        const int bp = tune_.block_point;
        launch(PH_ROUND, "build_begin", k_build_begin, dim3(1), dim3(256), d_st_, nseg_);
        const auto tb = bbox_targets();
        launch(PH_KICK_DRIFT, "kick_drift", k_kick_drift<R>, dim3(blocks_for(n_, bp)), dim3(bp), cur(), w_, move, tb);
        fold_parts(PH_KICK_DRIFT, tb, blocks_for(n_, bp));
        launch(PH_GRID, "grid", k_grid, dim3(1), dim3(kGridThreads), d_grid_, d_st_, (const SegPar*)d_par_, nseg_,
               tune_.cell_factor, capacity_, d_total_cells_);
        launch(PH_KEYS, "keys", k_keys<R>, dim3(blocks_for(n_, bp)), dim3(bp), cur(), w_, keys_[0], vals_[0],
               (const uint32_t*)d_total_cells_);
        int ka = 0;
        for (int p = 0; p < passes_; p++) {
            const int shift = 8 * p;
            const bool small = rs_items(n_) == 8;
            launch(PH_SORT, "rs_hist", small ? k_rs_hist<8> : k_rs_hist<16>, dim3(nb_sort_), dim3(kRsThreads),
                   (const uint32_t*)keys_[ka], n_, shift, hist_, nb_sort_);
            launch(PH_SORT, "rs_scan", k_rs_scan, dim3(256), dim3(256), hist_, nb_sort_, digit_totals_);
            launch(PH_SORT, "rs_scatter", small ? k_rs_scatter<8> : k_rs_scatter<16>, dim3(nb_sort_),
                   dim3(kRsThreads), (const uint32_t*)keys_[ka], (const uint32_t*)vals_[ka], keys_[1 - ka],
                   vals_[1 - ka], n_, shift, (const uint32_t*)hist_, nb_sort_, (const uint32_t*)digit_totals_);
            ka = 1 - ka;
        }
        // first slot of every cell (empty cells included), from the sorted keys
        launch(PH_SORT, "chunk_start", k_chunk_start, dim3(std::min<unsigned>(blocks_for(n_ + 1, 256), 4096u)),
               dim3(256), (const uint32_t*)keys_[ka], n_, chunk_lb_, (const uint32_t*)d_total_cells_);
        const unsigned ncs = std::min<unsigned>(blocks_for((int64_t)capacity_ + 1, 256 * kCellChunk), 1u << 20);
        launch(PH_SORT, "cell_start", k_cell_start, dim3(ncs), dim3(256), (const uint32_t*)keys_[ka], n_, w_.cstart,
               (const uint32_t*)chunk_lb_, (const uint32_t*)d_total_cells_);
        const auto th = hmax_target(0);
        launch(PH_PERMUTE, "permute", k_permute<R>, dim3(blocks_for(n_, bp)), dim3(bp), bufs_[cur_],
               bufs_[1 - cur_], w_, (const uint32_t*)keys_[ka], (const uint32_t*)vals_[ka], th, w_.nbr_cap > 0 ? 1 : 0);
        fold_parts(PH_PERMUTE, th, blocks_for(n_, bp));
        cur_ = 1 - cur_;
        built_ = true;
        records_ = false;
        lists_built_ = false;
        if (carry_all) {
            // Outside a step every per-particle field is state the caller may read
            // or feed back (System keeps them between calls), so carry them through
            // the re-sort as well.
            const uint32_t* perm = w_.perm;
            R* scratch = (R*)stage_;
            Work<R>& w = w_;
            for (R* f : {w.ax, w.ay, w.az, w.rho, w.P, w.cs, w.csig, w.Pf, w.Rart, w.dscale, w.sigmax,
                         w.eps_work}) {
                launch(PH_PERMUTE, "gather", k_gather<R>, dim3(blocks_for(n_, bp)), dim3(bp), (const R*)f, scratch,
                       perm, n_);
                HIPC(hipMemcpyAsync(f, scratch, n_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
            }
            HIPC(hipMemcpyAsync(w.dudt, w.dudt_old, n_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
            HIPC(hipMemcpyAsync(w.dSdt, w.dSdt_old, 6 * n_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
        }
    }

    // Up to four rounds of two iterations; rounds a segment does not need exit
    // at once on the device.  Round 0's first iteration builds the neighbour
    // lists that its second iteration and the force pass read.  Contexts with a
    // periodic segment run the pair kernels compiled with periodic images.
    void density() {
        if (periodic_) density_rounds<true>();
        else density_rounds<false>();
        launch(PH_ROUND, "round", k_round, dim3(1), dim3(256), d_st_, d_grid_, (const SegPar*)d_par_, nseg_, 4);
        records_ = true;
    }
    template <bool PER>
    void density_rounds() {
        const int bq = tune_.block_pair;
        const unsigned nb = blocks_for(n_, bq);
        const int lists = w_.nbr_cap > 0 ? 1 : 0;
        for (int r = 0; r < 4; r++) {
            launch(PH_ROUND, "round", k_round, dim3(1), dim3(256), d_st_, d_grid_, (const SegPar*)d_par_, nseg_, r);
            const auto th = hmax_target(r + 1);
            if constexpr (sizeof(R) == 4) {
                if (density64_) {
                    if (nseg_ == 1)
                        launch_lds(PH_DENSITY, "density", k_density_round_hp<PER, true>, dim3(nb), dim3(bq),
                                   dens_band_lds(bq), cur(), w_, r, tune_.xcd_remap, th);
                    else
                        launch_lds(PH_DENSITY, "density", k_density_round_hp<PER, false>, dim3(nb), dim3(bq),
                                   dens_band_lds(bq), cur(), w_, r, tune_.xcd_remap, th);
                    fold_parts(PH_DENSITY, th, nb);
                    continue;
                }
            }
            if (r == 0 && lists) {
                launch(PH_DENSITY_BUILD, "density_build", k_density_build<R, PER>, dim3(nb), dim3(bq), cur(), w_,
                       tune_.xcd_remap);
                launch(PH_DENSITY, "density", k_density<R, true, PER>, dim3(nb), dim3(bq), cur(), w_, r,
                       tune_.xcd_remap, lists, th);
            } else {
                const size_t lds = sizeof(R) == 8 ? dens_band_lds(bq) : 0;
                if (nseg_ == 1)
                    launch_lds(PH_DENSITY, "density", k_density_round<R, PER, true>, dim3(nb), dim3(bq), lds, cur(),
                               w_, r, tune_.xcd_remap, lists, th);
                else
                    launch_lds(PH_DENSITY, "density", k_density_round<R, PER, false>, dim3(nb), dim3(bq), lds,
                               cur(), w_, r, tune_.xcd_remap, lists, th);
            }
            fold_parts(PH_DENSITY, th, nb);
            if (r == 0 && lists) {
                launch(PH_DENSITY_LIST, "density_list", k_density_list<R, PER>, dim3(nb), dim3(bq), cur(), w_,
                       tune_.xcd_remap, th);
                fold_parts(PH_DENSITY, th, nb);
            }
        }
        lists_built_ = lists != 0;
    }

    void finalize_rounds() {
        launch(PH_ROUND, "round", k_round, dim3(1), dim3(256), d_st_, d_grid_, (const SegPar*)d_par_, nseg_, 4);
    }

    void force() {
        if (periodic_) force_pass<true>();
        else force_pass<false>();
        have_dt_ = true;
    }
    template <bool PER>
    void force_pass() {
        const int bq = tune_.block_pair;
        const auto td = dt_target();
        launch(PH_ROUND, "reset_dt", k_reset_dt, dim3(1), dim3(256), d_st_, nseg_);
        const unsigned nb = blocks_for(n_, bq);
        const int lists = lists_built_ ? 1 : 0;
        if (nseg_ == 1)
            launch(PH_FORCE, "force", k_force<R, PER, true>, dim3(nb), dim3(bq), cur(), w_, tune_.xcd_remap, lists, td);
        else
            launch(PH_FORCE, "force", k_force<R, PER, false>, dim3(nb), dim3(bq), cur(), w_, tune_.xcd_remap, lists,
                   td);
        if (lists) {
            fold_parts(PH_FORCE, td, nb);
            launch(PH_FORCE_LIST, "force_list", k_force_list<R, PER>, dim3(nb), dim3(bq), cur(), w_,
                   tune_.xcd_remap, td);
        }
        fold_parts(PH_FORCE, td, nb);
    }

    void finish(int damage, int kick) {
        const int bp = tune_.block_point;
        launch(PH_FINISH, "finish", k_finish<R>, dim3(blocks_for(n_, bp)), dim3(bp), cur(), w_, damage, kick);
    }

    void step(int mode, double dt) {
        if (mode == 0 && !have_dt_) throw Err{DSPH_E_STATE, "no device time step yet: call dsph_init first"};
        launch(PH_STEP_BEGIN, "step_begin", k_step_begin, dim3(1), dim3(256), d_st_, (const SegPar*)d_par_, nseg_,
               mode, dt, (const double*)d_segdt_);
        build(1, false);
        density();
        force();
        finish(1, 1);
    }

    // ---------------------------------------------------------------- API ops
    void init() {
        build(0, false);
        density();
        force();
    }
    void compute_density_h() {
        build(0, true);
        density();
    }
    void compute_forces() {
        if (!built_) {
            build(0, true);
            finalize_rounds();
        }
        if (!records_) {
            const int bp = tune_.block_point;
            if (density64_)
                launch(PH_EOS, "eos_records", k_eos_records<R, true>, dim3(blocks_for(n_, bp)), dim3(bp), cur(), w_);
            else
                launch(PH_EOS, "eos_records", k_eos_records<R, false>, dim3(blocks_for(n_, bp)), dim3(bp), cur(), w_);
            records_ = true;
        }
        force();
    }
    void grow_damage(double dt) {
        std::vector<double> d(nseg_, dt);
        HIPC(hipMemcpyAsync(d_segdt_, d.data(), nseg_ * 8, hipMemcpyHostToDevice, stream_));
        set_dt_cur_kernel();
        finish(1, 0);
    }
    void set_dt_cur_kernel() {
        // dt_cur := d_segdt without advancing time or the step counter
        launch(PH_STEP_BEGIN, "set_dt", k_set_dt, dim3(1), dim3(256), d_st_, nseg_, (const double*)d_segdt_);
    }

    void step_host(double dt) { step(dt > 0 ? 1 : 0, dt); }
    void step_dts(const double* dts) {
        HIPC(hipMemcpyAsync(d_segdt_, dts, nseg_ * 8, hipMemcpyHostToDevice, stream_));
        step(2, 0.0);
    }

    void run(int nsteps, double dt, bool use_graph) {
        const int mode = dt > 0 ? 1 : 0;
        if (mode == 0 && !have_dt_) throw Err{DSPH_E_STATE, "no device time step yet: call dsph_init first"};
        if (!use_graph || profiling_) {
            for (int s = 0; s < nsteps; s++) step(mode, dt);
            return;
        }
        auto key = std::make_tuple(cur_, mode, dt);
        hipGraphExec_t exec = nullptr;
        auto it = graphs_.find(key);
        if (it == graphs_.end()) {
            hipGraph_t g;
            capturing_ = true;
            HIPC(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try {
                step(mode, dt);
                step(mode, dt);
            } catch (...) {
                hipGraph_t tmp;
                (void)hipStreamEndCapture(stream_, &tmp);
                capturing_ = false;
                throw;
            }
            HIPC(hipStreamEndCapture(stream_, &g));
            capturing_ = false;
            HIPC(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
            HIPC(hipGraphDestroy(g));
            graphs_[key] = exec;
        } else {
            exec = it->second;
        }
        for (int s = 0; s + 1 < nsteps; s += 2) HIPC(hipGraphLaunch(exec, stream_));
        if (nsteps % 2) step(mode, dt);
    }
    void release_graphs() {
        for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
        graphs_.clear();
    }

    void sync() { HIPC(hipStreamSynchronize(stream_)); }

    std::vector<SegState> read_state() {
        std::vector<SegState> h(nseg_);
        HIPC(hipMemcpyAsync(h.data(), d_st_, nseg_ * sizeof(SegState), hipMemcpyDeviceToHost, stream_));
        sync();
        return h;
    }

    // Restore a segment's clock: simulated time, step count and the CFL step
    // the last force pass left (so a restarted run takes the same next step).
    void set_clock(int seg, double t, int64_t steps, double dt_next) {
        auto st = read_state();
        for (int s = 0; s < nseg_; s++) {
            if (seg >= 0 && s != seg) continue;
            st[s].t = t;
            st[s].steps = steps;
            unsigned long long b;
            std::memcpy(&b, &dt_next, 8);
            st[s].dt_next = b;
        }
        HIPC(hipMemcpyAsync(d_st_, st.data(), nseg_ * sizeof(SegState), hipMemcpyHostToDevice, stream_));
        sync();
        if (dt_next > 0.0) have_dt_ = true;
        release_graphs();
    }

    // The device's trouble flags of one segment as a status and a message (the
    // flags are sticky: once set, every later call of the context reports them).
    static int seg_status(const SegState& s, int seg, std::string* msg) {
        const std::string where = " (segment " + std::to_string(seg) + ")";
        if (s.bad & 1) { if (msg) *msg = "cell table capacity exceeded"; return DSPH_E_RANGE; }
        if (s.bad & 4) { if (msg) *msg = "non-finite particle position" + where; return DSPH_E_NUMERIC; }
        if (s.bad & 2) { if (msg) *msg = "non-finite or non-positive time step" + where; return DSPH_E_NUMERIC; }
        if (s.bad & 8) {
            if (msg) *msg = "a smoothing length far above 2 h_init reaches past two slice blocks; "
                            "pairs beyond them were not summed" + where;
            return DSPH_E_RANGE;
        }
        return DSPH_OK;
    }
    // Synchronises the stream, then raises the first segment's trouble flags.
    int check_health() {
        auto st = read_state();
        for (int s = 0; s < nseg_; s++) {
            std::string msg;
            const int rc = seg_status(st[s], s, &msg);
            if (rc != DSPH_OK) throw Err{rc, msg};
        }
        grow_table();
        return 0;
    }

    // Cell table growth.  The sort grid spans the particles' bounding box, which
    // a few fast particles can stretch far beyond the bulk; when the nominal
    // widths need more cells than the table holds, k_grid widens every cell
    // (correct, but every walk then visits more candidates).  After a call has
    // synchronised, the table is regrown to 1.25x what the last grid wanted, up
    // to kMaxCellsPerParticle cells per particle, so that later steps run at the
    // tuned widths (the autotuner grows it before timing a candidate and sizes
    // it for the chosen one).  The cells' order, hence each pair sum's order,
    // changes with the widths: a context's results depend on when its table
    // grew at round-off level only.  Returns true when the table was
    // reallocated.
    uint32_t wanted_cells() {
        uint32_t tc[2] = {0, 0};
        HIPC(hipMemcpyAsync(tc, d_total_cells_, sizeof(tc), hipMemcpyDeviceToHost, stream_));
        sync();
        return tc[1];
    }
    bool grow_table() {
        if (n_ <= 0 || !d_total_cells_) return false;
        const uint32_t wanted = wanted_cells();
        if (wanted <= capacity_) return false;
        return resize_table(wanted, false);
    }
    // A table of 1.25x `wanted` cells (at least the default 4 per particle, at
    // most kMaxCellsPerParticle); only ever smaller than the current one if
    // `shrink`.
    bool resize_table(uint32_t wanted, bool shrink) {
        const int cur = tune_.cell_capacity_per_particle;
        const double want = 1.25 * (double)wanted / (double)n_;
        const int per = (int)std::max<double>(kDefaultCellsPerParticle,
                                              std::min<double>(kMaxCellsPerParticle, std::ceil(want)));
        if (per == cur || (per < cur && !shrink)) return false;
        release_graphs();
        for (void* old : {(void*)w_.cstart, (void*)w_.cell_hmax, (void*)chunk_lb_}) free_ptr(old);
        tune_.cell_capacity_per_particle = per;
        try {
            alloc_table();
        } catch (const Err&) {
            // not enough memory: back to a table of the old size (coarser cells)
            (void)hipGetLastError();
            for (void* part : {(void*)w_.cstart, (void*)w_.cell_hmax, (void*)chunk_lb_}) free_ptr(part);
            tune_.cell_capacity_per_particle = cur;
            alloc_table();
            return false;
        }
        built_ = false;
        records_ = false;
        return true;
    }

    void timestep(double* out) {
        if (!have_dt_) throw Err{DSPH_E_STATE, "no force evaluation yet: call dsph_init first"};
        auto st = read_state();
        for (int s = 0; s < nseg_; s++) out[s] = bits_to_double(st[s].dt_next);
        check_health();
    }

    // ---------------------------------------------------------------- fields
    int width_of(int f) const {
        switch (f) {
            case DSPH_F_POS: case DSPH_F_VEL: case DSPH_F_ACC: return 3;
            case DSPH_F_S: case DSPH_F_DSDT: case DSPH_F_DSDT_OLD: return 6;
            default: return 1;
        }
    }
    // Device component arrays of a float field: FP64 (dbl) or R (rr).
    void comps(int f, std::vector<double*>& dbl, std::vector<R*>& rr) {
        Carried<R>& c = cur();
        Work<R>& w = w_;
        auto six = [&](R* base) { for (int k = 0; k < 6; k++) rr.push_back(base + (size_t)k * n_); };
        switch (f) {
            case DSPH_F_POS: dbl = {c.x, c.y, c.z}; break;
            case DSPH_F_VEL: dbl = {c.vx, c.vy, c.vz}; break;
            case DSPH_F_U: dbl = {c.u}; break;
            case DSPH_F_MASS: dbl = {c.mass}; break;
            case DSPH_F_ACC: rr = {w.ax, w.ay, w.az}; break;
            case DSPH_F_RHO: rr = {w.rho}; break;
            case DSPH_F_DUDT: rr = {w.dudt}; break;
            case DSPH_F_P: rr = {w.P}; break;
            case DSPH_F_CS: rr = {w.cs}; break;
            case DSPH_F_CSIG: rr = {w.csig}; break;
            case DSPH_F_HH: rr = {c.h}; break;
            case DSPH_F_EPS_WORK: rr = {w.eps_work}; break;
            case DSPH_F_D: rr = {c.D}; break;
            case DSPH_F_EPS_ACT: rr = {c.eps_act}; break;
            case DSPH_F_SIGMAX: rr = {w.sigmax}; break;
            case DSPH_F_PF: rr = {w.Pf}; break;
            case DSPH_F_RART: rr = {w.Rart}; break;
            case DSPH_F_DSCALE: rr = {w.dscale}; break;
            case DSPH_F_S: six(c.S); break;
            case DSPH_F_DSDT: six(w.dSdt); break;
            case DSPH_F_DUDT_OLD: rr = {w.dudt_old}; break;
            case DSPH_F_DSDT_OLD: six(w.dSdt_old); break;
            default: throw Err{DSPH_E_ARG, "unknown field " + std::to_string(f)};
        }
    }

    void get_field(int f, void* dst) {
        const unsigned nb = blocks_for(n_, 256);
        const uint32_t* orig = cur().orig;
        if (f == DSPH_F_MAT || f == DSPH_F_FIXED || f == DSPH_F_SEG) {
            const int kind = f == DSPH_F_MAT ? 0 : (f == DSPH_F_FIXED ? 1 : 2);
            launch(PH_IO, "export_info", k_export_info, dim3(nb), dim3(256), (const uint32_t*)cur().info,
                   (void*)stage_, kind, orig, n_);
            HIPC(hipMemcpyAsync(dst, stage_, n_ * (kind == 1 ? 1 : 4), hipMemcpyDeviceToHost, stream_));
            sync();
            return;
        }
        std::vector<double*> dbl;
        std::vector<R*> rr;
        comps(f, dbl, rr);
        const int width = width_of(f);
        for (int k = 0; k < width; k++) {
            if (!dbl.empty())
                launch(PH_IO, "export", k_export<double>, dim3(nb), dim3(256), (const double*)dbl[k], stage_, width, k,
                       orig, n_);
            else
                launch(PH_IO, "export", k_export<R>, dim3(nb), dim3(256), (const R*)rr[k], stage_, width, k, orig, n_);
        }
        HIPC(hipMemcpyAsync(dst, stage_, (size_t)n_ * width * 8, hipMemcpyDeviceToHost, stream_));
        sync();
    }

    void set_field(int f, const void* src) {
        const unsigned nb = blocks_for(n_, 256);
        const uint32_t* orig = cur().orig;
        if (f == DSPH_F_SEG) throw Err{DSPH_E_ARG, "the segment of a particle is fixed at upload"};
        if (f == DSPH_F_MAT || f == DSPH_F_FIXED) {
            const int kind = f == DSPH_F_MAT ? 0 : 1;
            if (kind == 0) {
                const int32_t* m = (const int32_t*)src;
                for (int64_t i = 0; i < n_; i++)
                    if (m[i] < 0 || m[i] >= (int)mats_.size()) throw Err{DSPH_E_ARG, "material index out of range"};
            }
            HIPC(hipMemcpyAsync(stage_, src, n_ * (kind == 1 ? 1 : 4), hipMemcpyHostToDevice, stream_));
            launch(PH_IO, "import_info", k_import_info, dim3(nb), dim3(256), (const void*)stage_, cur().info, kind,
                   orig, n_);
            sync();
            records_ = false;
            return;
        }
        std::vector<double*> dbl;
        std::vector<R*> rr;
        comps(f, dbl, rr);
        const int width = width_of(f);
        HIPC(hipMemcpyAsync(stage_, src, (size_t)n_ * width * 8, hipMemcpyHostToDevice, stream_));
        for (int k = 0; k < width; k++) {
            if (!dbl.empty())
                launch(PH_IO, "import", k_import<double>, dim3(nb), dim3(256), (const double*)stage_, dbl[k], width, k,
                       orig, n_);
            else
                launch(PH_IO, "import", k_import<R>, dim3(nb), dim3(256), (const double*)stage_, rr[k], width, k, orig,
                       n_);
        }
        sync();
        if (f == DSPH_F_POS) built_ = false;
        records_ = false;
        lists_built_ = false;      // host-written state: the force pass walks cells
    }

    // ---------------------------------------------------------------- diagnostics
    void setup_diag_chunks() {
        const int64_t C = 16384;
        std::vector<DiagChunk> ch;
        std::vector<int32_t> c0(nseg_ + 1, 0);
        for (int s = 0; s < nseg_; s++) {
            c0[s] = (int32_t)ch.size();
            for (int64_t b = seg_off_[s]; b < seg_off_[s + 1]; b += C)
                ch.push_back(DiagChunk{s, 0, b, std::min(seg_off_[s + 1], b + C)});
        }
        c0[nseg_] = (int32_t)ch.size();
        nchunks_ = (int)ch.size();
        d_chunks_ = dalloc<DiagChunk>(std::max(1, nchunks_));
        d_parts_ = dalloc<DiagPart>(std::max(1, nchunks_));
        d_chunk0_ = dalloc<int32_t>(nseg_ + 1);
        d_diag_ = dalloc<DiagPart>(nseg_);
        if (nchunks_)
            HIPC(hipMemcpyAsync(d_chunks_, ch.data(), ch.size() * sizeof(DiagChunk), hipMemcpyHostToDevice, stream_));
        HIPC(hipMemcpyAsync(d_chunk0_, c0.data(), c0.size() * 4, hipMemcpyHostToDevice, stream_));
    }

    void diagnostics(int seg, dsph_diag* out) {
        if (nchunks_ > 0)
            launch(PH_DIAG, "diag_partial", k_diag_partial<R>, dim3(nchunks_), dim3(256), cur(), w_,
                   (const DiagChunk*)d_chunks_, d_parts_);
        launch(PH_DIAG, "diag_final", k_diag_final, dim3(blocks_for(nseg_, 64)), dim3(64), (const DiagPart*)d_parts_,
               (const int32_t*)d_chunk0_, nseg_, d_diag_);
        std::vector<DiagPart> h(nseg_);
        HIPC(hipMemcpyAsync(h.data(), d_diag_, nseg_ * sizeof(DiagPart), hipMemcpyDeviceToHost, stream_));
        auto st = read_state();
        const DiagPart& d = h[seg];
        const SegState& s = st[seg];
        std::memset(out, 0, sizeof(*out));
        out->n = seg_off_[seg + 1] - seg_off_[seg];
        auto val = [](const KSum& k) { return k.s + k.c; };
        out->mass = val(d.sum[0]);
        out->e_kin = val(d.sum[1]);
        out->e_int = val(d.sum[2]);
        KSum tot = d.sum[1];
        {   // e_tot with the same compensated arithmetic
            const double t = tot.s + d.sum[2].s;
            const double bp = t - tot.s;
            tot.c += (tot.s - (t - bp)) + (d.sum[2].s - bp) + d.sum[2].c;
            tot.s = t;
        }
        out->e_tot = val(tot);
        for (int a = 0; a < 3; a++) out->mom[a] = val(d.sum[3 + a]);
        out->vmax = d.ext[0];
        out->umax = d.ext[1];
        out->hmin = -d.ext[2];
        out->hmax = d.ext[3];
        out->rho_min = -d.ext[4];
        out->rho_max = d.ext[5];
        out->dmax = d.ext[6];
        out->nonfinite = (int32_t)d.ext[7];
        out->t = s.t;
        out->dt_next = have_dt_ ? bits_to_double(s.dt_next) : 0.0;
        out->steps = s.steps;
        out->density_rounds = s.rounds;
        out->list_walk_steps = (int64_t)s.list_walk_steps;
        out->list_overflow = (int64_t)s.list_overflow;
        out->density_lanes = (int64_t)s.dens_lanes;
        out->density_rewalks = (int64_t)s.dens_rewalks;
        out->density_waves = (int64_t)s.dens_waves;
        out->density_rewalk_waves = (int64_t)s.dens_rewalk_waves;
        out->status = seg_status(s, seg, nullptr);
    }

    // ---------------------------------------------------------------- profile / tuning
    void set_profiling(bool on) {
        if (!on) drain_profile(true);
        profiling_ = on;
    }
    void get_profile(dsph_profile* out) {
        drain_profile(true);
        std::memset(out, 0, sizeof(*out));
        out->nphases = PH_COUNT;
        for (int p = 0; p < PH_COUNT; p++) {
            std::strncpy(out->name[p], kPhaseName[p], 23);
            out->total_ms[p] = prof_total_[p];
            out->launches[p] = prof_count_[p];
        }
    }
    void reset_profile() {
        drain_profile(false);
        for (auto& e : prof_total_) e = 0.0;
        for (auto& e : prof_count_) e = 0;
    }
    const dsph_tuning& tuning() const { return tune_; }
    void set_tuning(const dsph_tuning& t) {
        dsph_tuning nt = tune_;
        auto ok_block = [](int b) { return b == 64 || b == 128 || b == 256; };
        if (t.block_pair) { if (!ok_block(t.block_pair)) throw Err{DSPH_E_ARG, "block_pair must be 64, 128 or 256"}; nt.block_pair = t.block_pair; }
        if (t.block_point) { if (!ok_block(t.block_point)) throw Err{DSPH_E_ARG, "block_point must be 64, 128 or 256"}; nt.block_point = t.block_point; }
        if (t.cell_factor != 0.0) {
            // (a column at least h_init wide keeps every z range within two slice blocks)
            if (!(t.cell_factor >= 1.0 && t.cell_factor <= 8.0)) throw Err{DSPH_E_ARG, "cell_factor must be in [1, 8]"};
            nt.cell_factor = t.cell_factor;
        }
        nt.xcd_remap = t.xcd_remap ? 1 : 0;
        if (t.list_skin != 0.0) {
            if (!(t.list_skin > 0.0 && t.list_skin <= 0.25)) throw Err{DSPH_E_ARG, "list_skin must be in (0, 0.25]"};
            nt.list_skin = t.list_skin;
        }
        if (t.neighbor_capacity) {
            if (t.neighbor_capacity > 0 && (t.neighbor_capacity < 8 || t.neighbor_capacity > 4096))
                throw Err{DSPH_E_ARG, "neighbor_capacity must be in [8, 4096], or negative for no lists"};
            if (t.neighbor_capacity > 0 && density64_)
                throw Err{DSPH_E_ARG, "neighbour lists serve the FP32 density sums: use DSPH_MIXED_FAST or DSPH_FP64"};
            nt.neighbor_capacity = t.neighbor_capacity > 0 ? t.neighbor_capacity : -1;
        }
        const bool lists_changed = nt.neighbor_capacity != tune_.neighbor_capacity || nt.list_skin != tune_.list_skin;
        if (t.cell_capacity_per_particle && t.cell_capacity_per_particle != tune_.cell_capacity_per_particle) {
            if (t.cell_capacity_per_particle < 1 || t.cell_capacity_per_particle > kMaxCellsPerParticle)
                throw Err{DSPH_E_ARG, "cell_capacity_per_particle must be in [1, 64]"};
            nt.cell_capacity_per_particle = t.cell_capacity_per_particle;
            if (n_ > 0) {
                tune_ = nt;
                // the table is reallocated; everything else is kept
                release_graphs();
                for (void* old : {(void*)w_.cstart, (void*)w_.cell_hmax, (void*)chunk_lb_}) free_ptr(old);
                alloc_table();
                built_ = false;
                records_ = false;
            }
        }
        tune_ = nt;
        if (lists_changed && n_ > 0) {
            release_graphs();
            if (w_.nbr) free_ptr(w_.nbr);
            free_ptr(w_.nbr_cnt);
            alloc_lists();
        }
        release_graphs();
    }

    // state snapshot for the autotuner
    struct Snapshot {
        std::vector<void*> blobs;
    };
    std::vector<std::pair<void*, size_t>> state_regions() {
        std::vector<std::pair<void*, size_t>> r;
        Carried<R>& c = cur();
        auto addd = [&](double* p) { r.push_back({p, (size_t)n_ * 8}); };
        auto addr = [&](R* p, size_t k) { r.push_back({p, (size_t)n_ * k * sizeof(R)}); };
        addd(c.x); addd(c.y); addd(c.z); addd(c.vx); addd(c.vy); addd(c.vz); addd(c.u); addd(c.mass);
        addr(c.h, 1); addr(c.D, 1); addr(c.eps_act, 1); addr(c.S, 6);
        r.push_back({c.info, (size_t)n_ * 4});
        r.push_back({c.orig, (size_t)n_ * 4});
        addr(w_.ax, 1); addr(w_.ay, 1); addr(w_.az, 1); addr(w_.dudt, 1); addr(w_.dSdt, 6);
        addr(w_.rho, 1); addr(w_.P, 1); addr(w_.cs, 1); addr(w_.csig, 1); addr(w_.Pf, 1); addr(w_.Rart, 1);
        addr(w_.dscale, 1); addr(w_.sigmax, 1); addr(w_.eps_work, 1); addr(w_.dudt_old, 1); addr(w_.dSdt_old, 6);
        r.push_back({d_st_, (size_t)nseg_ * sizeof(SegState)});
        return r;
    }
    void autotune(int steps, dsph_tuning* chosen) {
        if (!have_dt_) throw Err{DSPH_E_STATE, "call dsph_init before dsph_autotune"};
        auto regions = state_regions();
        std::vector<void*> saved;
        for (auto& rg : regions) {
            void* p = nullptr;
            if (hipMalloc(&p, rg.second) != hipSuccess) {
                for (void* q : saved) (void)hipFree(q);
                throw Err{DSPH_E_NOMEM, "autotune snapshot allocation failed"};
            }
            saved.push_back(p);
            HIPC(hipMemcpyAsync(p, rg.first, rg.second, hipMemcpyDeviceToDevice, stream_));
        }
        const int saved_cur = cur_;
        const bool saved_prof = profiling_;
        const char* log_env = std::getenv("DSPH_AUTOTUNE_LOG");
        const bool log = log_env && *log_env && std::strcmp(log_env, "0") != 0;   // each candidate's time to stderr
        profiling_ = false;
        dsph_tuning best = tune_;
        double best_ms = 1e300;
        uint32_t best_wanted = 0;       // cells the fastest candidate's grid wanted
        const int blocks[3] = {64, 128, 256};
        // column widths in h_init; on lattice initial conditions the step time
        // varies sharply around 1.5 (how lattice lines fall into columns), hence
        // the close neighbours
        const double cells[6] = {1.0, 1.25, 1.45, 1.5, 1.55, 2.0};
        const int caps[2] = {-1, 128};          // cell walks, neighbour lists
        hipEvent_t e0, e1;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        auto restore = [&]() {
            // the snapshot was taken in buffer set saved_cur; copy back there
            if (cur_ != saved_cur) cur_ = saved_cur;
            auto reg2 = state_regions();
            for (size_t k = 0; k < reg2.size(); k++)
                HIPC(hipMemcpyAsync(reg2[k].first, saved[k], reg2[k].second, hipMemcpyDeviceToDevice, stream_));
            built_ = false;
            records_ = false;
        };
        auto set_lists = [&](int cap) {
            if (cap == tune_.neighbor_capacity) return;
            tune_.neighbor_capacity = cap;
            if (w_.nbr) free_ptr(w_.nbr);
            free_ptr(w_.nbr_cnt);
            alloc_lists();
        };
        for (int cap : caps)
        for (double cf : cells) {
            if (cap > 0 && density64_) continue;      // the FP64 density walks the cells
            for (int bq : blocks) {
                release_graphs();
                set_lists(cap);
                dsph_tuning t = tune_;
                t.block_pair = bq;
                t.cell_factor = cf;
                tune_ = t;
                restore();
                step(0, 0.0);           // warm-up (also rebuilds for this grid)
                sync();
                if (grow_table()) {     // time the candidate at its nominal widths
                    restore();
                    step(0, 0.0);
                }
                HIPC(hipEventRecord(e0, stream_));
                for (int s = 0; s < steps; s++) step(0, 0.0);
                HIPC(hipEventRecord(e1, stream_));
                HIPC(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPC(hipEventElapsedTime(&ms, e0, e1));
                const uint32_t wanted = wanted_cells();
                if (log)
                    std::fprintf(stderr, "dsph autotune: block_pair %d cell_factor %.2f neighbor_capacity %d: %.3f ms/step\n",
                                 bq, cf, cap, ms / steps);
                if (ms < best_ms) { best_ms = ms; best = t; best_wanted = wanted; }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        release_graphs();
        set_lists(best.neighbor_capacity);
        best.cell_capacity_per_particle = tune_.cell_capacity_per_particle;   // the table as grown
        tune_ = best;
        // size the table for the chosen widths (the narrowest candidates grew it
        // further; its size sets the number of radix passes)
        if (best_wanted) resize_table(best_wanted, true);
        restore();
        sync();
        for (void* q : saved) (void)hipFree(q);
        profiling_ = saved_prof;
        if (chosen) *chosen = tune_;   // best, with the table as sized above
    }

    // host-side Weibull seeding (sph.hpp:117-128) over one segment in IC order
    void seed_damage(int seg, double dx, uint32_t seed) {
        std::vector<int32_t> mat(n_);
        get_field(DSPH_F_MAT, mat.data());
        std::vector<double> eps(n_);
        get_field(DSPH_F_EPS_ACT, eps.data());
        const double V = dx * dx * dx;
        for (int s = 0; s < nseg_; s++) {
            if (seg >= 0 && s != seg) continue;
            std::mt19937 rng(seed);
            std::uniform_real_distribution<double> U(1e-9, 1.0);
            for (int64_t i = seg_off_[s]; i < seg_off_[s + 1]; i++) {
                const dsph_material& m = mats_[mat[i]];
                if (m.wk > 0 && m.wm > 0) eps[i] = std::pow(U(rng) / (m.wk * V), 1.0 / m.wm);
                else eps[i] = 1e30;
            }
        }
        set_field(DSPH_F_EPS_ACT, eps.data());
    }

    // host copy of the carried fields (used when more particles are appended)
    void download(HostParticles& hp) {
        hp.pos.resize(3 * n_); hp.vel.resize(3 * n_); hp.mass.resize(n_); hp.u.resize(n_);
        hp.h.resize(n_); hp.D.resize(n_); hp.eps_act.resize(n_); hp.S.resize(6 * n_);
        hp.mat.resize(n_); hp.seg.resize(n_); hp.fixed.resize(n_);
        get_field(DSPH_F_POS, hp.pos.data()); get_field(DSPH_F_VEL, hp.vel.data());
        get_field(DSPH_F_MASS, hp.mass.data()); get_field(DSPH_F_U, hp.u.data());
        get_field(DSPH_F_HH, hp.h.data()); get_field(DSPH_F_D, hp.D.data());
        get_field(DSPH_F_EPS_ACT, hp.eps_act.data()); get_field(DSPH_F_S, hp.S.data());
        get_field(DSPH_F_MAT, hp.mat.data()); get_field(DSPH_F_SEG, hp.seg.data());
        get_field(DSPH_F_FIXED, hp.fixed.data());
    }

    bool h_host_set_ = false;

  private:
    std::vector<dsph_params> segp_;
    std::vector<dsph_material> mats_;
    hipStream_t stream_;
    int nseg_ = 1;
    int64_t n_ = 0;
    std::vector<void*> allocs_;
    Carried<R> bufs_[2];
    int cur_ = 0;
    Work<R> w_{};
    uint32_t* keys_[2] = {nullptr, nullptr};
    uint32_t* vals_[2] = {nullptr, nullptr};
    uint32_t* hist_ = nullptr;
    uint32_t* digit_totals_ = nullptr;
    uint32_t* d_total_cells_ = nullptr;
    int nb_sort_ = 0, passes_ = 0, nparts_max_ = 0;
    uint32_t capacity_ = 0;
    uint32_t* chunk_lb_ = nullptr;   // first slot of every chunk of kStartChunk cells (k_chunk_start)
    MatT<R>* d_mats_ = nullptr;
    MatT<double>* d_mats64_ = nullptr;
    bool density64_ = false;
    SegPar* d_par_ = nullptr;
    SegGrid* d_grid_ = nullptr;
    SegState* d_st_ = nullptr;
    double* d_segdt_ = nullptr;
    double* stage_ = nullptr;
    std::vector<int64_t> seg_off_;
    DiagChunk* d_chunks_ = nullptr;
    DiagPart* d_parts_ = nullptr;
    DiagPart* d_diag_ = nullptr;
    int32_t* d_chunk0_ = nullptr;
    int nchunks_ = 0;
    dsph_tuning tune_{};
    bool built_ = false, records_ = false, have_dt_ = false;
    bool lists_built_ = false;     // the last density pass built neighbour lists
    bool periodic_ = false;        // some segment has a periodic axis
    bool profiling_ = false, capturing_ = false;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> ev_pool_;
    double prof_total_[PH_COUNT];
    int64_t prof_count_[PH_COUNT];
    std::map<std::tuple<int, int, double>, hipGraphExec_t> graphs_;
};

}  // namespace dsph

// ============================================================================ allocation-free sums
// Compensated FP64 mass, energy and momentum sums of caller-owned device arrays
// without any allocation: the per-chunk partials live in `scratch`, a
// file-scope device array of fixed size, one copy per GPU.  Host code never
// passes `scratch` itself to a kernel: its device address is looked up once per
// device with hipGetSymbolAddress and kept in g_scratch.
namespace dsph {

constexpr int kSumsChunk = 16384;          // particles per partial (the chunk of dsph_diagnostics)
constexpr int kSumsMaxChunks = 4096;       // partials the scratch array holds
constexpr int64_t kSumsMaxN = (int64_t)kSumsChunk * kSumsMaxChunks;   // 67,108,864 particles
constexpr int kSumsOut = 8;                // mass, e_kin, e_int, e_tot, mom x, y, z, non-finite count

struct SumsPart {
    KSum sum[kDiagSums];                   // mass, e_kin, e_int, mom x, y, z
    double nonfinite;
};

static __device__ __align__(256) SumsPart scratch[kSumsMaxChunks];   // 416 KiB per GPU

// One block per chunk of kSumsChunk particles: per-thread compensated sums over
// a strided share, then a fixed-order tree over the block (as k_diag_partial).
// Particles with a non-finite mass, velocity or u are skipped and counted.
__global__ __launch_bounds__(256) void k_sums_partial(const double* m, const double* vx, const double* vy,
                                                      const double* vz, const double* u, int64_t n,
                                                      SumsPart* parts) {
    __shared__ SumsPart sh[256];
    const int64_t beg = (int64_t)blockIdx.x * kSumsChunk;
    const int64_t end = beg + kSumsChunk < n ? beg + kSumsChunk : n;
    SumsPart acc;
    for (int k = 0; k < kDiagSums; k++) acc.sum[k] = KSum{0.0, 0.0};
    acc.nonfinite = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const double mi = m[i], x = vx[i], y = vy[i], z = vz[i], ui = u[i];
        const double v2 = x * x + y * y + z * z;
        if (!(isfinite(mi) && isfinite(v2) && isfinite(ui))) { acc.nonfinite += 1.0; continue; }
        ksum_add(acc.sum[0], mi);
        ksum_add(acc.sum[1], 0.5 * mi * v2);
        ksum_add(acc.sum[2], mi * ui);
        ksum_add(acc.sum[3], mi * x);
        ksum_add(acc.sum[4], mi * y);
        ksum_add(acc.sum[5], mi * z);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            SumsPart a = sh[threadIdx.x];
            const SumsPart& b = sh[threadIdx.x + off];
            for (int k = 0; k < kDiagSums; k++) a.sum[k] = ksum_merge(a.sum[k], b.sum[k]);
            a.nonfinite += b.nonfinite;
            sh[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}

// One thread folds the chunk partials in chunk order and writes the results.
__global__ void k_sums_final(const SumsPart* parts, int nchunks, double* out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    SumsPart acc;
    for (int k = 0; k < kDiagSums; k++) acc.sum[k] = KSum{0.0, 0.0};
    acc.nonfinite = 0.0;
    for (int q = 0; q < nchunks; q++) {
        for (int k = 0; k < kDiagSums; k++) acc.sum[k] = ksum_merge(acc.sum[k], parts[q].sum[k]);
        acc.nonfinite += parts[q].nonfinite;
    }
    KSum tot = acc.sum[1];
    ksum_add(tot, acc.sum[2].s);
    tot.c += acc.sum[2].c;
    out[0] = acc.sum[0].s + acc.sum[0].c;
    out[1] = acc.sum[1].s + acc.sum[1].c;
    out[2] = acc.sum[2].s + acc.sum[2].c;
    out[3] = tot.s + tot.c;
    for (int a = 0; a < 3; a++) out[4 + a] = acc.sum[3 + a].s + acc.sum[3 + a].c;
    out[7] = acc.nonfinite;
}

static std::mutex g_scratch_mu;
static std::map<int, SumsPart*> g_scratch;    // device ordinal -> device address of `scratch`

// Device address of `scratch` on the current device, looked up on the first
// use of each device and kept.
static SumsPart* scratch_on_current_device() {
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch.find(dev);
    if (it != g_scratch.end()) return it->second;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        throw Err{DSPH_E_HIP, std::string("hipGetSymbolAddress of the sums scratch failed: ") + hipGetErrorString(e)};
    g_scratch[dev] = (SumsPart*)ptr;
    return (SumsPart*)ptr;
}

}  // namespace dsph

// ============================================================================ C ABI
using namespace dsph;

struct dsph_ctx {
    int precision = DSPH_MIXED;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<dsph_params> segp;
    std::vector<dsph_material> mats;
    std::unique_ptr<Engine<float>> ef;
    std::unique_ptr<Engine<double>> ed;
    HostParticles host;          // staged particles not yet on the device
    bool committed = false;
    bool dirty = false;          // host staging holds particles not yet committed
};

namespace {

template <class F>
int guarded(F f) {
    try {
        return f();
    } catch (const Err& e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return DSPH_E_NOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return DSPH_E_STATE;
    }
}

template <class F>
auto with_engine(dsph_ctx* c, F f) {
    if (c->precision == DSPH_FP64) return f(*c->ed);
    return f(*c->ef);
}

void require(bool cond, int code, const std::string& msg) {
    if (!cond) throw Err{code, msg};
}

void validate_params(const dsph_params& p) {
    require(p.eta > 0 && p.h_init > 0 && p.cfl > 0, DSPH_E_ARG, "eta, h_init and cfl must be positive");
    require(p.alpha >= 0 && p.beta >= 0 && p.eps >= 0 && p.alpha_u >= 0 && p.eps_as >= 0, DSPH_E_ARG,
            "alpha, beta, eps, alpha_u and eps_as must be non-negative");
    for (int a = 0; a < 3; a++) {
        require(std::isfinite(p.lo[a]) && std::isfinite(p.hi[a]), DSPH_E_ARG, "domain bounds must be finite");
        require(!p.periodic[a] || p.hi[a] > p.lo[a], DSPH_E_ARG, "a periodic axis needs hi > lo");
    }
}

void ensure_committed(dsph_ctx* c) {
    if (c->dirty) {
        with_engine(c, [&](auto& e) { e.commit(c->host); return 0; });
        c->committed = true;
        c->dirty = false;
        HostParticles empty;
        std::swap(c->host, empty);
    }
    require(c->committed, DSPH_E_STATE, "no particles uploaded");
}

int create_impl(const dsph_params* params, int32_t nseg, const dsph_material* mats, int32_t nmat,
                int32_t precision, dsph_ctx** out) {
    require(out != nullptr && params != nullptr && mats != nullptr, DSPH_E_ARG, "null argument");
    require(nseg >= 1 && nseg <= 65535, DSPH_E_ARG, "segment count must be in [1, 65535]");
    require(nmat >= 1 && nmat <= 256, DSPH_E_ARG, "material count must be in [1, 256]");
    require(precision == DSPH_MIXED || precision == DSPH_FP64 || precision == DSPH_MIXED_FAST, DSPH_E_ARG,
            "precision must be DSPH_MIXED, DSPH_FP64 or DSPH_MIXED_FAST");
    for (int s = 0; s < nseg; s++) validate_params(params[s]);
    for (int m = 0; m < nmat; m++) {
        require(mats[m].type == DSPH_EOS_IDEAL || mats[m].type == DSPH_EOS_TILLOTSON, DSPH_E_ARG, "bad EOS type");
        if (mats[m].type == DSPH_EOS_TILLOTSON)
            require(mats[m].rho0 > 0 && mats[m].u0 > 0 && mats[m].ucv > mats[m].uiv, DSPH_E_ARG,
                    "Tillotson material needs rho0 > 0, u0 > 0 and ucv > uiv");
    }
    int ndev = 0;
    HIPC(hipGetDeviceCount(&ndev));
    require(ndev > 0, DSPH_E_HIP, "no HIP device");
    auto c = std::make_unique<dsph_ctx>();
    HIPC(hipGetDevice(&c->device));
    (void)scratch_on_current_device();      // set-up of the allocation-free sums on this device
    HIPC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->precision = precision;
    c->segp.assign(params, params + nseg);
    c->mats.assign(mats, mats + nmat);
    if (precision == DSPH_FP64) c->ed = std::make_unique<Engine<double>>(c->segp, c->mats, c->stream, false);
    else c->ef = std::make_unique<Engine<float>>(c->segp, c->mats, c->stream, precision == DSPH_MIXED);
    *out = c.release();
    return DSPH_OK;
}

}  // namespace

extern "C" {

const char* dsph_version(void) { return "dsph 0.2 (gfx950, abi 2)"; }
const char* dsph_last_error(void) { return g_last_error.c_str(); }

int dsph_create(const dsph_params* params, const dsph_material* mats, int32_t nmat, int32_t precision,
                dsph_ctx** out) {
    return guarded([&] { return create_impl(params, 1, mats, nmat, precision, out); });
}

int dsph_batch_create(const dsph_params* seg_params, int32_t nseg, const dsph_material* mats, int32_t nmat,
                      int32_t precision, dsph_ctx** out) {
    return guarded([&] { return create_impl(seg_params, nseg, mats, nmat, precision, out); });
}

void dsph_destroy(dsph_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx->ef.reset();
    ctx->ed.reset();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int dsph_upload(dsph_ctx* c, int64_t n, const double* pos3, const double* vel3, const double* mass, const double* u,
                const int32_t* mat, const uint8_t* fixed, const int32_t* seg) {
    return guarded([&] {
        require(c != nullptr, DSPH_E_ARG, "null context");
        require(n > 0 && pos3 && vel3 && mass && u && mat, DSPH_E_ARG, "n > 0 and pos3, vel3, mass, u, mat are required");
        HIPC(hipSetDevice(c->device));
        // Appending to particles already on the device copies them back to the host
        // staging first; a refused upload drops that copy again, so the context keeps
        // its device state as it was.
        const bool fetched = c->committed && !c->dirty;
        if (fetched) {
            with_engine(c, [&](auto& e) { e.download(c->host); return 0; });
            c->committed = false;
        }
        HostParticles& hp = c->host;
        const int nseg = (int)c->segp.size();
        const int nmat = (int)c->mats.size();
        int last_seg = hp.seg.empty() ? 0 : hp.seg.back();
        const double h_init_default = c->segp[0].h_init;
        try {
            require((int64_t)hp.mass.size() + n <= kMaxParticles, DSPH_E_RANGE,
                    "more than 134,217,664 particles in one context");
            for (int64_t i = 0; i < n; i++) {           // validate everything before appending anything
                const int s = seg ? seg[i] : 0;
                require(s >= 0 && s < nseg, DSPH_E_ARG, "segment index out of range");
                require(s >= last_seg, DSPH_E_ARG, "segment indices must be non-decreasing in IC order");
                last_seg = s;
                require(mat[i] >= 0 && mat[i] < nmat, DSPH_E_ARG, "material index out of range");
                require(mass[i] > 0 && std::isfinite(mass[i]), DSPH_E_ARG, "masses must be positive and finite");
                require(std::isfinite(u[i]), DSPH_E_ARG, "internal energies must be finite");
                for (int a = 0; a < 3; a++)
                    require(std::isfinite(pos3[3 * i + a]) && std::isfinite(vel3[3 * i + a]), DSPH_E_ARG,
                            "positions and velocities must be finite");
            }
        } catch (...) {
            if (fetched) {
                HostParticles empty;
                std::swap(c->host, empty);
                c->committed = true;
            }
            throw;
        }
        for (int64_t i = 0; i < n; i++) {
            const int s = seg ? seg[i] : 0;
            for (int a = 0; a < 3; a++) { hp.pos.push_back(pos3[3 * i + a]); hp.vel.push_back(vel3[3 * i + a]); }
            hp.mass.push_back(mass[i]);
            hp.u.push_back(u[i]);
            hp.mat.push_back(mat[i]);
            hp.fixed.push_back(fixed ? (fixed[i] ? 1 : 0) : 0);
            hp.seg.push_back(s);
            hp.h.push_back(seg ? c->segp[s].h_init : h_init_default);   // System::add: hh = h_init
            hp.D.push_back(0.0);
            hp.eps_act.push_back(1e30);
            for (int k = 0; k < 6; k++) hp.S.push_back(0.0);
        }
        c->dirty = true;
        ensure_committed(c);
        return DSPH_OK;
    });
}

int dsph_count(const dsph_ctx* c, int64_t* n_out, int32_t* nseg_out) {
    return guarded([&] {
        require(c != nullptr, DSPH_E_ARG, "null context");
        int64_t n = 0;
        if (c->committed) n = c->precision == DSPH_FP64 ? c->ed->n() : c->ef->n();
        if (n_out) *n_out = n;
        if (nseg_out) *nseg_out = (int32_t)c->segp.size();
        return DSPH_OK;
    });
}

#define CTX_OP(body)                                                        \
    return guarded([&] {                                                    \
        require(c != nullptr, DSPH_E_ARG, "null context");                  \
        HIPC(hipSetDevice(c->device));                                      \
        ensure_committed(c);                                                \
        return with_engine(c, [&](auto& e) { body; return (int)DSPH_OK; }); \
    })

int dsph_seed_damage(dsph_ctx* c, int32_t seg, double dx, uint32_t seed) {
    CTX_OP(require(seg < (int)c->segp.size(), DSPH_E_ARG, "segment out of range");
           require(dx > 0, DSPH_E_ARG, "dx must be positive"); e.seed_damage(seg, dx, seed));
}
int dsph_get_field(dsph_ctx* c, int32_t field, void* dst) {
    CTX_OP(require(dst != nullptr && field >= 0 && field < DSPH_NFIELDS, DSPH_E_ARG, "bad field or null buffer");
           e.get_field(field, dst));
}
int dsph_set_field(dsph_ctx* c, int32_t field, const void* src) {
    CTX_OP(require(src != nullptr && field >= 0 && field < DSPH_NFIELDS, DSPH_E_ARG, "bad field or null buffer");
           e.set_field(field, src));
}
int dsph_init(dsph_ctx* c) { CTX_OP(e.init(); e.check_health()); }
int dsph_compute_density_h(dsph_ctx* c) { CTX_OP(e.compute_density_h(); e.check_health()); }
int dsph_compute_forces(dsph_ctx* c) { CTX_OP(e.compute_forces(); e.check_health()); }
int dsph_grow_damage(dsph_ctx* c, double dt) {
    CTX_OP(require(dt >= 0 && std::isfinite(dt), DSPH_E_ARG, "dt must be finite and non-negative"); e.grow_damage(dt);
           e.sync());
}
int dsph_timestep(dsph_ctx* c, double* dt_out) {
    CTX_OP(require(dt_out != nullptr, DSPH_E_ARG, "null output"); e.timestep(dt_out));
}
int dsph_step(dsph_ctx* c, double dt) {
    CTX_OP(require(std::isfinite(dt), DSPH_E_ARG, "dt must be finite"); e.step_host(dt); e.check_health());
}
int dsph_step_dts(dsph_ctx* c, const double* dts) {
    CTX_OP(require(dts != nullptr, DSPH_E_ARG, "null dts");
           for (size_t s = 0; s < c->segp.size(); s++) require(dts[s] >= 0 && std::isfinite(dts[s]), DSPH_E_ARG,
                                                                  "per-segment dt must be finite and non-negative");
           e.step_dts(dts); e.check_health());
}
int dsph_run(dsph_ctx* c, int32_t nsteps, double dt, int32_t use_graph) {
    CTX_OP(require(nsteps >= 0 && std::isfinite(dt), DSPH_E_ARG, "nsteps must be >= 0 and dt finite");
           e.run(nsteps, dt, use_graph != 0); e.check_health());
}
int dsph_synchronize(dsph_ctx* c) { CTX_OP(e.sync(); e.check_health()); }
int dsph_set_clock(dsph_ctx* c, int32_t seg, double t, int64_t steps, double dt_next) {
    CTX_OP(require(seg < (int)c->segp.size(), DSPH_E_ARG, "segment out of range");
           require(std::isfinite(t) && steps >= 0 && std::isfinite(dt_next) && dt_next >= 0, DSPH_E_ARG,
                   "time, step count and dt must be finite and non-negative");
           e.set_clock(seg, t, steps, dt_next));
}
int dsph_diagnostics(dsph_ctx* c, int32_t seg, dsph_diag* out) {
    CTX_OP(require(out != nullptr && seg >= 0 && seg < (int)c->segp.size(), DSPH_E_ARG, "bad segment or null output");
           e.diagnostics(seg, out));
}
int dsph_set_profiling(dsph_ctx* c, int32_t on) { CTX_OP(e.set_profiling(on != 0)); }
int dsph_get_profile(dsph_ctx* c, dsph_profile* out) {
    CTX_OP(require(out != nullptr, DSPH_E_ARG, "null output"); e.get_profile(out));
}
int dsph_reset_profile(dsph_ctx* c) { CTX_OP(e.reset_profile()); }
int dsph_get_tuning(const dsph_ctx* cc, dsph_tuning* out) {
    dsph_ctx* c = const_cast<dsph_ctx*>(cc);
    return guarded([&] {
        require(c != nullptr && out != nullptr, DSPH_E_ARG, "null argument");
        if (c->precision == DSPH_FP64) *out = c->ed->tuning();
        else *out = c->ef->tuning();
        return DSPH_OK;
    });
}
int dsph_set_tuning(dsph_ctx* c, const dsph_tuning* t) {
    return guarded([&] {
        require(c != nullptr && t != nullptr, DSPH_E_ARG, "null argument");
        HIPC(hipSetDevice(c->device));
        if (c->precision == DSPH_FP64) c->ed->set_tuning(*t);
        else c->ef->set_tuning(*t);
        return DSPH_OK;
    });
}
int dsph_autotune(dsph_ctx* c, int32_t steps, dsph_tuning* chosen) {
    CTX_OP(require(steps >= 1, DSPH_E_ARG, "steps must be >= 1"); e.autotune(steps, chosen));
}

int dsph_energy_sums_noalloc(const double* d_mass, const double* d_vx, const double* d_vy, const double* d_vz,
                             const double* d_u, int64_t n, void* stream, double* d_out) {
    return guarded([&] {
        require(d_mass && d_vx && d_vy && d_vz && d_u && d_out, DSPH_E_ARG, "null device pointer");
        require(n >= 0, DSPH_E_ARG, "n must be non-negative");
        require(n <= kSumsMaxN, DSPH_E_RANGE,
                "n = " + std::to_string(n) + " exceeds the " + std::to_string(kSumsMaxN) +
                    " particles the allocation-free sums hold (" + std::to_string(kSumsMaxChunks) + " chunks of " +
                    std::to_string(kSumsChunk) + ")");
        SumsPart* parts = scratch_on_current_device();
        const int nchunks = n > 0 ? (int)((n + kSumsChunk - 1) / kSumsChunk) : 1;
        hipLaunchKernelGGL(k_sums_partial, dim3(nchunks), dim3(256), 0, (hipStream_t)stream, d_mass, d_vx, d_vy, d_vz,
                           d_u, n, parts);
        check_launch("sums_partial");
        hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(1), 0, (hipStream_t)stream, (const SumsPart*)parts, nchunks,
                           d_out);
        check_launch("sums_final");
        return DSPH_OK;
    });
}

}  // extern "C"
