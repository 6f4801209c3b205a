// System-shaped C++ wrapper over the dsph C ABI (include/dsph.h).
//
// Mirrors the reference solver's in-process interface, struct System in
// hydrocodes/sph/sph.hpp:72-380 (materials, parameters, set_domain, add,
// seed_damage, init, timestep, step and the phase methods), so a driver
// written against that interface compiles against this one.  Field access is
// by whole-array copies in insertion order instead of public std::vectors,
// because the state lives in device memory.  Errors throw dsph::Error: argument
// and state errors at once, and the device's trouble flags (non-finite state,
// a non-positive time step, capacity and reach limits) from every call that
// waits for the device (init, the phase methods, step, run, timestep and
// synchronize).
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "dsph.h"

namespace dsph {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != DSPH_OK) throw Error(rc, dsph_last_error());
}

struct Vec3d {
    double x = 0, y = 0, z = 0;
};

// Material factory with the argument order of Material::tillotson (eos.hpp:30-32).
inline dsph_material tillotson(double rho0, double A, double B, double a, double b, double al, double be,
                               double u0, double uiv, double ucv, double G = 0, double Y = 0, double wm = 0,
                               double wk = 0) {
    dsph_material m{};
    m.type = DSPH_EOS_TILLOTSON;
    m.gamma = 1.4;
    m.rho0 = rho0; m.A = A; m.B = B; m.a = a; m.b = b; m.alpha = al; m.beta = be;
    m.u0 = u0; m.uiv = uiv; m.ucv = ucv; m.G = G; m.Y = Y; m.wm = wm; m.wk = wk;
    m.Emod = 0.0;   // derived by the library as 9AG/(3A+G) (eos.hpp:37)
    return m;
}
inline dsph_material ideal_gas(double gamma) {
    dsph_material m{};
    m.type = DSPH_EOS_IDEAL;
    m.gamma = gamma;
    return m;
}
// Reference presets (eos.hpp:42-50).
inline dsph_material iron() { return tillotson(7800, 1.28e11, 1.05e11, 0.5, 1.5, 5, 5, 9.5e6, 2.4e6, 8.67e6, 7.75e10, 1.0e9); }
inline dsph_material basalt() {
    return tillotson(2700, 2.67e10, 2.67e10, 0.5, 1.5, 5, 5, 4.87e8, 4.72e6, 1.82e7, 2.27e10, 3.5e8, 16.0, 1.0e61);
}
inline dsph_material water() { return tillotson(998, 2.2e9, 9.94e9, 0.7, 0.15, 10, 5, 7.0e6, 4.19e5, 2.69e6); }

class System {
  public:
    // public parameters with the reference's names and defaults (sph.hpp:86-101)
    std::vector<dsph_material> materials;
    int cur_mat = 0;
    double eta = 1.3, h_init = 0.01, alpha = 1.0, beta = 2.0, eps = 0.01, alpha_u = 0.0, cfl = 0.2;
    double eps_as = 0.0;
    bool damage_on = false;
    double dlo[3] = {0, 0, 0}, dhi[3] = {0, 0, 0};
    bool per[3] = {false, false, false};
    double g = 0.0, damp = 1.0;                  // gpu_ic extras (gpu_ic.cpp:38-39)
    double t_end = 0.0;                          // > 0: device-driven runs stop exactly there
    int precision = DSPH_MIXED;

    System() = default;
    System(const System&) = delete;
    System& operator=(const System&) = delete;
    ~System() { if (ctx_) dsph_destroy(ctx_); }

    void set_domain(double xl, double xh, bool px, double yl, double yh, bool py, double zl, double zh, bool pz) {
        dlo[0] = xl; dhi[0] = xh; per[0] = px;
        dlo[1] = yl; dhi[1] = yh; per[1] = py;
        dlo[2] = zl; dhi[2] = zh; per[2] = pz;
    }

    // System::add (sph.hpp:103-111); particles are staged and uploaded on first use.
    void add(Vec3d p, Vec3d v, double m, double uu, bool fix = false) {
        pos_.insert(pos_.end(), {p.x, p.y, p.z});
        vel_.insert(vel_.end(), {v.x, v.y, v.z});
        mass_.push_back(m);
        u_.push_back(uu);
        mat_.push_back(cur_mat);
        fixed_.push_back(fix ? 1 : 0);
    }
    int64_t n() { flush(); int64_t k = 0; check(dsph_count(ctx_, &k, nullptr)); return k; }

    void seed_damage(double dx, unsigned seed = 12345) { flush(); check(dsph_seed_damage(ctx_, -1, dx, seed)); }
    void init() { flush(); check(dsph_init(ctx_)); }
    void compute_density_h() { flush(); check(dsph_compute_density_h(ctx_)); }
    void compute_forces() { flush(); check(dsph_compute_forces(ctx_)); }
    void grow_damage(double dt) { flush(); check(dsph_grow_damage(ctx_, dt)); }
    double timestep() { flush(); double dt = 0; check(dsph_timestep(ctx_, &dt)); return dt; }
    void step(double dt) { flush(); check(dsph_step(ctx_, dt)); }
    void run(int nsteps, double dt = 0.0, bool graph = true) { flush(); check(dsph_run(ctx_, nsteps, dt, graph)); }
    // wait for the device and raise its trouble flags (no reference counterpart:
    // the reference's calls return synchronously)
    void synchronize() { flush(); check(dsph_synchronize(ctx_)); }
    dsph_diag diagnostics() { flush(); dsph_diag d{}; check(dsph_diagnostics(ctx_, 0, &d)); return d; }

    // whole-field copies in insertion order (see dsph_field for widths and types)
    std::vector<double> get(int field, int width = 1) {
        flush();
        std::vector<double> out((size_t)n() * width);
        check(dsph_get_field(ctx_, field, out.data()));
        return out;
    }
    void set(int field, const std::vector<double>& v) { flush(); check(dsph_set_field(ctx_, field, v.data())); }
    std::vector<int32_t> get_mat() {
        flush();
        std::vector<int32_t> out((size_t)n());
        check(dsph_get_field(ctx_, DSPH_F_MAT, out.data()));
        return out;
    }
    std::vector<uint8_t> get_fixed() {
        flush();
        std::vector<uint8_t> out((size_t)n());
        check(dsph_get_field(ctx_, DSPH_F_FIXED, out.data()));
        return out;
    }
    // restore a checkpointed clock (no reference counterpart; see dsph_set_clock)
    void set_clock(double t, int64_t steps, double dt_next) { flush(); check(dsph_set_clock(ctx_, -1, t, steps, dt_next)); }
    dsph_ctx* handle() { flush(); return ctx_; }

  private:
    dsph_ctx* ctx_ = nullptr;
    std::vector<double> pos_, vel_, mass_, u_;
    std::vector<int32_t> mat_;
    std::vector<uint8_t> fixed_;

    void flush() {
        if (!ctx_) {
            dsph_params p{};
            p.eta = eta; p.h_init = h_init; p.alpha = alpha; p.beta = beta; p.eps = eps; p.alpha_u = alpha_u;
            p.cfl = cfl; p.eps_as = eps_as; p.damage_on = damage_on ? 1 : 0;
            for (int a = 0; a < 3; a++) { p.periodic[a] = per[a] ? 1 : 0; p.lo[a] = dlo[a]; p.hi[a] = dhi[a]; }
            p.g = g; p.damp = damp; p.t_end = t_end;
            check(dsph_create(&p, materials.data(), (int32_t)materials.size(), precision, &ctx_));
        }
        if (!mass_.empty()) {
            check(dsph_upload(ctx_, (int64_t)mass_.size(), pos_.data(), vel_.data(), mass_.data(), u_.data(),
                              mat_.data(), fixed_.data(), nullptr));
            pos_.clear(); vel_.clear(); mass_.clear(); u_.clear(); mat_.clear(); fixed_.clear();
        }
    }
};

}  // namespace dsph
