// Adapter exposing the reference solver (hydrocodes/sph/sph.hpp, compiled in
// place from the reference tree via -I) through the oracle C API declared in
// sph_oracle.h.  TEST INFRASTRUCTURE ONLY: it is built into oracle/_ref/ (never
// committed) and used to pin the project's own restatement bit for bit.
#include "sph_oracle.h"

#include <cstring>
#include <vector>

#include "sph.hpp"   // resolved against the reference's hydrocodes/sph directory

namespace {
Material to_material(const orc_material& m) {
    if (m.type == 0) return Material::ideal(m.gamma);
    return Material::tillotson(m.rho0, m.A, m.B, m.a, m.b, m.alpha, m.beta, m.u0, m.uiv, m.ucv,
                               m.G, m.Y, m.wm, m.wk);
}
}  // namespace

struct orc_sys {
    System sys;
};

extern "C" {

orc_sys* orc_create(const orc_params* p, const orc_material* mats, int32_t nmat) {
    orc_sys* s = new orc_sys();
    System& S = s->sys;
    for (int k = 0; k < nmat; k++) S.materials.push_back(to_material(mats[k]));
    S.eta = p->eta; S.h_init = p->h_init; S.h_max = p->h_init;
    S.alpha = p->alpha; S.beta = p->beta; S.eps = p->eps; S.alpha_u = p->alpha_u;
    S.cfl = p->cfl; S.eps_as = p->eps_as; S.damage_on = p->damage_on != 0;
    S.set_domain(p->lo[0], p->hi[0], p->periodic[0] != 0, p->lo[1], p->hi[1], p->periodic[1] != 0,
                 p->lo[2], p->hi[2], p->periodic[2] != 0);
    return s;
}

void orc_destroy(orc_sys* s) { delete s; }

void orc_add(orc_sys* s, int64_t n, const double* pos3, const double* vel3, const double* mass,
             const double* u, const int32_t* mat, const uint8_t* fixed) {
    System& S = s->sys;
    for (int64_t k = 0; k < n; k++) {
        S.cur_mat = mat[k];
        S.add({pos3[3 * k], pos3[3 * k + 1], pos3[3 * k + 2]},
              {vel3[3 * k], vel3[3 * k + 1], vel3[3 * k + 2]}, mass[k], u[k],
              fixed ? fixed[k] != 0 : false);
    }
}

int64_t orc_n(const orc_sys* s) { return s->sys.n; }
void orc_seed_damage(orc_sys* s, double dx, uint32_t seed) { s->sys.seed_damage(dx, seed); }
void orc_init(orc_sys* s) { s->sys.init(); }
double orc_timestep(orc_sys* s) { return s->sys.timestep(); }
void orc_step(orc_sys* s, double dt) { s->sys.step(dt); }
void orc_build_cells(orc_sys* s) { s->sys.build_cells(); }
void orc_compute_density_h(orc_sys* s, int32_t iters) { s->sys.compute_density_h(iters); }
void orc_compute_strain_stress(orc_sys* s) { s->sys.compute_strain_stress(); }
void orc_compute_forces(orc_sys* s) { s->sys.compute_forces(); }
void orc_grow_damage(orc_sys* s, double dt) { s->sys.grow_damage(dt); }
void orc_yield_limit(orc_sys* s, int64_t i) { s->sys.yield_limit((int)i); }

static std::vector<double>* scalar(System& S, int f) {
    switch (f) {
        case ORC_MASS: return &S.mass;     case ORC_RHO: return &S.rho;
        case ORC_U: return &S.u;           case ORC_DUDT: return &S.dudt;
        case ORC_P: return &S.P;           case ORC_CS: return &S.cs;
        case ORC_CSIG: return &S.csig;     case ORC_HH: return &S.hh;
        case ORC_EPS_WORK: return &S.eps_work;
        case ORC_D: return &S.D;           case ORC_EPS_ACT: return &S.eps_act;
        case ORC_SIGMAX: return &S.sigmax; case ORC_PF: return &S.Pf;
        case ORC_RART: return &S.Rart;     case ORC_DSCALE: return &S.dscale;
        default: return nullptr;
    }
}

int32_t orc_get(const orc_sys* cs, int32_t f, double* out) {
    System& S = const_cast<orc_sys*>(cs)->sys;
    int n = S.n;
    if (f == ORC_POS || f == ORC_VEL || f == ORC_ACC) {
        const std::vector<Vec3>& v = (f == ORC_POS) ? S.pos : (f == ORC_VEL) ? S.vel : S.acc;
        for (int i = 0; i < n; i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
    } else if (f == ORC_S || f == ORC_DSDT) {
        const std::vector<Sym3>& t = (f == ORC_S) ? S.S : S.dSdt;
        for (int i = 0; i < n; i++) {
            double* o = out + 6 * i;
            o[0] = t[i].xx; o[1] = t[i].yy; o[2] = t[i].zz; o[3] = t[i].xy; o[4] = t[i].xz; o[5] = t[i].yz;
        }
    } else if (f == ORC_MAT) {
        for (int i = 0; i < n; i++) out[i] = (double)S.mat[i];
    } else if (f == ORC_FIXED) {
        for (int i = 0; i < n; i++) out[i] = (double)S.fixed[i];
    } else {
        std::vector<double>* v = scalar(S, f);
        if (!v) return -1;
        std::memcpy(out, v->data(), sizeof(double) * n);
    }
    return 0;
}

int32_t orc_set(orc_sys* s, int32_t f, const double* in) {
    System& S = s->sys;
    int n = S.n;
    if (f == ORC_POS || f == ORC_VEL || f == ORC_ACC) {
        std::vector<Vec3>& v = (f == ORC_POS) ? S.pos : (f == ORC_VEL) ? S.vel : S.acc;
        for (int i = 0; i < n; i++) v[i] = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
    } else if (f == ORC_S || f == ORC_DSDT) {
        std::vector<Sym3>& t = (f == ORC_S) ? S.S : S.dSdt;
        for (int i = 0; i < n; i++) {
            const double* o = in + 6 * i;
            t[i].xx = o[0]; t[i].yy = o[1]; t[i].zz = o[2]; t[i].xy = o[3]; t[i].xz = o[4]; t[i].yz = o[5];
        }
    } else if (f == ORC_MAT) {
        for (int i = 0; i < n; i++) S.mat[i] = (int)in[i];
    } else if (f == ORC_FIXED) {
        for (int i = 0; i < n; i++) S.fixed[i] = in[i] != 0.0 ? 1 : 0;
    } else {
        std::vector<double>* v = scalar(S, f);
        if (!v) return -1;
        std::memcpy(v->data(), in, sizeof(double) * n);
    }
    return 0;
}

void orc_cell_info(const orc_sys* s, int32_t nc[3], double cw[3], double* h_max) {
    for (int a = 0; a < 3; a++) { nc[a] = s->sys.nc[a]; cw[a] = s->sys.cw[a]; }
    *h_max = s->sys.h_max;
}

double orc_kW(double r, double h) { return kW(r, h); }
double orc_kdWdr(double r, double h) { return kdWdr(r, h); }
double orc_eig_max(const double s6[6]) {
    Sym3 t;
    t.xx = s6[0]; t.yy = s6[1]; t.zz = s6[2]; t.xy = s6[3]; t.xz = s6[4]; t.yz = s6[5];
    return eig_max(t);
}
double orc_material_emod(const orc_material* m) { return to_material(*m).Emod; }
double orc_pressure(const orc_material* m, double rho, double u) { return to_material(*m).pressure(rho, u); }
double orc_sound_speed(const orc_material* m, double rho, double u) { return to_material(*m).sound_speed(rho, u); }

}  // extern "C"
