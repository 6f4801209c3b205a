/*
 * CPU FP64 oracle for the SPH impact step -- TEST INFRASTRUCTURE ONLY.
 *
 * This header declares one C API implemented twice:
 *   - oracle/sph_oracle.cpp : this project's FP64 restatement of the reference
 *     solver (hydrocodes/sph/sph.hpp:17-379, hydrocodes/common/eos.hpp:15-113);
 *   - oracle/ref_harness.cpp: a thin adapter over the reference's own headers,
 *     compiled in place from the reference tree into oracle/_ref/ (never copied).
 * Only tests/, the smoke check and bench.py's cpu_baseline leg may load either
 * library, and only as the checker / the timed CPU baseline.  The product path
 * (deorbit_amd) never links or calls anything here.
 *
 * Field layout conventions (all values float64, particles in insertion order):
 *   vectors (POS, VEL, ACC)   : n x 3  {x, y, z}
 *   symmetric 3x3 (S, DSDT)   : n x 6  {xx, yy, zz, xy, xz, yz}   (sph.hpp:31-33 order)
 *   MAT, FIXED                : n      stored as float64 of the integer value
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

/* Material record (eos.hpp:15-39).  Emod is derived exactly as eos.hpp:37. */
typedef struct {
    int32_t type;                 /* 0 = ideal gas, 1 = Tillotson          */
    double gamma;                 /* ideal gas only                        */
    double rho0, A, B, a, b, alpha, beta, u0, uiv, ucv;
    double G, Y, wm, wk;
} orc_material;

/* Solver parameters (sph.hpp:86-101). */
typedef struct {
    double eta, h_init, alpha, beta, eps, alpha_u, cfl, eps_as;
    int32_t damage_on;
    int32_t periodic[3];
    double lo[3], hi[3];
    int32_t nthreads;             /* 0 -> min(16, hardware_concurrency) (sph.hpp:58-60) */
} orc_params;

enum {
    ORC_POS = 0, ORC_VEL, ORC_ACC, ORC_MASS, ORC_RHO, ORC_U, ORC_DUDT, ORC_P, ORC_CS,
    ORC_CSIG, ORC_HH, ORC_EPS_WORK, ORC_D, ORC_EPS_ACT, ORC_SIGMAX, ORC_PF, ORC_RART,
    ORC_DSCALE, ORC_S, ORC_DSDT, ORC_MAT, ORC_FIXED, ORC_NFIELDS
};

orc_sys* orc_create(const orc_params* p, const orc_material* mats, int32_t nmat);
void     orc_destroy(orc_sys* s);
/* Append n particles (sph.hpp:103-111 per particle; mat[i] plays cur_mat). */
void     orc_add(orc_sys* s, int64_t n, const double* pos3, const double* vel3,
                 const double* mass, const double* u, const int32_t* mat, const uint8_t* fixed);
int64_t  orc_n(const orc_sys* s);
void     orc_seed_damage(orc_sys* s, double dx, uint32_t seed);       /* sph.hpp:117-128 */
void     orc_init(orc_sys* s);                                       /* sph.hpp:379     */
double   orc_timestep(orc_sys* s);                                   /* sph.hpp:350-355 */
void     orc_step(orc_sys* s, double dt);                            /* sph.hpp:357-377 */
void     orc_build_cells(orc_sys* s);                                /* sph.hpp:147-162 */
void     orc_compute_density_h(orc_sys* s, int32_t iters);           /* sph.hpp:184-208 */
void     orc_compute_strain_stress(orc_sys* s);                      /* sph.hpp:212-258 */
void     orc_compute_forces(orc_sys* s);                             /* sph.hpp:260-318 */
void     orc_grow_damage(orc_sys* s, double dt);                     /* sph.hpp:336-348 */
void     orc_yield_limit(orc_sys* s, int64_t i);                     /* sph.hpp:320-331 */
int32_t  orc_get(const orc_sys* s, int32_t field, double* out);
int32_t  orc_set(orc_sys* s, int32_t field, const double* in);
/* Cell grid of the last build_cells (sph.hpp:139-162). */
void     orc_cell_info(const orc_sys* s, int32_t nc[3], double cw[3], double* h_max);

/* Point functions. */
double   orc_kW(double r, double h);                                 /* sph.hpp:17-22   */
double   orc_kdWdr(double r, double h);                              /* sph.hpp:23-28   */
double   orc_eig_max(const double s6[6]);                            /* sph.hpp:41-54   */
double   orc_material_emod(const orc_material* m);                   /* eos.hpp:37      */
double   orc_pressure(const orc_material* m, double rho, double u);  /* eos.hpp:52-68   */
double   orc_sound_speed(const orc_material* m, double rho, double u); /* eos.hpp:95-113 */

#ifdef __cplusplus
}
#endif
