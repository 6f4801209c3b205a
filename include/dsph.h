/*
 * dsph -- MI355X-native (gfx950) 3D SPH impact step behind a C ABI.
 *
 * The entry points mirror the reference solver's public surface
 * (struct System in hydrocodes/sph/sph.hpp:72-380 and struct Material in
 * hydrocodes/common/eos.hpp:15-39) and replace the Metal kernel ABI driven by
 * hydrocodes/sph/gpu_ic.cpp:29-34,97-139.  Conventions:
 *   - SI units; host arrays are owned by the caller and copied in/out;
 *   - particle arrays are in IC (insertion) order on both sides of the ABI,
 *     whatever order the device keeps internally;
 *   - vectors are n x 3 {x,y,z}; symmetric tensors are n x 6
 *     {xx, yy, zz, xy, xz, yz} (the Sym3 order of sph.hpp:31-33);
 *   - every call returns 0 or a negative dsph_status; the text of the last
 *     error of the calling thread is dsph_last_error();
 *   - the kernels record trouble in per-segment device flags (cell table
 *     overflow, a non-finite position, a non-positive or non-finite time step,
 *     a smoothing length whose reach exceeds the cell layout).  The flags are
 *     sticky; every call that synchronises with the device (dsph_init, the
 *     phase methods, dsph_step, dsph_step_dts, dsph_run, dsph_timestep,
 *     dsph_synchronize) returns them as DSPH_E_RANGE or DSPH_E_NUMERIC, and
 *     dsph_diagnostics reports them per segment in dsph_diag.status;
 *   - a context is bound to one HIP device and one HIP stream; it is
 *     thread-compatible, not thread-safe.
 * No C++ exceptions cross this boundary.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSPH_ABI_VERSION 2

typedef struct dsph_ctx dsph_ctx;

typedef enum {
    DSPH_OK = 0,
    DSPH_E_ARG = -1,      /* invalid argument                                  */
    DSPH_E_STATE = -2,    /* call not valid in the current context state       */
    DSPH_E_HIP = -3,      /* HIP runtime error                                 */
    DSPH_E_NOMEM = -4,    /* device allocation failed                          */
    DSPH_E_RANGE = -5,    /* a size or capacity limit was exceeded             */
    DSPH_E_NUMERIC = -6,  /* non-finite state or non-positive time step found  */
    DSPH_E_IO = -7        /* file read/write failure                           */
} dsph_status;

/* DSPH_MIXED: FP32 force-pass pair sums, stress and sound speed; density pair
 *   terms, density, smoothing length and pressure in FP64 (the per-particle
 *   pressure above 10 MPa follows the FP64 reference to ~1e-6 relative after a
 *   step and within 1e-4 through ten steps);
 *   positions, velocities, internal energy, time step and reductions in FP64.
 * DSPH_FP64: every quantity in FP64 (the comparator path).
 * DSPH_MIXED_FAST: DSPH_MIXED with FP32 density sums and pressure as well
 *   (density within ~1e-6, pressure within ~1e-6 of max|P| but only ~5e-3 per
 *   particle where |rho/rho0 - 1| is small); the only mixed mode that can use
 *   neighbour lists.  It walks the neighbours twice per density round where
 *   the FP64 sums of DSPH_MIXED walk once, so it is at most a few percent
 *   faster. */
typedef enum { DSPH_MIXED = 0, DSPH_FP64 = 1, DSPH_MIXED_FAST = 2 } dsph_precision;
typedef enum { DSPH_EOS_IDEAL = 0, DSPH_EOS_TILLOTSON = 1 } dsph_eos_type;

/* Material record; replaces Material (eos.hpp:15-39) and the Metal GMat
 * (sph_force.metal:4).  Emod = 0 means "derive 9AG/(3A+G) as eos.hpp:37". */
typedef struct {
    int32_t type;                                   /* dsph_eos_type             */
    double gamma;                                   /* ideal gas                 */
    double rho0, A, B, a, b, alpha, beta, u0, uiv, ucv;   /* Tillotson          */
    double G, Y;                                    /* shear modulus, yield      */
    double wm, wk;                                  /* Weibull modulus, density  */
    double Emod;
} dsph_material;

/* Solver parameters; replaces the System fields of sph.hpp:86-101 plus the
 * gpu_ic.cpp:38-39,54-55 extras (gravity g along -z, velocity damping of
 * material 0 per step). */
typedef struct {
    double eta, h_init, alpha, beta, eps, alpha_u, cfl, eps_as;
    int32_t damage_on;
    int32_t periodic[3];
    double lo[3], hi[3];        /* set_domain bounds (sph.hpp:96-101)            */
    double g;                   /* 0 = off                                       */
    double damp;                /* 1 = off                                       */
    double t_end;               /* adaptive runs clamp dt to reach t_end; <=0 off */
} dsph_params;

/* Tunables picked by dsph_autotune (or set by hand). 0 = library default. */
typedef struct {
    int32_t block_pair;         /* threads per workgroup, neighbour kernels      */
    int32_t block_point;        /* threads per workgroup, per-particle kernels   */
    double cell_factor;         /* sort-column width = cell_factor * h_init (z slices: a quarter of it) */
    int32_t xcd_remap;          /* 1 = map consecutive tiles onto one XCD        */
    int32_t cell_capacity_per_particle;  /* cell table size / particle count; the
                                 * library grows it (up to 64) after a call whose
                                 * sort grid wanted more cells than the table held
                                 * (DSPH_AUTOTUNE_LOG=1 prints autotune candidates) */
    int32_t neighbor_capacity;  /* neighbour-list entries per particle; 0 = auto, <0 = no lists */
    double list_skin;           /* list radius = (1 + skin) x interaction radius */
} dsph_tuning;

/* Per-segment diagnostics, summed with compensated (Kahan/TwoSum) FP64. */
typedef struct {
    int64_t n;
    double mass, e_kin, e_int, e_tot;   /* cpu_bench.cpp:26 energy              */
    double mom[3];
    double vmax, umax, hmin, hmax, rho_min, rho_max, dmax;
    double t;                   /* simulated time of this segment               */
    double dt_next;             /* CFL step from the last force evaluation      */
    int64_t steps;
    int32_t density_rounds;     /* rounds run by the last density evaluation    */
    int32_t nonfinite;          /* count of non-finite pos/vel/u values         */
    int64_t list_walk_steps;    /* force passes that walked cells: an h outgrew the list skin */
    int64_t list_overflow;      /* particle neighbour lists that did not fit (cumulative)    */
    int32_t status;             /* DSPH_OK, or the dsph_status of the segment's device flags  */
    int32_t pad_;
    /* Density moment walks (cumulative; counted only when the environment sets
     * DSPH_COUNT_REWALKS=1 before the context uploads its particles): lanes and
     * waves of FP64-sum density rounds, and those whose second iteration walked
     * the cells again because h left the moments' band.  0 otherwise. */
    int64_t density_lanes, density_rewalks, density_waves, density_rewalk_waves;
} dsph_diag;

/* Per-phase device time from HIP events (profiling mode). */
#define DSPH_MAX_PHASES 24
typedef struct {
    int32_t nphases;
    char name[DSPH_MAX_PHASES][24];
    double total_ms[DSPH_MAX_PHASES];
    int64_t launches[DSPH_MAX_PHASES];
} dsph_profile;

/* Field identifiers for dsph_get_field / dsph_set_field.  Float fields travel
 * as float64 (n or n x 3 or n x 6), DSPH_F_MAT and DSPH_F_SEG as int32,
 * DSPH_F_FIXED as uint8. */
typedef enum {
    DSPH_F_POS = 0, DSPH_F_VEL, DSPH_F_ACC, DSPH_F_MASS, DSPH_F_RHO, DSPH_F_U, DSPH_F_DUDT,
    DSPH_F_P, DSPH_F_CS, DSPH_F_CSIG, DSPH_F_HH, DSPH_F_EPS_WORK, DSPH_F_D, DSPH_F_EPS_ACT,
    DSPH_F_SIGMAX, DSPH_F_PF, DSPH_F_RART, DSPH_F_DSCALE, DSPH_F_S, DSPH_F_DSDT,
    DSPH_F_MAT, DSPH_F_FIXED, DSPH_F_SEG, DSPH_F_DUDT_OLD, DSPH_F_DSDT_OLD, DSPH_NFIELDS
} dsph_field;

/* ---- lifetime --------------------------------------------------------- */
const char* dsph_version(void);
const char* dsph_last_error(void);

/* One impact: replaces `System S; S.materials = ...; S.eta = ...;
 * S.set_domain(...)` (sph.hpp:72-101; gpu_ic.cpp:54-67). */
int dsph_create(const dsph_params* params, const dsph_material* mats, int32_t nmat,
                int32_t precision, dsph_ctx** out);
/* A batch of nseg independent impacts stepped together (segment s keeps its
 * own params[s], domain, time step and density rounds).  No reference
 * counterpart: the reference runs one System per process (run_ic.cpp:218). */
int dsph_batch_create(const dsph_params* seg_params, int32_t nseg, const dsph_material* mats,
                      int32_t nmat, int32_t precision, dsph_ctx** out);
void dsph_destroy(dsph_ctx* ctx);

/* ---- state ------------------------------------------------------------ */
/* Append n particles; replaces System::add (sph.hpp:103-111) called n times
 * with cur_mat = mat[i].  seg may be NULL (all segment 0); otherwise it must
 * be non-decreasing.  Each call appends; state is (re)allocated on device.
 * A context holds at most 134,217,664 particles (DSPH_E_RANGE beyond). */
int dsph_upload(dsph_ctx* ctx, int64_t n, const double* pos3, const double* vel3,
                const double* mass, const double* u, const int32_t* mat, const uint8_t* fixed,
                const int32_t* seg);
int dsph_count(const dsph_ctx* ctx, int64_t* n_out, int32_t* nseg_out);
/* Weibull flaw seeding; replaces System::seed_damage (sph.hpp:117-128), run
 * per segment with the same host mt19937 stream semantics. seg < 0 = all. */
int dsph_seed_damage(dsph_ctx* ctx, int32_t seg, double dx, uint32_t seed);
/* Copy a whole field in IC order (see dsph_field for element types). */
int dsph_get_field(dsph_ctx* ctx, int32_t field, void* dst);
int dsph_set_field(dsph_ctx* ctx, int32_t field, const void* src);

/* ---- the step (System methods, sph.hpp:184-379) ------------------------ */
int dsph_init(dsph_ctx* ctx);                         /* System::init              */
int dsph_compute_density_h(dsph_ctx* ctx);            /* System::compute_density_h */
int dsph_compute_forces(dsph_ctx* ctx);               /* System::compute_forces    */
int dsph_grow_damage(dsph_ctx* ctx, double dt);       /* System::grow_damage       */
/* System::timestep; dt_out holds one value per segment. */
int dsph_timestep(dsph_ctx* ctx, double* dt_out);
/* System::step(dt). dt > 0: every segment steps by dt.  dt <= 0: each
 * segment uses its own device CFL step (no host round trip for the dt).
 * Waits for the step and returns the device flags. */
int dsph_step(dsph_ctx* ctx, double dt);
/* Per-segment host-chosen steps (nseg values). */
int dsph_step_dts(dsph_ctx* ctx, const double* dts);
/* nsteps steps with device-resident CFL dt (dt <= 0) or a fixed dt, replayed
 * from a captured hipGraph when use_graph != 0.  Replaces the host loop of
 * gpu_ic.cpp:128-144 without its per-step host reads.  Waits once, after the
 * last step, and returns the device flags. */
int dsph_run(dsph_ctx* ctx, int32_t nsteps, double dt, int32_t use_graph);
int dsph_synchronize(dsph_ctx* ctx);
/* Set the clock of segment seg (< 0: all): simulated time, step count and the
 * CFL step of the last force evaluation, as a checkpoint records them.  With
 * every field of DSPH_F_* restored as well (including ACC, DUDT and DSDT, the
 * derivatives the next step integrates with), the next dsph_step / dsph_run
 * continues bit for bit where the checkpointed run stopped.  Replaces the
 * restart path of run_ic.cpp:76-102 (which restores t from the header). */
int dsph_set_clock(dsph_ctx* ctx, int32_t seg, double t, int64_t steps, double dt_next);

/* ---- diagnostics, profiling, tuning ---------------------------------- */
int dsph_diagnostics(dsph_ctx* ctx, int32_t seg, dsph_diag* out);
int dsph_set_profiling(dsph_ctx* ctx, int32_t on);
int dsph_get_profile(dsph_ctx* ctx, dsph_profile* out);
int dsph_reset_profile(dsph_ctx* ctx);
int dsph_get_tuning(const dsph_ctx* ctx, dsph_tuning* out);
int dsph_set_tuning(dsph_ctx* ctx, const dsph_tuning* t);
/* Time `steps` steps for each candidate tuning (neighbour lists on or off x
 * cell widths x pair-kernel block sizes) on a scratch copy of the current
 * state, keep the fastest, and restore the state; the chosen tuning is
 * returned and applied. */
int dsph_autotune(dsph_ctx* ctx, int32_t steps, dsph_tuning* chosen);

/* ---- allocation-free sums -------------------------------------------- */
/* Compensated (TwoSum) FP64 sums over n particles of caller-owned device
 * arrays, enqueued on `stream` (a hipStream_t; NULL = the null stream) without
 * synchronising: d_out[0..7] (device memory) receives total mass, kinetic
 * energy sum(m v^2 / 2), internal energy sum(m u), total energy, momentum x, y,
 * z, and the number of particles skipped for a non-finite m, v or u.  The same
 * sums as dsph_diagnostics, over arrays instead of a context.
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  It allocates nothing and clears nothing: its per-chunk
 * partials live in a fixed device array of 4096 chunks of 16384 particles
 * (416 KiB per GPU, held until the process ends), so n is limited to
 * 67,108,864 (larger n returns DSPH_E_RANGE); every call on one GPU shares that
 * array, so calls that overlap on one GPU race.  The array's address on a
 * device is looked up by dsph_create on that device, or else by the first call
 * on that device (make that first call outside graph capture). */
int dsph_energy_sums_noalloc(const double* d_mass, const double* d_vx, const double* d_vy, const double* d_vz,
                             const double* d_u, int64_t n, void* stream, double* d_out);

#ifdef __cplusplus
}
#endif
