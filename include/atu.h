/*
 * atu.h -- C ABI of the MI355X dynamical-core step (RHS + RK4 + BC + red-black Poisson).
 *
 * Plain C: pointers, sizes and PODs only, no C++ or torch types, so a C++ host loop, a ctypes
 * binding and a cgo/JNI stub can all bind it. Every function returns 0 on success and a negative
 * atu_status otherwise; atu_last_error() then holds a message. No exception crosses this ABI.
 *
 * Host arrays handed to atu_set_field / atu_get_field are dense [im][jm][km] FP64, k (phi)
 * fastest -- the element order of the reference's jagged Array x[i][j][k] (lib/Array.h:31) once
 * each (i,j) row is packed. The context owns every device buffer; calls on one context are
 * serialised on its HIP stream and are not reentrant.
 *
 * Field naming follows the reference (planet/cUranusModel.h:1356-1373): u is RADIAL, v is the
 * theta (colatitude, +southward) component and w is ZONAL.
 */
#ifndef ATU_H
#define ATU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATU_ABI_VERSION 1

/* ---- status codes ---- */
typedef enum atu_status {
    ATU_OK = 0,
    ATU_EINVAL = -1,      /* bad argument (null pointer, unknown id, size out of range)   */
    ATU_EHIP = -2,        /* a HIP runtime call failed                                      */
    ATU_ENOMEM = -3,      /* device or host allocation failed                               */
    ATU_ENONFINITE = -4,  /* a non-finite value was detected where the step forbids it      */
    ATU_ENOCONV = -5,     /* the Poisson solve did not reach its tolerance within max_sweeps */
    ATU_EUNSUPPORTED = -6 /* option combination this build does not implement               */
} atu_status;

/* ---- fields ----
 * 0..18   the prognostic set P19 (planet/cUranusModel.h:1356-1373, 352-353)
 * 19..39  the 21 frozen per-step inputs the RHS reads at default knobs
 *         (planet/RHS_Uran_Turb.cpp:439,544,793-913,1011-1014)
 * 40..45  p_stat, p_dyn, p_hydro and the projection fields aux_u/v/w
 */
typedef enum atu_field {
    ATU_T = 0, ATU_U, ATU_V, ATU_W,
    ATU_CH4, ATU_CH4_CLOUD, ATU_CH4_ICE,
    ATU_H2O, ATU_H2O_CLOUD, ATU_H2O_ICE,
    ATU_H2S, ATU_H2S_CLOUD, ATU_H2S_ICE,
    ATU_NH3, ATU_NH3_CLOUD, ATU_NH3_ICE,
    ATU_NH4SH, ATU_TKE, ATU_DIS,                       /* 19 prognostic */
    ATU_NUE = 19, ATU_RHO_MIX, ATU_Q_PRECIP,
    ATU_SP_CH4, ATU_SP_CH4_CLOUD, ATU_SP_CH4_ICE,
    ATU_SP_H2O, ATU_SP_H2O_CLOUD, ATU_SP_H2O_ICE,
    ATU_SP_H2S, ATU_SP_H2S_CLOUD, ATU_SP_H2S_ICE,
    ATU_SP_NH3, ATU_SP_NH3_CLOUD, ATU_SP_NH3_ICE,
    ATU_MF_H2S, ATU_MF_NH3, ATU_MF_NH4SH,
    ATU_FLUXLIM_NH4SH, ATU_TKE_SOURCE, ATU_DIS_SOURCE, /* 21 frozen   */
    ATU_P_STAT = 40, ATU_P_DYN, ATU_P_HYDRO,
    ATU_AUX_U, ATU_AUX_V, ATU_AUX_W,
    ATU_NFIELDS = 46
} atu_field;

#define ATU_NPROG 19
#define ATU_NFROZEN 21

/* ---- grid ----
 * Coordinates are built by cumulative addition z[l] = z[l-1] + d (lib/Array_1D.cpp:50-61).
 * The reference grid: im,jm,km = 41,181,361 (planet/cUranusModel.h:867), r0 = 1, dr = 0.025,
 * the0 = phi0 = 0, dthe = dphi = 1/(180/pi) (planet/cUranusModel.cpp:339-351).
 * phi is periodic over km-1 distinct meridians; k = km-1 is the copy of k = 0.
 */
typedef struct atu_grid {
    int32_t im, jm, km;
    int32_t pad_;
    double r0, dr;
    double the0, dthe;
    double phi0, dphi;
} atu_grid;

/* ---- constants and default knobs consumed by the step (planet/cUranusModel.h, param.py) ---- */
typedef struct atu_params {
    double dt;                 /* nondimensional step, planet/cUranusModel.cpp:475          */
    double re, pr, g, omega, u_0, t_ref, L_atm;
    double Coriolis, centrifugal, buoyancy, chemical_reaction;
    double sc_ch4, sc_h2o, sc_h2s, sc_nh3, sc_nh4sh;
    double alpha_sponge;
    double rho_cond_nh4sh, r_p_nh4sh;
    double r_mix, R_mix, cp_mix, mue_mix;      /* mixture, after the T_INIT_ANCHOR rescale   */
    double r_h2, r_he, r_ch4;                  /* rescaled reference densities (cp(T) mix)   */
    double cp_h2, cp_he, cp_ch4;               /* constant specific heats, J/(kg K)          */
    double R_metric_km;                        /* metricRadius shift (planet/cUranusModel.h:641-647) */
    double sinthe_min;                         /* polar metric floor (planet/cUranusModel.h:675-688) */
    double t_min_K, t_max_K;                   /* RK clamps (planet/cUranusModel.h:116-117)  */
    double tke_max_m2s2, dis_min_nd;           /* planet/RungeKutta_Uran_Turb.cpp:210-211    */
    double turb_coupling, Pr_t, sig_k, sig_w;
    double precip_coupling;
    double p_dyn_cap;                          /* planet/PressureSolver.h:168                */
    double coriolis_scale;                     /* L_atm*1e3/u_0 with CORIOLIS_FULL            */
    int32_t cp_mode;                           /* CpH2 mode: 0 constant, 1 normal, 2 equilibrium */
    int32_t cp_gas;                            /* CpH2 CP_GAS: 1 = H2-He-CH4 gas mix          */
    int32_t turb_active;                       /* tke/dis transport on                        */
    int32_t lid_vw;                            /* ATURAN_BC_LID_VW: 0 Neumann, 2 linear, 3 cubic */
    int32_t rigid_lid;                         /* u = 0 on both radial planes                 */
    int32_t radius_positive;                   /* ATURAN_BC_RADIUS_POSITIVE: 0 off, 1 floor at 0 */
    int32_t species_clamp;                     /* ATURAN_SPECIES_CLAMP after the BC: 1 (the default,
                                                  planet/BC_Uran.h:110-128) zeroes every species value
                                                  that is not >= 0 (NaN included); 0 off            */
    int32_t pad_;
} atu_params;

/* ---- precision of a context ---- */
typedef enum atu_precision {
    ATU_PREC_FP64 = 0,   /* every array and every operation in FP64                          */
    ATU_PREC_MIXED = 1   /* FP32 stencil/stage states/smoother; FP64 y_n, accumulator,
                            residual, refinement correction and pole-row metric factors      */
} atu_precision;

/* ---- Poisson solve options ---- */
typedef enum atu_solve_mode {
    ATU_SOLVE_REFERENCE = 0, /* `sweeps` red-black sweeps with the wall BCs inside each sweep,
                                exactly planet/PressureSolver.h:318-757 (omega = 1 is the
                                reference's Gauss-Seidel; other omega is SOR)                 */
    ATU_SOLVE_CONVERGE = 1,  /* the same system with every wall folded in, iterated to
                                ||b - A p||_2 / ||b||_2 <= rel_tol, where b is the source less
                                its component along the all-Neumann left null vector; zebra
                                red-black radial line relaxation (the reference's PRESS_LINE)
                                plus an exact barotropic correction (its PRESS_BAROTROPIC) per
                                iteration; the interior mean of p is held. On a mixed context:
                                FP64 residual and accumulation, FP32 smoothing (iterative
                                refinement)                                                    */
    ATU_SOLVE_URAN = 2       /* PressureSolverUran::run (planet/PressureSolverUran.h:33-250), the
                                reference's default solver, `sweeps` calls: the radial walls of
                                aux, the defect-correction source src_scale div(rhs) + Lc(p_old)
                                from the last stage's rhs_u, rhs_v, rhs_w (kept by the context
                                when atu_keep_stage_rhs is on; atu_step turns it on for this
                                mode), one serial lexicographic Gauss-Seidel sweep (run as
                                hyperplanes i+j+k = const, which reproduces it exactly), and the
                                walls of p. FP64 contexts only.                                */
} atu_solve_mode;

/* Options OR'ed into atu_solver_opts.mode with ATU_SOLVE_REFERENCE: the reference's environment
 * switches <TAG>_PRESS_LINE and <TAG>_PRESS_BAROTROPIC (planet/PressureSolver.h:284-303), run after
 * the pointwise colour passes of every sweep and before its wall passes, in this order:
 *   ATU_SOLVE_PRESS_LINE        every (j, k) column of one colour ((j + k) & 1, colour 0 then 1)
 *                               relaxed exactly in r by the Thomas algorithm, the radial ends folded
 *                               with the 2-point Neumann relation; only finite values are written
 *                               (:497-570). The interior mean of p is held at its value before the
 *                               first sweep (:309-316, 748-755), summed in the serial (i, j, k) order
 *                               of the reference's reduction on one thread (its order on more threads
 *                               follows the thread count).
 *   ATU_SOLVE_PRESS_BAROTROPIC  the column-mean correction: column residuals, a DFT in phi, one
 *                               tridiagonal theta solve per mode (mode 0 made consistent, pinned at
 *                               j = 1, zero sin-weighted mean), the inverse DFT added to every level
 *                               (:572-708). The reference runs it only with a periodic seam and no
 *                               solid body, which is this model's configuration.
 * Any other bit, or either bit with another mode, is ATU_EINVAL. */
#define ATU_SOLVE_MODE_MASK 0xff
#define ATU_SOLVE_PRESS_LINE 0x100
#define ATU_SOLVE_PRESS_BAROTROPIC 0x200

typedef struct atu_solver_opts {
    int32_t mode;            /* atu_solve_mode, with ATU_SOLVE_REFERENCE optionally OR'ed with
                                ATU_SOLVE_PRESS_LINE and / or ATU_SOLVE_PRESS_BAROTROPIC        */
    int32_t sweeps;          /* REFERENCE: sweeps to run. CONVERGE on a mixed context: FP32
                                smoothing iterations per FP64 refinement (<= 0: 1)            */
    int32_t max_sweeps;      /* CONVERGE: upper bound on smoothing iterations                 */
    int32_t warm_start;      /* CONVERGE in atu_step: 1 starts each solve from the extrapolation
                                2 p_n - p_{n-1} of the last two solutions, 2 from
                                3 p_n - 3 p_{n-1} + p_{n-2} once three are held (the held mean
                                stays p_n's). The context forgets p_{n-1} on atu_set_field,
                                atu_hydro_prep, atu_rk4, atu_apply_bc, atu_pressure_solve and the
                                window calls; a caller that writes p_dyn through
                                atu_field_device_ptr passes 0 on its next atu_step. 0: from p_n  */
    double omega;            /* relaxation factor (1 = Gauss-Seidel)                          */
    double rel_tol;          /* CONVERGE: relative residual target (FP64)                     */
} atu_solver_opts;

typedef struct atu_solve_stats {
    int32_t sweeps_done;
    int32_t outer_iters;
    double rel_resid;        /* ||src - Lc p||_2 / ||src||_2 after the solve (FP64)           */
    double src_norm;         /* ||src||_2 over the interior                                   */
} atu_solve_stats;

typedef struct atu_ctx atu_ctx;

/* ---- grid and parameter defaults ---- */

/* The reference grid (refine = 1) or the refined shell (refine = 2: 81x361x721, dr/2, 0.5 deg).
 * Replaces the compile-time im, jm, km of planet/cUranusModel.h:867 and the coordinate set-up of
 * planet/cUranusModel.cpp:365-370. */
int atu_grid_default(int refine, atu_grid* out);

/* The constants of the default configuration, including the mixture properties of
 * ThermalPropertiesUran (planet/ChemistryUran.h:437-490) and the T_INIT_ANCHOR rescale of the
 * reference densities (planet/cUranusModel.cpp:504-511). */
int atu_params_default(const atu_grid* g, atu_params* out);

/* The H2 cp table of planet/CpH2.h:77-86 (mode 1 or 2; 2991 entries from 5 K in 0.5 K steps), in
 * J/(kg K). `n` receives the length; pass out = NULL to query it. */
int atu_cp_table(int mode, double* out, int* n);

/* cp(T) for lapse rates and the RHS compression term, CpH2::cp_lapse (planet/CpH2.h:128-134). */
double atu_cp_lapse(const atu_params* p, double T_K);

/* Deterministic synthetic initial state (SplitMix64 -> 53-bit uniforms -> Box-Muller, drawn in
 * (field, i, j, k) order). Writes all 46 fields into `fields` (ATU_NFIELDS consecutive dense
 * [im][jm][km] arrays). The domain BC is not applied: callers run atu_apply_bc once after upload.
 * No counterpart in the reference: it stands in for the reference's initialisation
 * (planet/InitVariables_Uran.cpp), which the benchmark replaces by this seeded state. */
int atu_synthetic_state(const atu_grid* g, const atu_params* p, uint64_t seed, double* fields);

/* ---- context ---- */

/* Create a context on the current HIP device. Allocates every device buffer: y_n, two stage
 * states and the accumulator for the 19 prognostic fields, the frozen inputs, p_stat, p_dyn,
 * p_hydro, aux, the solver source and the geometry and cp tables. Replaces the model-owned
 * arrays of cUranusModel::resetArrays (planet/cUranusModel.cpp:1416-1634). */
int atu_create(const atu_grid* g, const atu_params* p, int precision, atu_ctx** out);
void atu_destroy(atu_ctx* ctx);
const char* atu_last_error(const atu_ctx* ctx);   /* NULL ctx: the last error of a context-free
                                                     call (atu_create,
                                                     atu_pressure_residual_static)            */

/* Attach the context to a HIP stream (hipStream_t passed as void*; NULL = its own stream). */
int atu_set_stream(atu_ctx* ctx, void* hip_stream);
int atu_synchronize(atu_ctx* ctx);

/* Copy one field in or out (dense [im][jm][km] FP64 host array). atu_set_field on a prognostic
 * field writes the start-of-step state y_n and every stage buffer, the role restoreVar plays in
 * the reference (planet/cUranusModel.cpp:1638-1685). */
int atu_set_field(atu_ctx* ctx, int field, const double* host);
int atu_get_field(atu_ctx* ctx, int field, double* host);

/* The same from and to a dense device array of the context's device, rows of ld_k >= km FP64
 * elements ([im][jm][ld_k]): a caller's tensor in HBM, no host copy. Both return once the copy has
 * run on the context's stream; the caller makes sure its own writes of `dsrc` are complete before the
 * call (no counterpart in the reference: its arrays live on the host). */
int atu_set_field_device(atu_ctx* ctx, int field, const double* dsrc, int ld_k);
int atu_get_field_device(atu_ctx* ctx, int field, double* ddst, int ld_k);

/* ---- the step pieces ---- */

/* computeBuoyancyRefLevel + computeHydrostaticPressure (planet/RungeKutta_Uran_Turb.cpp:18-177):
 * the sin(theta)-weighted buoyancy level means and the trapezoid column scan into p_hydro.
 * `level_means` (optional, im doubles) receives buoy_ref_level. */
int atu_hydro_prep(atu_ctx* ctx, double* level_means);

/* cUranusModel::RungeKuttaUran (planet/RungeKutta_Uran_Turb.cpp:179-404) with RHSUran
 * (planet/RHS_Uran_Turb.cpp:20-1023) fused into its stages: four stages, the accumulator, the
 * clamps and the final update, in place on the prognostic fields. aux_u/v/w are those of the last
 * stage. Boundary planes are not written (the stencils read the start-of-step BC values). */
int atu_rk4(atu_ctx* ctx);

/* BC_Uran::bcRadius/bcTheta/bcPhi (planet/BC_Uran.h:140-173, planet/BoundaryConditions.h:152-518)
 * on the prognostic fields, then clampNegativeSpecies when params.species_clamp is set. */
int atu_apply_bc(atu_ctx* ctx);

/* prepareProjectionBoundaries (planet/Pressure_Uran.cpp:30-146) followed by the red-black solve
 * of PressureSolver<Planet>::run (planet/PressureSolver.h:145-766) for p_dyn. */
int atu_pressure_solve(atu_ctx* ctx, const atu_solver_opts* opts, atu_solve_stats* stats);

/* Keep (on != 0) or stop keeping rhs_u, rhs_v, rhs_w of the last RK stage in the context: the
 * reference's rhs_* arrays that its default solver reads (planet/cUranusModel.h:189-208). Off by
 * default; the stage kernel then does not write them. */
int atu_keep_stage_rhs(atu_ctx* ctx, int on);

/* One outer step as the benchmark loop orders it: hydro prep, RK4, BC (with clampNegativeSpecies
 * when params.species_clamp is set: the same values as atu_apply_bc's), Poisson, `nsteps` times.
 * `stats` (optional) receives the last solve's statistics. */
int atu_step(atu_ctx* ctx, int nsteps, const atu_solver_opts* opts, atu_solve_stats* stats);

/* Relative residual ||src - Lc p||_2 / ||src||_2 of the current p_dyn against the current source
 * (FP64; the source must have been prepared by a solve). */
int atu_pressure_residual(atu_ctx* ctx, double* rel_resid, double* src_norm);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): atu_pressure_residual's
 * result for caller-held device arrays p_dyn and src ([im][jm][ld_k] FP64, k fastest, ld_k >= km)
 * on grid g with parameters p, without a context. The same weights and kernels as a context's
 * residual, so the result is bit-identical to it for the same arrays. The weights and block partials
 * live in a static device array of 4 MiB per GPU (held until the process ends); 81x361x721 needs
 * 1.8 MB of it, and a grid needing more than 4 MiB is ATU_EINVAL, with the need stated in
 * atu_last_error(NULL). Allocates nothing; runs on `hip_stream` (hipStream_t as void*, NULL: the
 * default stream) and returns when the result is on the host. Calls are serialised by a lock; a
 * call overlapping another on one GPU would race on the array. */
int atu_pressure_residual_static(const atu_grid* g, const atu_params* p, const double* p_dyn, const double* src,
                                 int ld_k, void* hip_stream, double* rel_resid, double* src_norm);

/* Shapiro (1-2-1) wiggle damping of one prognostic field: AtomUtils::damp_wiggles
 * (lib/Utils.cpp:224-306) with i_surface = nullptr, as the reference's loop calls it after the RK
 * step on t (planet/cUranusModel.cpp:1218-1220) and on u, v, w (:1232-1236): `passes` times along
 * k (periodic over km-1 meridians, k = km-1 the copy of k = 0), along j (clamped), along i
 * (clamped), f + strength/4 (f[n-1] - 2 f + f[n+1]). The reference's loop applies the domain BC
 * after it (atu_apply_bc). */
int atu_damp_wiggles(atu_ctx* ctx, int field, int along_i, int along_j, int along_k, double strength, int passes);

/* ---- turbulence closure: Turbulence<cUranusModel> (planet/Turbulence.h, planet/TurbulenceUran.h) ----
 * Fills the eddy viscosity nue*, the production prod and the sources tke_source, dis_source of the
 * k-omega SST (Menter 1994, the default turb_model), k-omega (Wilcox) or k-epsilon (Chien) closure
 * from u, v, w, tke, dis: atu_turbulence_run is Turbulence::run (planet/Turbulence.h:306-332), the
 * reference loop's call after restoreVar (planet/cUranusModel.cpp:1270); atu_turbulence_init is
 * Turbulence::init (:282-303), which first seeds tke and dis. Both act on the context's fields
 * ATU_TKE, ATU_DIS, ATU_NUE, ATU_TKE_SOURCE, ATU_DIS_SOURCE and on four arrays of the closure kept
 * by the context (atu_turb_array): prod, tken, disn and the [jm][km] friction velocity vel_star.
 * The grid is Uranus's: no solid cells (is_solid false, surface_index 0), phi periodic, no
 * coordinate stretching, the closure's own boundaries off (bc_turb_active false). The buoyancy
 * term of the k equation (<TAG>_TURB_BUOY, default 0 on ATURAN) is not built. */
typedef enum atu_turb_model {
    ATU_TURB_K_OMEGA_SST = 0,   /* planet/Turbulence.h:1020-1107 (the default)                  */
    ATU_TURB_K_OMEGA = 1,       /* :929-1017                                                     */
    ATU_TURB_K_EPSILON = 2      /* :885-927                                                      */
} atu_turb_model;

/* The knobs of the closure the reference reads from its environment (ATURAN_<NAME>). */
typedef struct atu_turb_opts {
    int32_t model;              /* atu_turb_model (ATURAN_TURB_MODEL)                           */
    int32_t free_flow;          /* TURB_FREE: 1 no wall anywhere (the default), 0 the surface
                                   boundary-layer form above i = 0 (:131-142)                    */
    int32_t abl_taper;          /* ABL_TAPER: 1 tapers nue* to 0 above abl_height (:564-589)     */
    int32_t curvature;          /* TURB_CURV: 1 the curvature terms of the gradient tensor       */
    double k_bg;                /* TURB_K_BG [m2/s2], background k with free_flow (default 1)    */
    double l_bg;                /* TURB_L_BG [m], background eddy length (default L_atm/(im-1)) */
    double nue_max;             /* NUE_MAX [m2/s], eddy-viscosity ceiling (default 1e5)          */
    double vel_star_max;        /* VEL_STAR_MAX [m/s], u_tau ceiling (default sqrt(0.002) 200)   */
    double abl_height;          /* boundary-layer height [m] (planet/cUranusModel.h:364: 20000) */
} atu_turb_opts;

/* The reference's defaults (every knob unset). */
int atu_turb_opts_default(atu_turb_opts* out);
int atu_turbulence_init(atu_ctx* ctx, const atu_turb_opts* opts);
int atu_turbulence_run(atu_ctx* ctx, const atu_turb_opts* opts);

typedef enum atu_turb_array {
    ATU_TURB_PROD = 0, ATU_TURB_TKEN = 1, ATU_TURB_DISN = 2,   /* [im][jm][km]                   */
    ATU_TURB_VEL_STAR = 3                                      /* [jm][km]                       */
} atu_turb_array;
/* Host copies of the closure's own arrays (zero until the first closure call or a set). */
int atu_turb_set(atu_ctx* ctx, int which, const double* host);
int atu_turb_get(atu_ctx* ctx, int which, double* host);

/* ---- saturation adjustment: the reference loop's block of planet/cUranusModel.cpp:997-1033 ----
 * The four SaturationAdjustmentUran calls, H2O, H2S, NH3, CH4 in that order, each the shared
 * mixed-phase adjustment SaturationAdjustment<cUranusModel>::run (Tao et al. 1988 with the Newton
 * step, planet/SaturationAdjustment.h:256-668, dispatched by planet/SaturationAdjustmentUran.cpp:
 * 199-239), then ChemistryUran::EquilibriumNH4SH("loop") (planet/ChemistryUran.h:139-255). Acts on
 * t, p_stat (read), rho_mix (read: the model's rho_at with ATURAN_LOCAL_RHO 1), the vapour, cloud and
 * ice of each gas and nh3, h2s, nh4sh. Every cell is independent. */
typedef enum atu_gas_id { ATU_GAS_H2O = 0, ATU_GAS_H2S = 1, ATU_GAS_NH3 = 2, ATU_GAS_CH4 = 3 } atu_gas_id;

/* One gas's constants as the loop passes them (planet/cUranusModel.h:880-906, 1045-1192): the
 * melting and ice-onset temperatures [K], ep, the latent heats [J/kg] and the liquid and ice
 * vapour-pressure sets of ATPhys::saturation_vapour_pressure (planet/ATPhys.h:49-52). */
typedef struct atu_gas {
    double t_0, t_00, ep, lv, ls;
    double C, L0, R, del_alf, del_bet;
    double C_i, L0_i, del_alf_i, del_bet_i;
} atu_gas;

typedef struct atu_satadj_opts {
    atu_gas gas[4];             /* indexed by atu_gas_id                                          */
    double m_nh3, m_h2s, m_nh4sh;   /* molecular weights [kg/kmol] (planet/cUranusModel.h:997-999) */
    int32_t gas_mask;           /* bit n: adjust gas n                                            */
    int32_t nh4sh_eq;           /* ATURAN_NH4SH_EQ: the NH4SH equilibrium after the gases         */
    /* knobs of the shared routine (ATURAN_SATADJ_<NAME>), Uranus defaults in brackets:
       CONSERVE [1], NEWTON [1], EVAP [1], ICE_ENTRY [0], BERGERON [2], FREEZE [1], DRY_EXIT [1]  */
    int32_t conserve, newton, evap, ice_entry, bergeron, freeze, dry_exit;
    int32_t pad_;
} atu_satadj_opts;

/* The reference's defaults: all four gases, the equilibrium on, every knob unset. */
int atu_satadj_opts_default(atu_satadj_opts* out);
int atu_saturation_adjust(atu_ctx* ctx, const atu_satadj_opts* opts);

/* Device pointer of a field's storage (for zero-copy interop); `ld_k` receives the k stride in
 * elements, `elem_bytes` 8 or 4. The layout is [im][jm][ld_k]. */
int atu_field_device_ptr(atu_ctx* ctx, int field, void** dptr, int* ld_k, int* elem_bytes);

/* Library build information string (arch, flags, the HIP version it was built with). */
const char* atu_build_info(void);

/* HIP_VERSION the library was built with and hipRuntimeGetVersion of the runtime it runs on. The two
 * differ when another copy of libamdhip64 was loaded into the process first (a framework's bundled
 * runtime shares the soname); atu_create then refuses with ATU_EUNSUPPORTED. */
int atu_runtime_versions(int* built, int* running);

/* hipMemGetInfo of the current device: free and total bytes (for reporting the memory a context and
 * its window hold; no counterpart in the reference). */
int atu_device_memory(size_t* free_bytes, size_t* total_bytes);

/* A device-to-device copy of `bytes` (a multiple of 16) between two buffers it allocates, `launches`
 * times after one warm-up launch; *ms_per_launch receives the mean duration (HIP events). The
 * benchmark's measure of the HBM bandwidth a streaming kernel reaches on this device, and the
 * calibration of the hardware byte counters (kernel k_copy_probe). No counterpart in the reference. */
int atu_copy_probe(size_t bytes, int launches, double* ms_per_launch);

/* ---- gradients over a window of steps (no counterpart in the reference) ----
 * The loss of a window of `nsteps` outer steps (atu_step's order, FP64 context, converging solve)
 * is the misfit of the final zonal wind to a target jet,
 *     L = 1/(ni nj) sum_{i,j interior} ( 1/N sum_{k<N} w(i,j,k) - target_amp cos(3 lat_j) )^2,
 * lat_j = pi/2 - theta_j, N = km-1 meridians, ni = im-2, nj = jm-2 (w is non-dimensional: the
 * default target_amp 1.2 is 120 m/s at u_0 = 100 m/s). atu_window_forward runs the window from
 * the context's current state, keeping a checkpoint of the state (prognostic fields and p_dyn)
 * every `checkpoint_every` steps, and returns L; atu_window_backward then runs the discrete
 * adjoint of the same steps segment by segment (each segment recomputed from its checkpoint with
 * its stage states taped) and returns dloss * dL/dT0 and dloss * dL/dnue, T0 being the initial
 * temperature field and nue the eddy viscosity field (the derivative for one uniform change of
 * every cell of nue). The solve is differentiated as exact; both the forward solves of the
 * window and the adjoint solves run `solve_iters` iterations when it is > 0, else to rel_tol.
 * The state after either call is the window's final state. Needs rigid_lid = 1. On a mixed
 * context the window's steps (forward and recompute) are the mixed step and the adjoint solves the
 * mixed refinement; the stage states are taped widened to FP64 and the adjoint kernels are FP64. */
typedef struct atu_window_opts {
    int32_t nsteps;
    int32_t checkpoint_every;   /* <= 0: 10                                                  */
    int32_t solve_iters;        /* fixed iterations per solve (> 0), or 0: solve to rel_tol    */
    int32_t rb_sweeps;          /* 0: the solve is ATU_SOLVE_CONVERGE (solve_iters / rel_tol),
                                   differentiated as an exact solve. > 0: the solve is that many
                                   red-black sweeps of the reference (ATU_SOLVE_REFERENCE, omega
                                   1), differentiated exactly: the sweeps transposed in reverse
                                   order, each its walls folded back, then its black and its red
                                   pass with transposed coefficients                          */
    double rel_tol;             /* used when solve_iters == 0 and rb_sweeps == 0              */
    double target_amp;          /* <= 0: 1.2                                                  */
} atu_window_opts;

int atu_window_forward(atu_ctx* ctx, const atu_window_opts* opts, double* loss);
/* grad_t0: dense [im][jm][km] host array or NULL; grad_nue: scalar or NULL. The backward differentiates
 * the last atu_window_forward; it returns ATU_EINVAL when a frozen input of the window (fields 19..39)
 * or p_stat was written since (atu_set_field, atu_set_field_device, the turbulence closure), because
 * its recomputation would then read other inputs than the forward did. A caller that writes those
 * fields through atu_field_device_ptr runs the forward again itself. */
int atu_window_backward(atu_ctx* ctx, double dloss, double* grad_t0, double* grad_nue);
/* The same with grad_t0 a dense device array of rows of ld_k >= km elements. */
int atu_window_backward_device(atu_ctx* ctx, double dloss, double* grad_t0, int ld_k, double* grad_nue);
/* A number identifying the forward the next atu_window_backward would differentiate: it changes with
 * every atu_window_forward and is 0 when none is held or its inputs were written since (a binding
 * stores it with its forward and checks it before its backward). */
int atu_window_generation(const atu_ctx* ctx, uint64_t* generation);
/* Put the state (prognostic fields and p_dyn) back to the start of the last atu_window_forward,
 * from its first checkpoint on the device (for repeated evaluations such as finite differences:
 * rewind, change a field with atu_set_field, run the window again). */
int atu_window_rewind(atu_ctx* ctx);

/* ---- per-kernel timing (no counterpart in the reference; used by the benchmark) ----
 * Between atu_profile_begin and atu_profile_query every kernel launch of the context is bracketed
 * by HIP events recorded on the context's stream. atu_profile_query synchronises and returns the
 * summed duration and the launch count of one kernel class: "rk_stage", "rb_pass", "hydro", "bc",
 * "proj_src", "residual", "line", "baro", "adjoint", or "step" (each whole step of atu_step, from
 * its first launch to its last on the context's stream, gaps included). */
int atu_profile_begin(atu_ctx* ctx);
/* The same for the listed classes only (comma-separated names), leaving the other launches
 * unbracketed: an event pair between two kernels delays the second by a few microseconds. */
int atu_profile_begin_classes(atu_ctx* ctx, const char* classes);
int atu_profile_query(atu_ctx* ctx, const char* kernel_class, double* total_ms, int* launches);
int atu_profile_end(atu_ctx* ctx);
/* Bracket the launches of every `every`-th step of atu_step only (counted from atu_profile_begin;
 * default 1): a sample of the steps carries the events, the others run without their gaps. */
int atu_profile_sample(atu_ctx* ctx, int every);

#ifdef __cplusplus
}
#endif

#endif /* ATU_H */
