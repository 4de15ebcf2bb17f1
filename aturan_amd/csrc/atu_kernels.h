// Device-side argument blocks and launcher declarations of libaturan_amd.
#pragma once

#include <hip/hip_runtime.h>

#include "atu_internal.h"

namespace atu {

// Geometry tables in device memory (FP64), built by build_geometry().
struct DevGeo {
    const double *rm, *inv_rm, *inv_rm2, *two_inv_rm, *sponge, *num2;                 // [im]
    const double *costhe, *cotanthe, *sinthe_bf, *cor_hydro, *sin_weight, *pole_row;  // [jm]
    const double *inv_rmsinthe, *inv_rm2sinthe, *inv_rm2sinthe2, *costhe_inv_rm2sinthe,
        *v_metric, *num3, *den, *inv_rmsinthe_s;                                       // [im][jm]
    const double* cp_tab;                                                              // [cp_n]
    int cp_n;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2, num1;
};

// Shape of every 3-D device field: [im][jm][ldk], k fastest; fields of one buffer are `fs` apart.
struct Shape {
    int im, jm, km, ldk;
    size_t plane;   // jm * ldk
    size_t fs;      // im * jm * ldk
};

// One RK stage. `xs` is the state the stencils read; `yn` the FP64 start-of-step state; `acc` the
// FP64 running sum k1 + 2k2 + 2k3 + k4; `xnext` the next stage's input (stages 0-2); stage 3 writes
// y_{n+1} into `yn` (own cell only) and, when non-null, its FP32 mirror `ymirror`.
template <typename S>
struct StageArgs {
    const S* xs;
    const double* yn;
    double* yout;
    S* xnext;
    S* ymirror;
    double* acc;
    const S* frozen;     // 21 fields
    const double* pdyn;  // FP64 in both precisions: the pressure gradients are differences of
    const double* phydro;// nearby large values, formed in FP64 and rounded once
    S* aux;              // 3 fields, written at stage 3
    double* rhs3;        // when non-null, stage 3 also writes rhs_u, rhs_v, rhs_w (FP64 storage)
    // FP32 storage only: y_n rounded to FP32 (its mirror), from which stages 1-2 form the next
    // stage input (stage 0 reads it as its own stencil centre); the final update reads FP64 y_n
    const S* yn_lo;
    // stage 3 only: clampNegativeSpecies on the species it writes at the rows the BC does not read
    // (3 <= i <= im-4, 3 <= j <= jm-4), for a step whose BC is followed by the clamp of the other rows
    // (launch_species_clamp_shell): the same values as the clamp of every cell after the BC
    int sp_floor;
    Shape sh;
    DevGeo geo;
    StepConsts c;
};

// Launchers (stream-ordered, asynchronous): one RK stage over every interior cell.
hipError_t launch_rk_stage_f64(int stage, const StageArgs<double>& a, hipStream_t s);
hipError_t launch_rk_stage_mixed(int stage, const StageArgs<float>& a, hipStream_t s);

// Hydrostatic prep: level means (in a fixed tree order the oracle restates) into `level_out` [im],
// then the column scan into `phydro`. `rows` holds 2 im jm doubles (the row sums).
hipError_t launch_hydro(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t,
                        const double* v, const double* w, const double* pstat, double* rows, double* level_out,
                        double* phydro, hipStream_t s);
// The two halves: the level means (which read t only) and the column scan (t, v, w).
hipError_t launch_level(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t, const double* pstat,
                        double* rows, double* level_out, hipStream_t s);
hipError_t launch_hydro_scan(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t, const double* v,
                             const double* w, const double* pstat, const double* level, double* phydro,
                             hipStream_t s);

// Element-wise precision conversion of n values (FP32 mirrors of FP64 fields and back).
hipError_t launch_convert(const double* src, float* dst, size_t n, hipStream_t s);
hipError_t launch_convert(const float* src, double* dst, size_t n, hipStream_t s);
hipError_t launch_fill(float* dst, float v, size_t n, hipStream_t s, const int* done = nullptr);

// Domain BC on the fields of the path that the reference's lists name: the 17 prognostic fields of
// `y` (FP64; every value written also goes to each non-null mirror, FP64 `m64a`, `m64b`, FP32
// `m32a..c`) and the three reaction mass fluxes and the NH4SH flux limiter of `frozen` (with the
// optional FP32 mirror `frozen32`).
struct BcTargets {
    double* y;
    double* m64a;
    double* m64b;
    float* m32a;
    float* m32b;
    float* m32c;
    double* frozen;
    float* frozen32;
};
hipError_t launch_bc(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s, bool clamp = true);
// clampNegativeSpecies alone (launch_bc with clamp = false and this, in that order, is launch_bc)
hipError_t launch_species_clamp(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s);
// The same clamp on the rows the BC writes or reads only (i < 3 or i > im-4, or j < 3 or j > jm-4): the
// rest of a step's species were clamped by its last stage (StageArgs::sp_floor).
hipError_t launch_species_clamp_shell(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s);
// The same passes on the prognostic fields [f0, f1) only, and with `rest` the frozen fields and the
// field-specific forms (rigid lid, species floor, lid form of v and w, species clamp). The fields do
// not interact in the BC, so a pass over [0, 1) followed by one over [1, 17) with `rest` is the BC.
hipError_t launch_bc_part(const Shape& sh, const atu_params& p, const BcTargets& t, int f0, int f1, bool rest,
                          hipStream_t s, bool clamp = true);

// The clamped divergence source of the Poisson solve at interior cell (i, j, k) (x its offset, nk
// the ring of meridians): div(aux) with the central differences of the projection
// (planet/Pressure_Uran.cpp:30-146), non-finite -> 0, clamped to +-den * p_dyn_cap.
// The source at one cell from its six aux neighbours (u at i+-1, v at j+-1, w at k+-1 across the
// seam), formed as poisson_src_at forms it.
__device__ __forceinline__ double poisson_src_of(const DevGeo& g, const StepConsts& c, int i, size_t ij, double up,
                                                 double un, double vp, double vn, double we, double ww) {
    const double du_dr = (up - un) * g.inv_2dr;
    const double dv_dthe = (vp - vn) * g.inv_2dthe;
    const double dw_dphi = (we - ww) * g.inv_2dphi;
    double div_src = du_dr * 1.0 + dv_dthe * g.inv_rm[i] + dw_dphi * g.inv_rmsinthe_s[ij];
    const double src_max = g.den[ij] * c.p_dyn_cap;
    if (!isfinite(div_src)) div_src = 0.0;
    else if (div_src > src_max) div_src = src_max;
    else if (div_src < -src_max) div_src = -src_max;
    return div_src;
}
template <typename S>
__device__ __forceinline__ double poisson_src_at(const Shape& sh, const DevGeo& g, const StepConsts& c, const S* aux,
                                                 int i, int j, int k, size_t x) {
    const int nk = sh.km - 1;
    const size_t row = x - k;
    const size_t xw = row + (k == 0 ? nk - 1 : k - 1), xe = row + (k == nk - 1 ? 0 : k + 1);
    return poisson_src_of(g, c, i, (size_t)i * sh.jm + j, (double)aux[x + sh.plane], (double)aux[x - sh.plane],
                          (double)aux[sh.fs + x + sh.ldk], (double)aux[sh.fs + x - sh.ldk], (double)aux[2 * sh.fs + xe],
                          (double)aux[2 * sh.fs + xw]);
}

// Projection prep on aux (prepareProjectionBoundaries) and the clamped divergence source.
template <typename S>
hipError_t launch_proj_src(const Shape& sh, const DevGeo& g, const StepConsts& c, S* aux, double* src,
                           hipStream_t s, bool form_src = true);

// One colour pass of the red-black relaxation on p (FP64 or FP32 storage, FP64 source). The black
// pass (colour 1) also writes the radial and polar wall values and the seam copy that the
// reference's boundary passes form after each sweep.
// With form_walls 0 the black pass leaves the walls to a separate pass (launch_walls, launch_p_corners,
// launch_seam), for passes that run between the colour passes and the walls.
template <typename P>
hipError_t launch_rb_pass(const Shape& sh, const DevGeo& g, const StepConsts& c, P* p, const double* src,
                          int colour, double omega, hipStream_t s, int form_walls = 1);
// The four corner rows (i = 0, im-1 at j = 0, jm-1) and their seam copies, after the last sweep.
template <typename P>
hipError_t launch_p_corners(const Shape& sh, const StepConsts& c, P* p, hipStream_t s, const int* skip = nullptr);

// ---- the converging solve (atu_solve.hip) ----
struct DevSolve {
    const double *line_cp, *line_dn, *line_inv_dn, *mode_cp, *mode_dn, *ct, *st, *ctT, *stT;
    int ni, N, NM, J;
    double S2, yw, ncell, e;
};
enum { DSUM_P = 0, DSUM_YSRC = 1, DSUM_SQ = 2, DSUM_RESID2 = 3 };
// Device-side control of the converging solve. flags[CTL_DONE] set means converged: every kernel of
// an iteration given that flag returns at once, so the host can queue several iterations and read
// the outcome once. The residual norm's final sum updates the flags (see k_dsum_final).
enum { CTL_DONE = 0, CTL_ITERS = 1, CTL_OUTER = 2, CTL_NOTDONE = 3, CTL_N = 4 };
// What the host reads after a batch: written by the residual norm's final sum into host memory
// (pinned, mapped), so the host needs no copy, only to know that the batch has run.
struct SolveMailbox {
    int flags[CTL_N];
    double rr, bb;              // |r|^2, |b|^2
};
struct SolveCtl {
    int* flags = nullptr;       // CTL_N ints, or null: no control
    const double* bb = nullptr; // |b|^2
    double tol = 0.0;
    int iters = 0, outer = 0;   // added to the counters before the test (the work just done)
    SolveMailbox* mbox = nullptr;   // device address of the host mailbox, or null
    const int* skip = nullptr;      // when non-null and set, the sum's kernels return at once
};
// P: storage of the unknown, B: of the right-hand side, R: arithmetic of the line pass
template <typename P, typename B, typename R>
hipError_t launch_line(const Shape& sh, const DevGeo& g, const DevSolve& t, P* p, const B* b, int colour,
                       hipStream_t s, const int* done = nullptr);
template <typename P>
hipError_t launch_walls(const Shape& sh, P* p, hipStream_t s, const int* done = nullptr);
template <typename P>
hipError_t launch_seam(const Shape& sh, P* p, hipStream_t s, const int* skip = nullptr);
template <typename P, typename B>
hipError_t launch_baro(const Shape& sh, const DevGeo& g, const DevSolve& t, P* p, const B* b, double* R, double* Ac,
                       double* As, hipStream_t s, const int* done = nullptr);
// canonical-order interior sum of p (DSUM_P), y*v (DSUM_YSRC), v^2 (DSUM_SQ) or (v - A p)^2
// (DSUM_RESID2, which can also store the residual in FP32 into r_out); `rows` holds
// (im-2)*(jm-2) doubles; the total lands in *out (device)
template <typename P>
hipError_t launch_dsum(int mode, const Shape& sh, const DevGeo& g, const P* p, const double* v, double* rows,
                       double* out, hipStream_t s, float* r_out = nullptr, const SolveCtl& ctl = SolveCtl());
hipError_t launch_make_b(const Shape& sh, const double* src, const double* sc, double yw, double* b, hipStream_t s);
// The start of a converging solve in four launches: scal[0] = y . src and scal[1] = the interior sum
// of p (same orders as launch_dsum), the control flags cleared, b = src - scal[0] / yw and
// scal[2] = |b|^2. `rows`, `rows2`: (im-2)*(jm-2) doubles each. With `pprev` the first pass also
// stores p's interior there (moving the old pprev to `pprev2`, when given) and replaces p by
// 2 p - pprev (extrap 1) or 3 p - 3 pprev + pprev2 (extrap 2): the warm start.
hipError_t launch_solve_setup(const Shape& sh, double* p, const double* src, double* rows, double* rows2,
                              double* scal, double yw, double* b, int* flags, double* pprev, double* pprev2,
                              int extrap, hipStream_t s,
                              const void* aux = nullptr, bool aux_f32 = false, const DevGeo* g = nullptr,
                              const StepConsts* c = nullptr);
hipError_t launch_add_correction(const Shape& sh, const float* e, double* p, hipStream_t s, const int* done = nullptr);
// The end of a converging solve: launch_walls, launch_p_corners, launch_seam and
// launch_dsum(DSUM_P) in one pass plus the total (the same values); `rows` as launch_dsum's.
template <typename P>
hipError_t launch_post_dsum(const Shape& sh, P* p, double* rows, double* out, hipStream_t s, const int* skip = nullptr);
template <typename P>
hipError_t launch_shift_all(const Shape& sh, const double* sc, int slot_before, int slot_now, double ncell, P* p,
                            hipStream_t s, const int* skip = nullptr);

// ---- the red-black solver's PRESS_LINE and PRESS_BAROTROPIC options (atu_press_ref.hip) ----
struct DevPressRef {
    const double *num2, *num3, *line_cp, *line_dn, *abs_sin;
    double num1, a_mid, a_hi, A2, ys, sw;
    int ni;
};
// One colour ((j + k) & 1) of the radial line pass on p; `scratch` holds a field (its contents are
// not kept).
hipError_t launch_rline(const Shape& sh, const DevPressRef& r, double* p, const double* src, double* scratch,
                        int colour, hipStream_t s);
// The interior sum of p (i = 1..im-2, j = 1..jm-2, k = 0..km-2) in the serial (i, j, k) order, into
// *out (device).
hipError_t launch_serial_interior_sum(const Shape& sh, const double* p, double* out, hipStream_t s);
// The barotropic correction added to p: column residuals into R (J*N), their DFT into Ac / As
// ((NM+1)*J each), the per-mode theta solves, the inverse DFT added to every level. The DFT's
// cosines and sines and the modes' Thomas factors are the converging solve's tables `t`.
hipError_t launch_rbaro(const Shape& sh, const DevPressRef& r, const DevSolve& t, double* p, const double* src,
                        double* R, double* Ac, double* As, hipStream_t s);

// ---- the Shapiro wiggle filter (atu_damp.hip): AtomUtils::damp_wiggles on one field; `scratch`
// holds a field (its contents are not kept)
hipError_t launch_damp_wiggles(const Shape& sh, double* field, double* scratch, bool along_i, bool along_j,
                               bool along_k, double strength, int passes, hipStream_t s);

// ---- the turbulence closure (atu_turb.hip): Turbulence<cUranusModel>::init / run ----
struct TurbArgs {
    const double *u, *v, *w;
    double *tke, *dis, *nue, *tke_source, *dis_source;    // context fields
    double *prod, *tken, *disn;                           // the closure's own fields
    double* vel_star;                                     // [jm][ldk]
    const double *rm, *inv_rm, *ystar_sst, *yphys_ke, *ystar_ke, *taper, *karman_z, *pz;   // [im]
    const double *sin_pair, *cos_pair, *sin_alone;                                          // [jm]
    int model, free_flow, curvature, pad_;
    double inv_2dr, inv_2dthe, inv_2dphi, dr, dthe, dphi;
    double nue_max_nd, k_bg_free_nd, kbg_den, nue_air_nd, nue_air, u_0, L_m, re_turb, sqrt_bet_star;
    double sqrt_cd, vel_star_max;
    double init_tke, init_dis, inv_u02, nd_omega, nd_eps, c_m025, c_075, sqrt_cnue, karman_l;
};
hipError_t launch_turbulence(const Shape& sh, const TurbArgs& a, bool init, hipStream_t s);

// ---- the saturation-adjustment block (atu_satadj.hip) ----
struct SatAdjArgs {
    double* y;                  // FP64 prognostic masters (19 fields, `fs` apart)
    double *sa, *sb;            // FP64 stage buffers (FP64 context) or null
    float *y32, *sa32, *sb32;   // FP32 mirrors (mixed context) or null
    const double *p_stat, *rho_mix;
    size_t fs;
    atu_gas gas[4];
    int gas_mask, nh4sh_eq;
    int conserve, newton, evap, ice_entry, bergeron, freeze, dry_exit, pad_;
    double t_ref, cp_mix, r_mix, R_mix;
    double m_nh3, m_h2s, m_nh4sh, nh4sh_heat;
};
hipError_t launch_satadj(const Shape& sh, const SatAdjArgs& a, hipStream_t s);

// ---- PressureSolverUran (atu_uran.hip), the reference's default solver ----
struct UranTabs {
    const double *a_inv_rm, *a_num2, *m_num2;                                  // [im]
    const double *a_inv_rmsinthe, *a_num3, *a_denom, *m_num3, *m_inv_denom;    // [im][jm]
    double a_num1, m_num1, two_dr, two_dthe, two_dphi, src_scale;
};
// One PressureSolverUran::run: the radial walls of aux, the defect-correction source into `incr`
// (from rhs3 = rhs_u, rhs_v, rhs_w of the last stage and the current p), one lexicographic
// Gauss-Seidel sweep of p, and p's walls and seam.
hipError_t launch_uran_call(const Shape& sh, const UranTabs& t, double* aux, const double* rhs3, double* p,
                            double* incr, hipStream_t s);

// ---- the adjoint of the FP64 step (atu_adjoint.hip); 4-field blocks hold T, u, v, w ----
struct AdjStageArgs {
    Shape sh;
    DevGeo geo;
    StepConsts c;
    const double* xs;      // stage input X_s (4 fields)
    const double* pdyn;    // p_dyn read by the step
    const double* nue;     // frozen eddy viscosity
    const double* lk;      // lambda of the stage's tendencies (4 fields)
    const double* laux;    // lambda of aux (3 fields) at stage 3, else null
    const double* ls;      // lambda of the final accumulated sum (4 fields)
    double* lk_prev;       // lambda of the previous stage's tendencies = ws * ls + cs * lambda_X, or null
    double ws, cs;
    double* ly;            // += lambda_X (4 fields)
    double* lp;            // += lambda of p_dyn
    double* lph;           // += lambda of p_hydro
    double* lnue;          // += lambda of the eddy viscosity
};
hipError_t launch_rhs_adjoint(const AdjStageArgs& a, hipStream_t s);
// `parts`: 2 * kHydAdjParts * im doubles of scratch (the per-level partial sums)
constexpr int kHydAdjParts = 16;
hipError_t launch_hydro_adjoint(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* x4,
                                const double* pstat, const double* level, const double* lph, double* gref, double* ly,
                                double* parts,
                                hipStream_t s);
hipError_t launch_bc_adjoint(const Shape& sh, const atu_params& p, double* ly, hipStream_t s);
hipError_t launch_final_fold(const Shape& sh, const StepConsts& c, const double* tnext, const double* ly, double* ls,
                             double* lyn, hipStream_t s);
hipError_t launch_pwall_adjoint(const Shape& sh, const double* lp, double* lpi, hipStream_t s);
hipError_t launch_adj_rhs(const Shape& sh, const double* lpi, const double* sc, double ncell, double* b,
                          hipStream_t s);
hipError_t launch_adj_src(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* z, const double* sc,
                          double yw, const double* src, double* lsrc, hipStream_t s);
hipError_t launch_adj_div(const Shape& sh, const DevGeo& g, const double* lsrc, double* laux, hipStream_t s);
hipError_t launch_add_mean(const Shape& sh, const double* sc, double ncell, double* lp, hipStream_t s);
// The adjoint of `sweeps` red-black sweeps (ATU_SOLVE_REFERENCE at relaxation factor omega) and the
// source clamp: lp holds lambda of the solve's output on entry and of its initial p on return;
// lsrc receives lambda of the unclamped source (src: the clamped source the forward used); t1, t2
// are scratch fields.
hipError_t launch_rb_sweeps_adjoint(const Shape& sh, const DevGeo& g, const StepConsts& c, int sweeps, double omega,
                                    const double* src, double* lp, double* t1, double* t2, double* lsrc,
                                    hipStream_t s);
hipError_t launch_loss(const Shape& sh, const double* w, const double* target, double scale, double* rows, double* lw,
                       double* out, hipStream_t s);
hipError_t launch_sum(const double* v, size_t n, double* part, int nparts, double* out, hipStream_t s);
hipError_t launch_zero(double* v, size_t n, hipStream_t s);
// z = 3 z1 - 3 z2 + z3 over n values (2 z1 - z2 when z3 is null, z1 when z2 is null too)
hipError_t launch_extrapolate(double* z, const double* z1, const double* z2, const double* z3, size_t n,
                              hipStream_t s);

// Residual r = src - Lc p over the interior: writes per-block partial sums of r^2 and src^2 into
// `partials` (2 * nblocks doubles) and the fixed-order total into out[0] (r^2), out[1] (src^2).
int residual_partials_count(const Shape& sh);
template <typename P>
hipError_t launch_residual(const Shape& sh, const DevGeo& g, const P* p, const double* src,
                           double* r_out, double* partials, double* out, hipStream_t s);

}  // namespace atu
