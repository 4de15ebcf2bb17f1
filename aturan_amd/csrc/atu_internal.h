// Internal declarations shared by the host code and the HIP kernels of libaturan_amd.
#pragma once

#include "atu.h"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace atu {

// Host-built geometry. Every metric factor the stencils use is formed here once, in FP64, with
// the same operand order as the reference's CellGeometry (planet/RungeKutta_Uran_Turb.cpp:192-283)
// and solver weights (planet/PressureSolver.h:156-162,327-353), so the device reads bits the CPU
// would have computed. Per-(i,j) arrays are [im][jm].
struct Geometry {
    int im = 0, jm = 0, km = 0;
    std::vector<double> radz, thez;          // cumulative coordinates
    // per i
    std::vector<double> rm, inv_rm, inv_rm2, two_inv_rm, sponge, num2;
    // per j
    std::vector<double> sinthe, sinthe2, costhe, cotanthe, sinthe_bf, cor_hydro, sin_weight;
    std::vector<double> pole_row;            // 1 where the sin(theta) floor of the metric is active
    // per (i,j)
    std::vector<double> inv_rmsinthe, inv_rm2sinthe, inv_rm2sinthe2, costhe_inv_rm2sinthe,
        v_metric, num3, den, inv_rmsinthe_s;
    // scalars
    double inv_2dr = 0, inv_2dthe = 0, inv_2dphi = 0, inv_dr2 = 0, inv_dthe2 = 0, inv_dphi2 = 0;
    double num1 = 0;
};

void build_geometry(const atu_grid& g, const atu_params& p, Geometry& geo);
// The red-black sweep's weights (planet/PressureSolver.h:156-162, 327-353) written into caller
// arrays, allocating nothing: num2 [im], num3 and den [im][jm]; returns num1. build_geometry's
// num2, num3 and den are these.
double solver_weights(const atu_grid& g, const atu_params& p, double* num2, double* num3, double* den);

// Tables of the converging Poisson solve: Thomas factors of the radial lines (per (j, s), s = i-1,
// the walls folded), of the per-mode theta systems of the barotropic correction (per (m, jj)), the
// DFT cosines/sines in both layouts and the a-weighted column weight S2 = sum_i a_i / rm_i^2.
struct SolveTables {
    int ni = 0, N = 0, NM = 0, J = 0;
    std::vector<double> line_cp, line_dn;   // [jm][ni]
    std::vector<double> line_inv_dn;        // [jm][ni], 1 / line_dn (the FP32 line pass multiplies)
    std::vector<double> mode_cp, mode_dn;   // [NM+1][J]
    std::vector<double> ct, st;             // [NM+1][N]
    std::vector<double> ctT, stT;           // [N][NM+1]
    double S2 = 0.0, yw = 0.0, ncell = 0.0, e = 0.0;
};
void build_solve_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, SolveTables& t);

// The red-black solver's PRESS_LINE pass and PRESS_BAROTROPIC correction (planet/PressureSolver.h:
// 497-708), each weight formed with the reference's own expression: the line rows' num2 and num3
// (:522-524, a separate set from the pointwise pass's), the Thomas factors of every row j's folded
// radial system (:518-555; the same for every column of the row), the column weight
// A2 = sum_i 1 / rm^2 (:580) and the mode-0 weights |sin theta| with their sums (:657-669).
struct PressRefTables {
    int ni = 0;
    std::vector<double> num2;              // [im]
    std::vector<double> num3;              // [im][jm]
    std::vector<double> line_cp, line_dn;  // [jm][ni]
    std::vector<double> abs_sin;           // [J], |sin(theta_{jj+1})|
    double num1 = 0, a_mid = 0, a_hi = 0, A2 = 0, ys = 0, sw = 0;
};
void build_press_ref_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, PressRefTables& t);

// Weights of PressureSolverUran (planet/PressureSolverUran.h): the defect-correction source's set
// (a_*) and the sweep's set (m_*), each formed with the reference's own expressions.
struct UranTables {
    std::vector<double> a_inv_rm, a_num2, m_num2;                                  // [im]
    std::vector<double> a_inv_rmsinthe, a_num3, a_denom, m_num3, m_inv_denom;     // [im][jm]
    double a_num1 = 0, m_num1 = 0, two_dr = 0, two_dthe = 0, two_dphi = 0, src_scale = 0;
};
void build_uran_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, UranTables& u);

// Tables and scalars of the turbulence closure (planet/Turbulence.h) for one set of knobs, each
// formed with the reference's own expression: per level the metric radius, the wall distances of
// the SST and k-epsilon branches, the ABL taper of clamp_nue and init_fields' per-level factors;
// per row the sin / cos pair of compute_sources and the lone sin of compute_k_omega.
struct TurbTables {
    std::vector<double> rm, inv_rm, ystar_sst, yphys_ke, ystar_ke, taper, karman_z, pz;   // [im]
    std::vector<double> sin_pair, cos_pair, sin_alone;                                     // [jm]
    double nue_max_nd = 0, k_bg_free_nd = 0, kbg_den = 0, nue_air_nd = 0, nue_air = 0, u_0 = 0, L_m = 0;
    double re_turb = 0, sqrt_bet_star = 0, sqrt_cd = 0, vel_star_max = 0;
    double init_tke = 0, init_dis = 0, inv_u02 = 0, nd_omega = 0, nd_eps = 0, c_m025 = 0, c_075 = 0;
    double sqrt_cnue = 0, karman_l = 0;
};
// The knobs with every unset (<= 0) value replaced by the reference's default; false when one is
// out of range.
bool resolve_turb_opts(const atu_grid& g, const atu_params& p, const atu_turb_opts& in, atu_turb_opts& out);
void build_turb_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, const atu_turb_opts& o,
                       TurbTables& t);

// cp table of H2 for a CpH2 mode (1 normal, 2 equilibrium), J/(kg K).
const std::vector<double>& cp_h2_table(int mode);
double cp_lapse(const atu_params& p, double T_K);

// Derived per-run scalar constants of the RHS, formed in the reference's operand order.
struct StepConsts {
    double dt, c_half, c_full, dt6;
    double coeff_energy_p, scale_p, scale_Cor;
    double re_pr, re, inv_sig_k, inv_sig_w, inv_re;
    double sc_re_ch4, sc_re_h2o, sc_re_h2s, sc_re_nh3, sc_re_nh4sh;
    double precip_num, precip_den, precip_coupling, chemical_reaction;
    double stokes_c, rho_cond_nh4sh, g, mue_mix, u_0;
    double turb_coupling, Pr_t, two_omega, m_two_omega, Coriolis;
    double t_min, t_max, tke_max, dis_min;
    double cp_mix, t_ref, cp_a, cp_b, cp_den, r_h2;
    int cp_mode, cp_gas, turb_active;
    // hydro prep
    double hyd_bnum, nd_rel, dr;
    // solver
    double p_dyn_cap, c43, c13;
};
void build_step_consts(const atu_grid& g, const atu_params& p, StepConsts& c);

// Sets the message atu_last_error(NULL) returns (context-free calls); allocates nothing.
void set_free_error(const char* msg);

// Padded k stride of device fields: km rounded up to a multiple of 32 elements.
inline int padded_ldk(int km) { return (km + 31) / 32 * 32; }

}  // namespace atu
