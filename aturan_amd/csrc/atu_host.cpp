// Host side of libaturan_amd: grid and parameter defaults, the cp(T) table of H2, the geometry
// tables uploaded to the device, and the seeded synthetic initial state.
//
// Every formula that reaches the step is evaluated here in FP64 with the operand order of the
// reference expression it restates (cited per function), so the device reads the same bits the
// reference's CPU code forms inline.

#include "atu_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace atu {

// ---------------------------------------------------------------------------------------------
// cp of H2 (planet/CpH2.h:55-93): rotational level sum with theta_r = 87.57 K plus a vibrational
// term, tabulated at 0.5 K from 5 to 1500 K and linearly interpolated.
// ---------------------------------------------------------------------------------------------
namespace {
constexpr double kRgas = 8.314462618, kMH2 = 2.01588e-3, kThR = 87.57, kThV = 6332.0;
constexpr double kTlo = 5.0, kThi = 1500.0, kDT = 0.5;

void level_sums(double T, int J0, double wgt, double& Z, double& E1, double& E2) {
    for (int J = J0; J < 60; J += 2) {
        const double e = kThR * J * (J + 1.0);
        const double b = wgt * (2.0 * J + 1.0) * std::exp(-e / T);
        Z += b;
        E1 += b * e;
        E2 += b * e * e;
    }
}

double rot_heat(double T, double Z, double E1, double E2) {
    const double m1 = E1 / Z, m2 = E2 / Z;
    return (m2 - m1 * m1) / (T * T);
}

// cp/R of H2; mode 1 = normal (ortho:para frozen 3:1), 2 = equilibrium (planet/CpH2.h:55-75).
double cp_over_R(double T, int mode) {
    double c;
    if (mode == 1) {
        double Zp = 0, Ep = 0, Fp = 0, Zo = 0, Eo = 0, Fo = 0;
        level_sums(T, 0, 1.0, Zp, Ep, Fp);
        level_sums(T, 1, 1.0, Zo, Eo, Fo);
        c = 0.25 * rot_heat(T, Zp, Ep, Fp) + 0.75 * rot_heat(T, Zo, Eo, Fo);
    } else {
        double Z = 0, E1 = 0, E2 = 0;
        level_sums(T, 0, 1.0, Z, E1, E2);
        level_sums(T, 1, 3.0, Z, E1, E2);
        c = rot_heat(T, Z, E1, E2);
    }
    const double x = kThV / T;
    const double cv = (x < 700.0) ? x * x * std::exp(x) / ((std::exp(x) - 1.0) * (std::exp(x) - 1.0)) : 0.0;
    return 2.5 + c + cv;
}

std::vector<double> make_table(int mode) {
    const int n = int((kThi - kTlo) / kDT) + 1;
    std::vector<double> t(n);
    for (int i = 0; i < n; i++) t[i] = cp_over_R(kTlo + i * kDT, mode) * kRgas / kMH2;
    return t;
}
}  // namespace

const std::vector<double>& cp_h2_table(int mode) {
    static const std::vector<double> t1 = make_table(1);
    static const std::vector<double> t2 = make_table(2);
    return mode == 1 ? t1 : t2;
}

static double cp_h2_at(double T, int mode, double cp_const) {
    if (mode <= 0) return cp_const;
    const std::vector<double>& t = cp_h2_table(mode);
    const int n = (int)t.size();
    const double x = (std::min(std::max(T, kTlo), kThi) - kTlo) / kDT;
    const int i = std::min(n - 2, (int)x);
    const double f = x - i;
    return t[i] + f * (t[i + 1] - t[i]);
}

// CpH2::Lapse::cp with the model's dyn_on (planet/CpH2.h:128-134).
double cp_lapse(const atu_params& p, double T) {
    if (p.cp_mode == 0 && !p.cp_gas) return p.cp_mix;
    const double ch = cp_h2_at(T, p.cp_mode, p.cp_h2);
    if (p.cp_gas) return (p.r_h2 * ch + p.r_he * p.cp_he + p.r_ch4 * p.cp_ch4) / (p.r_h2 + p.r_he + p.r_ch4);
    return p.cp_mix + (p.r_h2 / p.r_mix) * (ch - p.cp_h2);
}

// ---------------------------------------------------------------------------------------------
// Geometry tables.
// ---------------------------------------------------------------------------------------------
void build_geometry(const atu_grid& g, const atu_params& p, Geometry& geo) {
    const int im = g.im, jm = g.jm, km = g.km;
    geo.im = im; geo.jm = jm; geo.km = km;
    geo.radz.assign(im, 0.0);
    geo.thez.assign(jm, 0.0);
    // Coordinates by cumulative addition, lib/Array_1D.cpp:50-61.
    geo.radz[0] = g.r0;
    for (int l = 1; l < im; l++) geo.radz[l] = geo.radz[l - 1] + g.dr;
    geo.thez[0] = g.the0;
    for (int l = 1; l < jm; l++) geo.thez[l] = geo.thez[l - 1] + g.dthe;

    geo.inv_2dr = 1.0 / (2.0 * g.dr);
    geo.inv_2dthe = 1.0 / (2.0 * g.dthe);
    geo.inv_2dphi = 1.0 / (2.0 * g.dphi);
    geo.inv_dr2 = 1.0 / (g.dr * g.dr);
    geo.inv_dthe2 = 1.0 / (g.dthe * g.dthe);
    geo.inv_dphi2 = 1.0 / (g.dphi * g.dphi);
    // num1 = exp_2_rm * inv_dr2 with exp_rm = 1 (no coordinate stretching), PressureSolver.h:350
    const double exp_rm = 1.0, exp_2_rm = exp_rm * exp_rm;
    geo.num1 = exp_2_rm * geo.inv_dr2;

    // per i: metricRadius (planet/cUranusModel.h:641-647), the sponge (RHS_Uran_Turb.cpp:594-595)
    geo.rm.assign(im, 0); geo.inv_rm.assign(im, 0); geo.inv_rm2.assign(im, 0);
    geo.two_inv_rm.assign(im, 0); geo.sponge.assign(im, 0); geo.num2.assign(im, 0);
    const double shift = (p.R_metric_km > 0.0) ? (p.R_metric_km / p.L_atm - 1.0) : 0.0;
    for (int i = 0; i < im; i++) {
        const double rm = (p.R_metric_km > 0.0) ? geo.radz[i] + shift : geo.radz[i];
        const double rm2 = rm * rm;
        geo.rm[i] = rm;
        geo.inv_rm[i] = 1.0 / rm;
        geo.inv_rm2[i] = 1.0 / rm2;
        geo.two_inv_rm[i] = 2.0 * geo.inv_rm[i];
        const double frac = std::max(0.0, (double)(i - (im - 1) * 3 / 4) / (double)((im - 1) / 4));
        geo.sponge[i] = p.alpha_sponge * frac * frac;
    }

    // per j: RK tables (RungeKutta_Uran_Turb.cpp:192-197), the body-force true sine
    // (RHS_Uran_Turb.cpp:224), the hydrostatic Coriolis factor (RungeKutta_Uran_Turb.cpp:132-133)
    // and the level-mean area weight (RungeKutta_Uran_Turb.cpp:25).
    geo.sinthe.assign(jm, 0); geo.sinthe2.assign(jm, 0); geo.costhe.assign(jm, 0);
    geo.cotanthe.assign(jm, 0); geo.sinthe_bf.assign(jm, 0); geo.cor_hydro.assign(jm, 0);
    geo.sin_weight.assign(jm, 0);
    geo.pole_row.assign(jm, 0);
    std::vector<double> sin_solver(jm);
    for (int j = 0; j < jm; j++) {
        // sin and cos of one angle: the reference's g++ build forms this pair (its loop at
        // planet/RungeKutta_Uran_Turb.cpp:195-196) with one glibc sincos call, whose last bit can
        // differ from separate sin and cos calls, so the pair is formed the same way here
        double sj, cj;
        sincos(geo.thez[j], &sj, &cj);
        const double s = std::max(p.sinthe_min, std::abs(sj));
        const double c = cj;
        geo.sinthe[j] = s;
        geo.sinthe2[j] = s * s;
        geo.costhe[j] = c;
        geo.cotanthe[j] = c / s;
        geo.sinthe_bf[j] = std::sqrt(std::max(0.0, 1.0 - c * c));
        geo.cor_hydro[j] = p.Coriolis * p.coriolis_scale * 2.0 * p.omega * std::abs(std::sin(geo.thez[j]));
        geo.sin_weight[j] = std::sin(geo.thez[j]);
        geo.pole_row[j] = (std::abs(std::sin(geo.thez[j])) < p.sinthe_min) ? 1.0 : 0.0;
        // the solver's own table, planet/PressureSolver.h:156-162
        double ss = std::sin(geo.thez[j]);
        if (ss < p.sinthe_min) ss = p.sinthe_min;
        if (ss == 0.0) ss = 1.0e-5;
        sin_solver[j] = ss;
    }

    const size_t nij = (size_t)im * jm;
    geo.inv_rmsinthe.assign(nij, 0); geo.inv_rm2sinthe.assign(nij, 0);
    geo.inv_rm2sinthe2.assign(nij, 0); geo.costhe_inv_rm2sinthe.assign(nij, 0);
    geo.v_metric.assign(nij, 0); geo.num3.assign(nij, 0); geo.den.assign(nij, 0);
    geo.inv_rmsinthe_s.assign(nij, 0);
    for (int i = 0; i < im; i++) {
        for (int j = 0; j < jm; j++) {
            const size_t q = (size_t)i * jm + j;
            const double rm = geo.rm[i];
            const double s = geo.sinthe[j];
            const double inv_rm2 = geo.inv_rm2[i];
            geo.inv_rmsinthe[q] = 1.0 / (rm * s);
            geo.inv_rm2sinthe[q] = inv_rm2 / s;
            geo.inv_rm2sinthe2[q] = inv_rm2 / geo.sinthe2[j];
            geo.costhe_inv_rm2sinthe[q] = geo.costhe[j] * geo.inv_rm2sinthe[q];
            geo.v_metric[q] = (1.0 + geo.costhe[j] / geo.sinthe2[j]) * inv_rm2;
            geo.inv_rmsinthe_s[q] = 1.0 / (rm * sin_solver[j]);
        }
    }
    solver_weights(g, p, geo.num2.data(), geo.num3.data(), geo.den.data());
}

double solver_weights(const atu_grid& g, const atu_params& p, double* num2, double* num3, double* den) {
    const int im = g.im, jm = g.jm;
    const double inv_dr2 = 1.0 / (g.dr * g.dr);
    const double inv_dthe2 = 1.0 / (g.dthe * g.dthe);
    const double inv_dphi2 = 1.0 / (g.dphi * g.dphi);
    const double exp_rm = 1.0, exp_2_rm = exp_rm * exp_rm;   // no coordinate stretching
    const double num1 = exp_2_rm * inv_dr2;
    const double shift = (p.R_metric_km > 0.0) ? (p.R_metric_km / p.L_atm - 1.0) : 0.0;
    double r = g.r0;                      // the coordinates by cumulative addition, as build_geometry
    for (int i = 0; i < im; i++) {
        if (i > 0) r = r + g.dr;
        const double rm = (p.R_metric_km > 0.0) ? r + shift : r;
        const double rm2 = rm * rm;
        const double inv_rm2 = 1.0 / rm2;
        num2[i] = inv_rm2 * inv_dthe2;
        double th = g.the0;
        for (int j = 0; j < jm; j++) {
            if (j > 0) th = th + g.dthe;
            double ss = std::sin(th);     // the solver's own sine table, PressureSolver.h:156-162
            if (ss < p.sinthe_min) ss = p.sinthe_min;
            if (ss == 0.0) ss = 1.0e-5;
            const double ss2 = ss * ss;
            const double n3 = (inv_rm2 / ss2) * inv_dphi2;
            num3[(size_t)i * jm + j] = n3;
            den[(size_t)i * jm + j] = 2.0 * num1 + 2.0 * num2[i] + 2.0 * n3;
        }
    }
    return num1;
}

// The converging solve works on the system of the red-black sweep with every wall folded in
// (p_wall = c43 p_1 - c13 p_2 in r and theta, planet/PressureSolver.h:718-732). Its radial lines
// are the reference's PRESS_LINE rows (:497-539) and its barotropic modes the PRESS_BAROTROPIC
// theta systems (:637-652), with the column sums weighted by the left null vector (3/2 at the end
// rows) so that the radial part drops out exactly.
void build_solve_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, SolveTables& t) {
    const int im = g.im, jm = g.jm, km = g.km;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const int n = im - 2;
    t.ni = n;
    t.line_cp.assign((size_t)jm * n, 0.0);
    t.line_dn.assign((size_t)jm * n, 0.0);
    t.line_inv_dn.assign((size_t)jm * n, 0.0);
    for (int j = 1; j < jm - 1; j++) {
        double* cp = &t.line_cp[(size_t)j * n];
        double* dn = &t.line_dn[(size_t)j * n];
        for (int s = 0; s < n; s++) {
            const int i = s + 1;
            const double num1 = geo.num1, num2 = geo.num2[i];
            double diag = -geo.den[(size_t)i * jm + j];
            if (j == 1 || j == jm - 2) diag = diag + num2 * c43;
            double a = num1, c = num1;
            if (s == 0) { diag = diag + num1 * c43; c = num1 - num1 * c13; a = 0.0; }
            if (s == n - 1) { diag = diag + num1 * c43; a = num1 - num1 * c13; c = 0.0; }
            if (n == 1) { a = 0.0; c = 0.0; }
            const double d = (s == 0) ? diag : diag - a * cp[s - 1];
            dn[s] = d;
            cp[s] = c / d;
            t.line_inv_dn[(size_t)j * n + s] = 1.0 / d;
        }
    }
    const int N = km - 1, NM = N / 2, J = jm - 2;
    t.N = N; t.NM = NM; t.J = J;
    t.ct.assign((size_t)(NM + 1) * N, 0.0);
    t.st.assign((size_t)(NM + 1) * N, 0.0);
    t.ctT.assign((size_t)(NM + 1) * N, 0.0);
    t.stT.assign((size_t)(NM + 1) * N, 0.0);
    for (int m = 0; m <= NM; m++)
        for (int k = 0; k < N; k++) {
            const double a = 2.0 * M_PI * (double)((long)m * k % N) / N;
            sincos(a, &t.st[(size_t)m * N + k], &t.ct[(size_t)m * N + k]);   // one call, as in the oracle
            t.ctT[(size_t)k * (NM + 1) + m] = t.ct[(size_t)m * N + k];
            t.stT[(size_t)k * (NM + 1) + m] = t.st[(size_t)m * N + k];
        }
    auto endw = [](int x, int lo, int hi) { return (x == lo || x == hi) ? 1.5 : 1.0; };
    double S2 = 0.0;
    for (int i = 1; i < im - 1; i++) {
        const double rm = geo.rm[i];
        S2 += endw(i, 1, im - 2) * (1.0 / (rm * rm));
    }
    t.S2 = S2;
    double sa = 0.0, sb = 0.0;
    for (int i = 1; i < im - 1; i++) sa += endw(i, 1, im - 2);
    for (int j = 1; j < jm - 1; j++) sb += endw(j, 1, jm - 2);
    t.yw = sa * sb * N;    // exact: every partial sum is a multiple of 1/4 below 2^53
    t.ncell = (double)n * (jm - 2) * N;
    std::vector<double> s2(jm);
    for (int j = 0; j < jm; j++) {
        double s = std::sin(geo.thez[j]);
        if (s < p.sinthe_min) s = p.sinthe_min;
        if (s == 0.0) s = 1.0e-5;
        s2[j] = s * s;
    }
    const double e = geo.inv_dthe2;
    t.e = e;
    t.mode_cp.assign((size_t)(NM + 1) * J, 0.0);
    t.mode_dn.assign((size_t)(NM + 1) * J, 0.0);
    for (int m = 0; m <= NM; m++) {
        const double sl = std::sin(M_PI * m / N);
        const double lam = 4.0 * sl * sl * geo.inv_dphi2;
        double prev_cp = 0.0;
        for (int jj = 0; jj < J; jj++) {
            double a = e, c = e, b = -2.0 * e - lam / s2[jj + 1];
            if (jj == 0) { a = 0.0; b = b + c43 * e; c = c - c13 * e; }
            if (jj == J - 1) { c = 0.0; b = b + c43 * e; a = a - c13 * e; }
            if (m == 0 && jj == 0) { b = 1.0; c = 0.0; }   // mode 0 is all-Neumann: q pinned at j = 1
            const double d = (jj == 0) ? b : b - a * prev_cp;
            t.mode_dn[(size_t)m * J + jj] = d;
            t.mode_cp[(size_t)m * J + jj] = c / d;
            prev_cp = t.mode_cp[(size_t)m * J + jj];
        }
    }
}

void build_press_ref_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, PressRefTables& t) {
    const int im = g.im, jm = g.jm, ni = im - 2, J = jm - 2;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    t.ni = ni;
    std::vector<double> st(jm);   // the solver's sine table, planet/PressureSolver.h:156-162
    for (int j = 0; j < jm; j++) {
        double s = std::sin(geo.thez[j]);
        if (s < p.sinthe_min) s = p.sinthe_min;
        if (s == 0.0) s = 1.0e-5;
        st[j] = s;
    }
    const double erm = 1.0;                       // no coordinate stretching
    t.num1 = erm * erm * geo.inv_dr2;
    t.num2.assign(im, 0.0);
    t.num3.assign((size_t)im * jm, 0.0);
    for (int i = 0; i < im; i++) {
        const double rm = geo.rm[i];
        t.num2[i] = geo.inv_dthe2 / (rm * rm);
        for (int j = 0; j < jm; j++) t.num3[(size_t)i * jm + j] = geo.inv_dphi2 / (rm * rm * st[j] * st[j]);
    }
    // a row's system: a = c = -num1, b = 2 (num1 + num2 + num3), the ends folded (:547-550)
    t.a_mid = -t.num1;
    t.a_hi = -t.num1 - (-t.num1) * c13;
    t.line_cp.assign((size_t)jm * ni, 0.0);
    t.line_dn.assign((size_t)jm * ni, 0.0);
    for (int j = 1; j < jm - 1; j++) {
        std::vector<double> aa(ni, -t.num1), bb(ni), cc(ni, -t.num1);
        for (int s = 0; s < ni; s++) {
            const size_t q = (size_t)(s + 1) * jm + j;
            bb[s] = 2.0 * (t.num1 + t.num2[s + 1] + t.num3[q]);
        }
        if (ni > 1) {
            bb[0] += aa[0] * c43;
            cc[0] -= aa[0] * c13;
            bb[ni - 1] += cc[ni - 1] * c43;
            aa[ni - 1] -= cc[ni - 1] * c13;
        }
        double* cp = &t.line_cp[(size_t)j * ni];
        double* dn = &t.line_dn[(size_t)j * ni];
        dn[0] = bb[0];
        cp[0] = cc[0] / bb[0];
        for (int s = 1; s < ni; s++) {
            dn[s] = bb[s] - aa[s] * cp[s - 1];
            cp[s] = cc[s] / dn[s];
        }
    }
    double A2 = 0.0;
    for (int i = 1; i < im - 1; i++) {
        const double rm = geo.rm[i];
        A2 += 1.0 / (rm * rm);
    }
    t.A2 = A2;
    t.abs_sin.assign(J > 0 ? J : 0, 0.0);
    double ys = 0.0, sw = 0.0;
    for (int jj = 0; jj < J; jj++) {
        const double y = (jj == 0 || jj == J - 1) ? 1.5 : 1.0;
        const double s_ = std::abs(std::sin(geo.thez[jj + 1]));
        t.abs_sin[jj] = s_;
        ys += y * s_;
    }
    for (int jj = 0; jj < J; jj++) sw += t.abs_sin[jj];
    t.ys = ys;
    t.sw = sw;
}

// PressureSolverUran::run's weights at PRESS_CONSISTENT (planet/PressureSolverUran.h:40-46, 67-68,
// 146-158 for the source, 180-195 for the sweep). Its sine table is a lone sin() per row, floored.
void build_uran_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, UranTables& u) {
    const int im = g.im, jm = g.jm;
    std::vector<double> sinthe(jm);
    for (int j = 0; j < jm; j++) {
        sinthe[j] = std::sin(geo.thez[j]);
        if (sinthe[j] < p.sinthe_min) sinthe[j] = p.sinthe_min;
    }
    const double dr = g.dr, dthe = g.dthe, dphi = g.dphi;
    const double inv_dr2 = 1.0 / (dr * dr), inv_dthe2 = 1.0 / (dthe * dthe), inv_dphi2 = 1.0 / (dphi * dphi);
    u.a_inv_rm.assign(im, 0); u.a_num2.assign(im, 0); u.m_num2.assign(im, 0);
    const size_t nij = (size_t)im * jm;
    u.a_inv_rmsinthe.assign(nij, 0); u.a_num3.assign(nij, 0); u.a_denom.assign(nij, 0);
    u.m_num3.assign(nij, 0); u.m_inv_denom.assign(nij, 0);
    u.a_num1 = 1.0 / (dr * dr);
    u.m_num1 = inv_dr2;
    u.two_dr = 2.0 * dr;
    u.two_dthe = 2.0 * dthe;
    u.two_dphi = 2.0 * dphi;
    u.src_scale = (p.r_mix * p.u_0 * p.u_0) / 1.0e5;
    for (int i = 0; i < im; i++) {
        const double rm = geo.rm[i];
        // the source's set
        const double inv_rm = 1.0 / rm;
        const double inv_rm2 = inv_rm * inv_rm;
        u.a_inv_rm[i] = inv_rm;
        u.a_num2[i] = inv_rm2 / (dthe * dthe);
        // the sweep's set
        const double rm2 = rm * rm;
        const double inv_rm2_m = 1.0 / rm2;
        u.m_num2[i] = inv_rm2_m * inv_dthe2;
        for (int j = 0; j < jm; j++) {
            const size_t q = (size_t)i * jm + j;
            const double s = sinthe[j];
            u.a_inv_rmsinthe[q] = 1.0 / (rm * s);
            const double inv_rm2sinthe2 = inv_rm2 / (s * s);
            u.a_num3[q] = inv_rm2sinthe2 / (dphi * dphi);
            u.a_denom[q] = 2.0 * (u.a_num1 + u.a_num2[i] + u.a_num3[q]);
            const double inv_rm2sinthe2_m = inv_rm2_m / (s * s);
            const double denom = 2.0 * inv_dr2 + 2.0 * inv_rm2_m * inv_dthe2 + 2.0 * inv_rm2sinthe2_m * inv_dphi2;
            u.m_inv_denom[q] = 1.0 / denom;
            u.m_num3[q] = inv_rm2sinthe2_m * inv_dphi2;
        }
    }
}

void build_step_consts(const atu_grid& g, const atu_params& p, StepConsts& c) {
    c.dt = p.dt;
    c.c_half = 0.5 * p.dt;
    c.c_full = p.dt;
    c.dt6 = 6.0;
    c.coeff_energy_p = p.u_0 * p.u_0 / (p.cp_mix * p.t_ref);
    c.scale_p = 1.0e5 / (p.r_mix * p.u_0 * p.u_0);
    c.scale_Cor = p.coriolis_scale;
    c.re_pr = p.re * p.pr;
    c.re = p.re;
    c.inv_re = 1.0 / p.re;
    c.inv_sig_k = p.sig_k;   // divisor, kept as a division on the device
    c.inv_sig_w = p.sig_w;
    c.sc_re_ch4 = p.sc_ch4 * p.re;
    c.sc_re_h2o = p.sc_h2o * p.re;
    c.sc_re_h2s = p.sc_h2s * p.re;
    c.sc_re_nh3 = p.sc_nh3 * p.re;
    c.sc_re_nh4sh = p.sc_nh4sh * p.re;
    c.precip_num = p.L_atm * 1.0e3;
    c.precip_den = p.r_mix * p.cp_mix * p.u_0 * p.t_ref;
    c.precip_coupling = p.precip_coupling;
    c.chemical_reaction = p.chemical_reaction;
    c.stokes_c = (2.0 / 9.0) * p.r_p_nh4sh * p.r_p_nh4sh;
    c.rho_cond_nh4sh = p.rho_cond_nh4sh;
    c.g = p.g;
    c.mue_mix = p.mue_mix;
    c.u_0 = p.u_0;
    c.turb_coupling = p.turb_coupling;
    c.Pr_t = p.Pr_t;
    c.two_omega = +2.0 * p.omega;
    c.m_two_omega = -2.0 * p.omega;
    c.Coriolis = p.Coriolis;
    c.t_min = p.t_min_K / p.t_ref;
    c.t_max = p.t_max_K / p.t_ref;
    c.tke_max = p.tke_max_m2s2 / (p.u_0 * p.u_0);
    c.dis_min = p.dis_min_nd;
    c.cp_mix = p.cp_mix;
    c.t_ref = p.t_ref;
    c.cp_a = p.r_he * p.cp_he;                          // cp_lapse numerator terms, kept apart
    c.cp_b = p.r_ch4 * p.cp_ch4;                        // so the sum keeps its operand order
    c.cp_den = p.r_h2 + p.r_he + p.r_ch4;
    c.r_h2 = p.r_h2;
    c.cp_mode = p.cp_mode;
    c.cp_gas = p.cp_gas;
    c.turb_active = p.turb_active;
    c.hyd_bnum = p.r_mix * p.R_mix;
    c.nd_rel = p.g * (p.L_atm * 1.0e3) / (p.u_0 * p.u_0);
    c.dr = g.dr;
    c.p_dyn_cap = p.p_dyn_cap;
    c.c43 = 4.0 / 3.0;
    c.c13 = 1.0 / 3.0;
}

}  // namespace atu

// =============================================================================================
// C ABI: defaults, cp table, synthetic state
// =============================================================================================
using namespace atu;

bool atu::resolve_turb_opts(const atu_grid& g, const atu_params& p, const atu_turb_opts& in, atu_turb_opts& o) {
    o = in;
    if (o.model < ATU_TURB_K_OMEGA_SST || o.model > ATU_TURB_K_EPSILON) return false;
    if (o.free_flow < 0 || o.free_flow > 1 || o.abl_taper < 0 || o.abl_taper > 1 || o.curvature < 0 ||
        o.curvature > 1)
        return false;
    // planet/Turbulence.h:150-151 (k_bg), 152-156 (L_bg = L_atm_m / (im-1)), 125-128 (nue_max),
    // 349-352 (u_tau ceiling); planet/cUranusModel.h:364 (abl_height)
    if (!(o.k_bg > 0.0)) o.k_bg = 1.0;
    if (!(o.l_bg > 0.0)) o.l_bg = (p.L_atm * 1.0e3) / double(g.im - 1);
    if (!(o.nue_max > 0.0)) o.nue_max = 1.0e5;
    if (!(o.vel_star_max > 0.0)) o.vel_star_max = std::sqrt(0.002) * 200.0;
    if (!(o.abl_height > 0.0)) o.abl_height = 20000.0;
    return true;
}

// planet/Turbulence.h, with the model's layer heights (planet/cUranusModel.h:1199-1209: FP32
// multiples of L_atm/(im-1), in km) and its accessors height_m (:262-265) and layer_thickness_m
// (planet/cUranusModel.h:372-375). The closure's constants: C_nue 0.028, Karman 0.42, dis_min 1e-10
// (:191-195); vel_star_ref 0.4 m/s, z_0 0.1 m, nue_gas 1.8e-5 m2/s (its constructor's defaults, :104-107).
void atu::build_turb_tables(const atu_grid& g, const atu_params& p, const Geometry& geo, const atu_turb_opts& o,
                            TurbTables& t) {
    const int im = g.im, jm = g.jm;
    const double C_nue = 0.028, Karman = 0.42, dis_min = 1.0e-10;
    const double nue_air = 1.8e-5, vel_star_ref = 0.4, z_0 = 0.1;
    const double free_wall = 1.0e20;
    std::vector<float> lh(im);
    const float h = (float)(p.L_atm / (im - 1));
    for (int i = 0; i < im; i++) lh[i] = (float)i * h;
    auto height_m = [&](int i) { return (double)((0 > i || i > im - 1) ? -1.0f : lh[i]) * 1.0e3; };
    auto layer_thickness_m = [&](int i) { return (i < 0 || i > im - 2) ? 0.0 : (double)(lh[i + 1] - lh[i]) * 1.0e3; };
    const double L_m = p.L_atm * 1.0e3;
    const double y_mount = height_m(0);
    t.rm.assign(im, 0); t.inv_rm.assign(im, 0); t.ystar_sst.assign(im, 0); t.yphys_ke.assign(im, 0);
    t.ystar_ke.assign(im, 0); t.taper.assign(im, 1.0); t.karman_z.assign(im, 0); t.pz.assign(im, 0);
    // clamp_nue's taper (:567-586)
    int i_abl = 0;
    while (i_abl < im - 1 && height_m(i_abl) <= o.abl_height) i_abl++;
    const double taper_dz = std::max(1.0, height_m(std::min(i_abl, im - 1)) - height_m(std::max(i_abl - 1, 0)));
    for (int i = 0; i < im; i++) {
        t.rm[i] = geo.rm[i];
        t.inv_rm[i] = 1.0 / geo.rm[i];
        const double dz_layer = std::max(1.0, layer_thickness_m(std::max(i - 1, 0)));
        const double y_sst = o.free_flow ? free_wall : std::max(height_m(i) - y_mount, dz_layer);   // :1040-1041
        t.ystar_sst[i] = y_sst / L_m;
        const double y_ke = o.free_flow ? free_wall : std::max(height_m(i) - y_mount, L_m);        // :896
        t.yphys_ke[i] = y_ke;
        t.ystar_ke[i] = y_ke / L_m;
        if (o.abl_taper) {
            const double z_agl = height_m(i) - height_m(0);
            const double f = (o.abl_height + taper_dz - z_agl) / taper_dz;
            t.taper[i] = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
        }
        // init_fields (:447-462)
        const double z_i = height_m(i) - y_mount;
        const double z_safe = o.free_flow ? free_wall : std::max(z_i, dz_layer);
        const double z_frac = std::min(z_i, o.abl_height) / o.abl_height;
        t.karman_z[i] = Karman * z_safe;
        t.pz[i] = (1.0 - z_frac) * (1.0 - z_frac);   // std::pow(1 - z_frac, 2), formed as the square by g++
    }
    t.sin_pair.assign(jm, 0); t.cos_pair.assign(jm, 0); t.sin_alone.assign(jm, 0);
    for (int j = 0; j < jm; j++) {
        // compute_sources forms the sin and cos of one angle (:680-682), which its g++ build takes
        // from one sincos call; compute_k_omega forms the sin alone (:944-945)
        double sj, cj;
        sincos(geo.thez[j], &sj, &cj);
        t.sin_pair[j] = (sj == 0.0) ? 1.0e-5 : sj;
        t.cos_pair[j] = cj;
    }
    for (int j = 0; j < jm; j++) {
        double sj = std::sin(geo.thez[j]);
        if (sj == 0.0) sj = 1.0e-5;
        t.sin_alone[j] = sj;
    }
    t.nue_max_nd = o.nue_max / (p.u_0 * L_m);
    t.k_bg_free_nd = o.k_bg / (p.u_0 * p.u_0);
    t.kbg_den = std::sqrt(C_nue) * p.u_0 * p.u_0;
    t.nue_air_nd = nue_air / (p.u_0 * L_m);
    t.nue_air = nue_air;
    t.u_0 = p.u_0;
    t.L_m = L_m;
    t.re_turb = vel_star_ref * z_0 / nue_air;
    t.sqrt_bet_star = std::sqrt(0.09);
    t.sqrt_cd = std::sqrt(0.002);
    t.vel_star_max = o.vel_star_max;
    t.inv_u02 = 1.0 / (p.u_0 * p.u_0);
    t.nd_omega = L_m / p.u_0;
    t.nd_eps = L_m / (p.u_0 * p.u_0 * p.u_0);
    t.c_m025 = std::pow(C_nue, -0.25);
    t.c_075 = std::pow(C_nue, 0.75);
    t.sqrt_cnue = std::sqrt(C_nue);
    t.karman_l = Karman * L_m;   // init_fields' Karman * L_bg without TURB_FREE
    // With TURB_FREE every interior cell of init_fields takes one value (:440-462)
    {
        const double k_bg = o.k_bg, L_bg = o.l_bg;
        const double om_bg = std::pow(C_nue, -0.25) * std::sqrt(k_bg) / (Karman * L_bg);
        const double eps_bg = std::pow(C_nue, 0.75) * std::pow(k_bg, 1.5) / (Karman * L_bg);
        const double k_phys = std::max(0.0, k_bg);
        double dis_nd;
        if (o.model == ATU_TURB_K_EPSILON) {
            const double eps_phys = std::pow(C_nue, 0.75) * std::pow(k_phys, 1.5) / (Karman * free_wall);
            dis_nd = std::max(eps_phys, eps_bg) * t.nd_eps;
        } else {
            const double om_phys = std::pow(C_nue, -0.25) * std::sqrt(k_phys) / (Karman * free_wall);
            dis_nd = std::max(om_phys, om_bg) * t.nd_omega;
        }
        t.init_tke = std::max(0.0, k_phys * t.inv_u02);
        t.init_dis = std::max(dis_min, dis_nd);
    }
}

extern "C" int atu_turb_opts_default(atu_turb_opts* o) {
    if (!o) return ATU_EINVAL;
    std::memset(o, 0, sizeof *o);
    o->model = ATU_TURB_K_OMEGA_SST;
    o->free_flow = 1;
    o->abl_taper = 0;
    o->curvature = 1;
    o->k_bg = 1.0;
    o->l_bg = 0.0;   // L_atm * 1e3 / (im-1) of the context's grid
    o->nue_max = 1.0e5;
    o->vel_star_max = std::sqrt(0.002) * 200.0;
    o->abl_height = 20000.0;
    return ATU_OK;
}

// The gas constants the reference loop passes to its four SaturationAdjustmentUran calls
// (planet/cUranusModel.cpp:998-1029 with planet/cUranusModel.h:880-906, 1094-1192) and the ice sets
// its dispatch picks (planet/SaturationAdjustmentUran.cpp:222-231; H2S takes its liquid set), with
// every knob unset (ATURAN_T00_LEGACY 0, ATURAN_LV_H2S unset, ATURAN_SVP_FIT 0).
extern "C" int atu_satadj_opts_default(atu_satadj_opts* o) {
    if (!o) return ATU_EINVAL;
    std::memset(o, 0, sizeof *o);
    o->gas[ATU_GAS_H2O] = atu_gas{273.15, 210.15, 8.1253, 2.5009e6, 2.8339e6, 25.096, 3148.2, 461.52, 0.0, -8.7e-3,
                                  28.418, 3567.3, 0.0, -8.7e-3};
    o->gas[ATU_GAS_H2S] = atu_gas{187.65, 140.0, 14.5192, 5.5e5, 7.4500e5, 17.064, 747.0, 243.96, 0.0, -2.9e-3,
                                  17.064, 747.0, 0.0, -2.9e-3};
    o->gas[ATU_GAS_NH3] = atu_gas{195.5, 140.0, 7.6752, 1.3720e6, 1.8320e6, 27.863, 2016.0, 488.21, -0.888, 0.0,
                                  34.948, 2692.5, -0.888, 0.0};
    o->gas[ATU_GAS_CH4] = atu_gas{90.69, 67.36, 7.6752, 5.11e5, 5.11e5, 1.627, 553.1, 518.28, 1.002, -4.1e-3,
                                  1.627, 553.1, 1.002, -4.1e-3};
    o->m_nh3 = 17.03052;
    o->m_h2s = 34.08088;
    o->m_nh4sh = 51.1114;
    o->gas_mask = 15;
    o->nh4sh_eq = 1;      // ATURAN_NH4SH_EQ, planet/ChemistryUran.h:134-137
    // planet/cUranusModel.h:578-589 and planet/SaturationAdjustment.h:286-287, 300-301
    o->conserve = 1;
    o->newton = 1;
    o->evap = 1;
    o->ice_entry = 0;
    o->bergeron = 2;
    o->freeze = 1;
    o->dry_exit = 1;
    return ATU_OK;
}

extern "C" int atu_grid_default(int refine, atu_grid* out) {
    if (!out || (refine != 1 && refine != 2)) return ATU_EINVAL;
    std::memset(out, 0, sizeof(*out));
    // planet/cUranusModel.cpp:339-351: pi180 = 180/pi, dthe = the_degree / pi180
    const double pi180 = 180.0 / M_PI;
    if (refine == 1) {
        out->im = 41; out->jm = 181; out->km = 361;
        out->dr = 0.025;
        out->dthe = 1.0 / pi180;
        out->dphi = 1.0 / pi180;
    } else {
        out->im = 81; out->jm = 361; out->km = 721;
        out->dr = 0.0125;
        out->dthe = 0.5 / pi180;
        out->dphi = 0.5 / pi180;
    }
    out->r0 = 1.0;
    out->the0 = 0.0;
    out->phi0 = 0.0;
    return ATU_OK;
}

extern "C" int atu_params_default(const atu_grid* g, atu_params* p) {
    if (!g || !p) return ATU_EINVAL;
    std::memset(p, 0, sizeof(*p));
    p->dt = 0.001;                                      // planet/cUranusModel.cpp:475
    p->re = 1000.0; p->pr = 0.69; p->g = 8.69; p->omega = 1.01e-4;   // param.py:59-88
    p->u_0 = 100.0; p->t_ref = 76.4; p->L_atm = 225.0;
    p->Coriolis = 1.0; p->centrifugal = 1.0; p->buoyancy = 1.0; p->chemical_reaction = 1.0;
    p->sc_ch4 = 0.99; p->sc_h2o = 0.61; p->sc_h2s = 0.94; p->sc_nh3 = 0.61; p->sc_nh4sh = 0.7;  // cUranusModel.h:1120-1124
    p->alpha_sponge = 10.0;                             // cUranusModel.h:1127
    p->rho_cond_nh4sh = 1170.0; p->r_p_nh4sh = 1.0e-5;  // cUranusModel.h:987,1022

    // Reference densities and species properties (planet/cUranusModel.h:1004-1089), then the
    // mass-weighted mixture of ThermalPropertiesUran (planet/ChemistryUran.h:437-490).
    double r_h2 = 0.864, r_he = 0.136, r_ch4 = 0.19, r_nh3 = 0.1, r_h2s = 0.3, r_h2o = 0.08, r_nh4sh = 0.007;
    const double R_h2 = 4124.2, R_he = 2077.1, R_nh3 = 488.21, R_nh4sh = 261.0, R_h2s = 243.96,
                 R_h2o = 461.52, R_ch4 = 518.28;
    const double mue_h2 = 0.84e-5, mue_he = 1.87e-5, mue_nh3 = 0.92e-5, mue_nh4sh = 0.99e-5,
                 mue_h2s = 1.3e-5, mue_h2o = 0.9e-5, mue_ch4 = 1.107e-5;
    const double cp_h2 = 14.32e3, cp_he = 5.19e3, cp_nh3 = 2.19e3, cp_nh4sh = 2.00e3, cp_h2s = 2.24e3,
                 cp_h2o = 1.93e3, cp_ch4 = 2.232e3;
    double r_mix = r_h2 + r_he + r_h2s + r_nh3 + r_nh4sh + r_h2o + r_ch4;
    const double cp_mix = (r_h2 * cp_h2 + r_h2s * cp_h2s + r_he * cp_he + r_nh3 * cp_nh3
                           + r_nh4sh * cp_nh4sh + r_h2o * cp_h2o + r_ch4 * cp_ch4) / r_mix;
    const double mue_mix = (r_h2 * mue_h2 + r_h2s * mue_h2s + r_he * mue_he + r_nh3 * mue_nh3
                            + r_nh4sh * mue_nh4sh + r_h2o * mue_h2o + r_ch4 * mue_ch4) / r_mix;
    const double R_mix = (r_h2 * R_h2 + r_he * R_he + r_h2o * R_h2o + r_h2s * R_h2s + r_nh3 * R_nh3
                          + r_nh4sh * R_nh4sh + r_ch4 * R_ch4)
                         / (r_h2 + r_he + r_h2o + r_h2s + r_nh3 + r_nh4sh + r_ch4);
    p->cp_h2 = cp_h2; p->cp_he = cp_he; p->cp_ch4 = cp_ch4;
    p->cp_mix = cp_mix; p->mue_mix = mue_mix; p->R_mix = R_mix;
    p->cp_mode = 2; p->cp_gas = 1;                      // CP_H2 2, CP_GAS 1 (cUranusModel.h:596-597)
    p->r_h2 = r_h2; p->r_he = r_he; p->r_ch4 = r_ch4; p->r_mix = r_mix;

    // T_INIT_ANCHOR (cUranusModel.cpp:504-511): the bottom temperature of the cp(T) adiabat
    // through 76 K at 1 bar (cUranusModel.h:149-156), then one factor on every density.
    const double R_ref = 3.61491, p_bottom = 17.3, t_1bar = 76.0;
    {
        const double Rk = R_ref * 1.0e3;
        double T = t_1bar;
        const int n = 2000;
        const double h = std::log(p_bottom) / n;
        for (int s = 0; s < n; s++) {
            const double k1 = T * Rk / cp_lapse(*p, T), Tm = T + 0.5 * h * k1;
            T += h * Tm * Rk / cp_lapse(*p, Tm);
        }
        const double f = p_bottom * 1.0e5 / (R_mix * T) / r_mix;
        r_h2 *= f; r_he *= f; r_ch4 *= f; r_nh3 *= f; r_h2s *= f; r_h2o *= f; r_nh4sh *= f;
        r_mix *= f;
    }
    p->r_h2 = r_h2; p->r_he = r_he; p->r_ch4 = r_ch4; p->r_mix = r_mix;

    p->R_metric_km = 25362.0;                           // cUranusModel.h:638
    p->sinthe_min = 0.4;
    p->t_min_K = 7.5; p->t_max_K = 2000.0;
    p->tke_max_m2s2 = 1000.0; p->dis_min_nd = 1.0e-10;
    p->turb_coupling = 1.0; p->Pr_t = 0.9; p->sig_k = 0.85; p->sig_w = 0.5;
    p->precip_coupling = 1.0;
    p->p_dyn_cap = 0.2;
    p->coriolis_scale = p->L_atm * 1.0e3 / p->u_0;      // CORIOLIS_FULL, cUranusModel.h:1498
    p->turb_active = 1;                                 // turb_model k_omega_SST (cUranusModel.cpp:482)
    p->lid_vw = 2;                                      // BC_LID_VW with GEO_BALANCE (BC_Uran.h:155)
    p->rigid_lid = 1;
    p->radius_positive = 1;                             // BC_Uran.h:70
    p->species_clamp = 1;                               // ATURAN_SPECIES_CLAMP, on by default (BC_Uran.h:111)
    (void)g;
    return ATU_OK;
}

extern "C" int atu_cp_table(int mode, double* out, int* n) {
    if ((mode != 1 && mode != 2) || !n) return ATU_EINVAL;
    const std::vector<double>& t = cp_h2_table(mode);
    *n = (int)t.size();
    if (out) std::memcpy(out, t.data(), t.size() * sizeof(double));
    return ATU_OK;
}

extern "C" double atu_cp_lapse(const atu_params* p, double T_K) {
    if (!p) return NAN;
    return cp_lapse(*p, T_K);
}

// ---------------------------------------------------------------------------------------------
// Synthetic state. SplitMix64 (Steele, Lea & Flood 2014) gives 64-bit words; the top 53 bits make
// a uniform in (0, 1]; Box-Muller turns two uniforms into one normal. Each field has its own
// stream (seed mixed with the field id) drawn in (i, j, k) order over the full array.
// ---------------------------------------------------------------------------------------------
namespace {
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }   // (0, 1]
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};
SplitMix64 stream(uint64_t seed, int field) {
    SplitMix64 r{seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(field + 1))};
    for (int w = 0; w < 4; w++) r.next();
    return r;
}
}  // namespace

extern "C" int atu_synthetic_state(const atu_grid* g, const atu_params* p, uint64_t seed, double* fields) {
    if (!g || !p || !fields || g->im < 4 || g->jm < 4 || g->km < 4) return ATU_EINVAL;
    const int im = g->im, jm = g->jm, km = g->km;
    const size_t n3 = (size_t)im * jm * km;
    std::memset(fields, 0, n3 * ATU_NFIELDS * sizeof(double));
    auto F = [&](int f) { return fields + (size_t)f * n3; };
    auto at = [&](int i, int j, int k) { return ((size_t)i * jm + j) * km + k; };

    Geometry geo;
    build_geometry(*g, *p, geo);
    const double t_eq = 60.65, t_pole = 49.15, lapse = 0.7, R_ref = 3.61491, p_bottom = 17.3;
    const double deg_per_row = (g->dthe * 180.0 / M_PI);

    // Height of level i in km: L_atm (rad.z - r0) / (rad.z[im-1] - r0) -- 5.625 i on the reference grid.
    std::vector<double> z_km(im);
    for (int i = 0; i < im; i++) z_km[i] = p->L_atm * (geo.radz[i] - g->r0) / (geo.radz[im - 1] - g->r0);

    // Temperature: the top profile T_top(j) = (t_pole - t_eq)(d^2/90^2 - 2d/90) + t_pole with d the
    // colatitude in degrees (planet/InitVariables_Uran.cpp:35-36), warming downward at 0.7 K/km.
    std::vector<double> T_top(jm);
    for (int j = 0; j < jm; j++) {
        const double d = j * deg_per_row;
        T_top[j] = (t_pole - t_eq) * (d * d / (90.0 * 90.0) - 2.0 * d / 90.0) + t_pole;
    }
    {
        SplitMix64 r = stream(seed, ATU_T);
        for (int i = 0; i < im; i++)
            for (int j = 0; j < jm; j++)
                for (int k = 0; k < km; k++)
                    F(ATU_T)[at(i, j, k)] = (T_top[j] + lapse * (p->L_atm - z_km[i])) / p->t_ref
                                           + (0.05 / p->t_ref) * r.normal();
    }
    // Zonal wind w = cos(3 lat) (1 + 0.1 N); v and the radial u ~ 0.001 N (0.1 m/s over u_0).
    {
        SplitMix64 r = stream(seed, ATU_W);
        for (int i = 0; i < im; i++)
            for (int j = 0; j < jm; j++) {
                const double lat = 0.5 * M_PI - geo.thez[j];
                const double c3 = std::cos(3.0 * lat);
                for (int k = 0; k < km; k++) F(ATU_W)[at(i, j, k)] = c3 * (1.0 + 0.1 * r.normal());
            }
    }
    {
        SplitMix64 r = stream(seed, ATU_V);
        for (size_t q = 0; q < n3; q++) F(ATU_V)[q] = 0.001 * r.normal();
    }
    {
        SplitMix64 r = stream(seed, ATU_U);
        for (size_t q = 0; q < n3; q++) F(ATU_U)[q] = 0.001 * r.normal();
    }
    // p_stat: exact hydrostatic pressure of the noise-free linear-T column from 17.3 bar at the
    // bottom, p *= (T1/T0)^(g dz / (R dT)) (planet/cUranusModel.h:171-180).
    for (int j = 0; j < jm; j++) {
        std::vector<double> col(im);
        double pr = p_bottom;
        col[0] = pr;
        for (int i = 1; i < im; i++) {
            const double T0 = T_top[j] + lapse * (p->L_atm - z_km[i - 1]);
            const double T1 = T_top[j] + lapse * (p->L_atm - z_km[i]);
            const double a = p->g * (z_km[i] - z_km[i - 1]) / R_ref;
            const double dT = T0 - T1;
            pr *= (std::fabs(dT) > 1.0e-9 * T0) ? std::pow(T1 / T0, a / dT) : std::exp(-a / T0);
            col[i] = pr;
        }
        for (int i = 0; i < im; i++)
            for (int k = 0; k < km; k++) F(ATU_P_STAT)[at(i, j, k)] = col[i];
    }
    for (size_t q = 0; q < n3; q++)
        F(ATU_RHO_MIX)[q] = F(ATU_P_STAT)[q] * 1.0e5 / (p->R_mix * F(ATU_T)[q] * p->t_ref);
    // CH4 = 0.02 exp(-z / 40 km); every other species 0.
    for (int i = 0; i < im; i++) {
        const double c = 0.02 * std::exp(-z_km[i] / 40.0);
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) F(ATU_CH4)[at(i, j, k)] = c;
    }
    // Turbulence: k = 1 m^2/s^2 background, omega_bg of a 5625 m length (planet/Turbulence.h:441-443),
    // and the seeded eddy viscosity nue* = 1.65e-4 (planet/RHS_Uran_Turb.cpp:433).
    const double C_nue = 0.09, karman = 0.41, L_bg = 5625.0, k_bg = 1.0;
    const double om_bg = std::pow(C_nue, -0.25) * std::sqrt(k_bg) / (karman * L_bg);
    const double tke0 = k_bg / (p->u_0 * p->u_0);
    const double dis0 = om_bg * (p->L_atm * 1.0e3) / p->u_0;
    for (size_t q = 0; q < n3; q++) {
        F(ATU_TKE)[q] = tke0;
        F(ATU_DIS)[q] = dis0;
        F(ATU_NUE)[q] = 1.65e-4;
    }
    return ATU_OK;
}
