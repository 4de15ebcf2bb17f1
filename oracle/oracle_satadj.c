/*
 * oracle_satadj.c -- TEST INFRASTRUCTURE ONLY. The plain-C FP64 restatement of the reference
 * loop's saturation-adjustment block (planet/cUranusModel.cpp:997-1033): the four
 * SaturationAdjustmentUran calls, each dispatching to the shared SaturationAdjustment<cUranusModel>
 * (planet/SaturationAdjustmentUran.cpp:199-239, planet/SaturationAdjustment.h:256-746), then
 * ChemistryUran::EquilibriumNH4SH("loop") (planet/ChemistryUran.h:139-255). The checker of the
 * library's atu_saturation_adjust; nothing in the product links or calls it.
 *
 * The model's answers: is_solid false, satadj_updates_pstat false (planet/cUranusModel.h:562),
 * rho_at = rho_mix where positive and finite, else r_mix (ATURAN_LOCAL_RHO default 1, :521-531),
 * no SeaMount cell. The knobs of the shared routine are arguments (their Uranus defaults:
 * CONSERVE 1, NEWTON 1, EVAP 1, ICE_ENTRY 0, BERGERON 2, FREEZE 1, DRY_EXIT 1); SATADJ_DIAG only
 * prints and is left out.
 *
 * Dense layout as oracle.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/atu.h"

#define IDX(i, j, k) (((size_t)(i) * jm + (j)) * km + (k))

static double smax(double a, double b) { return (a < b) ? b : a; }   /* std::max */
static double smin(double a, double b) { return (b < a) ? b : a; }   /* std::min */

/* ATPhys::saturation_vapour_pressure, planet/ATPhys.h:49-52 */
static double svp(double T, double C, double L0, double R, double del_alf, double del_bet) {
    return exp(C + (-L0 / T + del_alf * log(T) + del_bet * T) / (1e-3 * R));
}

typedef struct {
    int conserve, newton, evap, ice_entry, bergeron, freeze, dry_exit;
} knobs;

/* SaturationAdjustment<Planet>::run, planet/SaturationAdjustment.h:256-668, for one gas */
static void satadj_gas(const atu_grid* g, const atu_params* p, const knobs* kn, const atu_gas* gs, double* t,
                       const double* p_stat, const double* rho_mix, double* c, double* cloud, double* ice) {
    const int im = g->im, jm = g->jm, km = g->km;
    const double t_0 = gs->t_0, t_00 = gs->t_00, ep = gs->ep, lv = gs->lv, ls = gs->ls;
    const double C = gs->C, L0 = gs->L0, R = gs->R, del_alf = gs->del_alf, del_bet = gs->del_bet;
    const double C_i = gs->C_i, L0_i = gs->L0_i, del_alf_i = gs->del_alf_i, del_bet_i = gs->del_bet_i;
    const double t_range_inv = 1.0 / (t_0 - t_00);
    const int conserve = kn->conserve, newton = kn->newton, evap = kn->evap, ice_entry = kn->ice_entry;
    const int bergeron = kn->bergeron, dry_exit = kn->dry_exit;
    const int freeze_on = kn->freeze && conserve && t_00 > 0.0 && t_00 < t_0;
    const double Lf = ls - lv;
    const int iter_prec_end = 15;
    const double q_diff_min = 1.0e-3;
#define CND_OF(TT) ((TT) <= t_00 ? 0.0 : ((TT) >= t_0 ? 1.0 : ((TT) - t_00) * t_range_inv))
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < km; k++)
        for (int j = 0; j < jm; j++)
            for (int i = 0; i < im; i++) {
                const size_t q = IDX(i, j, k);
                double t_u = t[q] * p->t_ref;
                const double p_u = p_stat[q];
                if (!(t_u > 0.0) || !(p_u > 0.0)) continue;
                double rho_c;
                {
                    const double rho = rho_mix[q];
                    rho_c = (rho > 0.0 && isfinite(rho)) ? rho : p->r_mix;
                }
                if (t_u > t_0) {
                    if (conserve) {
                        if (ice[q] > 0.0) {
                            double dm = ice[q];
                            if (Lf > 0.0) {
                                const double hc = p->cp_mix * rho_c;
                                const double cap = hc * (t_u - t_0) / Lf;
                                if (dm > cap) dm = cap;
                                t_u -= Lf * dm / hc;
                                t[q] = t_u / p->t_ref;
                            }
                            ice[q] -= dm;
                            cloud[q] += dm;
                        }
                    } else {
                        ice[q] = 0.0;
                    }
                }
                if (freeze_on && t_u < t_00 && cloud[q] > 0.0) {
                    double dm = cloud[q];
                    if (Lf > 0.0) {
                        const double hc = p->cp_mix * rho_c;
                        const double cap = hc * (t_00 - t_u) / Lf;
                        if (dm > cap) dm = cap;
                        t_u += Lf * dm / hc;
                        t[q] = t_u / p->t_ref;
                    }
                    cloud[q] -= dm;
                    ice[q] += dm;
                }
                if (c[q] < 0.0) c[q] = 0.0;
                if (cloud[q] < 0.0) cloud[q] = 0.0;
                if (ice[q] < 0.0) ice[q] = 0.0;

                const double E_Rain_0 = svp(t_u, C, L0, R, del_alf, del_bet);
                const double q_Rain_0 = rho_c * ep * E_Rain_0 / p_u;
                if (q_Rain_0 <= 0.0) continue;
                double q_entry = q_Rain_0;
                if (ice_entry && cloud[q] <= 0.0 && ice[q] <= 0.0) {
                    const double cnd0 = CND_OF(t_u);
                    if (cnd0 < 1.0)
                        q_entry = cnd0 * q_Rain_0
                                  + (1.0 - cnd0) * rho_c * ep * svp(t_u, C_i, L0_i, R, del_alf_i, del_bet_i) / p_u;
                }
                if (bergeron >= 2 && cloud[q] <= 0.0 && ice[q] <= 0.0 && t_u < t_0)
                    q_entry = rho_c * ep * svp(t_u, C_i, L0_i, R, del_alf_i, del_bet_i) / p_u;
                if (c[q] <= q_entry) {
                    const double q_c0 = cloud[q], q_i0 = ice[q];
                    if (!evap || !conserve || (q_c0 <= 0.0 && q_i0 <= 0.0)) continue;
                    const double q_Ice_0 = rho_c * ep * svp(t_u, C_i, L0_i, R, del_alf_i, del_bet_i) / p_u;
                    double q_t0;
                    if (q_c0 > 0.0 && q_i0 > 0.0) q_t0 = (q_c0 * q_Rain_0 + q_i0 * q_Ice_0) / (q_c0 + q_i0);
                    else if (q_i0 <= 0.0) q_t0 = q_Rain_0;
                    else q_t0 = q_Ice_0;
                    if (fabs(c[q] - q_t0) <= q_diff_min * q_t0) continue;
                }

                double q_v_b = c[q], q_c_b = cloud[q], q_i_b = ice[q];
                double T = t_u;
                double q_v_hyp = q_v_b;
                double q_rain_now = q_Rain_0;
                for (int itr = 1; itr <= iter_prec_end; itr++) {
                    double CND = (T - t_00) * t_range_inv;
                    double DEP = (t_0 - T) * t_range_inv;
                    if (T <= t_00) { CND = 0.0; DEP = 1.0; }
                    if (T >= t_0) { CND = 1.0; DEP = 0.0; }
                    const int berg = bergeron != 0 && T < t_0 && q_v_b <= q_rain_now
                                     && (q_i_b > 0.0 || (bergeron >= 2 && q_c_b <= 0.0));
                    if (berg) {
                        CND = 0.0; DEP = 1.0;
                        if (q_c_b > 0.0) {
                            T -= lv * q_c_b / (p->cp_mix * rho_c);
                            q_v_b += q_c_b;
                            q_c_b = 0.0;
                        }
                    }
                    double d_q_v = q_v_hyp - q_v_b;
                    double d_q_c = -d_q_v * CND;
                    double d_q_i = -d_q_v * DEP;
                    if (conserve) {
                        if (d_q_v < -q_v_b) { d_q_v = -q_v_b; d_q_c = -d_q_v * CND; d_q_i = -d_q_v * DEP; }
                        if (q_c_b + d_q_c < 0.0) { const double ex = -(q_c_b + d_q_c); d_q_c = -q_c_b; d_q_i -= ex; }
                        if (q_i_b + d_q_i < 0.0) {
                            const double ex = -(q_i_b + d_q_i);
                            d_q_i = -q_i_b;
                            d_q_c -= ex;
                            if (q_c_b + d_q_c < 0.0) d_q_c = -q_c_b;
                        }
                        d_q_v = -(d_q_c + d_q_i);
                    }
                    T += (lv * d_q_c + ls * d_q_i) / (p->cp_mix * rho_c);
                    q_v_b += d_q_v;
                    q_c_b += d_q_c;
                    q_i_b += d_q_i;
                    if (q_v_b < 0.0) q_v_b = 0.0;
                    if (q_c_b < 0.0) q_c_b = 0.0;
                    if (q_i_b < 0.0) q_i_b = 0.0;

                    const double E_Rain = svp(T, C, L0, R, del_alf, del_bet);
                    const double E_Ice = svp(T, C_i, L0_i, R, del_alf_i, del_bet_i);
                    const double q_Rain = rho_c * ep * E_Rain / p_u;
                    const double q_Ice = rho_c * ep * E_Ice / p_u;
                    q_rain_now = q_Rain;
                    if (q_c_b > 0.0 && q_i_b > 0.0) q_v_hyp = (q_c_b * q_Rain + q_i_b * q_Ice) / (q_c_b + q_i_b);
                    else if (bergeron >= 2 && q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_v_hyp = q_Ice;
                    else if (ice_entry && q_c_b <= 0.0 && q_i_b <= 0.0) {
                        const double cn = CND_OF(T);
                        q_v_hyp = cn * q_Rain + (1.0 - cn) * q_Ice;
                    } else if (q_i_b == 0.0) q_v_hyp = q_Rain;
                    else q_v_hyp = q_Ice;

                    if (T >= t_0) {
                        if (conserve) {
                            if (q_i_b > 0.0) {
                                double dm = q_i_b;
                                if (Lf > 0.0) {
                                    const double hc = p->cp_mix * rho_c;
                                    const double cap = hc * (T - t_0) / Lf;
                                    if (dm > cap) dm = cap;
                                    T -= Lf * dm / hc;
                                }
                                q_i_b -= dm;
                                q_c_b += dm;
                            }
                        } else {
                            q_i_b = 0.0;
                        }
                    }
                    const double q_diff = fabs(q_v_b - q_v_hyp) / (q_v_hyp + 1e-20);
                    if (q_diff <= q_diff_min) break;
                    if (dry_exit && conserve && q_c_b <= 0.0 && q_i_b <= 0.0 && q_v_b < q_v_hyp) break;
                    if (newton) {
                        const double dT = 0.1;
                        const double q_Rain_d = rho_c * ep * svp(T + dT, C, L0, R, del_alf, del_bet) / p_u;
                        const double q_Ice_d = rho_c * ep * svp(T + dT, C_i, L0_i, R, del_alf_i, del_bet_i) / p_u;
                        double q_hyp_d;
                        if (q_c_b > 0.0 && q_i_b > 0.0) q_hyp_d = (q_c_b * q_Rain_d + q_i_b * q_Ice_d) / (q_c_b + q_i_b);
                        else if (bergeron >= 2 && q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_hyp_d = q_Ice_d;
                        else if (ice_entry && q_c_b <= 0.0 && q_i_b <= 0.0) {
                            const double cn = CND_OF(T);
                            q_hyp_d = cn * q_Rain_d + (1.0 - cn) * q_Ice_d;
                        } else if (q_i_b == 0.0) q_hyp_d = q_Rain_d;
                        else q_hyp_d = q_Ice_d;
                        double CNDn = (T - t_00) * t_range_inv;
                        if (T <= t_00) CNDn = 0.0;
                        if (T >= t_0) CNDn = 1.0;
                        if (berg) CNDn = 0.0;
                        const double G = (lv * CNDn + ls * (1.0 - CNDn)) / (p->cp_mix * rho_c) * (q_hyp_d - q_v_hyp) / dT;
                        double w = 1.0 / (1.0 + smax(G, 0.0));
                        if (w < 0.05) w = 0.05;
                        q_v_hyp = q_v_b + w * (q_v_hyp - q_v_b);
                    } else {
                        q_v_hyp = 0.5 * (q_v_hyp + q_v_b);
                    }
                }
                c[q] = q_v_b;
                cloud[q] = q_c_b;
                ice[q] = q_i_b;
                t[q] = T / p->t_ref;
            }
#undef CND_OF
    /* the clamp of :635-647 */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < (size_t)im * jm * km; q++) {
        if (c[q] <= 0.0) c[q] = 0.0;
        if (cloud[q] <= 0.0) cloud[q] = 0.0;
        if (ice[q] <= 0.0) ice[q] = 0.0;
    }
}

/* ChemistryUran::equilibrateNH4SHCell and EquilibriumNH4SH("loop"), planet/ChemistryUran.h:139-255 */
static double kp_nh4sh(double T) {
    const double atm = 101325.0;
    return pow(10.0, 14.82 - 4705.0 / T) * atm * atm;
}

static void nh4sh_loop(const atu_grid* g, const atu_params* p, double m_nh3, double m_h2s, double m_nh4sh,
                       double* t, const double* p_stat, double* nh3, double* h2s, double* nh4sh) {
    const int im = g->im, jm = g->jm, km = g->km;
    const double R_univ = 8314.46;
    const double L_f = 4705.0 * log(10.0) * 8314.46 / 51.1114;   /* nh4sh_formation_heat, :132 */
#pragma omp parallel for collapse(2) schedule(dynamic, 8)
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++)
            for (int i = 0; i < im; i++) {
                const size_t q = IDX(i, j, k);
                const double T0 = t[q] * p->t_ref;
                if (!(T0 > 0.0)) continue;
                double rho = p_stat[q] * 1e5 / (p->R_mix * T0);
                if (!(rho > 0.0) || !isfinite(rho)) rho = p->r_mix;
                const double a = smax(0.0, nh3[q]) / m_nh3;
                const double b = smax(0.0, h2s[q]) / m_h2s;
                const double s = smax(0.0, nh4sh[q]) / m_nh4sh;
                const double heat = L_f * m_nh4sh / (p->cp_mix * rho);
#define F_OF(x) ((a - (x)) * (b - (x)) * (R_univ * (T0 + heat * (x))) * (R_univ * (T0 + heat * (x))) - kp_nh4sh(T0 + heat * (x)))
                const double f0 = F_OF(0.0);
                double lo, hi;
                if (f0 > 0.0) {
                    lo = 0.0;
                    hi = smin(a, b);
                } else if (s > 0.0) {
                    if (F_OF(-s) <= 0.0) { lo = hi = -s; }
                    else { lo = -s; hi = 0.0; }
                } else continue;
                if (hi > lo)
                    for (int it = 0; it < 80; it++) {
                        const double xm = 0.5 * (lo + hi);
                        if (F_OF(xm) > 0.0) lo = xm; else hi = xm;
                        if (hi - lo <= 1.0e-15 * smax(a, smax(b, s))) break;
                    }
#undef F_OF
                const double x = (f0 > 0.0) ? lo : hi;
                if (x == 0.0) continue;
                nh3[q] = smax(0.0, nh3[q] - m_nh3 * x);
                h2s[q] = smax(0.0, h2s[q] - m_h2s * x);
                nh4sh[q] = smax(0.0, nh4sh[q] + m_nh4sh * x);
                const double dT = heat * x;
                t[q] = (T0 + dT) / p->t_ref;
            }
}

/* The block: the gases of `gas_mask` (bit n = atu_gas_id n) in the loop's order H2O, H2S, NH3, CH4,
 * then the NH4SH equilibrium with `nh4sh_eq`. knob order: conserve, newton, evap, ice_entry,
 * bergeron, freeze, dry_exit. */
int orc_satadj(const atu_grid* g, const atu_params* p, const atu_gas* gases, const int* kn_in, int gas_mask,
               int nh4sh_eq, double m_nh3, double m_h2s, double m_nh4sh, double* fields) {
    const size_t n3 = (size_t)g->im * g->jm * g->km;
    knobs kn = {kn_in[0], kn_in[1], kn_in[2], kn_in[3], kn_in[4], kn_in[5], kn_in[6]};
    double* F = fields;
    const int vap[4] = {ATU_H2O, ATU_H2S, ATU_NH3, ATU_CH4};
    for (int n = 0; n < 4; n++) {
        if (!(gas_mask & (1 << n))) continue;
        satadj_gas(g, p, &kn, &gases[n], F + (size_t)ATU_T * n3, F + (size_t)ATU_P_STAT * n3,
                   F + (size_t)ATU_RHO_MIX * n3, F + (size_t)vap[n] * n3, F + (size_t)(vap[n] + 1) * n3,
                   F + (size_t)(vap[n] + 2) * n3);
    }
    if (nh4sh_eq)
        nh4sh_loop(g, p, m_nh3, m_h2s, m_nh4sh, F + (size_t)ATU_T * n3, F + (size_t)ATU_P_STAT * n3,
                   F + (size_t)ATU_NH3 * n3, F + (size_t)ATU_H2S * n3, F + (size_t)ATU_NH4SH * n3);
    return 0;
}
