/*
 * oracle_turb.c -- TEST INFRASTRUCTURE ONLY. The plain-C FP64 restatement of the turbulence
 * closure Turbulence<cUranusModel> (planet/Turbulence.h, bound by planet/TurbulenceUran.h), the
 * checker of the library's atu_turbulence_init / atu_turbulence_run. Nothing in the product links
 * or calls it.
 *
 * It keeps the reference's pass structure: the clamp pass, the source loop, the clamp-out pass,
 * the wall copy, the eddy-viscosity clamp, the seam copy and the lid fill, each a loop of its own
 * over the reference's index ranges, in its order. The model's answers to the closure's questions
 * are Uranus's (planet/cUranusModel.h): surface_index 0 (:376), is_solid false (:511), no
 * coordinate stretching (:499), bc_turb_active false (:433), phi periodic (:412), the buoyancy
 * term off (turb_default_buoy 0, :411). With no solid cell every Neumann branch and
 * zero_land_cells are identities and are left out.
 *
 *   turb_vel_star      planet/Turbulence.h:336-379   compute_vel_star
 *   turb_init_fields                     :426-495   init_fields
 *   turb_wall                            :382-423   apply_wall_bc (the zero-gradient branch)
 *   turb_sources                         :613-882   compute_sources, TURB_CLAMP_PASS 1
 *   k_omega_sst / k_omega / k_epsilon   :1020-1107, :929-1017, :885-927
 *   turb_clamp_nue                       :564-589   clamp_nue
 *   turb_seam                            :542-561   seam_copy
 *   turb_lid                             :510-536   fill_lid_corners
 *   orc_turb                             :282-332   init() / run()
 *
 * Dense layout as oracle.c: fields[f * n3 + (i * jm + j) * km + k]; prod, tken, disn [im][jm][km];
 * vel_star [jm][km].
 */
#define _GNU_SOURCE   /* sincos */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/atu.h"

#define IDX(i, j, k) (((size_t)(i) * jm + (j)) * km + (k))

static double smax(double a, double b) { return (a < b) ? b : a; }   /* std::max */
static double smin(double a, double b) { return (b < a) ? b : a; }   /* std::min */
static double clamp3(double x, double lo, double hi) { return (x < lo) ? lo : ((x > hi) ? hi : x); }

/* The closure's constants (planet/Turbulence.h:104-107, 191-195) and its model's. */
static const double C_nue = 0.028, Karman = 0.42, dis_min = 1.0e-10;
static const double vel_star_ref = 0.4, z_0 = 0.1, nue_gas = 1.8e-5;

typedef struct {
    const atu_grid* g;
    const atu_params* p;
    atu_turb_opts o;
    int im, jm, km;
    double *t_, *u, *v, *w, *tke, *dis, *nue, *tke_source, *dis_source;
    double *prod, *tken, *disn, *vel_star;
    double *radz, *thez;
    float* lh;           /* m_layer_heights, km (planet/cUranusModel.h:1199-1209) */
    double L_m, u_0, re_turb;
} turb;

static double height_m(const turb* T, int i) {           /* Turbulence.h:262-265 */
    const float h = (0 > i || i > T->im - 1) ? -1.0f : T->lh[i];
    return (double)h * 1.0e3;
}
static double layer_thickness_m(const turb* T, int i) {  /* cUranusModel.h:372-375 */
    if (i < 0 || i > T->im - 2) return 0.0;
    return (double)(T->lh[i + 1] - T->lh[i]) * 1.0e3;
}
static double metric_radius(const turb* T, double r) {   /* cUranusModel.h:641-647 */
    if (!(T->p->R_metric_km > 0.0)) return r;
    return r + (T->p->R_metric_km / T->p->L_atm - 1.0);
}
static double nue_max_nd(const turb* T) { return T->o.nue_max / (T->u_0 * T->L_m); }
static double free_wall_dist(void) { return 1.0e20; }
static int phi_wrap(int kk, int km) {                     /* ATPhys.h:135-139, periodic */
    const int n = km - 1;
    return kk < 0 ? kk + n : (kk >= n ? kk - n : kk);
}

/* ---- compute_vel_star, Turbulence.h:336-379 ---- */
static void turb_vel_star(turb* T) {
    const int jm = T->jm, km = T->km;
    const double C_D = 0.002, vel_star_min = 0.05;
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++) {
            double* vs = &T->vel_star[(size_t)j * km + k];
            if (T->o.free_flow) { *vs = vel_star_min; continue; }
            const int i_surf = (0 + 1 < T->im - 1) ? 0 + 1 : T->im - 1;
            const double v_s = T->v[IDX(i_surf, j, k)];
            const double w_s = T->w[IDX(i_surf, j, k)];
            const double vw2 = v_s * v_s + w_s * w_s;
            if (!isfinite(vw2)) { *vs = vel_star_min; continue; }
            const double U_horiz = T->u_0 * sqrt(vw2);
            *vs = clamp3(sqrt(C_D) * U_horiz, vel_star_min, T->o.vel_star_max);
        }
}

/* ---- seam_copy, Turbulence.h:542-561 ---- */
static void turb_seam(turb* T) {
    const int im = T->im, jm = T->jm, km = T->km, kl = km - 1;
    double* f[] = {T->tke, T->dis, T->tken, T->disn, T->nue, T->prod, T->tke_source, T->dis_source};
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++)
            for (int a = 0; a < 8; a++) f[a][IDX(i, j, kl)] = f[a][IDX(i, j, 0)];
}

/* ---- init_fields, Turbulence.h:426-495 ---- */
static void turb_init_fields(turb* T) {
    const int im = T->im, jm = T->jm, km = T->km;
    const double inv_u02 = 1.0 / (T->u_0 * T->u_0);
    const double L_m = T->L_m;
    const double nd_omega = L_m / T->u_0;
    const double nd_eps = L_m / (T->u_0 * T->u_0 * T->u_0);
    const int k_first = 0;
    for (int j = 1; j < jm - 1; j++)
        for (int k = k_first; k < km - 1; k++) {
            const double vs = T->vel_star[(size_t)j * km + k];
            const double k_bg = T->o.free_flow ? T->o.k_bg : 0.01 * vs * vs / sqrt(C_nue);
            const double L_bg = T->o.free_flow ? T->o.l_bg : L_m;
            const double om_bg = pow(C_nue, -0.25) * sqrt(k_bg) / (Karman * L_bg);
            const double eps_bg = pow(C_nue, 0.75) * pow(k_bg, 1.5) / (Karman * L_bg);
            const double y_mount = height_m(T, 0);
            for (int i = 1; i < im - 1; i++) {
                const double z_i = height_m(T, i) - y_mount;
                const double dz_layer = smax(1.0, layer_thickness_m(T, (i - 1 > 0) ? i - 1 : 0));
                const double z_safe = T->o.free_flow ? free_wall_dist() : smax(z_i, dz_layer);
                const double z_frac = smin(z_i, T->o.abl_height) / T->o.abl_height;
                const double k_abl = T->o.free_flow ? 0.0 : vs * vs / sqrt(C_nue) * pow(1.0 - z_frac, 2);
                const double k_phys = smax(k_abl, k_bg);
                double dis_nd;
                if (T->o.model == ATU_TURB_K_EPSILON) {
                    const double eps_phys = pow(C_nue, 0.75) * pow(k_phys, 1.5) / (Karman * z_safe);
                    dis_nd = smax(eps_phys, eps_bg) * nd_eps;
                } else {
                    const double om_phys = pow(C_nue, -0.25) * sqrt(k_phys) / (Karman * z_safe);
                    dis_nd = smax(om_phys, om_bg) * nd_omega;
                }
                T->tke[IDX(i, j, k)] = smax(0.0, k_phys * inv_u02);
                T->dis[IDX(i, j, k)] = smax(dis_min, dis_nd);
                T->tken[IDX(i, j, k)] = T->tke[IDX(i, j, k)];
                T->disn[IDX(i, j, k)] = T->dis[IDX(i, j, k)];
            }
        }
    turb_seam(T);
}

/* ---- apply_wall_bc, Turbulence.h:382-423: no obstacle, so every column takes the zero-gradient
 * copy at i = 0 ---- */
static void turb_wall(turb* T) {
    const int jm = T->jm, km = T->km;
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++) {
            T->tke[IDX(0, j, k)] = T->tke[IDX(1, j, k)];
            T->dis[IDX(0, j, k)] = T->dis[IDX(1, j, k)];
            T->nue[IDX(0, j, k)] = T->nue[IDX(1, j, k)];
        }
}

/* ---- compute_k_omega_SST, Turbulence.h:1020-1107 ---- */
static void k_omega_sst(turb* T, int i, int j, int k, double y_mount, double rm, double sinthe, double dtkedr,
                        double ddisdr, double Omega) {
    const int jm = T->jm, km = T->km;
    const double a1 = 0.31, bet_star = 0.09, bet1 = 0.0333, bet2 = 0.0368, gam1 = 0.413, gam2 = 0.2,
                 sig_w2 = 1.168;
    const double L_m = T->L_m;
    const double nue_air_nd = nue_gas / (T->u_0 * L_m);
    const double dz_layer = smax(1.0, layer_thickness_m(T, (i - 1 > 0) ? i - 1 : 0));
    const double y_phys = T->o.free_flow ? free_wall_dist() : smax(height_m(T, i) - y_mount, dz_layer);
    const double y_star = y_phys / L_m;
    const double rmsinthe = rm * sinthe;
    const int kw = phi_wrap(k - 1, km), ke = phi_wrap(k + 1, km);
    const double* tke = T->tke;
    const double* dis = T->dis;
    const double tke_jm1 = tke[IDX(i, j - 1, k)], tke_jp1 = tke[IDX(i, j + 1, k)];
    const double dis_jm1 = dis[IDX(i, j - 1, k)], dis_jp1 = dis[IDX(i, j + 1, k)];
    const double tke_km1 = tke[IDX(i, j, kw)], tke_kp1 = tke[IDX(i, j, ke)];
    const double dis_km1 = dis[IDX(i, j, kw)], dis_kp1 = dis[IDX(i, j, ke)];
    const double dtkedr_neu = dtkedr, ddisdr_neu = ddisdr;
    const double dtkedthe_neu = (tke_jp1 - tke_jm1) / (rm * 2.0 * T->g->dthe);
    const double ddisdthe_neu = (dis_jp1 - dis_jm1) / (rm * 2.0 * T->g->dthe);
    const double dtkedphi_neu = (tke_kp1 - tke_km1) / (rmsinthe * 2.0 * T->g->dphi);
    const double ddisdphi_neu = (dis_kp1 - dis_km1) / (rmsinthe * 2.0 * T->g->dphi);
    const double grad_dot = dtkedr_neu * ddisdr_neu + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu;
    const double kc = tke[IDX(i, j, k)], dc = dis[IDX(i, j, k)];
    const double CD_kw = smax(2.0 * sig_w2 / smax(dc, 1.0e-20) * grad_dot, 1.0e-20);
    const double dis_safe = smax(dc, 1.0e-20);
    const double arg1 = smin(smax(sqrt(kc) / (bet_star * dis_safe * y_star),
                                  500.0 * nue_air_nd / (y_star * y_star * dis_safe)),
                             4.0 * sig_w2 * kc / (CD_kw * y_star * y_star));
    const double arg2 = smax(2.0 * sqrt(kc) / (bet_star * dis_safe * y_star),
                             500.0 * nue_air_nd / (y_star * y_star * dis_safe));
    const double F1 = tanh(pow(arg1, 4));
    const double F2 = tanh(pow(arg2, 2));
    T->nue[IDX(i, j, k)] = a1 * kc / smax(a1 * dis_safe, Omega * F2);
    const double tke_safe = smax(kc, 1.0e-20);
    const double P_k = smin(T->prod[IDX(i, j, k)], 20.0 * bet_star * tke_safe * dis_safe);
    const double Y_k = bet_star * tke_safe * dis_safe;
    const double P_w = (F1 * gam1 + (1.0 - F1) * gam2) * P_k * dis_safe / tke_safe;
    const double Y_w = (F1 * bet1 + (1.0 - F1) * bet2) * dis_safe * dis_safe;
    const double D_w = 2.0 * (1.0 - F1) * sig_w2 / dis_safe * grad_dot;
    T->tke_source[IDX(i, j, k)] = P_k - Y_k;
    T->dis_source[IDX(i, j, k)] = P_w - Y_w + D_w;
}

/* ---- compute_k_omega, Turbulence.h:929-1017 ---- */
static void k_omega(turb* T, int i, int j, int k, double dtkedr, double ddisdr, double Omega_mag) {
    const int jm = T->jm, km = T->km;
    const double bet_star = 0.09, gam = 0.52, C_lim = 0.875, bet_0 = 0.0708;
    const double rm = metric_radius(T, T->radz[i]);
    double sinthe = sin(T->thez[j]);
    if (sinthe == 0.0) sinthe = 1.0e-5;
    const double rmsinthe = rm * sinthe;
    const double dr = T->g->dr, dthe = T->g->dthe, dphi = T->g->dphi;
    const int kw = phi_wrap(k - 1, km), ke = phi_wrap(k + 1, km);
    const double *tke = T->tke, *dis = T->dis, *u = T->u, *v = T->v, *w = T->w;
    const double tke_jm1 = tke[IDX(i, j - 1, k)], tke_jp1 = tke[IDX(i, j + 1, k)];
    const double dis_jm1 = dis[IDX(i, j - 1, k)], dis_jp1 = dis[IDX(i, j + 1, k)];
    const double tke_km1 = tke[IDX(i, j, kw)], tke_kp1 = tke[IDX(i, j, ke)];
    const double dis_km1 = dis[IDX(i, j, kw)], dis_kp1 = dis[IDX(i, j, ke)];
    const double dtkedr_neu = dtkedr, ddisdr_neu = ddisdr;
    const double dtkedthe_neu = (tke_jp1 - tke_jm1) / (rm * 2.0 * dthe);
    const double ddisdthe_neu = (dis_jp1 - dis_jm1) / (rm * 2.0 * dthe);
    const double dtkedphi_neu = (tke_kp1 - tke_km1) / (rmsinthe * 2.0 * dphi);
    const double ddisdphi_neu = (dis_kp1 - dis_km1) / (rmsinthe * 2.0 * dphi);
    const double dudr = (u[IDX(i + 1, j, k)] - u[IDX(i - 1, j, k)]) / (2.0 * dr) / (rm + 1.0);
    const double dvdthe = (v[IDX(i, j + 1, k)] - v[IDX(i, j - 1, k)]) / (rm * 2.0 * dthe);
    const double dwdphi = (w[IDX(i, j, ke)] - w[IDX(i, j, kw)]) / (rmsinthe * 2.0 * dphi);
    const double dudthe_u = (u[IDX(i, j + 1, k)] - u[IDX(i, j - 1, k)]) / (rm * 2.0 * dthe);
    const double dudphi_u = (u[IDX(i, j, ke)] - u[IDX(i, j, kw)]) / (rmsinthe * 2.0 * dphi);
    const double dvdr_u = (v[IDX(i + 1, j, k)] - v[IDX(i - 1, j, k)]) / (2.0 * dr) / (rm + 1.0);
    const double dwdr_u = (w[IDX(i + 1, j, k)] - w[IDX(i - 1, j, k)]) / (2.0 * dr) / (rm + 1.0);
    const double dvdphi_u = (v[IDX(i, j, ke)] - v[IDX(i, j, kw)]) / (rmsinthe * 2.0 * dphi);
    const double dwdthe_u = (w[IDX(i, j + 1, k)] - w[IDX(i, j - 1, k)]) / (rm * 2.0 * dthe);
    const double S11 = dudr, S22 = dvdthe, S33 = dwdphi;
    const double S12 = 0.5 * (dudthe_u + dvdr_u);
    const double S13 = 0.5 * (dudphi_u + dwdr_u);
    const double S23 = 0.5 * (dvdphi_u + dwdthe_u);
    const double S_mag = sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
    const double kc = tke[IDX(i, j, k)], dc = dis[IDX(i, j, k)];
    const double w_hat = smax(dc, C_lim * S_mag / sqrt(bet_star));
    double sig_d = dtkedr_neu * ddisdr_neu + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu;
    sig_d = (sig_d <= 0.0) ? 0.0 : 0.125;
    const double chi_w = fabs(Omega_mag * Omega_mag * S_mag / pow(bet_star * smax(dc, 1.0e-20), 3));
    const double f_bet = (1.0 + 85.0 * chi_w) / (1.0 + 100.0 * chi_w);
    const double bet_wc = bet_0 * f_bet;
    T->nue[IDX(i, j, k)] = kc / smax(w_hat, 1.0e-20);
    const double P_k = smin(T->prod[IDX(i, j, k)], 20.0 * bet_star * kc * dc);
    const double Y_k = bet_star * kc * dc;
    const double P_w = gam * dc / smax(kc, 1.0e-20) * P_k;
    const double Y_w = bet_wc * dc * dc;
    const double D_w = sig_d / smax(dc, 1.0e-20)
                       * (dtkedr_neu * ddisdr_neu + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu);
    T->tke_source[IDX(i, j, k)] = P_k - Y_k;
    T->dis_source[IDX(i, j, k)] = P_w - Y_w + D_w;
}

/* ---- compute_k_epsilon, Turbulence.h:885-927 ---- */
static void k_epsilon(turb* T, int i, int j, int k, double y_mount) {
    const int jm = T->jm, km = T->km;
    const double C_eps_1 = 1.35, C_eps_2 = 1.80;
    const double L_m = T->L_m;
    const double y_phys = T->o.free_flow ? free_wall_dist() : smax(height_m(T, i) - y_mount, L_m);
    const double y_star = y_phys / L_m;
    const double kc = T->tke[IDX(i, j, k)], dc = T->dis[IDX(i, j, k)];
    T->nue[IDX(i, j, k)] = C_nue * kc * kc / smax(dc, dis_min);
    const double d_plus = y_phys * T->vel_star[(size_t)j * km + k] / nue_gas;
    const double f_nue = 1.0 - exp(-0.0115 * d_plus);
    const double Re_T = kc * kc * T->u_0 * L_m / smax(dc * nue_gas, dis_min * nue_gas);
    const double f_2 = 1.0 - 0.4 / 1.8 * exp(-Re_T * Re_T / 36.0);
    T->nue[IDX(i, j, k)] *= f_nue;
    const double y_star2 = y_star * y_star;
    const double L_k = -2.0 * kc / y_star2;
    const double L_w = -2.0 * dc / y_star2 * exp(-d_plus / 2.0);
    const double P_k = T->prod[IDX(i, j, k)];
    const double Y_k = dc;
    const double tke_safe = smax(kc, dis_min);
    const double P_w = C_eps_1 * dc / tke_safe * P_k;
    const double Y_w = C_eps_2 * f_2 * dc * dc / tke_safe;
    T->tke_source[IDX(i, j, k)] = P_k - Y_k + L_k / T->re_turb;
    T->dis_source[IDX(i, j, k)] = P_w - Y_w + L_w / T->re_turb;
}

static double k_bg_of(const turb* T, int j, int k) {       /* Turbulence.h:625-632 */
    const double vs = T->vel_star[(size_t)j * T->km + k];
    return T->o.free_flow ? T->o.k_bg / (T->u_0 * T->u_0)
                          : 0.01 * vs * vs / (sqrt(C_nue) * T->u_0 * T->u_0);
}

/* ---- compute_sources, Turbulence.h:613-882 (TURB_CLAMP_PASS 1: the clamps as passes) ---- */
static void turb_sources(turb* T) {
    const int im = T->im, jm = T->jm, km = T->km;
    const int k_first = 0;
    double *tke = T->tke, *dis = T->dis;
    const double *u = T->u, *v = T->v, *w = T->w;
    /* clamp_in, :633-635, as the pass before the loop (:649-655) */
    for (int j = 1; j < jm - 1; j++)
        for (int k = k_first; k < km - 1; k++)
            for (int i = 1; i < im - 1; i++) {
                tke[IDX(i, j, k)] = smax(0.0, tke[IDX(i, j, k)]);
                dis[IDX(i, j, k)] = smax(dis_min, dis[IDX(i, j, k)]);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j < jm - 1; j++)
        for (int k = k_first; k < km - 1; k++) {
            const int kw = phi_wrap(k - 1, km), ke = phi_wrap(k + 1, km);
            const double y_mount = height_m(T, 0);
            for (int i = 1; i < im - 1; i++) {
                const double rm = metric_radius(T, T->radz[i]);
                const double exp_rm = 1.0;   /* coord_stretching false */
                double sinthe = sin(T->thez[j]);
                if (sinthe == 0.0) sinthe = 1.0e-5;
                const double costhe = cos(T->thez[j]);
                const double cotanthe = costhe / sinthe;
                const double inv_rm = 1.0 / rm;
                const double rmsinthe = rm * sinthe;
                const double inv_2dr = 1.0 / (2.0 * T->g->dr);
                const double inv_2dthe = 1.0 / (2.0 * T->g->dthe);
                const double inv_2dphi = 1.0 / (2.0 * T->g->dphi);
                const double inv_rm2dthe = inv_2dthe / rm;
                const double inv_rmsinthe2dphi = inv_2dphi / rmsinthe;

                const double dudr = (u[IDX(i + 1, j, k)] - u[IDX(i - 1, j, k)]) * inv_2dr * exp_rm;
                const double dvdr = (v[IDX(i + 1, j, k)] - v[IDX(i - 1, j, k)]) * inv_2dr * exp_rm;
                const double dwdr = (w[IDX(i + 1, j, k)] - w[IDX(i - 1, j, k)]) * inv_2dr * exp_rm;
                const double dudthe = (u[IDX(i, j + 1, k)] - u[IDX(i, j - 1, k)]) * inv_rm2dthe;
                const double dvdthe = (v[IDX(i, j + 1, k)] - v[IDX(i, j - 1, k)]) * inv_rm2dthe;
                const double dwdthe = (w[IDX(i, j + 1, k)] - w[IDX(i, j - 1, k)]) * inv_rm2dthe;
                const double dudphi = (u[IDX(i, j, ke)] - u[IDX(i, j, kw)]) * inv_rmsinthe2dphi;
                const double dvdphi = (v[IDX(i, j, ke)] - v[IDX(i, j, kw)]) * inv_rmsinthe2dphi;
                const double dwdphi = (w[IDX(i, j, ke)] - w[IDX(i, j, kw)]) * inv_rmsinthe2dphi;

                const double dtkedr = (tke[IDX(i + 1, j, k)] - tke[IDX(i - 1, j, k)]) * inv_2dr * exp_rm;
                const double ddisdr = (dis[IDX(i + 1, j, k)] - dis[IDX(i - 1, j, k)]) * inv_2dr * exp_rm;

                double cnue;
                {
                    const double dis_here = smax(dis[IDX(i, j, k)], dis_min);
                    const double tke_here = smax(tke[IDX(i, j, k)], 0.0);
                    cnue = (T->o.model == ATU_TURB_K_EPSILON) ? C_nue * tke_here * tke_here / dis_here
                                                              : tke_here / dis_here;
                    cnue = smin(cnue, nue_max_nd(T));
                }
                const int curv = T->o.curvature;
                const double u_r = curv ? u[IDX(i, j, k)] * inv_rm : 0.0;
                const double v_r = curv ? v[IDX(i, j, k)] * inv_rm : 0.0;
                const double w_r = curv ? w[IDX(i, j, k)] * inv_rm : 0.0;
                const double v_cot_r = v_r * cotanthe;
                const double w_cot_r = w_r * cotanthe;
                const double g_rr = dudr, g_rt = dudthe - v_r, g_rp = dudphi - w_r;
                const double g_tr = dvdr, g_tt = dvdthe + u_r, g_tp = dvdphi - w_cot_r;
                const double g_pr = dwdr, g_pt = dwdthe, g_pp = dwdphi + u_r + v_cot_r;
                const double der = 0.66667 * (g_rr + g_tt + g_pp);
                const double kc = tke[IDX(i, j, k)];
                T->prod[IDX(i, j, k)] = smax(0.0, (cnue * (2.0 * g_rr - der) - 0.66667 * kc) * g_rr
                                                      + (cnue * (g_rt + g_tr)) * g_rt
                                                      + (cnue * (g_rp + g_pr)) * g_rp
                                                      + (cnue * (2.0 * g_tt - der) - 0.66667 * kc) * g_tt
                                                      + (cnue * (g_tr + g_rt)) * g_tr
                                                      + (cnue * (g_tp + g_pt)) * g_tp
                                                      + (cnue * (2.0 * g_pp - der) - 0.66667 * kc) * g_pp
                                                      + (cnue * (g_pr + g_rp)) * g_pr
                                                      + (cnue * (g_pt + g_tp)) * g_pt);
                const double W12 = g_rt - g_tr, W13 = g_rp - g_pr, W23 = g_tp - g_pt;
                const double Omega = sqrt(W12 * W12 + W13 * W13 + W23 * W23);
                if (T->o.model == ATU_TURB_K_EPSILON)
                    k_epsilon(T, i, j, k, y_mount);
                else if (T->o.model == ATU_TURB_K_OMEGA)
                    k_omega(T, i, j, k, dtkedr, ddisdr, Omega);
                else
                    k_omega_sst(T, i, j, k, y_mount, rm, sinthe, dtkedr, ddisdr, Omega);
            }
        }
    /* clamp_out, :636-647, as the pass after the loop (:870-878) */
    const double nmax = nue_max_nd(T);
    for (int j = 1; j < jm - 1; j++)
        for (int k = k_first; k < km - 1; k++) {
            const double k_bg_nd = k_bg_of(T, j, k);
            for (int i = 1; i < im - 1; i++) {
                const size_t q = IDX(i, j, k);
                tke[q] = smax(k_bg_nd, tke[q]);
                dis[q] = smax(dis_min, dis[q]);
                T->nue[q] = clamp3(T->nue[q], 0.0, nmax);
                const double dis_src_max = 20.0 * dis[q] * dis[q];
                T->dis_source[q] = clamp3(T->dis_source[q], -dis_src_max, dis_src_max);
                const double tke_src_max = 20.0 * tke[q] * dis[q];
                T->tke_source[q] = clamp3(T->tke_source[q], -tke_src_max, tke_src_max);
            }
        }
}

/* ---- clamp_nue, Turbulence.h:564-589 ---- */
static void turb_clamp_nue(turb* T) {
    const int im = T->im, jm = T->jm, km = T->km;
    const double nmax = nue_max_nd(T);
    const int taper_on = T->o.abl_taper != 0;
    int i_abl = 0;
    while (i_abl < im - 1 && height_m(T, i_abl) <= T->o.abl_height) i_abl++;
    const double taper_dz = smax(1.0, height_m(T, (i_abl < im - 1) ? i_abl : im - 1)
                                          - height_m(T, (i_abl - 1 > 0) ? i_abl - 1 : 0));
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) {
                double f = 1.0;
                if (taper_on) {
                    const double z_agl = height_m(T, i) - height_m(T, 0);
                    f = clamp3((T->o.abl_height + taper_dz - z_agl) / taper_dz, 0.0, 1.0);
                }
                T->nue[IDX(i, j, k)] = clamp3(T->nue[IDX(i, j, k)], 0.0, nmax) * f;
            }
}

/* ---- fill_lid_corners, Turbulence.h:510-536 (bc_turb_active false: the zero-gradient copy) ---- */
static void turb_lid(turb* T) {
    if (!T->o.free_flow) return;
    const int im = T->im, jm = T->jm, km = T->km, il = im - 1;
    double* f[] = {T->tke, T->dis, T->tken, T->disn, T->nue, T->prod, T->tke_source, T->dis_source};
    for (int j = 1; j < jm - 1; j++)
        for (int k = 0; k < km; k++)
            for (int a = 0; a < 8; a++) f[a][IDX(il, j, k)] = f[a][IDX(il - 1, j, k)];
    for (int i = 0; i < im; i++)
        for (int k = 0; k < km; k++)
            for (int a = 0; a < 8; a++) {
                f[a][IDX(i, 0, k)] = f[a][IDX(i, 1, k)];
                f[a][IDX(i, jm - 1, k)] = f[a][IDX(i, jm - 2, k)];
            }
}

/* init() (Turbulence.h:282-303) with `init`, else run() (:306-332). Options as atu_turb_opts with
 * the defaults already filled in. Returns 0, or 1 for an unsupported argument. */
int orc_turb(const atu_grid* g, const atu_params* p, const atu_turb_opts* o, int init, double* fields,
             double* prod, double* tken, double* disn, double* vel_star) {
    if (o->model < 0 || o->model > 2 || g->km < 4 || g->jm < 3 || g->im < 3) return 1;
    turb T;
    memset(&T, 0, sizeof T);
    T.g = g; T.p = p; T.o = *o;
    T.im = g->im; T.jm = g->jm; T.km = g->km;
    const size_t n3 = (size_t)g->im * g->jm * g->km;
    T.t_ = fields + (size_t)ATU_T * n3;
    T.u = fields + (size_t)ATU_U * n3;
    T.v = fields + (size_t)ATU_V * n3;
    T.w = fields + (size_t)ATU_W * n3;
    T.tke = fields + (size_t)ATU_TKE * n3;
    T.dis = fields + (size_t)ATU_DIS * n3;
    T.nue = fields + (size_t)ATU_NUE * n3;
    T.tke_source = fields + (size_t)ATU_TKE_SOURCE * n3;
    T.dis_source = fields + (size_t)ATU_DIS_SOURCE * n3;
    T.prod = prod; T.tken = tken; T.disn = disn; T.vel_star = vel_star;
    T.radz = malloc(sizeof(double) * g->im);
    T.thez = malloc(sizeof(double) * g->jm);
    T.lh = malloc(sizeof(float) * g->im);
    T.radz[0] = g->r0;
    for (int l = 1; l < g->im; l++) T.radz[l] = T.radz[l - 1] + g->dr;
    T.thez[0] = g->the0;
    for (int l = 1; l < g->jm; l++) T.thez[l] = T.thez[l - 1] + g->dthe;
    const float h = (float)(p->L_atm / (g->im - 1));
    for (int i = 0; i < g->im; i++) T.lh[i] = (float)i * h;
    T.L_m = p->L_atm * 1.0e3;
    T.u_0 = p->u_0;
    T.re_turb = vel_star_ref * z_0 / nue_gas;
    turb_vel_star(&T);
    if (init) {
        turb_init_fields(&T);
        turb_wall(&T);
        turb_sources(&T);
    } else {
        turb_sources(&T);
        turb_wall(&T);
    }
    turb_clamp_nue(&T);
    turb_seam(&T);
    turb_lid(&T);
    free(T.radz); free(T.thez); free(T.lh);
    return 0;
}
