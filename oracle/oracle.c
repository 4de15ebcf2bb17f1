/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY. A plain-C, FP64, OpenMP restatement of the reference's
 * dynamical-core step, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker of the HIP library. Nothing in the product links or calls it.
 *
 * It follows the reference's own structure rather than the GPU design: RHSUran writes rhs_* and
 * aux_* arrays per cell (pass A), a second pass folds them into acc_* and the next stage input from
 * the n-copies (pass B), the final update clamps, and the pressure solver runs its wall passes as
 * separate loops after each sweep. Citations are to files under the reference tree.
 *
 *   orc_geometry      lib/Array_1D.cpp:50-61, planet/RungeKutta_Uran_Turb.cpp:192-283,
 *                     planet/cUranusModel.h:641-647
 *   orc_cp_h2 / table planet/CpH2.h:55-93, cp_lapse :128-134
 *   rhs_cell          planet/RHS_Uran_Turb.cpp:20-1023 at the default knobs
 *   orc_rk4           planet/RungeKutta_Uran_Turb.cpp:179-404
 *   orc_hydro         planet/RungeKutta_Uran_Turb.cpp:18-177
 *   orc_bc            planet/BoundaryConditions.h:152-518, planet/BC_Uran.h:91-171, BC_Uran.cpp
 *   orc_pressure      planet/Pressure_Uran.cpp:30-146, planet/PressureSolver.h:145-766 (plus omega)
 *
 * Dense layout: fields[f * n3 + (i * jm + j) * km + k], f in the atu_field order, n3 = im*jm*km.
 */
#define _GNU_SOURCE   /* sincos */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/atu.h"

#define IDX(i, j, k) (((size_t)(i) * jm + (j)) * km + (k))
#ifndef M_PI
#define M_PI 3.14159265358979323846   /* glibc's value, as the reference's <cmath> has it */
#endif

static double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */
static double dclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static void halving_tree(double* t, int w);
#define ORC_PRESS_LINE 0x100
#define ORC_PRESS_BAROTROPIC 0x200

/* ------------------------------------------------------------------------------------------- */
/* cp of H2 and cp_lapse                                                                        */
/* ------------------------------------------------------------------------------------------- */
static double cpR(double T, int mode) {
    const double thr = 87.57, thv = 6332.0;
    double Zp = 0, Ep = 0, Fp = 0, Zo = 0, Eo = 0, Fo = 0, c;
    if (mode == 1) {
        for (int J = 0; J < 60; J += 2) {
            double e = thr * J * (J + 1.0), b = 1.0 * (2.0 * J + 1.0) * exp(-e / T);
            Zp += b; Ep += b * e; Fp += b * e * e;
        }
        for (int J = 1; J < 60; J += 2) {
            double e = thr * J * (J + 1.0), b = 1.0 * (2.0 * J + 1.0) * exp(-e / T);
            Zo += b; Eo += b * e; Fo += b * e * e;
        }
        double m1p = Ep / Zp, m2p = Fp / Zp, m1o = Eo / Zo, m2o = Fo / Zo;
        c = 0.25 * ((m2p - m1p * m1p) / (T * T)) + 0.75 * ((m2o - m1o * m1o) / (T * T));
    } else {
        double Z = 0, E1 = 0, E2 = 0;
        for (int J = 0; J < 60; J += 2) {
            double e = thr * J * (J + 1.0), b = 1.0 * (2.0 * J + 1.0) * exp(-e / T);
            Z += b; E1 += b * e; E2 += b * e * e;
        }
        for (int J = 1; J < 60; J += 2) {
            double e = thr * J * (J + 1.0), b = 3.0 * (2.0 * J + 1.0) * exp(-e / T);
            Z += b; E1 += b * e; E2 += b * e * e;
        }
        double m1 = E1 / Z, m2 = E2 / Z;
        c = (m2 - m1 * m1) / (T * T);
    }
    double x = thv / T;
    double cv = (x < 700.0) ? x * x * exp(x) / ((exp(x) - 1.0) * (exp(x) - 1.0)) : 0.0;
    return 2.5 + c + cv;
}

#define CP_N 2991
static double cp_tab[3][CP_N];
static int cp_ready[3];

static void cp_build(int mode) {
    if (cp_ready[mode]) return;
#pragma omp critical(orc_cp)
    {
        if (!cp_ready[mode]) {
            for (int i = 0; i < CP_N; i++) cp_tab[mode][i] = cpR(5.0 + i * 0.5, mode) * 8.314462618 / 2.01588e-3;
            cp_ready[mode] = 1;
        }
    }
}

int orc_cp_table(int mode, double* out) {
    if (mode != 1 && mode != 2) return -1;
    cp_build(mode);
    memcpy(out, cp_tab[mode], sizeof(double) * CP_N);
    return CP_N;
}

double orc_cp_lapse(const atu_params* p, double T) {
    if (p->cp_mode == 0 && !p->cp_gas) return p->cp_mix;
    double ch = p->cp_h2;
    if (p->cp_mode > 0) {
        cp_build(p->cp_mode);
        const double* t = cp_tab[p->cp_mode];
        const double lo = (T < 5.0) ? 5.0 : T;              /* std::max(T, T_LO) */
        const double x = (((1500.0 < lo) ? 1500.0 : lo) - 5.0) / 0.5;   /* std::min(.., T_HI) */
        int i = (int)x;
        if (CP_N - 2 < i) i = CP_N - 2;
        double f = x - i;
        ch = t[i] + f * (t[i + 1] - t[i]);
    }
    if (p->cp_gas) return (p->r_h2 * ch + p->r_he * p->cp_he + p->r_ch4 * p->cp_ch4) / (p->r_h2 + p->r_he + p->r_ch4);
    return p->cp_mix + (p->r_h2 / p->r_mix) * (ch - p->cp_h2);
}

/* ------------------------------------------------------------------------------------------- */
/* geometry                                                                                      */
/* ------------------------------------------------------------------------------------------- */
typedef struct {
    double rm, rm2, sinthe, sinthe2, costhe, cotanthe, inv_rm, inv_rm2, inv_rmsinthe, inv_rm2sinthe,
        inv_rm2sinthe2, costhe_inv_rm2sinthe;
} cellgeo;

typedef struct {
    int im, jm, km;
    double *radz, *thez, *sin_tbl, *cos_tbl;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
} orc_geo;

static double metric_radius(const atu_params* p, double r) {
    if (!(p->R_metric_km > 0.0)) return r;
    return r + (p->R_metric_km / p->L_atm - 1.0);
}

static void geo_init(orc_geo* G, const atu_grid* g, const atu_params* p) {
    G->im = g->im; G->jm = g->jm; G->km = g->km;
    G->radz = malloc(sizeof(double) * g->im);
    G->thez = malloc(sizeof(double) * g->jm);
    G->sin_tbl = malloc(sizeof(double) * g->jm);
    G->cos_tbl = malloc(sizeof(double) * g->jm);
    G->radz[0] = g->r0;
    for (int l = 1; l < g->im; l++) G->radz[l] = G->radz[l - 1] + g->dr;
    G->thez[0] = g->the0;
    for (int l = 1; l < g->jm; l++) G->thez[l] = G->thez[l - 1] + g->dthe;
    for (int j = 0; j < g->jm; j++) {
        /* one sincos call for the pair, as the reference's g++ build forms it (its -O2 folds the
         * sin and cos of one angle at planet/RungeKutta_Uran_Turb.cpp:195-196 into sincos) */
        double sj, cj;
        sincos(G->thez[j], &sj, &cj);
        G->sin_tbl[j] = dmax(p->sinthe_min, fabs(sj));
        G->cos_tbl[j] = cj;
    }
    G->inv_2dr = 1.0 / (2.0 * g->dr);
    G->inv_2dthe = 1.0 / (2.0 * g->dthe);
    G->inv_2dphi = 1.0 / (2.0 * g->dphi);
    G->inv_dr2 = 1.0 / (g->dr * g->dr);
    G->inv_dthe2 = 1.0 / (g->dthe * g->dthe);
    G->inv_dphi2 = 1.0 / (g->dphi * g->dphi);
}

static void geo_free(orc_geo* G) {
    free(G->radz); free(G->thez); free(G->sin_tbl); free(G->cos_tbl);
}

static cellgeo cell_geo(const orc_geo* G, const atu_params* p, int i, int j) {
    cellgeo c;
    c.rm = metric_radius(p, G->radz[i]);
    c.rm2 = c.rm * c.rm;
    c.sinthe = G->sin_tbl[j];
    c.sinthe2 = c.sinthe * c.sinthe;
    c.costhe = G->cos_tbl[j];
    c.cotanthe = c.costhe / c.sinthe;
    c.inv_rm = 1.0 / c.rm;
    c.inv_rm2 = 1.0 / c.rm2;
    c.inv_rmsinthe = 1.0 / (c.rm * c.sinthe);
    c.inv_rm2sinthe = c.inv_rm2 / c.sinthe;
    c.inv_rm2sinthe2 = c.inv_rm2 / c.sinthe2;
    c.costhe_inv_rm2sinthe = c.costhe * c.inv_rm2sinthe;
    return c;
}

/* ------------------------------------------------------------------------------------------- */
/* RHS of one cell. `st` points at the 19 stage fields, `fz` at the 21 frozen fields, `out` at   */
/* 22 output fields (19 rhs + aux_u, aux_v, aux_w).                                              */
/* ------------------------------------------------------------------------------------------- */
typedef struct {
    const double* st[ATU_NPROG];
    const double* fz[ATU_NFROZEN];
    const double *pdyn, *phydro;
    double* out[ATU_NPROG + 3];
} rhs_io;

static void rhs_cell(const orc_geo* G, const atu_params* P, const rhs_io* io, int i, int j, int k,
                     const cellgeo* geo) {
    const int jm = G->jm, km = G->km, im = G->im;
    const int n = km - 1;
    const int kw = (k - 1 < 0) ? k - 1 + n : k - 1;
    const int ke = (k + 1 >= n) ? k + 1 - n : k + 1;
    const size_t c = IDX(i, j, k);
    const size_t cip = IDX(i + 1, j, k), cim = IDX(i - 1, j, k);
    const size_t cjp = IDX(i, j + 1, k), cjm = IDX(i, j - 1, k);
    const size_t cke = IDX(i, j, ke), ckw = IDX(i, j, kw);

    double d1r[ATU_NPROG], d2r[ATU_NPROG], d1t[ATU_NPROG], d2t[ATU_NPROG], d1p[ATU_NPROG], d2p[ATU_NPROG];
    for (int f = 0; f < ATU_NPROG; f++) {
        const double* F = io->st[f];
        d1r[f] = (F[cip] - F[cim]) * G->inv_2dr;
        d2r[f] = (F[cip] - 2.0 * F[c] + F[cim]) * G->inv_dr2;
        d1t[f] = (F[cjp] - F[cjm]) * G->inv_2dthe;
        d2t[f] = (F[cjp] - 2.0 * F[c] + F[cjm]) * G->inv_dthe2;
        d1p[f] = (F[cke] - F[ckw]) * G->inv_2dphi;
        d2p[f] = (F[cke] - 2.0 * F[c] + F[ckw]) * G->inv_dphi2;
    }
    const double* pd = io->pdyn;
    const double dpdr = (pd[cip] - pd[cim]) * G->inv_2dr;
    const double dpdthe = (pd[cjp] - pd[cjm]) * G->inv_2dthe;
    const double dpdphi = (pd[cke] - pd[ckw]) * G->inv_2dphi;

    const double u = io->st[ATU_U][c], v = io->st[ATU_V][c], w = io->st[ATU_W][c];
    const double T = io->st[ATU_T][c];

    /* body forces: true sine (CORIOLIS_TRUE_SINE), radial Coriolis and centrifugal inert */
    const double sinthe_bf = sqrt(dmax(0.0, 1.0 - geo->costhe * geo->costhe));
    const double Coriolis_the = -2.0 * P->omega * geo->costhe * w;
    const double Coriolis_phi = +2.0 * P->omega * (+1.0 * geo->costhe * v + sinthe_bf * u);

    const double coeff_energy_p = P->u_0 * P->u_0 / (P->cp_mix * P->t_ref);
    const double scale_p = 1.0e5 / (P->r_mix * P->u_0 * P->u_0);
    const double scale_Cor = P->coriolis_scale;

    const double v_invrm = v * geo->inv_rm;
    const double w_invrs = w * geo->inv_rmsinthe;
    double pressure_t = coeff_energy_p * scale_p * (u * dpdr + v_invrm * dpdthe + w_invrs * dpdphi);
    if ((P->cp_mode > 0 || P->cp_gas) && T > 0.0) pressure_t *= P->cp_mix / orc_cp_lapse(P, T * P->t_ref);

    double tr[ATU_NPROG], df[ATU_NPROG];
    const double two_inv_rm = 2.0 * geo->inv_rm;
    const double v_metric = (1.0 + geo->costhe / geo->sinthe2) * geo->inv_rm2;
    for (int f = 0; f < ATU_NPROG; f++) {
        tr[f] = u * d1r[f] + v_invrm * d1t[f] + w_invrs * d1p[f];
        df[f] = d2r[f] + d1r[f] * two_inv_rm + d2t[f] * geo->inv_rm2 + d1t[f] * geo->costhe_inv_rm2sinthe
                + d2p[f] * geo->inv_rm2sinthe2;
    }
    /* momentum transport with metric terms (METRIC_SPLIT: the radial metric term is in p_hydro) */
    const double transport_v = tr[ATU_V] + (u * v - w * w * geo->cotanthe) * geo->inv_rm;
    const double transport_w = tr[ATU_W] + (w * u + v * w * geo->cotanthe) * geo->inv_rm;
    const double diffusion_u = d2r[ATU_U] + 2.0 * u * geo->inv_rm2 + d2t[ATU_U] * geo->inv_rm2
                               + 4.0 * d1r[ATU_U] * geo->inv_rm + d1t[ATU_U] * geo->costhe_inv_rm2sinthe
                               + d2p[ATU_U] * geo->inv_rm2sinthe2;
    const double diffusion_v = d2r[ATU_V] + d1r[ATU_V] * two_inv_rm + d2t[ATU_V] * geo->inv_rm2
                               + d1t[ATU_V] * geo->costhe_inv_rm2sinthe - v_metric * v
                               + d2p[ATU_V] * geo->inv_rm2sinthe2 + 2.0 * d1t[ATU_U] * geo->inv_rm2
                               - d1p[ATU_W] * 2.0 * geo->costhe * geo->inv_rm2sinthe2;
    const double diffusion_w = d2r[ATU_W] + d1r[ATU_W] * two_inv_rm + d2t[ATU_W] * geo->inv_rm2
                               + d1t[ATU_W] * geo->costhe_inv_rm2sinthe - v_metric * w
                               + d2p[ATU_W] * geo->inv_rm2sinthe2 + 2.0 * d1p[ATU_U] * geo->inv_rm2sinthe
                               + d1p[ATU_V] * 2.0 * geo->costhe * geo->inv_rm2sinthe2;

    const double dpdr_term = dpdr * scale_p;
    const double dpdthe_term = dpdthe * scale_p * geo->inv_rm;
    const double dpdphi_term = dpdphi * scale_p * geo->inv_rmsinthe;

    const double re_eff = P->re / (1.0 + 2.5 * 0.0);
    const double nue = io->fz[ATU_NUE - ATU_NUE][c];
    const double nue_t = (P->turb_coupling != 0.0 && isfinite(nue)) ? P->turb_coupling * dmax(0.0, nue) : 0.0;
    const double nue_t_s = nue_t / P->Pr_t;

    double precip_t = P->precip_coupling * io->fz[ATU_Q_PRECIP - ATU_NUE][c] * (P->L_atm * 1.0e3)
                      / (P->r_mix * P->cp_mix * P->u_0 * P->t_ref);
    if (!isfinite(precip_t)) precip_t = 0.0;
    else if (precip_t > 0.5) precip_t = 0.5;
    else if (precip_t < -0.5) precip_t = -0.5;

    io->out[ATU_T][c] = +pressure_t - tr[ATU_T] + df[ATU_T] / (P->re * P->pr) + df[ATU_T] * nue_t_s + precip_t;

    const double frac_sp = dmax(0.0, (double)(i - (im - 1) * 3 / 4) / (double)((im - 1) / 4));
    const double sponge = P->alpha_sponge * frac_sp * frac_sp;

    const double dphdthe_term = (io->phydro[cjp] - io->phydro[cjm]) * G->inv_2dthe * geo->inv_rm;
    const double dphdphi_term = (io->phydro[cke] - io->phydro[ckw]) * G->inv_2dphi * geo->inv_rmsinthe;

    const double rhs_u = -dpdr_term - tr[ATU_U] + diffusion_u / re_eff + diffusion_u * nue_t - sponge * u;
    const double rhs_v = -dpdthe_term - dphdthe_term - transport_v + diffusion_v / re_eff + diffusion_v * nue_t
                         - P->Coriolis * scale_Cor * Coriolis_the;
    const double rhs_w = -dpdphi_term - dphdphi_term - transport_w + diffusion_w / re_eff + diffusion_w * nue_t
                         - P->Coriolis * scale_Cor * Coriolis_phi;
    io->out[ATU_U][c] = rhs_u;
    io->out[ATU_V][c] = rhs_v;
    io->out[ATU_W][c] = rhs_w;

    /* species: f, Schmidt number, precipitation source, reaction mass flux (or -1) */
    static const int sp_f[12] = {ATU_CH4, ATU_CH4_CLOUD, ATU_CH4_ICE, ATU_H2O, ATU_H2O_CLOUD, ATU_H2O_ICE,
                                 ATU_H2S, ATU_H2S_CLOUD, ATU_H2S_ICE, ATU_NH3, ATU_NH3_CLOUD, ATU_NH3_ICE};
    static const int sp_s[12] = {ATU_SP_CH4, ATU_SP_CH4_CLOUD, ATU_SP_CH4_ICE, ATU_SP_H2O, ATU_SP_H2O_CLOUD,
                                 ATU_SP_H2O_ICE, ATU_SP_H2S, ATU_SP_H2S_CLOUD, ATU_SP_H2S_ICE, ATU_SP_NH3,
                                 ATU_SP_NH3_CLOUD, ATU_SP_NH3_ICE};
    const double sc[12] = {P->sc_ch4, P->sc_ch4, P->sc_ch4, P->sc_h2o, P->sc_h2o, P->sc_h2o,
                           P->sc_h2s, P->sc_h2s, P->sc_h2s, P->sc_nh3, P->sc_nh3, P->sc_nh3};
    for (int s = 0; s < 12; s++) {
        const int f = sp_f[s];
        double r = -tr[f] + df[f] / (sc[s] * P->re) + df[f] * nue_t_s;
        if (f == ATU_H2S) r = r + P->chemical_reaction * io->fz[ATU_MF_H2S - ATU_NUE][c];
        if (f == ATU_NH3) r = r + P->chemical_reaction * io->fz[ATU_MF_NH3 - ATU_NUE][c];
        r = r + P->precip_coupling * io->fz[sp_s[s] - ATU_NUE][c];
        io->out[f][c] = r;
    }
    const double v_stokes = (2.0 / 9.0) * P->r_p_nh4sh * P->r_p_nh4sh
                            * (P->rho_cond_nh4sh - io->fz[ATU_RHO_MIX - ATU_NUE][c]) * P->g / P->mue_mix;
    io->out[ATU_NH4SH][c] = -tr[ATU_NH4SH] + io->fz[ATU_FLUXLIM_NH4SH - ATU_NUE][c]
                            + df[ATU_NH4SH] / (P->sc_nh4sh * P->re) + df[ATU_NH4SH] * nue_t_s
                            + P->chemical_reaction * io->fz[ATU_MF_NH4SH - ATU_NUE][c]
                            + (v_stokes / P->u_0) * d1r[ATU_NH4SH];

    if (P->turb_active) {
        const double sig_k = P->sig_k, sig_w = P->sig_w;
        for (int q = 0; q < 2; q++) {
            const int f = q == 0 ? ATU_TKE : ATU_DIS;
            const double transport = u * d1r[f] + v * d1t[f] * geo->inv_rm + w * d1p[f] * geo->inv_rmsinthe;
            const double diffusion = d2r[f] + d1r[f] * 2.0 * geo->inv_rm + d2t[f] * geo->inv_rm2
                                     + d1t[f] * geo->costhe_inv_rm2sinthe + d2p[f] * geo->inv_rm2sinthe2;
            const double src = io->fz[(q == 0 ? ATU_TKE_SOURCE : ATU_DIS_SOURCE) - ATU_NUE][c];
            io->out[f][c] = -transport + diffusion * (1.0 / P->re + nue / (q == 0 ? sig_k : sig_w)) + src;
        }
    } else {
        io->out[ATU_TKE][c] = 0.0;
        io->out[ATU_DIS][c] = 0.0;
    }
    io->out[ATU_NPROG + 0][c] = rhs_u + dpdr_term;
    io->out[ATU_NPROG + 1][c] = rhs_v + dpdthe_term;
    io->out[ATU_NPROG + 2][c] = rhs_w + dpdphi_term;
}

static void set_io(rhs_io* io, double* fields, double* const* stage, double* out, size_t n3) {
    for (int f = 0; f < ATU_NPROG; f++) io->st[f] = stage[f];
    for (int f = 0; f < ATU_NFROZEN; f++) io->fz[f] = fields + (size_t)(ATU_NUE + f) * n3;
    io->pdyn = fields + (size_t)ATU_P_DYN * n3;
    io->phydro = fields + (size_t)ATU_P_HYDRO * n3;
    for (int f = 0; f < ATU_NPROG + 3; f++) io->out[f] = out + (size_t)f * n3;
}

/* One RHS evaluation on the current state into out (22 dense fields: 19 rhs + aux u, v, w). */
int orc_rhs(const atu_grid* g, const atu_params* p, double* fields, double* out) {
    orc_geo G;
    geo_init(&G, g, p);
    const int im = g->im, jm = g->jm, km = g->km;
    const size_t n3 = (size_t)im * jm * km;
    double* stage[ATU_NPROG];
    for (int f = 0; f < ATU_NPROG; f++) stage[f] = fields + (size_t)f * n3;
    rhs_io io;
    set_io(&io, fields, stage, out, n3);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++) {
            cellgeo cg = cell_geo(&G, p, i, j);
            for (int k = 0; k < km - 1; k++) rhs_cell(&G, p, &io, i, j, k, &cg);
        }
    geo_free(&G);
    return 0;
}

/* ------------------------------------------------------------------------------------------- */
/* RK4 (planet/RungeKutta_Uran_Turb.cpp:179-404); aux_* of the last stage land in fields, and, when */
/* rhs3 is given, rhs_u, rhs_v, rhs_w of the last stage (3 dense fields, 0 off the interior).   */
/* ------------------------------------------------------------------------------------------- */
int orc_rk4_rhs(const atu_grid* g, const atu_params* p, double* fields, double* rhs3) {
    orc_geo G;
    geo_init(&G, g, p);
    const int im = g->im, jm = g->jm, km = g->km;
    const size_t n3 = (size_t)im * jm * km;
    double* yn = malloc(sizeof(double) * n3 * ATU_NPROG);
    double* acc = malloc(sizeof(double) * n3 * ATU_NPROG);
    double* rhs = malloc(sizeof(double) * n3 * (ATU_NPROG + 3));
    if (!yn || !acc || !rhs) { free(yn); free(acc); free(rhs); geo_free(&G); return -3; }
    memcpy(yn, fields, sizeof(double) * n3 * ATU_NPROG);          /* restoreVar */
    memset(rhs, 0, sizeof(double) * n3 * (ATU_NPROG + 3));
    double* stage[ATU_NPROG];
    for (int f = 0; f < ATU_NPROG; f++) stage[f] = fields + (size_t)f * n3;
    rhs_io io;
    set_io(&io, fields, stage, rhs, n3);
    /* aux: keep the model's arrays (their planes carry the previous projection prep) */
    for (int q = 0; q < 3; q++)
        memcpy(rhs + (size_t)(ATU_NPROG + q) * n3, fields + (size_t)(ATU_AUX_U + q) * n3, sizeof(double) * n3);

    const double dt = p->dt;
    const double tke_max = p->tke_max_m2s2 / (p->u_0 * p->u_0), dis_min = p->dis_min_nd;
    const double t_min = p->t_min_K / p->t_ref, t_max = p->t_max_K / p->t_ref;
    for (int stage_n = 0; stage_n < 4; stage_n++) {
        const double c_in = (stage_n == 0 || stage_n == 1) ? 0.5 * dt : (stage_n == 2 ? dt : 0.0);
        const double wgt = (stage_n == 0 || stage_n == 3) ? 1.0 : 2.0;
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++) {
                cellgeo cg = cell_geo(&G, p, i, j);
                for (int k = 0; k < km - 1; k++) rhs_cell(&G, p, &io, i, j, k, &cg);
            }
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++)
                for (int k = 0; k < km - 1; k++) {
                    const size_t c = IDX(i, j, k);
                    for (int f = 0; f < ATU_NPROG; f++) {
                        if (!p->turb_active && (f == ATU_TKE || f == ATU_DIS)) continue;
                        double* A = acc + (size_t)f * n3;
                        const double r = rhs[(size_t)f * n3 + c];
                        A[c] = (stage_n == 0) ? wgt * r : A[c] + wgt * r;
                        if (stage_n < 3) {
                            double y = yn[(size_t)f * n3 + c] + c_in * r;
                            if (f == ATU_TKE) y = dclamp(y, 0.0, tke_max);
                            if (f == ATU_DIS) y = dmax(y, dis_min);
                            fields[(size_t)f * n3 + c] = y;
                        }
                    }
                }
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < km - 1; k++) {
                const size_t c = IDX(i, j, k);
                for (int f = 0; f < ATU_NPROG; f++) {
                    if (!p->turb_active && (f == ATU_TKE || f == ATU_DIS)) continue;
                    double y = yn[(size_t)f * n3 + c] + dt * acc[(size_t)f * n3 + c] / 6.0;
                    if (f == ATU_T) y = dclamp(y, t_min, t_max);
                    if (f == ATU_TKE) y = dclamp(dmax(y, 0.0), 0.0, tke_max);
                    if (f == ATU_DIS) y = dmax(y, dis_min);
                    fields[(size_t)f * n3 + c] = y;
                }
            }
    for (int q = 0; q < 3; q++)
        memcpy(fields + (size_t)(ATU_AUX_U + q) * n3, rhs + (size_t)(ATU_NPROG + q) * n3, sizeof(double) * n3);
    if (rhs3)
        for (int q = 0; q < 3; q++)
            memcpy(rhs3 + (size_t)q * n3, rhs + (size_t)(ATU_U + q) * n3, sizeof(double) * n3);
    free(yn); free(acc); free(rhs);
    geo_free(&G);
    return 0;
}

int orc_rk4(const atu_grid* g, const atu_params* p, double* fields) { return orc_rk4_rhs(g, p, fields, NULL); }

/* ------------------------------------------------------------------------------------------- */
/* hydrostatic prep (planet/RungeKutta_Uran_Turb.cpp:18-177)                                    */
/* ------------------------------------------------------------------------------------------- */
int orc_hydro(const atu_grid* g, const atu_params* p, double* fields, double* level) {
    orc_geo G;
    geo_init(&G, g, p);
    const int im = g->im, jm = g->jm, km = g->km;
    const size_t n3 = (size_t)im * jm * km;
    const double* t = fields + (size_t)ATU_T * n3;
    const double* v = fields + (size_t)ATU_V * n3;
    const double* w = fields + (size_t)ATU_W * n3;
    const double* ps = fields + (size_t)ATU_P_STAT * n3;
    double* ph = fields + (size_t)ATU_P_HYDRO * n3;
    double* ref = malloc(sizeof(double) * im);
    /* The level means (:18-37) with each sum in the fixed order the GPU uses (its k_level_tree),
     * instead of the reference's serial (j, k) order: per (i, j) row the km terms in chunks of 64
     * consecutive k, each chunk reduced by the halving tree (a k past the row, or a skipped cell --
     * T <= 0 or a non-finite b, which the reference skips -- is +0.0), the chunk results added in
     * sequence to 0.0; the jm row sums dealt to 64 partials (partial t adds rows t, t + 64, ... in
     * sequence to 0.0) reduced by the halving tree; the weights the same way. */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < im; i++) {
        double* rs = malloc(sizeof(double) * 2 * jm);
        for (int j = 0; j < jm; j++) {
            const double wgt = sin(G.thez[j]);
            double s = 0.0, ws = 0.0;
            for (int k0 = 0; k0 < km; k0 += 64) {
                double tv[64], wv[64];
                for (int l = 0; l < 64; l++) {
                    const int k = k0 + l;
                    tv[l] = 0.0;
                    wv[l] = 0.0;
                    if (k >= km) continue;
                    const double T = t[IDX(i, j, k)];
                    if (!(T > 0.0)) continue;
                    const double b = p->g * ps[IDX(i, j, k)] / (p->r_mix * p->R_mix * T * p->t_ref);
                    if (!isfinite(b)) continue;
                    tv[l] = wgt * b;
                    wv[l] = wgt;
                }
                halving_tree(tv, 64);
                halving_tree(wv, 64);
                s = s + tv[0];
                ws = ws + wv[0];
            }
            rs[j] = s;
            rs[jm + j] = ws;
        }
        double pt[64], pw[64];
        for (int l = 0; l < 64; l++) {
            double a = 0.0, b = 0.0;
            for (int q = l; q < jm; q += 64) {
                a = a + rs[q];
                b = b + rs[jm + q];
            }
            pt[l] = a;
            pw[l] = b;
        }
        free(rs);
        halving_tree(pt, 64);
        halving_tree(pw, 64);
        ref[i] = (pw[0] > 0.0) ? pt[0] / pw[0] : 0.0;
    }
    const double nd_rel = p->g * (p->L_atm * 1.0e3) / (p->u_0 * p->u_0);
    const double dr = g->dr;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++) {
            const double cor = p->Coriolis * p->coriolis_scale * 2.0 * p->omega * fabs(sin(G.thez[j]));
            double fprev = 0.0;
            for (int i = 0; i < im; i++) {
                const size_t x = IDX(i, j, k);
                double buoy = 0.0;
                if (t[x] > 0.0) {
                    const double b = p->g * ps[x] / (p->r_mix * p->R_mix * t[x] * p->t_ref) - ref[i];
                    const double f = -nd_rel * 1.0 * p->buoyancy * b / ref[i];
                    buoy = isfinite(f) ? f : 0.0;
                }
                const double fi = buoy + (v[x] * v[x] + w[x] * w[x]) / metric_radius(p, G.radz[i]) + cor * w[x];
                if (i == 0) ph[x] = 0.0;
                else ph[x] = ph[IDX(i - 1, j, k)] + 0.5 * (fprev + fi) * dr;
                fprev = fi;
            }
        }
    if (level) memcpy(level, ref, sizeof(double) * im);
    free(ref);
    geo_free(&G);
    return 0;
}

/* ------------------------------------------------------------------------------------------- */
/* domain BC on t, u, v, w and the 13 species                                                    */
/* ------------------------------------------------------------------------------------------- */
/* The fields of the path that the reference's BC lists name (planet/BC_Uran.cpp): t, u, v, w, the 13
 * species, the three reaction mass fluxes and the NH4SH flux limiter. */
static const int bc_list[21] = {ATU_T, ATU_U, ATU_V, ATU_W, ATU_CH4, ATU_CH4_CLOUD, ATU_CH4_ICE, ATU_H2O,
                                ATU_H2O_CLOUD, ATU_H2O_ICE, ATU_H2S, ATU_H2S_CLOUD, ATU_H2S_ICE, ATU_NH3,
                                ATU_NH3_CLOUD, ATU_NH3_ICE, ATU_NH4SH, ATU_MF_H2S, ATU_MF_NH3, ATU_MF_NH4SH,
                                ATU_FLUXLIM_NH4SH};
static int zero_at_poles(int f) {
    return f == ATU_V || f == ATU_W || f == ATU_MF_H2S || f == ATU_MF_NH3 || f == ATU_MF_NH4SH ||
           f == ATU_FLUXLIM_NH4SH;
}

int orc_bc(const atu_grid* g, const atu_params* p, double* fields) {
    const int im = g->im, jm = g->jm, km = g->km;
    const size_t n3 = (size_t)im * jm * km;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    /* radius list (bc_fields_radius, restricted to the fields of the path) */
#pragma omp parallel for
    for (int j = 0; j < jm; j++)
        for (int k = 0; k < km; k++)
            for (int q = 0; q < 21; q++) {
                const int f = bc_list[q];
                double* F = fields + (size_t)f * n3;
                F[IDX(0, j, k)] = c43 * F[IDX(1, j, k)] - c13 * F[IDX(2, j, k)];
                F[IDX(im - 1, j, k)] = c43 * F[IDX(im - 2, j, k)] - c13 * F[IDX(im - 3, j, k)];
            }
    if (p->rigid_lid) {
        double* U = fields + (size_t)ATU_U * n3;
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++) { U[IDX(0, j, k)] = 0.0; U[IDX(im - 1, j, k)] = 0.0; }
    }
    if (p->radius_positive) {
        for (int j = 0; j < jm; j++)
            for (int k = 0; k < km; k++)
                for (int f = ATU_CH4; f <= ATU_NH4SH; f++) {
                    double* F = fields + (size_t)f * n3;
                    if (F[IDX(0, j, k)] < 0.0)
                        F[IDX(0, j, k)] = (p->radius_positive == 2) ? dmax(0.0, F[IDX(1, j, k)]) : 0.0;
                    if (F[IDX(im - 1, j, k)] < 0.0)
                        F[IDX(im - 1, j, k)] = (p->radius_positive == 2) ? dmax(0.0, F[IDX(im - 2, j, k)]) : 0.0;
                }
    }
    if (p->lid_vw == 2 || p->lid_vw == 3) {
        const int a = im - 2, b = im - 3, c = im - 4, l = im - 1;
        for (int q = 0; q < 2; q++) {
            double* F = fields + (size_t)(q == 0 ? ATU_V : ATU_W) * n3;
            for (int j = 0; j < jm; j++)
                for (int k = 0; k < km; k++)
                    F[IDX(l, j, k)] = (p->lid_vw == 2) ? 2.0 * F[IDX(a, j, k)] - F[IDX(b, j, k)]
                                                       : F[IDX(c, j, k)] - 3.0 * F[IDX(b, j, k)] + 3.0 * F[IDX(a, j, k)];
        }
    }
    /* poles: v, w, the mass fluxes and the flux limiter zeroed (bc_fields_theta_zero); t, u and
     * the species 2-point Neumann (bc_fields_theta_extrap) */
#pragma omp parallel for
    for (int k = 0; k < km; k++)
        for (int i = 0; i < im; i++) {
            for (int q = 0; q < 21; q++) {
                const int f = bc_list[q];
                if (!zero_at_poles(f)) continue;
                double* F = fields + (size_t)f * n3;
                F[IDX(i, 0, k)] = 0.0;
                F[IDX(i, jm - 1, k)] = 0.0;
            }
            for (int q = 0; q < 21; q++) {
                const int f = bc_list[q];
                if (zero_at_poles(f)) continue;
                double* F = fields + (size_t)f * n3;
                F[IDX(i, 0, k)] = c43 * F[IDX(i, 1, k)] - c13 * F[IDX(i, 2, k)];
                F[IDX(i, jm - 1, k)] = c43 * F[IDX(i, jm - 2, k)] - c13 * F[IDX(i, jm - 3, k)];
            }
        }
    /* periodic seam (bc_fields_phi) */
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++)
            for (int q = 0; q < 21; q++) {
                double* F = fields + (size_t)bc_list[q] * n3;
                F[IDX(i, j, km - 1)] = F[IDX(i, j, 0)];
            }
    if (p->species_clamp) {
        for (int f = ATU_CH4; f <= ATU_NH4SH; f++) {
            double* F = fields + (size_t)f * n3;
            for (size_t q = 0; q < n3; q++)
                if (!(F[q] >= 0.0)) F[q] = 0.0;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------- */
/* pressure: projection prep, source, red-black sweeps with the wall passes after each sweep     */
/* ------------------------------------------------------------------------------------------- */
typedef struct {
    double *num2, *num3, *den, *inv_rm, *inv_rmsinthe, num1;
} orc_pgeo;

static void pgeo_init(orc_pgeo* q, const orc_geo* G, const atu_params* p) {
    const int im = G->im, jm = G->jm;
    q->num2 = malloc(sizeof(double) * im);
    q->inv_rm = malloc(sizeof(double) * im);
    q->num3 = malloc(sizeof(double) * im * jm);
    q->den = malloc(sizeof(double) * im * jm);
    q->inv_rmsinthe = malloc(sizeof(double) * im * jm);
    double* st = malloc(sizeof(double) * jm);
    for (int j = 0; j < jm; j++) {
        st[j] = sin(G->thez[j]);
        if (st[j] < p->sinthe_min) st[j] = p->sinthe_min;
        if (st[j] == 0.0) st[j] = 1.0e-5;
    }
    const double exp_rm = 1.0, exp_2_rm = exp_rm * exp_rm;
    q->num1 = exp_2_rm * G->inv_dr2;
    for (int i = 0; i < im; i++) {
        const double rm = metric_radius(p, G->radz[i]);
        const double rm2 = rm * rm;
        const double inv_rm2 = 1.0 / rm2;
        q->inv_rm[i] = 1.0 / rm;
        q->num2[i] = inv_rm2 * G->inv_dthe2;
        for (int j = 0; j < jm; j++) {
            const double s2 = st[j] * st[j];
            const double n3 = (inv_rm2 / s2) * G->inv_dphi2;
            q->num3[i * jm + j] = n3;
            q->den[i * jm + j] = 2.0 * q->num1 + 2.0 * q->num2[i] + 2.0 * n3;
            q->inv_rmsinthe[i * jm + j] = 1.0 / (rm * st[j]);
        }
    }
    free(st);
}

static void pgeo_free(orc_pgeo* q) {
    free(q->num2); free(q->num3); free(q->den); free(q->inv_rm); free(q->inv_rmsinthe);
}

/* prepareProjectionBoundaries with the rigid lid, then the clamped div(aux) source into src. */
static void proj_prep_src(const orc_geo* G, const orc_pgeo* q, const atu_params* p, double* fields, double* src) {
    const int im = G->im, jm = G->jm, km = G->km;
    const size_t n3 = (size_t)im * jm * km;
    double* au = fields + (size_t)ATU_AUX_U * n3;
    double* av = fields + (size_t)ATU_AUX_V * n3;
    double* aw = fields + (size_t)ATU_AUX_W * n3;
    for (int j = 1; j < jm - 1; j++)
        for (int k = 0; k < km - 1; k++) {
            au[IDX(0, j, k)] = 0.0;
            au[IDX(im - 1, j, k)] = 0.0;
            double* A[2] = {av, aw};
            for (int s = 0; s < 2; s++) {
                double* a = A[s];
                a[IDX(0, j, k)] = a[IDX(3, j, k)] - 3.0 * a[IDX(2, j, k)] + 3.0 * a[IDX(1, j, k)];
                a[IDX(im - 1, j, k)] = a[IDX(im - 4, j, k)] - 3.0 * a[IDX(im - 3, j, k)] + 3.0 * a[IDX(im - 2, j, k)];
            }
        }
    for (int k = 0; k < km - 1; k++)
        for (int i = 1; i < im - 1; i++) {
            au[IDX(i, 0, k)] = au[IDX(i, 3, k)] - 3.0 * au[IDX(i, 2, k)] + 3.0 * au[IDX(i, 1, k)];
            av[IDX(i, 0, k)] = 0.0;
            aw[IDX(i, 0, k)] = 0.0;
            au[IDX(i, jm - 1, k)] = au[IDX(i, jm - 4, k)] - 3.0 * au[IDX(i, jm - 3, k)] + 3.0 * au[IDX(i, jm - 2, k)];
            av[IDX(i, jm - 1, k)] = 0.0;
            aw[IDX(i, jm - 1, k)] = 0.0;
        }
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++) {
            au[IDX(i, j, km - 1)] = au[IDX(i, j, 0)];
            av[IDX(i, j, km - 1)] = av[IDX(i, j, 0)];
            aw[IDX(i, j, km - 1)] = aw[IDX(i, j, 0)];
        }
    const int n = km - 1;
    memset(src, 0, sizeof(double) * n3);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < n; k++) {
                const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
                const double du_dr = (au[IDX(i + 1, j, k)] - au[IDX(i - 1, j, k)]) * G->inv_2dr;
                const double dv_dthe = (av[IDX(i, j + 1, k)] - av[IDX(i, j - 1, k)]) * G->inv_2dthe;
                const double dw_dphi = (aw[IDX(i, j, ke)] - aw[IDX(i, j, kw)]) * G->inv_2dphi;
                double div_src = du_dr * 1.0 + dv_dthe * q->inv_rm[i] + dw_dphi * q->inv_rmsinthe[i * jm + j];
                const double src_max = q->den[i * jm + j] * p->p_dyn_cap;
                if (!isfinite(div_src)) div_src = 0.0;
                else if (div_src > src_max) div_src = src_max;
                else if (div_src < -src_max) div_src = -src_max;
                src[IDX(i, j, k)] = div_src;
            }
}

static void wall_passes(int im, int jm, int km, double* pd) {
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    for (int k = 0; k < km; k++)
        for (int j = 0; j < jm; j++) {
            pd[IDX(0, j, k)] = c43 * pd[IDX(1, j, k)] - c13 * pd[IDX(2, j, k)];
            pd[IDX(im - 1, j, k)] = c43 * pd[IDX(im - 2, j, k)] - c13 * pd[IDX(im - 3, j, k)];
        }
    for (int k = 0; k < km; k++)
        for (int i = 0; i < im; i++) {
            pd[IDX(i, 0, k)] = c43 * pd[IDX(i, 1, k)] - c13 * pd[IDX(i, 2, k)];
            pd[IDX(i, jm - 1, k)] = c43 * pd[IDX(i, jm - 2, k)] - c13 * pd[IDX(i, jm - 3, k)];
        }
    for (int i = 0; i < im; i++)
        for (int j = 0; j < jm; j++) pd[IDX(i, j, km - 1)] = pd[IDX(i, j, 0)];
}

/* PRESS_LINE (planet/PressureSolver.h:497-570): every (j, k) column of one colour ((j + k) & 1)
 * relaxed exactly in r by the Thomas algorithm, the theta / phi neighbours at their current values,
 * the two radial end rows folded with the 2-point Neumann relation; the column's weights are the
 * line pass's own expressions (:522-525). */
static void ref_line_pass(const orc_geo* G, const atu_params* p, const double* st, const double* src, double* pd) {
    const int im = G->im, jm = G->jm, km = G->km, n = km - 1;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel
        {
            double* aa = malloc(sizeof(double) * im);
            double* bb = malloc(sizeof(double) * im);
            double* cc = malloc(sizeof(double) * im);
            double* dd = malloc(sizeof(double) * im);
            double* cp = malloc(sizeof(double) * im);
            double* dp = malloc(sizeof(double) * im);
#pragma omp for collapse(2) schedule(static)
            for (int j = 1; j < jm - 1; j++)
                for (int k = 0; k < n; k++) {
                    if (((j + k) & 1) != colour) continue;
                    const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
                    for (int i = 1; i < im - 1; i++) {
                        const double rm = metric_radius(p, G->radz[i]);
                        const double erm = 1.0;
                        const double num1 = erm * erm * G->inv_dr2;
                        const double num2 = G->inv_dthe2 / (rm * rm);
                        const double num3 = G->inv_dphi2 / (rm * rm * st[j] * st[j]);
                        const double acc_h = num2 * (pd[IDX(i, j + 1, k)] + pd[IDX(i, j - 1, k)])
                                             + num3 * (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)]);
                        const double den = 2.0 * (num1 + num2 + num3);
                        aa[i] = -num1;
                        bb[i] = den;
                        cc[i] = -num1;
                        dd[i] = acc_h - src[IDX(i, j, k)];
                    }
                    const int lo = 1, hi = im - 2;
                    if (hi > lo) {
                        bb[lo] += aa[lo] * c43;
                        cc[lo] -= aa[lo] * c13;
                        bb[hi] += cc[hi] * c43;
                        aa[hi] -= cc[hi] * c13;
                    }
                    cp[lo] = cc[lo] / bb[lo];
                    dp[lo] = dd[lo] / bb[lo];
                    for (int q = lo + 1; q <= hi; q++) {
                        const double dn = bb[q] - aa[q] * cp[q - 1];
                        cp[q] = cc[q] / dn;
                        dp[q] = (dd[q] - aa[q] * dp[q - 1]) / dn;
                    }
                    double x = dp[hi];
                    if (isfinite(x)) pd[IDX(hi, j, k)] = x;
                    for (int q = hi - 1; q >= lo; q--) {
                        x = dp[q] - cp[q] * x;
                        if (isfinite(x)) pd[IDX(q, j, k)] = x;
                    }
                }
            free(aa); free(bb); free(cc); free(dd); free(cp); free(dp);
        }
    }
}

/* PRESS_BAROTROPIC (planet/PressureSolver.h:572-708): the column sums of the residual (src - Lc p,
 * the end rows read as the wall pass would set them from the interior) divided by the sum of
 * 1/rm^2, a DFT in phi (modes m > 0 of round-off size dropped), per mode a tridiagonal solve in theta
 * with folded ends (mode 0: the consistency shift, q pinned at j = 1, zero sin-weighted mean), the
 * inverse DFT added to every level of the column. Every sum in the reference's serial order. */
static void ref_baro_pass(const orc_geo* G, const atu_params* p, const double* st, const double* src, double* pd) {
    const int im = G->im, jm = G->jm, km = G->km, N = km - 1, J = jm - 2, NM = N / 2;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    double A2 = 0.0;
    for (int i = 1; i < im - 1; i++) {
        const double rm = metric_radius(p, G->radz[i]);
        A2 += 1.0 / (rm * rm);
    }
    double* R = malloc(sizeof(double) * (size_t)J * N);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < jm - 1; j++)
        for (int k = 0; k < N; k++) {
            const int kw = (k == 0) ? N - 1 : k - 1, ke = (k == N - 1) ? 0 : k + 1;
            double sum = 0.0;
            for (int i = 1; i < im - 1; i++) {
                const double rm = metric_radius(p, G->radz[i]);
                const double erm = 1.0;
                const double num1 = erm * erm * G->inv_dr2;
                const double num2 = G->inv_dthe2 / (rm * rm);
                const double num3 = G->inv_dphi2 / (rm * rm * st[j] * st[j]);
#define PJ(jn) ((jn) == 0 ? c43 * pd[IDX(i, 1, k)] - c13 * pd[IDX(i, 2, k)] \
                : ((jn) == jm - 1 ? c43 * pd[IDX(i, jm - 2, k)] - c13 * pd[IDX(i, jm - 3, k)] : pd[IDX(i, jn, k)]))
#define PI(in) ((in) == 0 ? c43 * pd[IDX(1, j, k)] - c13 * pd[IDX(2, j, k)] \
                : ((in) == im - 1 ? c43 * pd[IDX(im - 2, j, k)] - c13 * pd[IDX(im - 3, j, k)] : pd[IDX(in, j, k)]))
                const double p0 = pd[IDX(i, j, k)];
                const double lc = num1 * (PI(i + 1) + PI(i - 1) - 2.0 * p0) + num2 * (PJ(j + 1) + PJ(j - 1) - 2.0 * p0)
                                  + num3 * (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)] - 2.0 * p0);
#undef PJ
#undef PI
                sum += src[IDX(i, j, k)] - lc;
            }
            R[(size_t)(j - 1) * N + k] = sum / A2;
        }
    double* ct = malloc(sizeof(double) * (size_t)(NM + 1) * N);
    double* sn = malloc(sizeof(double) * (size_t)(NM + 1) * N);
    for (int mm = 0; mm <= NM; mm++)
        for (int k = 0; k < N; k++) {
            const double a = 2.0 * M_PI * (double)((long)mm * k % N) / N;
            ct[(size_t)mm * N + k] = cos(a);
            sn[(size_t)mm * N + k] = sin(a);
        }
    double* Ac = malloc(sizeof(double) * (size_t)(NM + 1) * J);
    double* As = malloc(sizeof(double) * (size_t)(NM + 1) * J);
#pragma omp parallel for collapse(2) schedule(static)
    for (int mm = 0; mm <= NM; mm++)
        for (int jj = 0; jj < J; jj++) {
            double sc = 0.0, ss = 0.0;
            const double* r = R + (size_t)jj * N;
            for (int k = 0; k < N; k++) {
                sc += r[k] * ct[(size_t)mm * N + k];
                ss += r[k] * sn[(size_t)mm * N + k];
            }
            if (mm > 0) {
                double ra = 0.0;
                for (int k = 0; k < N; k++) ra += fabs(r[k]);
                const double tol = 1.0e-12 * ra;
                if (fabs(sc) < tol) sc = 0.0;
                if (fabs(ss) < tol) ss = 0.0;
            }
            Ac[(size_t)mm * J + jj] = sc;
            As[(size_t)mm * J + jj] = ss;
        }
    const double e = G->inv_dthe2;
#pragma omp parallel
    {
        double* a = malloc(sizeof(double) * J);
        double* b = malloc(sizeof(double) * J);
        double* c = malloc(sizeof(double) * J);
        double* cp = malloc(sizeof(double) * J);
        double* dp = malloc(sizeof(double) * J);
#pragma omp for schedule(static)
        for (int mm = 0; mm <= NM; mm++) {
            const double sl = sin(M_PI * mm / N);
            const double lam = 4.0 * sl * sl * G->inv_dphi2;
            for (int jj = 0; jj < J; jj++) {
                const double s_ = st[jj + 1];
                a[jj] = e;
                c[jj] = e;
                b[jj] = -2.0 * e - lam / (s_ * s_);
            }
            a[0] = 0.0;
            b[0] += c43 * e;
            c[0] -= c13 * e;
            c[J - 1] = 0.0;
            b[J - 1] += c43 * e;
            a[J - 1] -= c13 * e;
            for (int part = 0; part < 2; part++) {
                if (mm == 0 && part == 1) continue;
                if (2 * mm == N && part == 1) continue;
                double* d = (part == 0 ? Ac : As) + (size_t)mm * J;
                double b0 = b[0], c0 = c[0];
                if (mm == 0) {
                    double yr = 0.0, ys = 0.0;
                    for (int jj = 0; jj < J; jj++) {
                        const double y = (jj == 0 || jj == J - 1) ? 1.5 : 1.0;
                        const double s_ = fabs(sin(G->thez[jj + 1]));
                        yr += y * d[jj];
                        ys += y * s_;
                    }
                    for (int jj = 0; jj < J; jj++) d[jj] -= yr / ys * fabs(sin(G->thez[jj + 1]));
                    b0 = 1.0;
                    c0 = 0.0;
                    d[0] = 0.0;
                }
                cp[0] = c0 / b0;
                dp[0] = d[0] / b0;
                for (int jj = 1; jj < J; jj++) {
                    const double dn = b[jj] - a[jj] * cp[jj - 1];
                    cp[jj] = c[jj] / dn;
                    dp[jj] = (d[jj] - a[jj] * dp[jj - 1]) / dn;
                }
                d[J - 1] = dp[J - 1];
                for (int jj = J - 2; jj >= 0; jj--) d[jj] = dp[jj] - cp[jj] * d[jj + 1];
                if (mm == 0) {
                    double sq = 0.0, sw = 0.0;
                    for (int jj = 0; jj < J; jj++) {
                        const double s_ = fabs(sin(G->thez[jj + 1]));
                        sq += s_ * d[jj];
                        sw += s_;
                    }
                    for (int jj = 0; jj < J; jj++) d[jj] -= sq / sw;
                }
            }
        }
        free(a); free(b); free(c); free(cp); free(dp);
    }
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < J; jj++)
        for (int k = 0; k < N; k++) {
            double q = 0.0;
            for (int mm = 0; mm <= NM; mm++) {
                const int edge = (mm == 0 || 2 * mm == N);
                const double w_ = edge ? 1.0 : 2.0;
                q += w_ * (Ac[(size_t)mm * J + jj] * ct[(size_t)mm * N + k]
                           + (edge ? 0.0 : As[(size_t)mm * J + jj] * sn[(size_t)mm * N + k]));
            }
            q /= N;
            for (int i = 0; i < im; i++) pd[IDX(i, jj + 1, k)] += q;
        }
    free(R); free(ct); free(sn); free(Ac); free(As);
}

/* The interior mean of p (i = 1..im-2, j = 1..jm-2, k = 0..km-2) summed in (i, j, k) order, as the
 * reference's reduction does on one thread (planet/PressureSolver.h:309-316). */
static double interior_mean_serial(int im, int jm, int km, const double* pd) {
    double s = 0.0;
    long n = 0;
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < km - 1; k++) {
                s += pd[IDX(i, j, k)];
                n++;
            }
    return n ? s / n : 0.0;
}

/* `sweeps` red-black sweeps at relaxation factor omega; src_out (optional, dense) receives the
 * clamped source. flags: ORC_PRESS_LINE and / or ORC_PRESS_BAROTROPIC, the reference's PRESS_LINE and
 * PRESS_BAROTROPIC after the pointwise passes of every sweep, before its wall passes; with
 * PRESS_LINE the interior mean of p is held after every sweep (:748-755). */
int orc_pressure_flags(const atu_grid* g, const atu_params* p, double* fields, int sweeps, double omega, int flags,
                       double* src_out) {
    orc_geo G;
    geo_init(&G, g, p);
    orc_pgeo q;
    pgeo_init(&q, &G, p);
    const int im = g->im, jm = g->jm, km = g->km, n = km - 1;
    const size_t n3 = (size_t)im * jm * km;
    double* src = src_out ? src_out : malloc(sizeof(double) * n3);
    proj_prep_src(&G, &q, p, fields, src);
    double* pd = fields + (size_t)ATU_P_DYN * n3;
    double* st = malloc(sizeof(double) * jm);
    for (int j = 0; j < jm; j++) {
        st[j] = sin(G.thez[j]);
        if (st[j] < p->sinthe_min) st[j] = p->sinthe_min;
        if (st[j] == 0.0) st[j] = 1.0e-5;
    }
    const int line = (flags & ORC_PRESS_LINE) != 0, baro = (flags & ORC_PRESS_BAROTROPIC) != 0;
    const double mean0 = line ? interior_mean_serial(im, jm, km, pd) : 0.0;
    if (sweeps < 1) sweeps = 1;
    for (int s = 0; s < sweeps; s++) {
        for (int colour = 0; colour < 2; colour++) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int i = 1; i < im - 1; i++)
                for (int j = 1; j < jm - 1; j++) {
                    const double num1 = q.num1, num2 = q.num2[i], num3 = q.num3[i * jm + j], den = q.den[i * jm + j];
                    for (int k = 0; k < n; k++) {
                        if (((i + j + k) & 1) != colour) continue;
                        const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
                        const double acc = (pd[IDX(i + 1, j, k)] + pd[IDX(i - 1, j, k)]) * num1
                                           + (pd[IDX(i, j + 1, k)] + pd[IDX(i, j - 1, k)]) * num2
                                           + (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)]) * num3;
                        double pn = (acc - src[IDX(i, j, k)]) / den;
                        if (omega != 1.0) pn = (1.0 - omega) * pd[IDX(i, j, k)] + omega * pn;
                        pd[IDX(i, j, k)] = pn;
                    }
                }
        }
        if (line) ref_line_pass(&G, p, st, src, pd);
        if (baro) ref_baro_pass(&G, p, st, src, pd);
        wall_passes(im, jm, km, pd);
        if (line) {
            const double shift = interior_mean_serial(im, jm, km, pd) - mean0;
            for (size_t x = 0; x < n3; x++) pd[x] -= shift;
        }
    }
    free(st);
    if (!src_out) free(src);
    pgeo_free(&q);
    geo_free(&G);
    return 0;
}

int orc_pressure(const atu_grid* g, const atu_params* p, double* fields, int sweeps, double omega, double* src_out) {
    return orc_pressure_flags(g, p, fields, sweeps, omega, 0, src_out);
}

/* ------------------------------------------------------------------------------------------- */
/* Converging solve of the same folded-wall system (the library's ATU_SOLVE_CONVERGE): zebra     */
/* red-black radial line relaxation (the reference's PRESS_LINE, planet/PressureSolver.h:284-303, */
/* 497-570, with the polar walls folded too) and an exact barotropic correction of the            */
/* a-weighted column residual by a DFT in phi and a tridiagonal solve in theta per mode (the      */
/* reference's PRESS_BAROTROPIC, :49-67, 572-708). The source is made consistent with the         */
/* all-Neumann left null vector y = a_i b_j (3/2 at the first and last interior row in i and j,   */
/* 1 elsewhere) by a uniform shift; the interior mean of p is held (as PRESS_LINE does, :748-755).*/
/* Every sum runs in a fixed sequential order shared with the GPU kernels.                        */
/* ------------------------------------------------------------------------------------------- */
static double endw(int x, int lo, int hi) { return (x == lo || x == hi) ? 1.5 : 1.0; }

/* The canonical order of every interior sum of the converging solve, shared with the GPU: per
 * (i, j) row, chunks of 64 consecutive k reduced by a halving tree (v[l] += v[l + h], h = 32..1,
 * missing k as 0) and the chunks of a row added in sequence; the row sums, numbered q = (i-1) nj +
 * (j-1), are dealt to 256 partial sums (partial t adds rows t, t + 256, ... in sequence), each 64
 * consecutive partials are reduced by the halving tree, and the 4 results by the halving tree
 * (h = 2, 1). `v` is a dense field; only interior cells are read. */
#define DSUM_PARTS 256
static void halving_tree(double* t, int w) {
    for (int h = w / 2; h > 0; h >>= 1)
        for (int l = 0; l < h; l++) t[l] += t[l + h];
}

static double dsum_interior(int im, int jm, int km, const double* v) {
    const int n = km - 1;
    const int nch = (n + 63) / 64;
    const int ni = im - 2, nj = jm - 2;
    double* rows = (double*)malloc(sizeof(double) * (size_t)ni * nj);
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++) {
            double row = 0.0;
            for (int c = 0; c < nch; c++) {
                double t[64];
                for (int l = 0; l < 64; l++) {
                    const int k = 64 * c + l;
                    t[l] = (k < n) ? v[IDX(i, j, k)] : 0.0;
                }
                halving_tree(t, 64);
                row += t[0];
            }
            rows[(size_t)(i - 1) * nj + (j - 1)] = row;
        }
    double part[DSUM_PARTS], wsum[DSUM_PARTS / 64];
    for (int t = 0; t < DSUM_PARTS; t++) {
        double s = 0.0;
        for (size_t q = t; q < (size_t)ni * nj; q += DSUM_PARTS) s += rows[q];
        part[t] = s;
    }
    free(rows);
    for (int w = 0; w < DSUM_PARTS / 64; w++) {
        halving_tree(part + 64 * w, 64);
        wsum[w] = part[64 * w];
    }
    halving_tree(wsum, DSUM_PARTS / 64);
    return wsum[0];
}

/* Thomas factors of one column (depend on i and j only): cp[i], dn[i] for i = 1..im-2. */
static void line_factors(const orc_pgeo* q, int im, int jm, int j, double* cp, double* dn) {
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const int n = im - 2;
    for (int s = 0; s < n; s++) {
        const int i = s + 1;
        const double num1 = q->num1, num2 = q->num2[i];
        double diag = -q->den[i * jm + j];
        if (j == 1 || j == jm - 2) diag = diag + num2 * c43;
        double a = num1, c = num1;
        if (s == 0) { diag = diag + num1 * c43; c = num1 - num1 * c13; a = 0.0; }
        if (s == n - 1) { diag = diag + num1 * c43; a = num1 - num1 * c13; c = 0.0; }
        if (n == 1) { a = 0.0; c = 0.0; }
        const double d = (s == 0) ? diag : diag - a * cp[s - 1];
        dn[s] = d;
        cp[s] = c / d;
    }
}

static void line_pass(const orc_pgeo* q, int im, int jm, int km, double* pd, const double* b, int colour,
                      const double* cpt, const double* dnt) {
    const double c13 = 1.0 / 3.0;
    const int n = km - 1, ni = im - 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j < jm - 1; j++)
        for (int k = 0; k < n; k++) {
            if (((j + k) & 1) != colour) continue;
            const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
            const double* cp = cpt + (size_t)j * ni;
            const double* dn = dnt + (size_t)j * ni;
            double dprev = 0.0;
            for (int s = 0; s < ni; s++) {
                const int i = s + 1;
                const double num1 = q->num1, num2 = q->num2[i], num3 = q->num3[i * jm + j];
                double th;
                if (j == 1) th = (num2 - num2 * c13) * pd[IDX(i, 2, k)];
                else if (j == jm - 2) th = (num2 - num2 * c13) * pd[IDX(i, jm - 3, k)];
                else th = (pd[IDX(i, j + 1, k)] + pd[IDX(i, j - 1, k)]) * num2;
                const double rhs = b[IDX(i, j, k)] - th - (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)]) * num3;
                const double a = (s == 0) ? 0.0 : ((s == ni - 1) ? num1 - num1 * c13 : num1);
                const double dp = (rhs - a * dprev) / dn[s];
                pd[IDX(i, j, k)] = dp;
                dprev = dp;
            }
            double x = pd[IDX(ni, j, k)];
            for (int s = ni - 2; s >= 0; s--) {
                x = pd[IDX(s + 1, j, k)] - cp[s] * x;
                pd[IDX(s + 1, j, k)] = x;
            }
        }
}

/* residual r = b - A p with the stored (folded) walls, written into r (interior) */
static void resid_field(const orc_pgeo* q, int im, int jm, int km, const double* pd, const double* b, double* r) {
    const int n = km - 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < n; k++) {
                const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
                const double acc = (pd[IDX(i + 1, j, k)] + pd[IDX(i - 1, j, k)]) * q->num1
                                   + (pd[IDX(i, j + 1, k)] + pd[IDX(i, j - 1, k)]) * q->num2[i]
                                   + (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)]) * q->num3[i * jm + j];
                r[IDX(i, j, k)] = b[IDX(i, j, k)] - (acc - q->den[i * jm + j] * pd[IDX(i, j, k)]);
            }
}

typedef struct {
    int N, NM, J;
    double *ct, *st;        /* [NM+1][N] */
    double *cpm, *dnm;      /* per mode Thomas factors [NM+1][J] */
    double S2;
} orc_baro;

static void baro_init(orc_baro* B, const orc_geo* G, const orc_pgeo* q, const atu_params* p) {
    const int im = G->im, jm = G->jm, km = G->km;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    B->N = km - 1;
    B->NM = B->N / 2;
    B->J = jm - 2;
    const int N = B->N, NM = B->NM, J = B->J;
    B->ct = malloc(sizeof(double) * (NM + 1) * N);
    B->st = malloc(sizeof(double) * (NM + 1) * N);
    for (int m = 0; m <= NM; m++)
        for (int k = 0; k < N; k++) {
            const double a = 2.0 * M_PI * (double)((long)m * k % N) / N;
            sincos(a, &B->st[(size_t)m * N + k], &B->ct[(size_t)m * N + k]);
        }
    double S2 = 0.0;
    for (int i = 1; i < im - 1; i++) {
        const double rm = metric_radius(p, G->radz[i]);
        S2 += endw(i, 1, im - 2) * (1.0 / (rm * rm));
    }
    B->S2 = S2;
    /* the solver's own sine table */
    double* s2 = malloc(sizeof(double) * jm);
    for (int j = 0; j < jm; j++) {
        double s = sin(G->thez[j]);
        if (s < p->sinthe_min) s = p->sinthe_min;
        if (s == 0.0) s = 1.0e-5;
        s2[j] = s * s;
    }
    B->cpm = malloc(sizeof(double) * (NM + 1) * J);
    B->dnm = malloc(sizeof(double) * (NM + 1) * J);
    const double e = G->inv_dthe2;
    for (int m = 0; m <= NM; m++) {
        const double sl = sin(M_PI * m / N);
        const double lam = 4.0 * sl * sl * G->inv_dphi2;
        double prev_cp = 0.0;
        for (int jj = 0; jj < J; jj++) {
            double a = e, c = e, b = -2.0 * e - lam / s2[jj + 1];
            if (jj == 0) { a = 0.0; b = b + c43 * e; c = c - c13 * e; }
            if (jj == J - 1) { c = 0.0; b = b + c43 * e; a = a - c13 * e; }
            if (m == 0 && jj == 0) { b = 1.0; c = 0.0; }   /* mode 0: q pinned at j = 1 */
            const double d = (jj == 0) ? b : b - a * prev_cp;
            B->dnm[(size_t)m * J + jj] = d;
            B->cpm[(size_t)m * J + jj] = c / d;
            prev_cp = B->cpm[(size_t)m * J + jj];
        }
    }
    (void)q;
    free(s2);
}

static void baro_free(orc_baro* B) { free(B->ct); free(B->st); free(B->cpm); free(B->dnm); }

/* The sum of 16 partials by a halving tree: p[l] += p[l + h] for h = 8, 4, 2, 1 (the order of a
 * 16-lane shuffle-down reduction). */
static double tree16(double* p) {
    for (int h = 8; h > 0; h >>= 1)
        for (int l = 0; l < h; l++) p[l] = p[l] + p[l + h];
    return p[0];
}

/* one barotropic correction: r is a scratch field */
static void baro_correct(const orc_baro* B, const orc_geo* G, const orc_pgeo* q, double* pd, const double* b,
                         double* r) {
    const int im = G->im, jm = G->jm, km = G->km, N = B->N, NM = B->NM, J = B->J;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0, e = G->inv_dthe2;
    resid_field(q, im, jm, km, pd, b, r);
    double* R = malloc(sizeof(double) * J * N);
    double* Ac = malloc(sizeof(double) * (NM + 1) * J);
    double* As = malloc(sizeof(double) * (NM + 1) * J);
#pragma omp parallel for collapse(2) schedule(static)
    for (int jj = 0; jj < J; jj++)
        for (int k = 0; k < N; k++) {
            double s = 0.0;
            for (int i = 1; i < im - 1; i++) s += endw(i, 1, im - 2) * r[IDX(i, jj + 1, k)];
            R[(size_t)jj * N + k] = s / B->S2;
        }
    /* the DFTs' sums in the device's order: 16 partials (partial l adds the terms l, l + 16, ... in
       sequence from 0), then a halving tree over the 16 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int m = 0; m <= NM; m++)
        for (int jj = 0; jj < J; jj++) {
            double pc[16] = {0.0}, ps[16] = {0.0};
            for (int k = 0; k < N; k++) {
                pc[k & 15] += R[(size_t)jj * N + k] * B->ct[(size_t)m * N + k];
                ps[k & 15] += R[(size_t)jj * N + k] * B->st[(size_t)m * N + k];
            }
            Ac[(size_t)m * J + jj] = tree16(pc);
            As[(size_t)m * J + jj] = tree16(ps);
        }
#pragma omp parallel for schedule(static)
    for (int t = 0; t < 2 * (NM + 1); t++) {
        const int m = t >> 1, part = t & 1;
        double* d = (part == 0 ? Ac : As) + (size_t)m * J;
        if (part == 1 && (m == 0 || 2 * m == N)) { for (int jj = 0; jj < J; jj++) d[jj] = 0.0; continue; }
        const double* cp = B->cpm + (size_t)m * J;
        const double* dn = B->dnm + (size_t)m * J;
        double prev = 0.0;
        for (int jj = 0; jj < J; jj++) {
            double a = (jj == 0) ? 0.0 : e;
            if (jj == J - 1) a = e - c13 * e;
            double rhs = d[jj];
            if (m == 0 && jj == 0) rhs = 0.0;
            const double dp = (rhs - a * prev) * (1.0 / dn[jj]);
            d[jj] = dp;
            prev = dp;
        }
        double x = d[J - 1];
        for (int jj = J - 2; jj >= 0; jj--) {
            x = d[jj] - cp[jj] * x;
            d[jj] = x;
        }
    }
    (void)c43;
#pragma omp parallel for collapse(2) schedule(static)
    for (int jj = 0; jj < J; jj++)
        for (int k = 0; k < N; k++) {
            double pt[16] = {0.0};
            for (int m = 0; m <= NM; m++) {
                const double a = Ac[(size_t)m * J + jj], c = B->ct[(size_t)m * N + k];
                const double tm = (m == 0 || m == NM) ? a * c
                                                      : 2.0 * (a * c + As[(size_t)m * J + jj] * B->st[(size_t)m * N + k]);
                pt[m & 15] += tm;
            }
            const double qv = tree16(pt) / N;
            for (int i = 0; i < im; i++) pd[IDX(i, jj + 1, k)] += qv;
        }
    free(R); free(Ac); free(As);
}

/* `pprev` (optional, interior cells): the previous solution. When given, it receives p as the
 * solve finds it (and `pprev2`, when given, the previous pprev); with `extrap` 1 the iterations
 * start from 2 p - pprev instead of p, with 2 from 3 p - 3 pprev + pprev2 (the held mean is p's
 * before the extrapolation). */
static int converge_impl(const atu_grid* g, const atu_params* p, double* fields, int max_iters, double rel_tol,
                         int* iters_out, double* rel_out, double* src_out, double* pprev, double* pprev2,
                         int extrap) {
    orc_geo G;
    geo_init(&G, g, p);
    orc_pgeo q;
    pgeo_init(&q, &G, p);
    orc_baro B;
    baro_init(&B, &G, &q, p);
    const int im = g->im, jm = g->jm, km = g->km, n = km - 1, ni = im - 2;
    const size_t n3 = (size_t)im * jm * km;
    double* src = src_out ? src_out : malloc(sizeof(double) * n3);
    double* b = malloc(sizeof(double) * n3);
    double* r = malloc(sizeof(double) * n3);
    proj_prep_src(&G, &q, p, fields, src);
    double* pd = fields + (size_t)ATU_P_DYN * n3;
    /* consistency shift and the mean to hold */
    memset(r, 0, sizeof(double) * n3);
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < n; k++) r[IDX(i, j, k)] = (endw(i, 1, im - 2) * endw(j, 1, jm - 2)) * src[IDX(i, j, k)];
    const double ys = dsum_interior(im, jm, km, r);
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < n; k++) r[IDX(i, j, k)] = endw(i, 1, im - 2) * endw(j, 1, jm - 2);
    const double yw = dsum_interior(im, jm, km, r);
    const double shift = ys / yw;
    const double ncell = (double)ni * (jm - 2) * n;
    const double mean0 = dsum_interior(im, jm, km, pd) / ncell;
    if (pprev)
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++)
                for (int k = 0; k < n; k++) {
                    const double a = pd[IDX(i, j, k)];
                    const double a1 = pprev[IDX(i, j, k)];
                    if (extrap == 2) pd[IDX(i, j, k)] = 3.0 * a - 3.0 * a1 + pprev2[IDX(i, j, k)];
                    else if (extrap) pd[IDX(i, j, k)] = 2.0 * a - a1;
                    if (pprev2) pprev2[IDX(i, j, k)] = a1;
                    pprev[IDX(i, j, k)] = a;
                }
    memset(b, 0, sizeof(double) * n3);
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < n; k++) {
                b[IDX(i, j, k)] = src[IDX(i, j, k)] - shift;
                r[IDX(i, j, k)] = b[IDX(i, j, k)] * b[IDX(i, j, k)];
            }
    const double bnorm = sqrt(dsum_interior(im, jm, km, r));
    double* cpt = malloc(sizeof(double) * jm * ni);
    double* dnt = malloc(sizeof(double) * jm * ni);
    for (int j = 1; j < jm - 1; j++) line_factors(&q, im, jm, j, cpt + (size_t)j * ni, dnt + (size_t)j * ni);
    int it = 0;
    double rel = INFINITY;
    while (it < max_iters) {
        line_pass(&q, im, jm, km, pd, b, 0, cpt, dnt);
        line_pass(&q, im, jm, km, pd, b, 1, cpt, dnt);
        wall_passes(im, jm, km, pd);
        baro_correct(&B, &G, &q, pd, b, r);
        wall_passes(im, jm, km, pd);
        it++;
        resid_field(&q, im, jm, km, pd, b, r);
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++)
                for (int k = 0; k < n; k++) r[IDX(i, j, k)] = r[IDX(i, j, k)] * r[IDX(i, j, k)];
        const double rr = dsum_interior(im, jm, km, r);
        rel = bnorm > 0.0 ? sqrt(rr) / bnorm : sqrt(rr);
        if (rel <= rel_tol) break;
    }
    /* hold the interior mean */
    const double dm = dsum_interior(im, jm, km, pd) / ncell - mean0;
    for (size_t x = 0; x < n3; x++) pd[x] -= dm;
    if (iters_out) *iters_out = it;
    if (rel_out) *rel_out = rel;
    free(cpt); free(dnt); free(b); free(r);
    if (!src_out) free(src);
    baro_free(&B);
    pgeo_free(&q);
    geo_free(&G);
    return 0;
}

int orc_pressure_converge(const atu_grid* g, const atu_params* p, double* fields, int max_iters, double rel_tol,
                          int* iters_out, double* rel_out, double* src_out) {
    return converge_impl(g, p, fields, max_iters, rel_tol, iters_out, rel_out, src_out, NULL, NULL, 0);
}

/* ------------------------------------------------------------------------------------------- */
/* PressureSolverUran::run (planet/PressureSolverUran.h:33-250), the reference's default solver,   */
/* at its defaults: periodic phi (k = 0..km-2 relaxed, km-1 the copy), PRESS_CONSISTENT (metric    */
/* radius, src_scale = r_mix u0^2 / 1e5) and PRESS_INCR (the defect-correction source              */
/* src_scale div(rhs) + Lc(p_old) over the last stage's rhs_*). One call: the radial walls of aux, */
/* the source, one serial lexicographic Gauss-Seidel sweep, the walls of p. `calls` calls in a row. */
/* ------------------------------------------------------------------------------------------- */
int orc_pressure_uran(const atu_grid* g, const atu_params* p, double* fields, const double* rhs3, int calls) {
    orc_geo G;
    geo_init(&G, g, p);
    const int im = g->im, jm = g->jm, km = g->km, n = km - 1;
    const size_t n3 = (size_t)im * jm * km;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    double* au = fields + (size_t)ATU_AUX_U * n3;
    double* av = fields + (size_t)ATU_AUX_V * n3;
    double* aw = fields + (size_t)ATU_AUX_W * n3;
    double* pd = fields + (size_t)ATU_P_DYN * n3;
    const double* ru_ = rhs3;
    const double* rv_ = rhs3 + n3;
    const double* rw_ = rhs3 + 2 * n3;
    double* sinthe = malloc(sizeof(double) * jm);
    double* incr = malloc(sizeof(double) * n3);
    for (int j = 0; j < jm; j++) {
        sinthe[j] = sin(G.thez[j]);
        if (sinthe[j] < p->sinthe_min) sinthe[j] = p->sinthe_min;
    }
    const double src_scale = (p->r_mix * p->u_0 * p->u_0) / 1.0e5;
    const double dr = g->dr, dthe = g->dthe, dphi = g->dphi;
    const double inv_2dr = 1.0 / (2.0 * dr), inv_2dthe = 1.0 / (2.0 * dthe), inv_2dphi = 1.0 / (2.0 * dphi);
    const double inv_dr2 = 1.0 / (dr * dr), inv_dthe2 = 1.0 / (dthe * dthe), inv_dphi2 = 1.0 / (dphi * dphi);
    (void)inv_2dr; (void)inv_2dthe; (void)inv_2dphi;
    for (int call = 0; call < calls; call++) {
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < km - 1; k++) {
                double* A[3] = {au, av, aw};
                for (int q = 0; q < 3; q++) {
                    A[q][IDX(0, j, k)] = c43 * A[q][IDX(1, j, k)] - c13 * A[q][IDX(2, j, k)];
                    A[q][IDX(im - 1, j, k)] = c43 * A[q][IDX(im - 2, j, k)] - c13 * A[q][IDX(im - 3, j, k)];
                }
            }
        /* the defect-correction source, from the p_dyn before this call's sweep */
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++) {
                const double rm = metric_radius(p, G.radz[i]);
                const double inv_rm = 1.0 / rm;
                const double inv_rmsinthe = 1.0 / (rm * sinthe[j]);
                const double inv_rm2 = inv_rm * inv_rm;
                const double inv_rm2sinthe2 = inv_rm2 / (sinthe[j] * sinthe[j]);
                const double num1 = 1.0 / (dr * dr), num2 = inv_rm2 / (dthe * dthe),
                             num3 = inv_rm2sinthe2 / (dphi * dphi);
                const double denom = 2.0 * (num1 + num2 + num3);
                for (int k = 0; k < n; k++) {
                    const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
                    const double up = (i + 1 == im - 1) ? c43 * ru_[IDX(im - 2, j, k)] - c13 * ru_[IDX(im - 3, j, k)]
                                                        : ru_[IDX(i + 1, j, k)];
                    const double um = (i - 1 == 0) ? c43 * ru_[IDX(1, j, k)] - c13 * ru_[IDX(2, j, k)]
                                                   : ru_[IDX(i - 1, j, k)];
                    const double div_rhs = (up - um) / (2.0 * dr)
                                           + (rv_[IDX(i, j + 1, k)] - rv_[IDX(i, j - 1, k)]) / (2.0 * dthe) * inv_rm
                                           + (rw_[IDX(i, j, ke)] - rw_[IDX(i, j, kw)]) / (2.0 * dphi) * inv_rmsinthe;
                    const double lc_old = (pd[IDX(i + 1, j, k)] + pd[IDX(i - 1, j, k)]) * num1
                                          + (pd[IDX(i, j + 1, k)] + pd[IDX(i, j - 1, k)]) * num2
                                          + (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)]) * num3
                                          - denom * pd[IDX(i, j, k)];
                    incr[IDX(i, j, k)] = src_scale * div_rhs + lc_old;
                }
            }
        /* one lexicographic Gauss-Seidel sweep, in place */
        for (int i = 1; i < im - 1; i++)
            for (int j = 1; j < jm - 1; j++) {
                const double rm = metric_radius(p, G.radz[i]);
                const double rm2 = rm * rm;
                const double s = sinthe[j];
                const double inv_rm2 = 1.0 / rm2;
                const double inv_rm2sinthe2 = inv_rm2 / (s * s);
                const double denom = 2.0 * inv_dr2 + 2.0 * inv_rm2 * inv_dthe2 + 2.0 * inv_rm2sinthe2 * inv_dphi2;
                const double inv_denom = 1.0 / denom;
                const double num1 = inv_dr2, num2 = inv_rm2 * inv_dthe2, num3 = inv_rm2sinthe2 * inv_dphi2;
                for (int k = 0; k < n; k++) {
                    const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
                    pd[IDX(i, j, k)] = ((pd[IDX(i + 1, j, k)] + pd[IDX(i - 1, j, k)]) * num1
                                        + (pd[IDX(i, j + 1, k)] + pd[IDX(i, j - 1, k)]) * num2
                                        + (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)]) * num3
                                        - incr[IDX(i, j, k)]) * inv_denom;
                }
            }
        for (int k = 0; k < km; k++)
            for (int j = 0; j < jm; j++) {
                pd[IDX(0, j, k)] = c43 * pd[IDX(1, j, k)] - c13 * pd[IDX(2, j, k)];
                pd[IDX(im - 1, j, k)] = c43 * pd[IDX(im - 2, j, k)] - c13 * pd[IDX(im - 3, j, k)];
            }
        for (int k = 0; k < km; k++)
            for (int i = 0; i < im; i++) {
                pd[IDX(i, 0, k)] = c43 * pd[IDX(i, 1, k)] - c13 * pd[IDX(i, 2, k)];
                pd[IDX(i, jm - 1, k)] = c43 * pd[IDX(i, jm - 2, k)] - c13 * pd[IDX(i, jm - 3, k)];
            }
        for (int i = 0; i < im; i++)
            for (int j = 0; j < jm; j++) pd[IDX(i, j, km - 1)] = pd[IDX(i, j, 0)];
    }
    free(sinthe);
    free(incr);
    geo_free(&G);
    return 0;
}

/* ||src - Lc p||_2 and ||src||_2 over the interior, Lc p = acc - den p with the stored walls. */
int orc_residual(const atu_grid* g, const atu_params* p, const double* fields, const double* src, double* rnorm,
                 double* snorm) {
    orc_geo G;
    geo_init(&G, g, p);
    orc_pgeo q;
    pgeo_init(&q, &G, p);
    const int im = g->im, jm = g->jm, km = g->km, n = km - 1;
    const size_t n3 = (size_t)im * jm * km;
    const double* pd = fields + (size_t)ATU_P_DYN * n3;
    double rr = 0.0, ss = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : rr, ss) schedule(static)
    for (int i = 1; i < im - 1; i++)
        for (int j = 1; j < jm - 1; j++)
            for (int k = 0; k < n; k++) {
                const int kw = (k == 0) ? n - 1 : k - 1, ke = (k == n - 1) ? 0 : k + 1;
                const double acc = (pd[IDX(i + 1, j, k)] + pd[IDX(i - 1, j, k)]) * q.num1
                                   + (pd[IDX(i, j + 1, k)] + pd[IDX(i, j - 1, k)]) * q.num2[i]
                                   + (pd[IDX(i, j, ke)] + pd[IDX(i, j, kw)]) * q.num3[i * jm + j];
                const double s = src[IDX(i, j, k)];
                const double r = s - (acc - q.den[i * jm + j] * pd[IDX(i, j, k)]);
                rr += r * r;
                ss += s * s;
            }
    *rnorm = sqrt(rr);
    *snorm = sqrt(ss);
    pgeo_free(&q);
    geo_free(&G);
    return 0;
}

/* The benchmark loop: hydro prep, RK4, BC, `sweeps` red-black sweeps at omega, nsteps times. */
int orc_step(const atu_grid* g, const atu_params* p, double* fields, int nsteps, int sweeps, double omega) {
    for (int s = 0; s < nsteps; s++) {
        int e;
        if ((e = orc_hydro(g, p, fields, NULL))) return e;
        if ((e = orc_rk4(g, p, fields))) return e;
        if ((e = orc_bc(g, p, fields))) return e;
        if ((e = orc_pressure(g, p, fields, sweeps, omega, NULL))) return e;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------- */
/* Shapiro (1-2-1) wiggle damping, AtomUtils::damp_wiggles (lib/Utils.cpp:224-306), without a fluid */
/* floor (i_surface = nullptr, as every call in the reference's loop) and with the periodic seam   */
/* the model sets (planet/cUranusModel.cpp:459): per pass along k (k = 0..km-2 over the ring of    */
/* km-1 meridians, then k = km-1 the copy of k = 0), along j (clamped at the poles), along i       */
/* (clamped), each from a copy of the field: f + strength/4 (f[n-1] - 2 f + f[n+1]).              */
/* ------------------------------------------------------------------------------------------- */
int orc_damp_wiggles(const atu_grid* g, double* fields, int field, int along_i, int along_j, int along_k,
                     double strength, int passes) {
    const int im = g->im, jm = g->jm, km = g->km;
    const size_t n3 = (size_t)im * jm * km;
    if (field < 0 || field >= ATU_NFIELDS || km < 3) return 1;
    double* f = fields + (size_t)field * n3;
    double* tmp = malloc(sizeof(double) * n3);
    const double coeff = strength * 0.25;
    for (int pass = 0; pass < passes; pass++) {
        if (along_k) {
            memcpy(tmp, f, sizeof(double) * n3);
#pragma omp parallel for collapse(2) schedule(static)
            for (int i = 0; i < im; i++)
                for (int j = 0; j < jm; j++) {
                    for (int k = 0; k < km - 1; k++) {
                        const int km1 = k == 0 ? km - 2 : k - 1, kp1 = k == km - 2 ? 0 : k + 1;
                        f[IDX(i, j, k)] = tmp[IDX(i, j, k)] + coeff * (tmp[IDX(i, j, km1)] - 2.0 * tmp[IDX(i, j, k)]
                                                                      + tmp[IDX(i, j, kp1)]);
                    }
                    f[IDX(i, j, km - 1)] = f[IDX(i, j, 0)];
                }
        }
        if (along_j) {
            memcpy(tmp, f, sizeof(double) * n3);
#pragma omp parallel for collapse(2) schedule(static)
            for (int i = 0; i < im; i++)
                for (int j = 0; j < jm; j++) {
                    const int jm1 = j > 0 ? j - 1 : 0, jp1 = j < jm - 1 ? j + 1 : jm - 1;
                    for (int k = 0; k < km; k++)
                        f[IDX(i, j, k)] = tmp[IDX(i, j, k)] + coeff * (tmp[IDX(i, jm1, k)] - 2.0 * tmp[IDX(i, j, k)]
                                                                      + tmp[IDX(i, jp1, k)]);
                }
        }
        if (along_i) {
            memcpy(tmp, f, sizeof(double) * n3);
#pragma omp parallel for collapse(2) schedule(static)
            for (int j = 0; j < jm; j++)
                for (int k = 0; k < km; k++)
                    for (int i = 0; i < im; i++) {
                        const int im1 = i > 0 ? i - 1 : 0, ip1 = i < im - 1 ? i + 1 : im - 1;
                        f[IDX(i, j, k)] = tmp[IDX(i, j, k)] + coeff * (tmp[IDX(im1, j, k)] - 2.0 * tmp[IDX(i, j, k)]
                                                                      + tmp[IDX(ip1, j, k)]);
                    }
        }
    }
    free(tmp);
    return 0;
}

/* The same outer step with the converging solve; with `warm` 1, every solve after the first starts
 * from the linear extrapolation 2 p_n - p_{n-1} of the last two solutions; with 2, every solve after
 * the second from the quadratic one 3 p_n - 3 p_{n-1} + p_{n-2} of the last three (the second from
 * the linear one). iters_out / rel_out: the last solve's. */
int orc_step_converge(const atu_grid* g, const atu_params* p, double* fields, int nsteps, int max_iters,
                      double rel_tol, int warm, int* iters_out, double* rel_out) {
    const size_t n3 = (size_t)g->im * g->jm * g->km;
    double* pprev = warm ? calloc(n3, sizeof(double)) : NULL;
    double* pprev2 = warm == 2 ? calloc(n3, sizeof(double)) : NULL;
    int e = 0;
    for (int s = 0; s < nsteps && !e; s++) {
        if ((e = orc_hydro(g, p, fields, NULL)) || (e = orc_rk4(g, p, fields)) || (e = orc_bc(g, p, fields))) break;
        const int mode = !warm ? 0 : (s == 0 ? 0 : ((warm == 2 && s >= 2) ? 2 : 1));
        e = converge_impl(g, p, fields, max_iters, rel_tol, iters_out, rel_out, NULL, pprev, pprev2, mode);
    }
    free(pprev);
    free(pprev2);
    return e;
}
