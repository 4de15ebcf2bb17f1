// Fused RK4 stage kernel: RHSUran (planet/RHS_Uran_Turb.cpp:20-1023) evaluated per cell and folded
// at once into the accumulator and the next stage's input (RungeKuttaUran pass A + pass B + final,
// planet/RungeKutta_Uran_Turb.cpp:242-403). The reference's rhs_* arrays are never materialised:
// each tendency lives in registers between its stencil and its fold.
//
// Each expression keeps the reference's operand order and association, and the library is
// compiled with -ffp-contract=off, so the FP64 instantiation forms the same IEEE results as the
// reference's CPU code. Terms whose coefficient is zero at the default knobs (thermal mass flux,
// radiation, buoyancy in rhs_u with the hydrostatic split, radial Coriolis with CORIOLIS_SPLIT,
// centrifugal with CENTRIFUGAL_MOM=0) add an exact zero in the reference and are left out.
//
// Thread layout: one thread per interior cell; a 64-lane wavefront spans 64 consecutive k of one
// (i, j) row, so every field load is a coalesced 512 B (FP64) or 256 B (FP32) row segment and the
// per-(i,j) geometry is wave-uniform.

#include "atu_kernels.h"

namespace atu {
namespace {

enum { CL_NONE = 0, CL_T = 1, CL_TKE = 2, CL_DIS = 3 };

template <typename R>
struct D7 {
    R c, d1r, d2r, d1t, d2t, d1p, d2p;
};

// Central first and second differences in r, theta, phi (planet/RHS_Uran_Turb.cpp:93-176).
template <typename R, typename S>
__device__ __forceinline__ D7<R> stencil(const S* __restrict__ f, size_t c, size_t cw, size_t ce,
                                         size_t plane, int ldk, R i2dr, R idr2, R i2dt, R idt2, R i2dp,
                                         R idp2) {
    const R x0 = (R)f[c];
    const R rp = (R)f[c + plane], rn = (R)f[c - plane];
    const R tp = (R)f[c + ldk], tn = (R)f[c - ldk];
    const R pe = (R)f[ce], pw = (R)f[cw];
    D7<R> d;
    d.c = x0;
    d.d1r = (rp - rn) * i2dr;
    d.d2r = (rp - R(2.0) * x0 + rn) * idr2;
    d.d1t = (tp - tn) * i2dt;
    d.d2t = (tp - R(2.0) * x0 + tn) * idt2;
    d.d1p = (pe - pw) * i2dp;
    d.d2p = (pe - R(2.0) * x0 + pw) * idp2;
    return d;
}

// (a - b) * s in FP64, rounded once to R. For R = double this is the reference's (a - b) * s.
template <typename R>
__device__ __forceinline__ R fp64_diff(double a, double b, double s) {
    return (R)((a - b) * s);
}

// Streams touched once per launch (the accumulator, y_n, the next stage input, the step output)
// bypass the caches' retention so that the stencil rows shared by neighbouring tiles stay in them.
template <typename T>
__device__ __forceinline__ T nt_ld(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void nt_st(T* p, T v) { __builtin_nontemporal_store(v, p); }

// Fold one tendency into the accumulator and form the next input (pass B) or the step (final).
template <int STAGE, int CLAMP, typename S>
__device__ __forceinline__ void fold(const StageArgs<S>& a, int f, size_t c, double y0, double r) {
    const size_t q = (size_t)f * a.sh.fs + c;
    if (STAGE < 3) {
        if (STAGE == 0)
            nt_st(a.acc + q, 1.0 * r);
        else
            nt_st(a.acc + q, nt_ld(a.acc + q) + 2.0 * r);
        const double cin = (STAGE == 2) ? a.c.c_full : a.c.c_half;
        double y = y0 + cin * r;
        if (CLAMP == CL_TKE) y = y < 0.0 ? 0.0 : (y > a.c.tke_max ? a.c.tke_max : y);
        if (CLAMP == CL_DIS) y = (y < a.c.dis_min) ? a.c.dis_min : y;
        nt_st(a.xnext + q, (S)y);
    } else {
        const double s = nt_ld(a.acc + q) + 1.0 * r;
        double y = y0 + a.c.dt * s / 6.0;
        if (CLAMP == CL_T) y = y < a.c.t_min ? a.c.t_min : (y > a.c.t_max ? a.c.t_max : y);
        if (CLAMP == CL_TKE) {
            y = (y < 0.0) ? 0.0 : y;
            y = y < 0.0 ? 0.0 : (y > a.c.tke_max ? a.c.tke_max : y);
        }
        if (CLAMP == CL_DIS) y = (y < a.c.dis_min) ? a.c.dis_min : y;
        nt_st(a.yout + q, y);
        if (a.ymirror) nt_st(a.ymirror + q, (S)y);
    }
}

// cp_lapse(T) / cp_mix ratio input: CpH2::Lapse::cp (planet/CpH2.h:88-93,128-134) on the
// host-built H2 table.
template <typename S>
__device__ __forceinline__ double cp_lapse_dev(const StageArgs<S>& a, double T) {
    const double* t = a.geo.cp_tab;
    const int n = a.geo.cp_n;
    const double Tc = (T < 5.0 ? 5.0 : T);
    const double x = ((Tc > 1500.0 ? 1500.0 : Tc) - 5.0) / 0.5;
    const int ix = (int)x;
    const int i = (n - 2 < ix) ? n - 2 : ix;
    const double f = x - i;
    const double ch = t[i] + f * (t[i + 1] - t[i]);
    return (a.c.r_h2 * ch + a.c.cp_a + a.c.cp_b) / a.c.cp_den;
}

// The tendencies of one cell in arithmetic precision R from stencil storage S, folded in FP64.
// PART 0: every field; PART 1: t only; PART 2: every field but t (t's tendency reads no other
// tendency, so the last stage can update t first and the rest after); PART 3: every field but v
// and w, PART 4: v and w (only their tendencies read p_hydro, so the first stage can start before
// the hydrostatic prep has finished). Each field's fold reads and writes that field alone.
template <typename R, typename S, int STAGE, int PART>
__device__ __forceinline__ void rk_cell(const StageArgs<S>& a, int i, int j, int k) {
    const int nk = a.sh.km - 1;
    const int ldk = a.sh.ldk;
    const size_t plane = a.sh.plane, fs = a.sh.fs;
    const size_t row = ((size_t)i * a.sh.jm + j) * ldk;
    const size_t c = row + k;
    const size_t cw = row + (k == 0 ? nk - 1 : k - 1);   // phi ring over km-1 meridians
    const size_t ce = row + (k == nk - 1 ? 0 : k + 1);

    const DevGeo& G = a.geo;
    const size_t ij = (size_t)i * a.sh.jm + j;
    const R inv_rm = (R)G.inv_rm[i], inv_rm2 = (R)G.inv_rm2[i], two_inv_rm = (R)G.two_inv_rm[i];
    const R sponge = (R)G.sponge[i];
    const R costhe = (R)G.costhe[j], cotanthe = (R)G.cotanthe[j], sinthe_bf = (R)G.sinthe_bf[j];
    const R inv_rmsinthe = (R)G.inv_rmsinthe[ij], inv_rm2sinthe = (R)G.inv_rm2sinthe[ij];
    const R inv_rm2sinthe2 = (R)G.inv_rm2sinthe2[ij], cirs = (R)G.costhe_inv_rm2sinthe[ij];
    const R v_metric = (R)G.v_metric[ij];
    const R i2dr = (R)G.inv_2dr, idr2 = (R)G.inv_dr2, i2dt = (R)G.inv_2dthe, idt2 = (R)G.inv_dthe2;
    const R i2dp = (R)G.inv_2dphi, idp2 = (R)G.inv_dphi2;
    const StepConsts& C = a.c;

    auto ST = [&](int f) {
        return stencil<R, S>(a.xs + (size_t)f * fs, c, cw, ce, plane, ldk, i2dr, idr2, i2dt, idt2, i2dp, idp2);
    };
    // y_n of the fold: the next stage input of the FP32 path is rounded to FP32 anyway, so it is
    // formed from y_n's FP32 mirror (at stage 0 the stencil centre itself); FP64 y_n elsewhere
    auto Y0 = [&](int f, R center) -> double {
        if (STAGE == 0) return (double)center;
        if (STAGE < 3 && sizeof(S) == sizeof(float)) return (double)nt_ld(a.yn_lo + (size_t)f * fs + c);
        return nt_ld(a.yn + (size_t)f * fs + c);
    };
    auto FZ = [&](int f) -> R { return (R)nt_ld(a.frozen + (size_t)(f - ATU_NUE) * fs + c); };

    // velocities and the pressure gradient
    D7<R> U{}, V{}, W{};
    if constexpr (PART != 1) {
        U = ST(ATU_U);
        V = ST(ATU_V);
        W = ST(ATU_W);
    } else {   // t's tendency reads the velocities at the cell only
        U.c = (R)a.xs[(size_t)ATU_U * fs + c];
        V.c = (R)a.xs[(size_t)ATU_V * fs + c];
        W.c = (R)a.xs[(size_t)ATU_W * fs + c];
    }
    const R u0 = U.c, v0 = V.c, w0 = W.c;
    // p_dyn (|p| ~ 20) varies slowly: its differences are formed in FP64 and rounded once to R,
    // since an FP32 difference of two such values keeps few significant bits
    const double* pd = a.pdyn;
    const R dpdr = fp64_diff<R>(pd[c + plane], pd[c - plane], G.inv_2dr);
    const R dpdthe = fp64_diff<R>(pd[c + ldk], pd[c - ldk], G.inv_2dthe);
    const R dpdphi = fp64_diff<R>(pd[ce], pd[cw], G.inv_2dphi);
    const R scale_p = (R)C.scale_p;
    const R v_invrm = v0 * inv_rm;
    const R w_invrs = w0 * inv_rmsinthe;
    const R dpdr_term = dpdr * scale_p;
    const R dpdthe_term = dpdthe * scale_p * inv_rm;
    const R dpdphi_term = dpdphi * scale_p * inv_rmsinthe;

    // eddy viscosity (planet/RHS_Uran_Turb.cpp:436-441)
    const R nue = FZ(ATU_NUE);
    const R nue_t = ((R)C.turb_coupling != R(0.0) && isfinite(nue))
                        ? (R)C.turb_coupling * (R(0.0) < nue ? nue : R(0.0)) : R(0.0);
    const R nue_t_s = nue_t / (R)C.Pr_t;
    const R re = (R)C.re;

    // ---- temperature (:280-296, :298, :333-334, :540-557) ----
    if constexpr (PART == 0 || PART == 1 || PART == 3) {
        const D7<R> T = ST(ATU_T);
        R pressure_t = (R)C.coeff_energy_p * scale_p * (u0 * dpdr + v_invrm * dpdthe + w_invrs * dpdphi);
        if ((C.cp_mode != 0 || C.cp_gas != 0) && T.c > R(0.0))
            pressure_t *= (R)(C.cp_mix / cp_lapse_dev(a, (double)T.c * C.t_ref));
        const R transport_t = u0 * T.d1r + v_invrm * T.d1t + w_invrs * T.d1p;
        const R diffusion_t = T.d2r + T.d1r * two_inv_rm + T.d2t * inv_rm2 + T.d1t * cirs + T.d2p * inv_rm2sinthe2;
        R precip_t = (R)C.precip_coupling * FZ(ATU_Q_PRECIP) * (R)C.precip_num / (R)C.precip_den;
        if (!isfinite(precip_t)) precip_t = R(0.0);
        else if (precip_t > R(0.5)) precip_t = R(0.5);
        else if (precip_t < R(-0.5)) precip_t = R(-0.5);
        const R rhs_t = +pressure_t - transport_t + diffusion_t / (R)C.re_pr + diffusion_t * nue_t_s + precip_t;
        fold<STAGE, CL_T>(a, ATU_T, c, Y0(ATU_T, T.c), (double)rhs_t);
    }
    if constexpr (PART == 1) return;

    // ---- radial velocity u (:303-304, :336-338, :723-732) ----
    if constexpr (PART != 4) {
        const R transport_u = u0 * U.d1r + v_invrm * U.d1t + w_invrs * U.d1p;
        const R diffusion_u = U.d2r + R(2.0) * u0 * inv_rm2 + U.d2t * inv_rm2 + R(4.0) * U.d1r * inv_rm
                              + U.d1t * cirs + U.d2p * inv_rm2sinthe2;
        const R rhs_u = -dpdr_term - transport_u + diffusion_u / re + diffusion_u * nue_t - sponge * u0;
        fold<STAGE, CL_NONE>(a, ATU_U, c, Y0(ATU_U, u0), (double)rhs_u);
        if (STAGE == 3) {
            nt_st(a.aux + c, (S)(rhs_u + dpdr_term));
            if (a.rhs3) nt_st(a.rhs3 + c, (double)rhs_u);
        }
    }
    const R cor_sc = (R)C.Coriolis * (R)C.scale_Cor;
    const double* ph = a.phydro;
    if constexpr (PART != 3) {
        // ---- meridional velocity v (:305-306, :339-343, :719, :759-765) ----
        {
            const R transport_v = u0 * V.d1r + v_invrm * V.d1t + w_invrs * V.d1p
                                  + (u0 * v0 - w0 * w0 * cotanthe) * inv_rm;
            const R diffusion_v = V.d2r + V.d1r * two_inv_rm + V.d2t * inv_rm2 + V.d1t * cirs - v_metric * v0
                                  + V.d2p * inv_rm2sinthe2 + R(2.0) * U.d1t * inv_rm2
                                  - W.d1p * R(2.0) * costhe * inv_rm2sinthe2;
            const R dphdthe_term = fp64_diff<R>(ph[c + ldk], ph[c - ldk], G.inv_2dthe) * inv_rm;
            const R Coriolis_the = (R)C.m_two_omega * costhe * w0;
            const R rhs_v = -dpdthe_term - dphdthe_term - transport_v + diffusion_v / re + diffusion_v * nue_t
                            - cor_sc * Coriolis_the;
            fold<STAGE, CL_NONE>(a, ATU_V, c, Y0(ATU_V, v0), (double)rhs_v);
            if (STAGE == 3) {
                nt_st(a.aux + fs + c, (S)(rhs_v + dpdthe_term));
                if (a.rhs3) nt_st(a.rhs3 + fs + c, (double)rhs_v);
            }
        }
        // ---- zonal velocity w (:307-308, :344-348, :720, :780-788) ----
        {
            const R transport_w = u0 * W.d1r + v_invrm * W.d1t + w_invrs * W.d1p
                                  + (w0 * u0 + v0 * w0 * cotanthe) * inv_rm;
            const R diffusion_w = W.d2r + W.d1r * two_inv_rm + W.d2t * inv_rm2 + W.d1t * cirs - v_metric * w0
                                  + W.d2p * inv_rm2sinthe2 + R(2.0) * U.d1p * inv_rm2sinthe
                                  + V.d1p * R(2.0) * costhe * inv_rm2sinthe2;
            const R dphdphi_term = fp64_diff<R>(ph[ce], ph[cw], G.inv_2dphi) * inv_rmsinthe;
            const R Coriolis_phi = (R)C.two_omega * (R(1.0) * costhe * v0 + sinthe_bf * u0);
            const R rhs_w = -dpdphi_term - dphdphi_term - transport_w + diffusion_w / re + diffusion_w * nue_t
                            - cor_sc * Coriolis_phi;
            fold<STAGE, CL_NONE>(a, ATU_W, c, Y0(ATU_W, w0), (double)rhs_w);
            if (STAGE == 3) {
                nt_st(a.aux + 2 * fs + c, (S)(rhs_w + dpdphi_term));
                if (a.rhs3) nt_st(a.rhs3 + 2 * fs + c, (double)rhs_w);
            }
        }

    }
    if constexpr (PART == 4) return;

    // ---- species (:310-326, :350-379, :790-913) ----
    const R pc = (R)C.precip_coupling, cr = (R)C.chemical_reaction;
    auto species = [&](int f, R sc_re, int sp, int mf) {
        const D7<R> X = ST(f);
        const R transport = u0 * X.d1r + v_invrm * X.d1t + w_invrs * X.d1p;
        const R diffusion = X.d2r + X.d1r * two_inv_rm + X.d2t * inv_rm2 + X.d1t * cirs + X.d2p * inv_rm2sinthe2;
        R rhs = -transport + diffusion / sc_re + diffusion * nue_t_s;
        if (mf >= 0) rhs = rhs + cr * FZ(mf);
        rhs = rhs + pc * FZ(sp);
        fold<STAGE, CL_NONE>(a, f, c, Y0(f, X.c), (double)rhs);
    };
    species(ATU_CH4, (R)C.sc_re_ch4, ATU_SP_CH4, -1);
    species(ATU_CH4_CLOUD, (R)C.sc_re_ch4, ATU_SP_CH4_CLOUD, -1);
    species(ATU_CH4_ICE, (R)C.sc_re_ch4, ATU_SP_CH4_ICE, -1);
    species(ATU_H2O, (R)C.sc_re_h2o, ATU_SP_H2O, -1);
    species(ATU_H2O_CLOUD, (R)C.sc_re_h2o, ATU_SP_H2O_CLOUD, -1);
    species(ATU_H2O_ICE, (R)C.sc_re_h2o, ATU_SP_H2O_ICE, -1);
    species(ATU_H2S, (R)C.sc_re_h2s, ATU_SP_H2S, ATU_MF_H2S);
    species(ATU_H2S_CLOUD, (R)C.sc_re_h2s, ATU_SP_H2S_CLOUD, -1);
    species(ATU_H2S_ICE, (R)C.sc_re_h2s, ATU_SP_H2S_ICE, -1);
    species(ATU_NH3, (R)C.sc_re_nh3, ATU_SP_NH3, ATU_MF_NH3);
    species(ATU_NH3_CLOUD, (R)C.sc_re_nh3, ATU_SP_NH3_CLOUD, -1);
    species(ATU_NH3_ICE, (R)C.sc_re_nh3, ATU_SP_NH3_ICE, -1);
    {   // NH4SH: flux limiter, reaction and Stokes settling (:852-913)
        const D7<R> X = ST(ATU_NH4SH);
        const R transport = u0 * X.d1r + v_invrm * X.d1t + w_invrs * X.d1p;
        const R diffusion = X.d2r + X.d1r * two_inv_rm + X.d2t * inv_rm2 + X.d1t * cirs + X.d2p * inv_rm2sinthe2;
        const R v_stokes = (R)C.stokes_c * ((R)C.rho_cond_nh4sh - FZ(ATU_RHO_MIX)) * (R)C.g / (R)C.mue_mix;
        const R rhs = -transport + FZ(ATU_FLUXLIM_NH4SH) + diffusion / (R)C.sc_re_nh4sh + diffusion * nue_t_s
                      + cr * FZ(ATU_MF_NH4SH) + (v_stokes / (R)C.u_0) * X.d1r;
        fold<STAGE, CL_NONE>(a, ATU_NH4SH, c, Y0(ATU_NH4SH, X.c), (double)rhs);
    }

    // ---- k* and dis* transport (:989-1018); untouched with the closure off ----
    if (C.turb_active) {
        const R inv_re = R(1.0) / re;
        {
            const D7<R> X = ST(ATU_TKE);
            const R transport = u0 * X.d1r + v0 * X.d1t * inv_rm + w0 * X.d1p * inv_rmsinthe;
            const R diffusion = X.d2r + X.d1r * R(2.0) * inv_rm + X.d2t * inv_rm2 + X.d1t * cirs
                                + X.d2p * inv_rm2sinthe2;
            const R rhs = -transport + diffusion * (inv_re + nue / (R)C.inv_sig_k) + FZ(ATU_TKE_SOURCE);
            fold<STAGE, CL_TKE>(a, ATU_TKE, c, Y0(ATU_TKE, X.c), (double)rhs);
        }
        {
            const D7<R> X = ST(ATU_DIS);
            const R transport = u0 * X.d1r + v0 * X.d1t * inv_rm + w0 * X.d1p * inv_rmsinthe;
            const R diffusion = X.d2r + X.d1r * R(2.0) * inv_rm + X.d2t * inv_rm2 + X.d1t * cirs
                                + X.d2p * inv_rm2sinthe2;
            const R rhs = -transport + diffusion * (inv_re + nue / (R)C.inv_sig_w) + FZ(ATU_DIS_SOURCE);
            fold<STAGE, CL_DIS>(a, ATU_DIS, c, Y0(ATU_DIS, X.c), (double)rhs);
        }
    }
}

// FP64 storage: FP64 arithmetic. FP32 storage (the mixed path): FP32 arithmetic, except on the
// rows next to the poles where the sin(theta) floor of the metric is active, which run the same
// stencil in FP64 arithmetic (j is uniform across a wavefront, so the branch does not diverge).
template <typename S, int STAGE, int PART>
__device__ __forceinline__ void cell(const StageArgs<S>& a, int i, int j, int k) {
    if constexpr (sizeof(S) == sizeof(double)) {
        rk_cell<double, S, STAGE, PART>(a, i, j, k);
    } else {
        if (a.geo.pole_row[j] != 0.0)
            rk_cell<double, S, STAGE, PART>(a, i, j, k);
        else
            rk_cell<float, S, STAGE, PART>(a, i, j, k);
    }
}

// One thread per interior cell; a block is 64 k x TJ j x TI i (launched as 64 x 4 x 1: taller
// blocks, 4x4, 2x4, 4x2 and 8x2, cut the fetched bytes by up to 11% but ran 1-6% slower).
// BAND 0: every interior row, the arithmetic picked per row; BAND 1: rows [jb0, jb1) in FP32;
// BAND 2: the rows outside [jb0, jb1) in FP64 (blockIdx.y counts the rows of both bands).
template <typename S, int STAGE, int TJ, int TI, int BAND, int PART>
__global__ void __launch_bounds__(64 * TJ * TI) k_rk_stage(const StageArgs<S> a) {
    const int nk = a.sh.km - 1;
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int r = blockIdx.y * TJ + threadIdx.y;
    const int i = 1 + blockIdx.z * TI + threadIdx.z;
    if (k >= nk || i > a.sh.im - 2) return;
    if constexpr (BAND == 0) {
        const int j = 1 + r;
        if (j > a.sh.jm - 2) return;
        cell<S, STAGE, PART>(a, i, j, k);
    } else if constexpr (BAND == 1) {
        const int j = a.jb0 + r;
        if (j >= a.jb1) return;
        rk_cell<float, S, STAGE, PART>(a, i, j, k);
    } else {
        const int nlo = a.jb0 - 1;
        const int j = r < nlo ? 1 + r : a.jb1 + (r - nlo);
        if (j > a.sh.jm - 2) return;
        rk_cell<double, S, STAGE, PART>(a, i, j, k);
    }
}

template <typename S, int TJ, int TI, int BAND>
hipError_t launch_rows(int stage, int part, int rows, const StageArgs<S>& a, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    const int nk = a.sh.km - 1;
    const dim3 block(64, TJ, TI);
    const dim3 grid((nk + 63) / 64, (rows + TJ - 1) / TJ, (a.sh.im - 2 + TI - 1) / TI);
    switch (stage) {
        case 0:
            if (part == 3)
                hipLaunchKernelGGL((k_rk_stage<S, 0, TJ, TI, BAND, 3>), grid, block, 0, s, a);
            else if (part == 4)
                hipLaunchKernelGGL((k_rk_stage<S, 0, TJ, TI, BAND, 4>), grid, block, 0, s, a);
            else
                hipLaunchKernelGGL((k_rk_stage<S, 0, TJ, TI, BAND, 0>), grid, block, 0, s, a);
            break;
        case 1: hipLaunchKernelGGL((k_rk_stage<S, 1, TJ, TI, BAND, 0>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_rk_stage<S, 2, TJ, TI, BAND, 0>), grid, block, 0, s, a); break;
        default:
            if (part == 1)
                hipLaunchKernelGGL((k_rk_stage<S, 3, TJ, TI, BAND, 1>), grid, block, 0, s, a);
            else if (part == 2)
                hipLaunchKernelGGL((k_rk_stage<S, 3, TJ, TI, BAND, 2>), grid, block, 0, s, a);
            else
                hipLaunchKernelGGL((k_rk_stage<S, 3, TJ, TI, BAND, 0>), grid, block, 0, s, a);
            break;
    }
    return hipGetLastError();
}

// FP64 storage: one launch. FP32 storage with banded pole rows: the FP32 rows and the FP64 bands as
// two launches (each kernel holds one arithmetic path, so neither carries the other's registers).
template <typename S>
hipError_t launch_stage(int stage, int part, const StageArgs<S>& a, hipStream_t s) {
    const int interior = a.sh.jm - 2;
    if constexpr (sizeof(S) == sizeof(double)) {
        return launch_rows<S, 4, 1, 0>(stage, part, interior, a, s);
    } else {
        if (a.jb0 < 1 || a.jb1 < a.jb0 || a.jb1 > a.sh.jm - 1)
            return launch_rows<S, 4, 1, 0>(stage, part, interior, a, s);
        hipError_t e = launch_rows<S, 4, 1, 1>(stage, part, a.jb1 - a.jb0, a, s);
        if (e != hipSuccess) return e;
        return launch_rows<S, 4, 1, 2>(stage, part, interior - (a.jb1 - a.jb0), a, s);
    }
}

}  // namespace

hipError_t launch_rk_stage_f64(int stage, const StageArgs<double>& a, hipStream_t s, int part) {
    return launch_stage<double>(stage, part, a, s);
}
hipError_t launch_rk_stage_mixed(int stage, const StageArgs<float>& a, hipStream_t s, int part) {
    return launch_stage<float>(stage, part, a, s);
}

}  // namespace atu
