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
// The arrays a stage touches once per cell (the frozen inputs, y_n, the accumulator, the outputs) are
// read and written with non-temporal hints, so that they do not push the stencil rows out of the
// caches.

#include "atu_kernels.h"

namespace atu {
namespace {

enum { CL_NONE = 0, CL_T = 1, CL_TKE = 2, CL_DIS = 3 };

template <typename R>
struct D7 {
    R c, d1r, d2r, d1t, d2t, d1p, d2p;
};

template <typename R>
struct Inv {
    R i2dr, idr2, i2dt, idt2, i2dp, idp2;
};

// Central first and second differences in r, theta, phi from the seven values of the stencil
// (planet/RHS_Uran_Turb.cpp:93-176): rp, rn the radial neighbours (i+1, i-1), tp, tn the polar ones
// (j+1, j-1), pe, pw the zonal ones (k+1, k-1 on the ring of km-1 meridians).
template <typename R>
__device__ __forceinline__ D7<R> d7(R x0, R rp, R rn, R tp, R tn, R pe, R pw, const Inv<R>& v) {
    D7<R> d;
    d.c = x0;
    d.d1r = (rp - rn) * v.i2dr;
    d.d2r = (rp - R(2.0) * x0 + rn) * v.idr2;
    d.d1t = (tp - tn) * v.i2dt;
    d.d2t = (tp - R(2.0) * x0 + tn) * v.idt2;
    d.d1p = (pe - pw) * v.i2dp;
    d.d2p = (pe - R(2.0) * x0 + pw) * v.idp2;
    return d;
}

// (a - b) * s in FP64, rounded once to R. For R = double this is the reference's (a - b) * s.
template <typename R>
__device__ __forceinline__ R fp64_diff(double a, double b, double s) {
    return (R)((a - b) * s);
}

// Streams touched once per launch (the accumulator, y_n, the frozen inputs, the next stage input,
// the step output, aux) bypass the caches' retention.
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

// A metric factor that carries 1/sin(theta) (1/(rm sin), 1/(rm^2 sin), 1/(rm^2 sin^2),
// cos/(rm^2 sin), cot, the v metric): x * f in R, except in FP32 arithmetic on a row where the
// sin(theta) floor of the metric is active, where the product is formed in FP64 from the FP64 factor
// and rounded once. For R = double both are the reference's x * f.
template <typename R>
struct MF {
    R f;
    double fd;
};
template <typename R>
__device__ __forceinline__ R mf(R x, const MF<R>& m, bool pole) {
    if constexpr (sizeof(R) == sizeof(double)) {
        return x * m.f;
    } else {
        return pole ? (R)((double)x * m.fd) : x * m.f;
    }
}

// The tendencies of one interior cell (i, j, k) in arithmetic precision R, folded in FP64; `c` is the
// cell's offset in a field, NB supplies the stencils:
//   nb.st(f): the D7 of prognostic field f;
//   nb.pd_r(), pd_t(), pd_p(): p_dyn's neighbour pairs (i+1, i-1), (j+1, j-1), (k+1, k-1), FP64;
//   nb.ph_t(), ph_p(): p_hydro's polar and zonal neighbour pairs, FP64.
// `pole`: the row's sin(theta) floor is active (FP32 arithmetic forms its 1/sin factors in FP64).
// The cell's tendencies fall in two groups that read nothing of each other: tend_dyn (t, u, v, w and
// aux: the velocity and t stencils, p_dyn, p_hydro) and tend_sca (the 13 species, tke, dis: their own
// stencils and the velocities at the cell).
template <typename R, typename S, int STAGE, typename NB>
__device__ __forceinline__ void tend_dyn(const StageArgs<S>& a, int i, int j, size_t c, const NB& nb, bool pole) {
    const DevGeo& G = a.geo;
    const size_t fs = a.sh.fs;
    const size_t ij = (size_t)i * a.sh.jm + j;
    const R inv_rm = (R)G.inv_rm[i], inv_rm2 = (R)G.inv_rm2[i], two_inv_rm = (R)G.two_inv_rm[i];
    const R sponge = (R)G.sponge[i];
    const R costhe = (R)G.costhe[j], sinthe_bf = (R)G.sinthe_bf[j];
    const MF<R> cot{(R)G.cotanthe[j], G.cotanthe[j]};
    const MF<R> irs{(R)G.inv_rmsinthe[ij], G.inv_rmsinthe[ij]};
    const MF<R> ir2s{(R)G.inv_rm2sinthe[ij], G.inv_rm2sinthe[ij]};
    const MF<R> ir2s2{(R)G.inv_rm2sinthe2[ij], G.inv_rm2sinthe2[ij]};
    const MF<R> cirs{(R)G.costhe_inv_rm2sinthe[ij], G.costhe_inv_rm2sinthe[ij]};
    const MF<R> vmet{(R)G.v_metric[ij], G.v_metric[ij]};
    const StepConsts& C = a.c;

    // y_n of the fold: the next stage input of the FP32 path is rounded to FP32 anyway, so it is
    // formed from y_n's FP32 mirror (at stage 0 the stencil centre itself); FP64 y_n elsewhere
    auto Y0 = [&](int f, R center) -> double {
        if (STAGE == 0) return (double)center;
        if (STAGE < 3 && sizeof(S) == sizeof(float)) return (double)nt_ld(a.yn_lo + (size_t)f * fs + c);
        return nt_ld(a.yn + (size_t)f * fs + c);
    };
    auto FZ = [&](int f) -> R { return (R)nt_ld(a.frozen + (size_t)(f - ATU_NUE) * fs + c); };

    // velocities and the pressure gradient
    const D7<R> U = nb.st(ATU_U), V = nb.st(ATU_V), W = nb.st(ATU_W);
    const R u0 = U.c, v0 = V.c, w0 = W.c;
    // p_dyn (|p| ~ 20) varies slowly: its differences are formed in FP64 and rounded once to R,
    // since an FP32 difference of two such values keeps few significant bits
    const double2 pr = nb.pd_r(), pt = nb.pd_t(), pp = nb.pd_p();
    const R dpdr = fp64_diff<R>(pr.x, pr.y, G.inv_2dr);
    const R dpdthe = fp64_diff<R>(pt.x, pt.y, G.inv_2dthe);
    const R dpdphi = fp64_diff<R>(pp.x, pp.y, G.inv_2dphi);
    const R scale_p = (R)C.scale_p;
    const R v_invrm = v0 * inv_rm;
    const R w_invrs = mf(w0, irs, pole);
    const R dpdr_term = dpdr * scale_p;
    const R dpdthe_term = dpdthe * scale_p * inv_rm;
    const R dpdphi_term = mf(dpdphi * scale_p, irs, pole);

    // eddy viscosity (planet/RHS_Uran_Turb.cpp:436-441)
    const R nue = FZ(ATU_NUE);
    const R nue_t = ((R)C.turb_coupling != R(0.0) && isfinite(nue))
                        ? (R)C.turb_coupling * (R(0.0) < nue ? nue : R(0.0)) : R(0.0);
    const R nue_t_s = nue_t / (R)C.Pr_t;
    const R re = (R)C.re;

    // ---- temperature (:280-296, :298, :333-334, :540-557) ----
    {
        const D7<R> T = nb.st(ATU_T);
        R pressure_t = (R)C.coeff_energy_p * scale_p * (u0 * dpdr + v_invrm * dpdthe + w_invrs * dpdphi);
        if ((C.cp_mode != 0 || C.cp_gas != 0) && T.c > R(0.0))
            pressure_t *= (R)(C.cp_mix / cp_lapse_dev(a, (double)T.c * C.t_ref));
        const R transport_t = u0 * T.d1r + v_invrm * T.d1t + w_invrs * T.d1p;
        const R diffusion_t = T.d2r + T.d1r * two_inv_rm + T.d2t * inv_rm2 + mf(T.d1t, cirs, pole)
                              + mf(T.d2p, ir2s2, pole);
        R precip_t = (R)C.precip_coupling * FZ(ATU_Q_PRECIP) * (R)C.precip_num / (R)C.precip_den;
        if (!isfinite(precip_t)) precip_t = R(0.0);
        else if (precip_t > R(0.5)) precip_t = R(0.5);
        else if (precip_t < R(-0.5)) precip_t = R(-0.5);
        const R rhs_t = +pressure_t - transport_t + diffusion_t / (R)C.re_pr + diffusion_t * nue_t_s + precip_t;
        fold<STAGE, CL_T>(a, ATU_T, c, Y0(ATU_T, T.c), (double)rhs_t);
    }

    // ---- radial velocity u (:303-304, :336-338, :723-732) ----
    {
        const R transport_u = u0 * U.d1r + v_invrm * U.d1t + w_invrs * U.d1p;
        const R diffusion_u = U.d2r + R(2.0) * u0 * inv_rm2 + U.d2t * inv_rm2 + R(4.0) * U.d1r * inv_rm
                              + mf(U.d1t, cirs, pole) + mf(U.d2p, ir2s2, pole);
        const R rhs_u = -dpdr_term - transport_u + diffusion_u / re + diffusion_u * nue_t - sponge * u0;
        fold<STAGE, CL_NONE>(a, ATU_U, c, Y0(ATU_U, u0), (double)rhs_u);
        if (STAGE == 3) {
            nt_st(a.aux + c, (S)(rhs_u + dpdr_term));
            if (a.rhs3) nt_st(a.rhs3 + c, (double)rhs_u);
        }
    }
    const R cor_sc = (R)C.Coriolis * (R)C.scale_Cor;
    // ---- meridional velocity v (:305-306, :339-343, :719, :759-765) ----
    {
        const R transport_v = u0 * V.d1r + v_invrm * V.d1t + w_invrs * V.d1p + (u0 * v0 - mf(w0 * w0, cot, pole)) * inv_rm;
        const R diffusion_v = V.d2r + V.d1r * two_inv_rm + V.d2t * inv_rm2 + mf(V.d1t, cirs, pole) - mf(v0, vmet, pole)
                              + mf(V.d2p, ir2s2, pole) + R(2.0) * U.d1t * inv_rm2
                              - mf(W.d1p * R(2.0) * costhe, ir2s2, pole);
        const double2 h = nb.ph_t();
        const R dphdthe_term = fp64_diff<R>(h.x, h.y, G.inv_2dthe) * inv_rm;
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
        const R transport_w = u0 * W.d1r + v_invrm * W.d1t + w_invrs * W.d1p + (w0 * u0 + mf(v0 * w0, cot, pole)) * inv_rm;
        const R diffusion_w = W.d2r + W.d1r * two_inv_rm + W.d2t * inv_rm2 + mf(W.d1t, cirs, pole) - mf(w0, vmet, pole)
                              + mf(W.d2p, ir2s2, pole) + mf(R(2.0) * U.d1p, ir2s, pole)
                              + mf(V.d1p * R(2.0) * costhe, ir2s2, pole);
        const double2 h = nb.ph_p();
        const R dphdphi_term = mf(fp64_diff<R>(h.x, h.y, G.inv_2dphi), irs, pole);
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

template <typename R, typename S, int STAGE, typename NB>
__device__ __forceinline__ void tend_sca(const StageArgs<S>& a, int i, int j, size_t c, const NB& nb, bool pole) {
    const DevGeo& G = a.geo;
    const size_t fs = a.sh.fs;
    const size_t ij = (size_t)i * a.sh.jm + j;
    const R inv_rm = (R)G.inv_rm[i], inv_rm2 = (R)G.inv_rm2[i], two_inv_rm = (R)G.two_inv_rm[i];
    const MF<R> irs{(R)G.inv_rmsinthe[ij], G.inv_rmsinthe[ij]};
    const MF<R> ir2s2{(R)G.inv_rm2sinthe2[ij], G.inv_rm2sinthe2[ij]};
    const MF<R> cirs{(R)G.costhe_inv_rm2sinthe[ij], G.costhe_inv_rm2sinthe[ij]};
    const StepConsts& C = a.c;

    // y_n of the fold: the next stage input of the FP32 path is rounded to FP32 anyway, so it is
    // formed from y_n's FP32 mirror (at stage 0 the stencil centre itself); FP64 y_n elsewhere
    auto Y0 = [&](int f, R center) -> double {
        if (STAGE == 0) return (double)center;
        if (STAGE < 3 && sizeof(S) == sizeof(float)) return (double)nt_ld(a.yn_lo + (size_t)f * fs + c);
        return nt_ld(a.yn + (size_t)f * fs + c);
    };
    auto FZ = [&](int f) -> R { return (R)nt_ld(a.frozen + (size_t)(f - ATU_NUE) * fs + c); };

    // the velocities at the cell (the stage input) and the eddy viscosity
    const R u0 = (R)a.xs[(size_t)ATU_U * fs + c], v0 = (R)a.xs[(size_t)ATU_V * fs + c];
    const R w0 = (R)a.xs[(size_t)ATU_W * fs + c];
    const R v_invrm = v0 * inv_rm;
    const R w_invrs = mf(w0, irs, pole);
    const R nue = FZ(ATU_NUE);
    const R nue_t = ((R)C.turb_coupling != R(0.0) && isfinite(nue))
                        ? (R)C.turb_coupling * (R(0.0) < nue ? nue : R(0.0)) : R(0.0);
    const R nue_t_s = nue_t / (R)C.Pr_t;
    const R re = (R)C.re;

    // ---- species (:310-326, :350-379, :790-913) ----
    const R pc = (R)C.precip_coupling, cr = (R)C.chemical_reaction;
    auto species = [&](int f, R sc_re, int sp, int mfl) {
        const D7<R> X = nb.st(f);
        const R transport = u0 * X.d1r + v_invrm * X.d1t + w_invrs * X.d1p;
        const R diffusion = X.d2r + X.d1r * two_inv_rm + X.d2t * inv_rm2 + mf(X.d1t, cirs, pole)
                            + mf(X.d2p, ir2s2, pole);
        R rhs = -transport + diffusion / sc_re + diffusion * nue_t_s;
        if (mfl >= 0) rhs = rhs + cr * FZ(mfl);
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
        const D7<R> X = nb.st(ATU_NH4SH);
        const R transport = u0 * X.d1r + v_invrm * X.d1t + w_invrs * X.d1p;
        const R diffusion = X.d2r + X.d1r * two_inv_rm + X.d2t * inv_rm2 + mf(X.d1t, cirs, pole)
                            + mf(X.d2p, ir2s2, pole);
        const R v_stokes = (R)C.stokes_c * ((R)C.rho_cond_nh4sh - FZ(ATU_RHO_MIX)) * (R)C.g / (R)C.mue_mix;
        const R rhs = -transport + FZ(ATU_FLUXLIM_NH4SH) + diffusion / (R)C.sc_re_nh4sh + diffusion * nue_t_s
                      + cr * FZ(ATU_MF_NH4SH) + (v_stokes / (R)C.u_0) * X.d1r;
        fold<STAGE, CL_NONE>(a, ATU_NH4SH, c, Y0(ATU_NH4SH, X.c), (double)rhs);
    }

    // ---- k* and dis* transport (:989-1018); untouched with the closure off ----
    if (C.turb_active) {
        const R inv_re = R(1.0) / re;
        {
            const D7<R> X = nb.st(ATU_TKE);
            const R transport = u0 * X.d1r + v0 * X.d1t * inv_rm + mf(w0 * X.d1p, irs, pole);
            const R diffusion = X.d2r + X.d1r * R(2.0) * inv_rm + X.d2t * inv_rm2 + mf(X.d1t, cirs, pole)
                                + mf(X.d2p, ir2s2, pole);
            const R rhs = -transport + diffusion * (inv_re + nue / (R)C.inv_sig_k) + FZ(ATU_TKE_SOURCE);
            fold<STAGE, CL_TKE>(a, ATU_TKE, c, Y0(ATU_TKE, X.c), (double)rhs);
        }
            {
            const D7<R> X = nb.st(ATU_DIS);
            const R transport = u0 * X.d1r + v0 * X.d1t * inv_rm + mf(w0 * X.d1p, irs, pole);
            const R diffusion = X.d2r + X.d1r * R(2.0) * inv_rm + X.d2t * inv_rm2 + mf(X.d1t, cirs, pole)
                                + mf(X.d2p, ir2s2, pole);
            const R rhs = -transport + diffusion * (inv_re + nue / (R)C.inv_sig_w) + FZ(ATU_DIS_SOURCE);
            fold<STAGE, CL_DIS>(a, ATU_DIS, c, Y0(ATU_DIS, X.c), (double)rhs);
        }
    }
}

// ================================ the stage kernel ================================
// One thread per interior cell; a block is 64 consecutive k (one wavefront per row) by TJ rows j at one
// level i, so every field load is a coalesced row segment (512 B in FP64, 256 B in FP32) and the
// per-(i, j) geometry is wave-uniform. The stencil values come straight from global memory: the
// neighbouring rows and levels that other blocks also read are served by the caches.
template <typename R, typename S>
struct NbGlobal {
    const S* xs;
    const double* pd;
    const double* ph;
    size_t c, cw, ce, plane, fs;
    int ldk;
    Inv<R> v;
    __device__ __forceinline__ D7<R> st(int f) const {
        const S* p = xs + (size_t)f * fs;
        return d7<R>((R)p[c], (R)p[c + plane], (R)p[c - plane], (R)p[c + ldk], (R)p[c - ldk], (R)p[ce], (R)p[cw], v);
    }
    __device__ __forceinline__ double2 pd_r() const { return make_double2(pd[c + plane], pd[c - plane]); }
    __device__ __forceinline__ double2 pd_t() const { return make_double2(pd[c + ldk], pd[c - ldk]); }
    __device__ __forceinline__ double2 pd_p() const { return make_double2(pd[ce], pd[cw]); }
    __device__ __forceinline__ double2 ph_t() const { return make_double2(ph[c + ldk], ph[c - ldk]); }
    __device__ __forceinline__ double2 ph_p() const { return make_double2(ph[ce], ph[cw]); }
};

template <typename R, typename S>
__device__ __forceinline__ Inv<R> inv_of(const StageArgs<S>& a) {
    const DevGeo& G = a.geo;
    return Inv<R>{(R)G.inv_2dr, (R)G.inv_dr2, (R)G.inv_2dthe, (R)G.inv_dthe2, (R)G.inv_2dphi, (R)G.inv_dphi2};
}

template <typename S, int STAGE, int TJ>
__global__ void __launch_bounds__(64 * TJ) k_rk_stage(const StageArgs<S> a) {
    using R = S;   // FP64 storage: FP64 arithmetic; FP32 storage: FP32 arithmetic
    const int nk = a.sh.km - 1;
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = 1 + blockIdx.y * TJ + threadIdx.y;
    const int i = 1 + blockIdx.z;
    if (k >= nk || j > a.sh.jm - 2) return;
    const int ldk = a.sh.ldk;
    const size_t row = ((size_t)i * a.sh.jm + j) * ldk;
    const size_t c = row + k;
    const size_t cw = row + (k == 0 ? nk - 1 : k - 1);   // phi ring over km-1 meridians
    const size_t ce = row + (k == nk - 1 ? 0 : k + 1);
    const bool pole = a.geo.pole_row[j] != 0.0;
    const NbGlobal<R, S> nb{a.xs, a.pdyn, a.phydro, c, cw, ce, a.sh.plane, a.sh.fs, ldk, inv_of<R>(a)};
    tend_dyn<R, S, STAGE>(a, i, j, c, nb, pole);
    tend_sca<R, S, STAGE>(a, i, j, c, nb, pole);
}

// blocks of 64 k x 4 rows (taller blocks, 64 x 8 and 64 x 12, and the XCD-aware block orders that
// were measured against it ran slower)
// The same stage with the arrays that are never written through another pointer passed as restrict
// kernel arguments (the stage input, the next stage input, the accumulator, the frozen inputs, the
// pressures, aux), so that the compiler may issue a field's loads ahead of the previous field's
// stores. Measured: the FP32 stage 0.335 -> 0.325 ms, the FP64 stage 0.517 -> 0.530 ms (the FP64
// kernel's schedule gets longer), so only the mixed path launches it. y_n is left out: the last stage
// writes y_{n+1} over it.
template <typename S, int STAGE, int TJ>
__global__ void __launch_bounds__(64 * TJ) k_rk_stage_noalias(const StageArgs<S> a0, const S* __restrict__ xs,
                                                         S* __restrict__ xnext, double* __restrict__ acc,
                                                         const S* __restrict__ frozen,
                                                         const double* __restrict__ pdyn,
                                                         const double* __restrict__ phydro, S* __restrict__ aux) {
    using R = S;
    StageArgs<S> a = a0;
    a.xs = xs;
    a.xnext = xnext;
    a.acc = acc;
    a.frozen = frozen;
    a.pdyn = pdyn;
    a.phydro = phydro;
    a.aux = aux;
    const int nk = a.sh.km - 1;
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = 1 + blockIdx.y * TJ + threadIdx.y;
    const int i = 1 + blockIdx.z;
    if (k >= nk || j > a.sh.jm - 2) return;
    const int ldk = a.sh.ldk;
    const size_t row = ((size_t)i * a.sh.jm + j) * ldk;
    const size_t c = row + k;
    const size_t cw = row + (k == 0 ? nk - 1 : k - 1);
    const size_t ce = row + (k == nk - 1 ? 0 : k + 1);
    const bool pole = a.geo.pole_row[j] != 0.0;
    const NbGlobal<R, S> nb{a.xs, a.pdyn, a.phydro, c, cw, ce, a.sh.plane, a.sh.fs, ldk, inv_of<R>(a)};
    tend_dyn<R, S, STAGE>(a, i, j, c, nb, pole);
    tend_sca<R, S, STAGE>(a, i, j, c, nb, pole);
}

constexpr int kTJ = 4;

template <typename S>
hipError_t launch_stage(int stage, const StageArgs<S>& a, hipStream_t s) {
    const int nk = a.sh.km - 1;
    const dim3 block(64, kTJ, 1);
    const dim3 grid((nk + 63) / 64, (a.sh.jm - 2 + kTJ - 1) / kTJ, a.sh.im - 2);
    switch (stage) {
        case 0: hipLaunchKernelGGL((k_rk_stage<S, 0, kTJ>), grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_rk_stage<S, 1, kTJ>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_rk_stage<S, 2, kTJ>), grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL((k_rk_stage<S, 3, kTJ>), grid, block, 0, s, a); break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_rk_stage_f64(int stage, const StageArgs<double>& a, hipStream_t s) {
    return launch_stage<double>(stage, a, s);
}
hipError_t launch_rk_stage_mixed(int stage, const StageArgs<float>& a, hipStream_t s) {
    const int nk = a.sh.km - 1;
    const dim3 block(64, kTJ, 1);
    const dim3 grid((nk + 63) / 64, (a.sh.jm - 2 + kTJ - 1) / kTJ, a.sh.im - 2);
#define ATU_NOALIAS(ST)                                                                                          \
    hipLaunchKernelGGL((k_rk_stage_noalias<float, ST, kTJ>), grid, block, 0, s, a, a.xs, a.xnext, a.acc, a.frozen, \
                       a.pdyn, a.phydro, a.aux)
    switch (stage) {
        case 0: ATU_NOALIAS(0); break;
        case 1: ATU_NOALIAS(1); break;
        case 2: ATU_NOALIAS(2); break;
        default: ATU_NOALIAS(3); break;
    }
#undef ATU_NOALIAS
    return hipGetLastError();
}

}  // namespace atu
