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
// caches (FP64 stage 0.4885 -> 0.4735 ms for the ring's loads).

#include "atu_kernels.h"

namespace atu {
namespace {

enum { CL_NONE = 0, CL_T = 1, CL_TKE = 2, CL_DIS = 3, CL_SP = 4 };

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

// Fold one tendency into the accumulator and form the next input (pass B) or the step (final). The
// source `nb` supplies the accumulator's value and the store targets. A species (CL_SP) at the final
// update with `floor` set takes clampNegativeSpecies' value (planet/BC_Uran.h:110-128): not >= 0, NaN
// included, becomes 0.
template <int STAGE, int CLAMP, typename S, typename NB>
__device__ __forceinline__ void fold(const StageArgs<S>& a, NB& nb, int f, size_t c, double y0, double r,
                                     bool floor = false) {
    const size_t q = (size_t)f * a.sh.fs + c;
    if (STAGE < 3) {
        if (STAGE == 0)
            nt_st(nb.out(a.acc + q), 1.0 * r);
        else
            nt_st(nb.out(a.acc + q), nb.acc(a, f, q) + 2.0 * r);
        const double cin = (STAGE == 2) ? a.c.c_full : a.c.c_half;
        double y = y0 + cin * r;
        if (CLAMP == CL_TKE) y = y < 0.0 ? 0.0 : (y > a.c.tke_max ? a.c.tke_max : y);
        if (CLAMP == CL_DIS) y = (y < a.c.dis_min) ? a.c.dis_min : y;
        nt_st(nb.out(a.xnext + q), (S)y);
    } else {
        const double s = nb.acc(a, f, q) + 1.0 * r;
        double y = y0 + a.c.dt * s / 6.0;
        if (CLAMP == CL_T) y = y < a.c.t_min ? a.c.t_min : (y > a.c.t_max ? a.c.t_max : y);
        if (CLAMP == CL_TKE) {
            y = (y < 0.0) ? 0.0 : y;
            y = y < 0.0 ? 0.0 : (y > a.c.tke_max ? a.c.tke_max : y);
        }
        if (CLAMP == CL_DIS) y = (y < a.c.dis_min) ? a.c.dis_min : y;
        if (CLAMP == CL_SP && floor) y = (y >= 0.0) ? y : 0.0;
        nt_st(nb.out(a.yout + q), y);
        if (a.ymirror) nt_st(nb.out(a.ymirror + q), (S)y);
    }
}

// cp_lapse(T) / cp_mix ratio input: CpH2::Lapse::cp (planet/CpH2.h:88-93,128-134) on the
// host-built H2 table: the table interval of T, then the interpolation on it.
__device__ __forceinline__ double cp_x(double T) {
    const double Tc = (T < 5.0 ? 5.0 : T);
    return ((Tc > 1500.0 ? 1500.0 : Tc) - 5.0) / 0.5;
}
__device__ __forceinline__ int cp_i(double x, int n) {
    const int ix = (int)x;
    return (n - 2 < ix) ? n - 2 : ix;
}
__device__ __forceinline__ double cp_value(const StepConsts& c, double x, int i, double t0, double t1) {
    const double f = x - i;
    const double ch = t0 + f * (t1 - t0);
    return (c.r_h2 * ch + c.cp_a + c.cp_b) / c.cp_den;
}
template <typename S>
__device__ __forceinline__ double cp_lapse_dev(const StageArgs<S>& a, double T) {
    const double* t = a.geo.cp_tab;
    const double x = cp_x(T);
    const int i = cp_i(x, a.geo.cp_n);
    return cp_value(a.c, x, i, t[i], t[i + 1]);
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

// The metric factors of one row (i, j), read once per wave before any load or store of the stage (so
// that they come through the scalar cache); the 1/sin(theta) factors carry their FP64 values too.
template <typename R>
struct CellGeo {
    R inv_rm, inv_rm2, two_inv_rm, sponge, costhe, sinthe_bf;
    MF<R> cot, irs, ir2s, ir2s2, cirs, vmet;
};
template <typename R, typename S>
__device__ __forceinline__ CellGeo<R> cell_geo(const StageArgs<S>& a, int i, int j) {
    const DevGeo& G = a.geo;
    const size_t ij = (size_t)i * a.sh.jm + j;
    CellGeo<R> g;
    g.inv_rm = (R)G.inv_rm[i];
    g.inv_rm2 = (R)G.inv_rm2[i];
    g.two_inv_rm = (R)G.two_inv_rm[i];
    g.sponge = (R)G.sponge[i];
    g.costhe = (R)G.costhe[j];
    g.sinthe_bf = (R)G.sinthe_bf[j];
    g.cot = MF<R>{(R)G.cotanthe[j], G.cotanthe[j]};
    g.irs = MF<R>{(R)G.inv_rmsinthe[ij], G.inv_rmsinthe[ij]};
    g.ir2s = MF<R>{(R)G.inv_rm2sinthe[ij], G.inv_rm2sinthe[ij]};
    g.ir2s2 = MF<R>{(R)G.inv_rm2sinthe2[ij], G.inv_rm2sinthe2[ij]};
    g.cirs = MF<R>{(R)G.costhe_inv_rm2sinthe[ij], G.costhe_inv_rm2sinthe[ij]};
    g.vmet = MF<R>{(R)G.v_metric[ij], G.v_metric[ij]};
    return g;
}

// The tendencies of one interior cell (i, j, k) in arithmetic precision R, folded in FP64; `c` is the
// cell's offset in a field. The source NB supplies every value the cell reads:
//   nb.st(f): the D7 of prognostic field f;
//   nb.pd_r(), pd_t(), pd_p(): p_dyn's neighbour pairs (i+1, i-1), (j+1, j-1), (k+1, k-1), FP64;
//   nb.dph_t(), dph_p(): the FP64 differences of p_hydro's polar and zonal neighbour pairs;
//   nb.y0<STAGE>(a, f, c, centre): y_n of the fold (the next stage input of the FP32 path is rounded
//     to FP32 anyway, so it is formed from y_n's FP32 mirror; at stage 0 the stencil centre itself);
//   nb.fz(a, f, c): frozen input f; nb.xc(a, f, c): the stage input at the cell;
//   nb.acc(a, f, q), nb.out(p): the accumulator's value and the store targets (inside fold).
// nb.item<N>() marks where the values of the source's N-th group are read (a no-op for sources that
// read global memory directly); every value of group N is read before item<N + 1>().
// `pole`: the row's sin(theta) floor is active (FP32 arithmetic forms its 1/sin factors in FP64).
// The cell's tendencies fall in two groups that read nothing of each other: tend_dyn (t, u, v, w and
// aux: the velocity and t stencils, p_dyn, p_hydro) and tend_sca (the 13 species, tke, dis: their own
// stencils and the velocities at the cell).
template <typename R, typename S, int STAGE, typename NB>
__device__ __forceinline__ void tend_dyn(const StageArgs<S>& a, const CellGeo<R>& g, size_t c, NB& nb, bool pole) {
    const DevGeo& G = a.geo;
    const size_t fs = a.sh.fs;
    const R inv_rm = g.inv_rm, inv_rm2 = g.inv_rm2, two_inv_rm = g.two_inv_rm;
    const R sponge = g.sponge;
    const R costhe = g.costhe, sinthe_bf = g.sinthe_bf;
    const MF<R> cot = g.cot, irs = g.irs, ir2s = g.ir2s, ir2s2 = g.ir2s2, cirs = g.cirs, vmet = g.vmet;
    const StepConsts& C = a.c;

    // velocities and the pressure gradient
    nb.template item<0>();
    const D7<R> U = nb.st(ATU_U);
    nb.template item<1>();
    const D7<R> V = nb.st(ATU_V);
    nb.template item<2>();
    const D7<R> W = nb.st(ATU_W);
    const R u0 = U.c, v0 = V.c, w0 = W.c;
    // p_dyn (|p| ~ 20) varies slowly: its differences are formed in FP64 and rounded once to R,
    // since an FP32 difference of two such values keeps few significant bits
    nb.template item<3>();
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
    const R nue = nb.fz(a, ATU_NUE, c);
    const R nue_t = ((R)C.turb_coupling != R(0.0) && isfinite(nue))
                        ? (R)C.turb_coupling * (R(0.0) < nue ? nue : R(0.0)) : R(0.0);
    const R nue_t_s = nue_t / (R)C.Pr_t;
    const R re = (R)C.re;

    // ---- temperature (:280-296, :298, :333-334, :540-557) ----
    {
        nb.template item<4>();
        const D7<R> T = nb.st(ATU_T);
        // cp(T) of the pressure work: the source may start its table reads here and finish them below
        const bool cp_on = (C.cp_mode != 0 || C.cp_gas != 0) && T.c > R(0.0);
        if (cp_on) nb.cp_request(a, (double)T.c * C.t_ref);
        const R transport_t = u0 * T.d1r + v_invrm * T.d1t + w_invrs * T.d1p;
        const R diffusion_t = T.d2r + T.d1r * two_inv_rm + T.d2t * inv_rm2 + mf(T.d1t, cirs, pole)
                              + mf(T.d2p, ir2s2, pole);
        nb.template item<5>();
        R precip_t = (R)C.precip_coupling * nb.fz(a, ATU_Q_PRECIP, c) * (R)C.precip_num / (R)C.precip_den;
        if (!isfinite(precip_t)) precip_t = R(0.0);
        else if (precip_t > R(0.5)) precip_t = R(0.5);
        else if (precip_t < R(-0.5)) precip_t = R(-0.5);
        R pressure_t = (R)C.coeff_energy_p * scale_p * (u0 * dpdr + v_invrm * dpdthe + w_invrs * dpdphi);
        if (cp_on) pressure_t *= (R)(C.cp_mix / nb.cp_lapse(a, (double)T.c * C.t_ref));
        const R rhs_t = +pressure_t - transport_t + diffusion_t / (R)C.re_pr + diffusion_t * nue_t_s + precip_t;
        fold<STAGE, CL_T>(a, nb, ATU_T, c, nb.template y0<STAGE>(a, ATU_T, c, T.c), (double)rhs_t);
    }

    // ---- radial velocity u (:303-304, :336-338, :723-732) ----
    {
        const R transport_u = u0 * U.d1r + v_invrm * U.d1t + w_invrs * U.d1p;
        const R diffusion_u = U.d2r + R(2.0) * u0 * inv_rm2 + U.d2t * inv_rm2 + R(4.0) * U.d1r * inv_rm
                              + mf(U.d1t, cirs, pole) + mf(U.d2p, ir2s2, pole);
        const R rhs_u = -dpdr_term - transport_u + diffusion_u / re + diffusion_u * nue_t - sponge * u0;
        fold<STAGE, CL_NONE>(a, nb, ATU_U, c, nb.template y0<STAGE>(a, ATU_U, c, u0), (double)rhs_u);
        if (STAGE == 3) {
            nt_st(nb.out(a.aux + c), (S)(rhs_u + dpdr_term));
            if (a.rhs3) nt_st(nb.out(a.rhs3 + c), (double)rhs_u);
        }
    }
    const R cor_sc = (R)C.Coriolis * (R)C.scale_Cor;
    // ---- meridional velocity v (:305-306, :339-343, :719, :759-765) ----
    {
        const R transport_v = u0 * V.d1r + v_invrm * V.d1t + w_invrs * V.d1p + (u0 * v0 - mf(w0 * w0, cot, pole)) * inv_rm;
        const R diffusion_v = V.d2r + V.d1r * two_inv_rm + V.d2t * inv_rm2 + mf(V.d1t, cirs, pole) - mf(v0, vmet, pole)
                              + mf(V.d2p, ir2s2, pole) + R(2.0) * U.d1t * inv_rm2
                              - mf(W.d1p * R(2.0) * costhe, ir2s2, pole);
        const R dphdthe_term = (R)(nb.dph_t() * G.inv_2dthe) * inv_rm;
        const R Coriolis_the = (R)C.m_two_omega * costhe * w0;
        const R rhs_v = -dpdthe_term - dphdthe_term - transport_v + diffusion_v / re + diffusion_v * nue_t
                        - cor_sc * Coriolis_the;
        fold<STAGE, CL_NONE>(a, nb, ATU_V, c, nb.template y0<STAGE>(a, ATU_V, c, v0), (double)rhs_v);
        if (STAGE == 3) {
            nt_st(nb.out(a.aux + fs + c), (S)(rhs_v + dpdthe_term));
            if (a.rhs3) nt_st(nb.out(a.rhs3 + fs + c), (double)rhs_v);
        }
    }
    // ---- zonal velocity w (:307-308, :344-348, :720, :780-788) ----
    {
        const R transport_w = u0 * W.d1r + v_invrm * W.d1t + w_invrs * W.d1p + (w0 * u0 + mf(v0 * w0, cot, pole)) * inv_rm;
        const R diffusion_w = W.d2r + W.d1r * two_inv_rm + W.d2t * inv_rm2 + mf(W.d1t, cirs, pole) - mf(w0, vmet, pole)
                              + mf(W.d2p, ir2s2, pole) + mf(R(2.0) * U.d1p, ir2s, pole)
                              + mf(V.d1p * R(2.0) * costhe, ir2s2, pole);
        const R dphdphi_term = mf((R)(nb.dph_p() * G.inv_2dphi), irs, pole);
        const R Coriolis_phi = (R)C.two_omega * (R(1.0) * costhe * v0 + sinthe_bf * u0);
        const R rhs_w = -dpdphi_term - dphdphi_term - transport_w + diffusion_w / re + diffusion_w * nue_t
                        - cor_sc * Coriolis_phi;
        fold<STAGE, CL_NONE>(a, nb, ATU_W, c, nb.template y0<STAGE>(a, ATU_W, c, w0), (double)rhs_w);
        if (STAGE == 3) {
            nt_st(nb.out(a.aux + 2 * fs + c), (S)(rhs_w + dpdphi_term));
            if (a.rhs3) nt_st(nb.out(a.rhs3 + 2 * fs + c), (double)rhs_w);
        }
    }
}

template <typename R, typename S, int STAGE, typename NB>
__device__ __forceinline__ void tend_sca(const StageArgs<S>& a, const CellGeo<R>& g, size_t c, NB& nb, bool pole,
                                         bool floor) {
    const R inv_rm = g.inv_rm, inv_rm2 = g.inv_rm2, two_inv_rm = g.two_inv_rm;
    const MF<R> irs = g.irs, ir2s2 = g.ir2s2, cirs = g.cirs;
    const StepConsts& C = a.c;

    // the velocities at the cell (the stage input) and the eddy viscosity
    const R u0 = nb.xc(a, ATU_U, c), v0 = nb.xc(a, ATU_V, c);
    const R w0 = nb.xc(a, ATU_W, c);
    const R v_invrm = v0 * inv_rm;
    const R w_invrs = mf(w0, irs, pole);
    const R nue = nb.fz(a, ATU_NUE, c);
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
        if (mfl >= 0) rhs = rhs + cr * nb.fz(a, mfl, c);
        rhs = rhs + pc * nb.fz(a, sp, c);
        fold<STAGE, CL_SP>(a, nb, f, c, nb.template y0<STAGE>(a, f, c, X.c), (double)rhs, floor);
        };
    nb.template item<6>();
    species(ATU_CH4, (R)C.sc_re_ch4, ATU_SP_CH4, -1);
    nb.template item<7>();
    species(ATU_CH4_CLOUD, (R)C.sc_re_ch4, ATU_SP_CH4_CLOUD, -1);
    nb.template item<8>();
    species(ATU_CH4_ICE, (R)C.sc_re_ch4, ATU_SP_CH4_ICE, -1);
    nb.template item<9>();
    species(ATU_H2O, (R)C.sc_re_h2o, ATU_SP_H2O, -1);
    nb.template item<10>();
    species(ATU_H2O_CLOUD, (R)C.sc_re_h2o, ATU_SP_H2O_CLOUD, -1);
    nb.template item<11>();
    species(ATU_H2O_ICE, (R)C.sc_re_h2o, ATU_SP_H2O_ICE, -1);
    nb.template item<12>();
    species(ATU_H2S, (R)C.sc_re_h2s, ATU_SP_H2S, ATU_MF_H2S);
    nb.template item<13>();
    species(ATU_H2S_CLOUD, (R)C.sc_re_h2s, ATU_SP_H2S_CLOUD, -1);
    nb.template item<14>();
    species(ATU_H2S_ICE, (R)C.sc_re_h2s, ATU_SP_H2S_ICE, -1);
    nb.template item<15>();
    species(ATU_NH3, (R)C.sc_re_nh3, ATU_SP_NH3, ATU_MF_NH3);
    nb.template item<16>();
    species(ATU_NH3_CLOUD, (R)C.sc_re_nh3, ATU_SP_NH3_CLOUD, -1);
    nb.template item<17>();
    species(ATU_NH3_ICE, (R)C.sc_re_nh3, ATU_SP_NH3_ICE, -1);
    {   // NH4SH: flux limiter, reaction and Stokes settling (:852-913)
        nb.template item<18>();
        const D7<R> X = nb.st(ATU_NH4SH);
        const R transport = u0 * X.d1r + v_invrm * X.d1t + w_invrs * X.d1p;
        const R diffusion = X.d2r + X.d1r * two_inv_rm + X.d2t * inv_rm2 + mf(X.d1t, cirs, pole)
                            + mf(X.d2p, ir2s2, pole);
        const R v_stokes = (R)C.stokes_c * ((R)C.rho_cond_nh4sh - nb.fz(a, ATU_RHO_MIX, c)) * (R)C.g / (R)C.mue_mix;
        const R rhs = -transport + nb.fz(a, ATU_FLUXLIM_NH4SH, c) + diffusion / (R)C.sc_re_nh4sh + diffusion * nue_t_s
                      + cr * nb.fz(a, ATU_MF_NH4SH, c) + (v_stokes / (R)C.u_0) * X.d1r;
        fold<STAGE, CL_SP>(a, nb, ATU_NH4SH, c, nb.template y0<STAGE>(a, ATU_NH4SH, c, X.c), (double)rhs, floor);
    }

    // ---- k* and dis* transport (:989-1018); untouched with the closure off ----
    const R inv_re = R(1.0) / re;
    nb.template item<19>();
    if (C.turb_active) {
        const D7<R> X = nb.st(ATU_TKE);
        const R transport = u0 * X.d1r + v0 * X.d1t * inv_rm + mf(w0 * X.d1p, irs, pole);
        const R diffusion = X.d2r + X.d1r * R(2.0) * inv_rm + X.d2t * inv_rm2 + mf(X.d1t, cirs, pole)
                            + mf(X.d2p, ir2s2, pole);
        const R rhs = -transport + diffusion * (inv_re + nue / (R)C.inv_sig_k) + nb.fz(a, ATU_TKE_SOURCE, c);
        fold<STAGE, CL_TKE>(a, nb, ATU_TKE, c, nb.template y0<STAGE>(a, ATU_TKE, c, X.c), (double)rhs);
    }
    nb.template item<20>();
    if (C.turb_active) {
        const D7<R> X = nb.st(ATU_DIS);
        const R transport = u0 * X.d1r + v0 * X.d1t * inv_rm + mf(w0 * X.d1p, irs, pole);
        const R diffusion = X.d2r + X.d1r * R(2.0) * inv_rm + X.d2t * inv_rm2 + mf(X.d1t, cirs, pole)
                            + mf(X.d2p, ir2s2, pole);
        const R rhs = -transport + diffusion * (inv_re + nue / (R)C.inv_sig_w) + nb.fz(a, ATU_DIS_SOURCE, c);
        fold<STAGE, CL_DIS>(a, nb, ATU_DIS, c, nb.template y0<STAGE>(a, ATU_DIS, c, X.c), (double)rhs);
    }
}

// ================================ the stage kernels ================================
// One thread per interior cell; a block is 64 consecutive k (one wavefront per row) by 4 rows j at one
// level i, so every row segment is coalesced and the per-(i, j) geometry is wave-uniform. Two sources
// of the values: NbGlobal reads them straight from global memory (the mixed path), NbRing through a
// per-wave LDS ring (the FP64 path).
template <typename R, typename S>
struct NbGlobal {
    const S* xs;
    const double* pd;
    const double* ph;
    size_t c, cw, ce, plane, fs;
    int ldk;
    Inv<R> v;
    template <int N>
    __device__ __forceinline__ void item() const {}
    __device__ __forceinline__ D7<R> st(int f) const {
        const S* p = xs + (size_t)f * fs;
        return d7<R>((R)p[c], (R)p[c + plane], (R)p[c - plane], (R)p[c + ldk], (R)p[c - ldk], (R)p[ce], (R)p[cw], v);
    }
    __device__ __forceinline__ double2 pd_r() const { return make_double2(pd[c + plane], pd[c - plane]); }
    __device__ __forceinline__ double2 pd_t() const { return make_double2(pd[c + ldk], pd[c - ldk]); }
    __device__ __forceinline__ double2 pd_p() const { return make_double2(pd[ce], pd[cw]); }
    __device__ __forceinline__ double dph_t() const { return ph[c + ldk] - ph[c - ldk]; }
    __device__ __forceinline__ double dph_p() const { return ph[ce] - ph[cw]; }
    template <int STAGE>
    __device__ __forceinline__ double y0(const StageArgs<S>& a, int f, size_t cc, R centre) const {
        if (STAGE == 0) return (double)centre;
        if (STAGE < 3 && sizeof(S) == sizeof(float)) return (double)nt_ld(a.yn_lo + (size_t)f * fs + cc);
        return nt_ld(a.yn + (size_t)f * fs + cc);
    }
    __device__ __forceinline__ R fz(const StageArgs<S>& a, int f, size_t cc) const {
        return (R)nt_ld(a.frozen + (size_t)(f - ATU_NUE) * fs + cc);
    }
    __device__ __forceinline__ R xc(const StageArgs<S>& a, int f, size_t cc) const { return (R)a.xs[(size_t)f * fs + cc]; }
    __device__ __forceinline__ double acc(const StageArgs<S>& a, int, size_t q) const { return nt_ld(a.acc + q); }
    template <typename T>
    __device__ __forceinline__ T* out(T* p) const { return p; }
    __device__ __forceinline__ void cp_request(const StageArgs<S>&, double) const {}
    __device__ __forceinline__ double cp_lapse(const StageArgs<S>& a, double T) const { return cp_lapse_dev(a, T); }
};

// ---- the LDS-ring source (FP64) ----
// One wave owns 64 consecutive k of one row (i, j). Every array value the wave's cells read comes into a
// per-wave ring of two LDS slots by global_load_lds_dwordx4 (16 bytes per lane, 1 KiB per
// instruction, no VGPRs held while in flight): while the tendencies of one group of arrays are formed
// from one slot, the next group's loads fill the other. A slot is five segments of 128 doubles; a
// segment is filled by one instruction, lanes 0-31 loading one 64-cell row segment (two cells per
// lane) and lanes 32-63 another:
//   R1: the rows at i+1 | i-1;  R2: the rows at j+1 | j-1;
//   C: the row itself (cells 0-63), the zonal halo (cell 65: k0-1 on the ring of km-1 meridians; cell 66:
//      the east neighbour of the segment's last cell), the other lanes repeat the row;
//   P0, P1: two once-touched rows at the cells each (y_n | accumulator, or two frozen inputs).
// Groups, in the order the tendencies read them (stages 1-3; stage 0 reads neither y_n nor the
// accumulator):
//   0 U: R1 R2 C, P0 (nue | nue)   1 V: R1 R2 C   2 W: R1 R2 C
//   3 pressures: p_dyn R1 R2 C, p_hydro R2 C   4 T: R1 R2 C (+ P0 (q_precip | q_precip) at stages 1-3)
//   5 (y_n | acc) of t, u, v, w in R1 R2 C P0 (stage 0: P0 (q_precip | q_precip)); P1 the cp table pairs
//   6-17 the twelve species: R1 R2 C, P0 (sp | mf of H2S and NH3, mf_nh4sh at 17, else sp), P1 (y_n | acc)
//   18 NH4SH: R1 R2 C, P0 (rho_mix | fluxlim), P1 (y_n | acc)
//   19 TKE: R1 R2 C, P0 (tke_source | tke_source), P1 (y_n | acc)
//   20 DIS: R1 R2 C, P0 (y_n | acc), P1 (dis_source | dis_source)
// Values a later group needs (nue, mf_nh4sh, q_precip, the p_hydro differences, u, v, w at the cell)
// are read into registers while their group is current.
// The compiler does not order LDS reads after these loads, so the ring waits explicitly: entering
// group N, the wave issues group N+1 into the other slot (after s_waitcnt lgkmcnt(0), so that the
// slot's earlier LDS reads have completed) and then waits with s_waitcnt vmcnt(n), n = the number of
// group N+1's load instructions: vector memory operations complete in issue order, so group N's loads
// are then done. Any other vector memory operation the compiler places after group N's loads (stores,
// the cp table loads) only raises the count in flight, so n stays a safe bound; the asm statements
// carry memory clobbers, so no load, store or LDS read moves across them.
constexpr int kRingSeg = 128;
constexpr int kRingSlot = 5 * kRingSeg;   // doubles
constexpr int kRingGroups = 21;
enum { SEG_R1 = 0, SEG_R2 = kRingSeg, SEG_C = 2 * kRingSeg, SEG_P0 = 3 * kRingSeg, SEG_P1 = 4 * kRingSeg };

// Load instructions of group g at stage STAGE.
template <int STAGE>
__host__ __device__ constexpr int ring_ops(int g) {
    return g == 0 ? 4 : g <= 2 ? 3 : g == 3 ? 5 : g == 4 ? (STAGE > 0 ? 4 : 3) : g == 5 ? (STAGE > 0 ? 4 : 1)
         : g <= 20 ? (STAGE > 0 ? 5 : 4) : 0;
}

template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N <= 5, "ring groups hold at most five loads");
    if constexpr (N == 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    else if constexpr (N == 1) asm volatile("s_waitcnt vmcnt(1)" ::: "memory");
    else if constexpr (N == 2) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
    else if constexpr (N == 3) asm volatile("s_waitcnt vmcnt(3)" ::: "memory");
    else if constexpr (N == 4) asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(5)" ::: "memory");
}

typedef __attribute__((address_space(3))) void* lds_void_ptr;
template <int CPOL = 0>
__device__ __forceinline__ void ring_load(const double* g, double* l) {
    __builtin_amdgcn_global_load_lds((const void*)g, (lds_void_ptr)l, 16, 0, CPOL);
}

// Invalid lanes (k >= km-1 in a row's last segment) form values from whatever the ring holds and store
// them here instead of into a field.
__device__ double g_ring_sink[64];

template <int STAGE>
struct NbRing {
    const StageArgs<double>& a;
    double* ring;            // this wave's two slots
    const double* cur;       // the current group's slot
    int o_p, d_r1, d_r2, o_c;   // per lane: the pair's offset in a field, the R1 / R2 steps, C's offset
    int lane, iw, ie;        // the lane and its LDS indices of the zonal neighbours in segment C
    bool hi, valid;
    Inv<double> v;
    double nue_v, mf4_v, qp_v, u0_v, v0_v, w0_v, dpht_v, dphp_v;

    __device__ __forceinline__ double* slot(int g) const { return ring + (g & 1) * kRingSlot; }
    __device__ __forceinline__ const double* fld(const double* base, int f) const { return base + (size_t)f * a.sh.fs; }
    __device__ __forceinline__ void stencil(const double* p, double* s) const {
        ring_load(p + (o_p + d_r1), s + SEG_R1);
        ring_load(p + (o_p + d_r2), s + SEG_R2);
        ring_load(p + o_c, s + SEG_C);
    }
    // (lo | hi) pair of rows at the cells
    __device__ __forceinline__ void pair(const double* lo, const double* hi_p, double* s) const {
        ring_load<2>((hi ? hi_p : lo) + o_p, s);
    }
    __device__ __forceinline__ void ya(int f, double* s) const { pair(fld(a.yn, f), fld(a.acc, f), s); }
    __device__ __forceinline__ const double* fzp(int f) const { return fld(a.frozen, f - ATU_NUE); }

    template <int G>
    __device__ __forceinline__ void issue() const {
        double* s = slot(G);
        constexpr bool Y = STAGE > 0;
        if constexpr (G == 0) {
            stencil(fld(a.xs, ATU_U), s);
            pair(fzp(ATU_NUE), fzp(ATU_NUE), s + SEG_P0);
        } else if constexpr (G == 1) {
            stencil(fld(a.xs, ATU_V), s);
        } else if constexpr (G == 2) {
            stencil(fld(a.xs, ATU_W), s);
        } else if constexpr (G == 3) {
            stencil(a.pdyn, s);
            ring_load(a.phydro + (o_p + d_r2), s + SEG_P0);
            ring_load(a.phydro + o_c, s + SEG_P1);
        } else if constexpr (G == 4) {
            stencil(fld(a.xs, ATU_T), s);
            if constexpr (Y) pair(fzp(ATU_Q_PRECIP), fzp(ATU_Q_PRECIP), s + SEG_P0);
        } else if constexpr (G == 5) {
            if constexpr (Y) {
                ya(ATU_T, s + SEG_R1);
                ya(ATU_U, s + SEG_R2);
                ya(ATU_V, s + SEG_C);
                ya(ATU_W, s + SEG_P0);
            } else {
                pair(fzp(ATU_Q_PRECIP), fzp(ATU_Q_PRECIP), s + SEG_P0);
            }
        } else if constexpr (G <= 17) {
            constexpr int f = ATU_CH4 + (G - 6);
            constexpr int sp = ATU_SP_CH4 + (G - 6);
            constexpr int hf = f == ATU_H2S ? ATU_MF_H2S : f == ATU_NH3 ? ATU_MF_NH3 : f == ATU_NH3_ICE ? ATU_MF_NH4SH : sp;
            stencil(fld(a.xs, f), s);
            pair(fzp(sp), fzp(hf), s + SEG_P0);
            if constexpr (Y) ya(f, s + SEG_P1);
        } else if constexpr (G == 18) {
            stencil(fld(a.xs, ATU_NH4SH), s);
            pair(fzp(ATU_RHO_MIX), fzp(ATU_FLUXLIM_NH4SH), s + SEG_P0);
            if constexpr (Y) ya(ATU_NH4SH, s + SEG_P1);
        } else if constexpr (G == 19) {
            stencil(fld(a.xs, ATU_TKE), s);
            pair(fzp(ATU_TKE_SOURCE), fzp(ATU_TKE_SOURCE), s + SEG_P0);
            if constexpr (Y) ya(ATU_TKE, s + SEG_P1);
        } else {
            stencil(fld(a.xs, ATU_DIS), s);
            if constexpr (Y) ya(ATU_DIS, s + SEG_P0);
            pair(fzp(ATU_DIS_SOURCE), fzp(ATU_DIS_SOURCE), s + SEG_P1);
        }
    }

    template <int G>
    __device__ __forceinline__ void item() {
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        if constexpr (G + 1 < kRingGroups) {
            issue<G + 1>();
            wait_vm<ring_ops<STAGE>(G + 1)>();
        } else {
            wait_vm<0>();
        }
        cur = slot(G);
        if constexpr (G == 0) {
            nue_v = cur[SEG_P0 + lane];
            u0_v = cur[SEG_C + lane];
        } else if constexpr (G == 1) {
            v0_v = cur[SEG_C + lane];
        } else if constexpr (G == 2) {
            w0_v = cur[SEG_C + lane];
        } else if constexpr (G == 3) {
            dpht_v = cur[SEG_P0 + lane] - cur[SEG_P0 + 64 + lane];
            dphp_v = cur[SEG_P1 + ie] - cur[SEG_P1 + iw];
        } else if constexpr (G == 4) {
            if constexpr (STAGE > 0) qp_v = cur[SEG_P0 + lane];
        } else if constexpr (G == 5) {
            if constexpr (STAGE == 0) qp_v = cur[SEG_P0 + lane];
        } else if constexpr (G == 17) {
            mf4_v = cur[SEG_P0 + 64 + lane];
        }
    }

    __device__ __forceinline__ D7<double> st(int) const {
        const double* s = cur;
        return d7<double>(s[SEG_C + lane], s[SEG_R1 + lane], s[SEG_R1 + 64 + lane], s[SEG_R2 + lane],
                          s[SEG_R2 + 64 + lane], s[SEG_C + ie], s[SEG_C + iw], v);
    }
    __device__ __forceinline__ double2 pd_r() const { return make_double2(cur[SEG_R1 + lane], cur[SEG_R1 + 64 + lane]); }
    __device__ __forceinline__ double2 pd_t() const { return make_double2(cur[SEG_R2 + lane], cur[SEG_R2 + 64 + lane]); }
    __device__ __forceinline__ double2 pd_p() const { return make_double2(cur[SEG_C + ie], cur[SEG_C + iw]); }
    __device__ __forceinline__ double dph_t() const { return dpht_v; }
    __device__ __forceinline__ double dph_p() const { return dphp_v; }
    // the (y_n | acc) segment of field f in the current group
    __device__ __forceinline__ int ya_seg(int f) const {
        return f == ATU_T ? SEG_R1 : f == ATU_U ? SEG_R2 : f == ATU_V ? SEG_C : f == ATU_W ? SEG_P0
             : f == ATU_DIS ? SEG_P0 : SEG_P1;
    }
    template <int ST>
    __device__ __forceinline__ double y0(const StageArgs<double>&, int f, size_t, double centre) const {
        if (ST == 0) return centre;
        return cur[ya_seg(f) + lane];
    }
    __device__ __forceinline__ double acc(const StageArgs<double>&, int f, size_t) const {
        return cur[ya_seg(f) + 64 + lane];
    }
    __device__ __forceinline__ double fz(const StageArgs<double>&, int f, size_t) const {
        if (f == ATU_NUE) return nue_v;
        if (f == ATU_MF_NH4SH) return mf4_v;
        if (f == ATU_Q_PRECIP) return qp_v;
        if (f == ATU_DIS_SOURCE) return cur[SEG_P1 + lane];
        // the low row of the current group's frozen pair, or the high one (mf of H2S / NH3, the flux limiter)
        const bool high = f == ATU_MF_H2S || f == ATU_MF_NH3 || f == ATU_FLUXLIM_NH4SH;
        return cur[SEG_P0 + (high ? 64 : 0) + lane];
    }
    __device__ __forceinline__ double xc(const StageArgs<double>&, int f, size_t) const {
        return f == ATU_U ? u0_v : f == ATU_V ? v0_v : w0_v;
    }
    template <typename T>
    __device__ __forceinline__ T* out(T* p) const { return valid ? p : (T*)(g_ring_sink + lane); }
    // The cp table pair (t[i], t[i + 1]) of the cell, requested during group 4 as four 4-byte loads per
    // lane into segment P1 of group 5's slot (free in every stage: group 5 fills R1 R2 C P0 at most),
    // and read when group 5 is current (the loads precede group 6's, so item<5>'s wait covers them).
    // Only lanes with cp_on request; only they read.
    __device__ __forceinline__ void cp_request(const StageArgs<double>& aa, double T) const {
        const int i = cp_i(cp_x(T), aa.geo.cp_n);
        const char* src = (const char*)(aa.geo.cp_tab + i);
        unsigned* d = (unsigned*)(slot(5) + SEG_P1);
        __builtin_amdgcn_global_load_lds((const void*)(src + 0), (lds_void_ptr)(d + 0), 4, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(src + 4), (lds_void_ptr)(d + 64), 4, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(src + 8), (lds_void_ptr)(d + 128), 4, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(src + 12), (lds_void_ptr)(d + 192), 4, 0, 0);
    }
    __device__ __forceinline__ double cp_lapse(const StageArgs<double>& aa, double T) const {
        const double x = cp_x(T);
        const int i = cp_i(x, aa.geo.cp_n);
        const unsigned* d = (const unsigned*)(cur + SEG_P1);
        const double t0 = __hiloint2double((int)d[64 + lane], (int)d[lane]);
        const double t1 = __hiloint2double((int)d[192 + lane], (int)d[128 + lane]);
        return cp_value(aa.c, x, i, t0, t1);
    }
};

template <typename R, typename S>
__device__ __forceinline__ Inv<R> inv_of(const StageArgs<S>& a) {
    const DevGeo& G = a.geo;
    return Inv<R>{(R)G.inv_2dr, (R)G.inv_dr2, (R)G.inv_2dthe, (R)G.inv_dthe2, (R)G.inv_2dphi, (R)G.inv_dphi2};
}

// The rows at which the final update takes the species clamp (StageArgs::sp_floor): those the BC
// neither writes nor reads (its NEUMANN forms read i = 1, 2, im-3, im-2 and j = 1, 2, jm-3, jm-2;
// the seam copy takes k = 0 of the row, which this clamp has then already covered).
template <int STAGE, typename S>
__device__ __forceinline__ bool sp_floor_row(const StageArgs<S>& a, int i, int j) {
    return STAGE == 3 && a.sp_floor && i >= 3 && i <= a.sh.im - 4 && j >= 3 && j <= a.sh.jm - 4;
}

// The FP64 stage with the ring source: a block is 64 k x 4 rows at one level (one wave per row), each
// wave with its own ring (2 slots x 5 KiB), no barriers. Needs km-1 and ldk even (the halo and row
// segments are loaded as 16-byte pairs); launch_rk_stage_f64 checks that.
template <int STAGE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) k_rk_stage_ring(const StageArgs<double> a) {
    __shared__ double lds[4 * 2 * kRingSlot];
    const int nk = a.sh.km - 1;
    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int k0 = blockIdx.x * 64;
    const int j = 1 + blockIdx.y * 4 + w;
    const int i = 1 + blockIdx.z;
    if (j > a.sh.jm - 2) return;   // the whole wave: no load of the ring is issued
    const int ldk = a.sh.ldk;
    const size_t row = ((size_t)i * a.sh.jm + j) * ldk;
    const int k = k0 + lane;
    const bool valid = k < nk;
    const bool hi = lane >= 32;
    const int m = lane & 31;
    const int col = min(k0 + 2 * m, ldk - 2);            // a 16-byte pair inside the row
    const int wk0 = k0 == 0 ? nk - 2 : k0 - 2;           // the pair ending at the west neighbour of k0
    const int ek0 = k0 + 64 >= nk ? 0 : k0 + 64;         // the pair starting at the east neighbour
    NbRing<STAGE> nb{a};
    nb.ring = lds + (size_t)w * 2 * kRingSlot;
    nb.cur = nb.ring;
    nb.o_p = (int)(row + col);
    nb.d_r1 = hi ? -(int)a.sh.plane : (int)a.sh.plane;
    nb.d_r2 = hi ? -ldk : ldk;
    nb.o_c = (int)row + (lane == 32 ? wk0 : lane == 33 ? ek0 : (hi ? min(k0 + 2 * (m - 2), ldk - 2) : col));
    nb.lane = lane;
    nb.hi = hi;
    nb.valid = valid;
    nb.iw = lane == 0 ? 65 : lane - 1;
    nb.ie = (lane == 63 || k == nk - 1) ? 66 : lane + 1;
    nb.v = inv_of<double>(a);
    const size_t c = row + (valid ? k : 0);
    const bool pole = a.geo.pole_row[j] != 0.0;
    const CellGeo<double> g = cell_geo<double>(a, i, j);
    const bool floor = sp_floor_row<STAGE>(a, i, j);
    nb.template issue<0>();
    tend_dyn<double, double, STAGE>(a, g, c, nb, pole);
    tend_sca<double, double, STAGE>(a, g, c, nb, pole, floor);
}

// The mixed (FP32 storage) stage with the direct source: a block is 64 k x 4 rows at one level, one
// thread per cell. The arrays that are never written through another pointer are passed as restrict
// kernel arguments (the stage input, the next stage input, the accumulator, the frozen inputs, the
// pressures, aux), so that the compiler may issue a field's loads ahead of the previous field's
// stores (0.335 -> 0.325 ms per stage). y_n is left out: the last stage writes y_{n+1} over it.
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
    NbGlobal<R, S> nb{a.xs, a.pdyn, a.phydro, c, cw, ce, a.sh.plane, a.sh.fs, ldk, inv_of<R>(a)};
    const CellGeo<R> g = cell_geo<R>(a, i, j);
    tend_dyn<R, S, STAGE>(a, g, c, nb, pole);
    tend_sca<R, S, STAGE>(a, g, c, nb, pole, sp_floor_row<STAGE>(a, i, j));
}

constexpr int kTJ = 4;

}  // namespace

hipError_t launch_rk_stage_f64(int stage, const StageArgs<double>& a, hipStream_t s) {
    const int nk = a.sh.km - 1;
    // the ring loads 16-byte pairs; atu_create accepts only an even ring of meridians, ldk is a
    // multiple of 32
    if ((nk & 1) || (a.sh.ldk & 1)) return hipErrorInvalidValue;
    const dim3 block(64, 4, 1);
    const dim3 grid((nk + 63) / 64, (a.sh.jm - 2 + 3) / 4, a.sh.im - 2);
    switch (stage) {
        case 0: hipLaunchKernelGGL(k_rk_stage_ring<0>, grid, block, 0, s, a); break;
        case 1: hipLaunchKernelGGL(k_rk_stage_ring<1>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_rk_stage_ring<2>, grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL(k_rk_stage_ring<3>, grid, block, 0, s, a); break;
    }
    return hipGetLastError();
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
