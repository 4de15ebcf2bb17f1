// The turbulence closure Turbulence<cUranusModel> (planet/Turbulence.h) on the device: the eddy
// viscosity, the production and the k / omega (or k / epsilon) sources from u, v, w, tke, dis.
//
// The reference runs seven passes (run(), planet/Turbulence.h:306-332): the friction velocity, a
// clamp pass over the interior, the source loop, a clamp-out pass, the wall copy at i = 0, the
// eddy-viscosity clamp over every cell, the seam copy and the lid fill. Here:
//   k_turb_vel_star   compute_vel_star (:336-379)
//   k_turb_sources    the clamp pass applied to every tke / dis value as it is read (the clamp is
//                     idempotent, so a neighbour read before or after its own clamp gives the same
//                     value), the source loop (:657-869), the clamp-out of nue and the two sources
//                     (:636-647) and clamp_nue of nue; in run() the threads of i = 1 also write the
//                     wall copy of i = 0 (:414-418), which only they read
//   k_turb_clamp      the clamp-out of tke and dis (deferred: neighbours read them in the loop) and
//                     clamp_nue of the interior columns' end cells
//   k_turb_edge_wall, k_turb_edge_clamp   the columns outside the loop (the pole rows and the seam
//                     meridian): the wall copy, then clamp_nue
//   k_turb_seam, k_turb_lid   seam_copy (:542-561) and fill_lid_corners (:510-536)
// init() (:282-303) adds k_turb_init (init_fields, :426-495), a seam copy and a wall copy before
// the sources, and takes no wall copy after them.
//
// Every expression keeps the reference's operand order (the library is built without FP
// contraction), so the results are the reference's bits wherever the transcendental functions
// it calls (pow, tanh, exp with arguments away from their trivial ranges) are not involved; the
// default configuration (k-omega SST, TURB_FREE 1) is bit-exact.

#include "atu_kernels.h"

namespace atu {
namespace {

__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }   // std::max
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }   // std::min
__device__ __forceinline__ double clamp3(double x, double lo, double hi) {
    return (x < lo) ? lo : ((x > hi) ? hi : x);
}

constexpr double kDisMin = 1.0e-10;    // Turbulence::dis_min
constexpr double kCnue = 0.028;        // Turbulence::C_nue

// The clamp pass (:633-635) of an interior cell, applied on read.
__device__ __forceinline__ double tke_in(double x) { return smax(0.0, x); }
__device__ __forceinline__ double dis_in(double x) { return smax(kDisMin, x); }

__global__ void __launch_bounds__(256) k_turb_vel_star(Shape sh, TurbArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= sh.km) return;
    const double vel_star_min = 0.05;
    double vs;
    if (a.free_flow) {
        vs = vel_star_min;
    } else {
        const size_t q = sh.plane + (size_t)j * sh.ldk + k;   // i_surf = min(surface_index + 1, im-1) = 1
        const double v_s = a.v[q], w_s = a.w[q];
        const double vw2 = v_s * v_s + w_s * w_s;
        if (!isfinite(vw2)) {
            vs = vel_star_min;
        } else {
            const double U_horiz = a.u_0 * sqrt(vw2);
            vs = clamp3(a.sqrt_cd * U_horiz, vel_star_min, a.vel_star_max);
        }
    }
    a.vel_star[(size_t)j * sh.ldk + k] = vs;
}

// init_fields (:426-495) over the interior; with TURB_FREE every cell takes the background values
// the host formed.
__global__ void __launch_bounds__(256) k_turb_init(Shape sh, TurbArgs a) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y + 1, i = blockIdx.z + 1;
    if (k >= sh.km - 1 || j > sh.jm - 2) return;
    const size_t q = (size_t)i * sh.plane + (size_t)j * sh.ldk + k;
    double tke, dis;
    if (a.free_flow) {
        tke = a.init_tke;
        dis = a.init_dis;
    } else {
        const double vs = a.vel_star[(size_t)j * sh.ldk + k];
        const double k_bg = 0.01 * vs * vs / a.sqrt_cnue;
        const double om_bg = a.c_m025 * sqrt(k_bg) / (a.karman_l);
        const double k_abl = vs * vs / a.sqrt_cnue * a.pz[i];
        const double k_phys = smax(k_abl, k_bg);
        double dis_nd;
        if (a.model == ATU_TURB_K_EPSILON) {
            const double eps_bg = a.c_075 * pow(k_bg, 1.5) / (a.karman_l);
            const double eps_phys = a.c_075 * pow(k_phys, 1.5) / a.karman_z[i];
            dis_nd = smax(eps_phys, eps_bg) * a.nd_eps;
        } else {
            const double om_phys = a.c_m025 * sqrt(k_phys) / a.karman_z[i];
            dis_nd = smax(om_phys, om_bg) * a.nd_omega;
        }
        tke = smax(0.0, k_phys * a.inv_u02);
        dis = smax(kDisMin, dis_nd);
    }
    a.tke[q] = tke;
    a.dis[q] = dis;
    a.tken[q] = tke;
    a.disn[q] = dis;
}

// apply_wall_bc's zero-gradient copy at i = 0 (:414-418), every column.
__global__ void __launch_bounds__(256) k_turb_wall(Shape sh, TurbArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= sh.km) return;
    const size_t q0 = (size_t)j * sh.ldk + k, q1 = sh.plane + q0;
    a.tke[q0] = a.tke[q1];
    a.dis[q0] = a.dis[q1];
    a.nue[q0] = a.nue[q1];
}

struct SrcOut {
    double nue, tke_src, dis_src;
};

// compute_k_omega_SST (:1020-1107) with no solid cell.
__device__ __forceinline__ SrcOut sst_cell(const TurbArgs& a, int i, double rm, double sinthe, double kc, double dc,
                                           double prod, double dtkedr, double ddisdr, double tke_jm1, double tke_jp1,
                                           double dis_jm1, double dis_jp1, double tke_km1, double tke_kp1,
                                           double dis_km1, double dis_kp1, double Omega) {
    const double a1 = 0.31, bet_star = 0.09, bet1 = 0.0333, bet2 = 0.0368, gam1 = 0.413, gam2 = 0.2, sig_w2 = 1.168;
    const double y_star = a.ystar_sst[i];
    const double rmsinthe = rm * sinthe;
    const double dtkedthe_neu = (tke_jp1 - tke_jm1) / (rm * 2.0 * a.dthe);
    const double ddisdthe_neu = (dis_jp1 - dis_jm1) / (rm * 2.0 * a.dthe);
    const double dtkedphi_neu = (tke_kp1 - tke_km1) / (rmsinthe * 2.0 * a.dphi);
    const double ddisdphi_neu = (dis_kp1 - dis_km1) / (rmsinthe * 2.0 * a.dphi);
    const double grad_dot = dtkedr * ddisdr + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu;
    const double CD_kw = smax(2.0 * sig_w2 / smax(dc, 1.0e-20) * grad_dot, 1.0e-20);
    const double dis_safe = smax(dc, 1.0e-20);
    const double arg1 = smin(smax(sqrt(kc) / (bet_star * dis_safe * y_star),
                                  500.0 * a.nue_air_nd / (y_star * y_star * dis_safe)),
                             4.0 * sig_w2 * kc / (CD_kw * y_star * y_star));
    const double arg2 = smax(2.0 * sqrt(kc) / (bet_star * dis_safe * y_star),
                             500.0 * a.nue_air_nd / (y_star * y_star * dis_safe));
    // F1 = tanh(arg1^4) enters only as F1 gam1 + (1 - F1) gam2, F1 bet1 + (1 - F1) bet2 and
    // 2 (1 - F1). Below arg1 = 2^-15, F1 <= arg1^4 < 2^-60 is under half an ulp of 1, of gam2 and of
    // bet2, so each of those rounds to its value at F1 = 0: the pow and the tanh are skipped there
    // (away from walls, e.g. the free-flow default, arg1 ~ 1e-16) with the same bits.
    const double F1 = (arg1 < 0x1p-15) ? 0.0 : tanh(pow(arg1, 4.0));
    // F2 = tanh(y), y = std::pow(arg2, 2), which g++ and clang both form as arg2 * arg2. F2 <= y, so
    // where Omega y is at most half of a1 dis_safe the larger of a1 dis_safe and Omega F2 is
    // a1 dis_safe whatever the last bit of F2: the tanh is skipped there.
    const double y2 = arg2 * arg2;
    const double lhs = a1 * dis_safe;
    const double rhs = (Omega * y2 <= 0.5 * lhs) ? 0.0 : Omega * tanh(y2);
    SrcOut o;
    o.nue = a1 * kc / smax(lhs, rhs);
    const double tke_safe = smax(kc, 1.0e-20);
    const double P_k = smin(prod, 20.0 * bet_star * tke_safe * dis_safe);
    const double Y_k = bet_star * tke_safe * dis_safe;
    const double P_w = (F1 * gam1 + (1.0 - F1) * gam2) * P_k * dis_safe / tke_safe;
    const double Y_w = (F1 * bet1 + (1.0 - F1) * bet2) * dis_safe * dis_safe;
    const double D_w = 2.0 * (1.0 - F1) * sig_w2 / dis_safe * grad_dot;
    o.tke_src = P_k - Y_k;
    o.dis_src = P_w - Y_w + D_w;
    return o;
}

// compute_k_omega (:929-1017) with no solid cell; its own sin(theta) (sin alone, floored at 1e-5)
// and (rm + 1) divisions.
__device__ __forceinline__ SrcOut komega_cell(const TurbArgs& a, const Shape& sh, int i, int j, int k, int kw, int ke,
                                              double kc, double dc, double prod, double dtkedr, double ddisdr,
                                              double tke_jm1, double tke_jp1, double dis_jm1, double dis_jp1,
                                              double tke_km1, double tke_kp1, double dis_km1, double dis_kp1,
                                              double Omega_mag) {
    const double bet_star = 0.09, gam = 0.52, C_lim = 0.875, bet_0 = 0.0708;
    const double rm = a.rm[i];
    const double sinthe = a.sin_alone[j];
    const double rmsinthe = rm * sinthe;
    const double dr = a.dr, dthe = a.dthe, dphi = a.dphi;
    const size_t c = (size_t)i * sh.plane + (size_t)j * sh.ldk;
    const double *u = a.u, *v = a.v, *w = a.w;
    const double dtkedthe_neu = (tke_jp1 - tke_jm1) / (rm * 2.0 * dthe);
    const double ddisdthe_neu = (dis_jp1 - dis_jm1) / (rm * 2.0 * dthe);
    const double dtkedphi_neu = (tke_kp1 - tke_km1) / (rmsinthe * 2.0 * dphi);
    const double ddisdphi_neu = (dis_kp1 - dis_km1) / (rmsinthe * 2.0 * dphi);
    const double dudr = (u[c + sh.plane + k] - u[c - sh.plane + k]) / (2.0 * dr) / (rm + 1.0);
    const double dvdthe = (v[c + sh.ldk + k] - v[c - sh.ldk + k]) / (rm * 2.0 * dthe);
    const double dwdphi = (w[c + ke] - w[c + kw]) / (rmsinthe * 2.0 * dphi);
    const double dudthe_u = (u[c + sh.ldk + k] - u[c - sh.ldk + k]) / (rm * 2.0 * dthe);
    const double dudphi_u = (u[c + ke] - u[c + kw]) / (rmsinthe * 2.0 * dphi);
    const double dvdr_u = (v[c + sh.plane + k] - v[c - sh.plane + k]) / (2.0 * dr) / (rm + 1.0);
    const double dwdr_u = (w[c + sh.plane + k] - w[c - sh.plane + k]) / (2.0 * dr) / (rm + 1.0);
    const double dvdphi_u = (v[c + ke] - v[c + kw]) / (rmsinthe * 2.0 * dphi);
    const double dwdthe_u = (w[c + sh.ldk + k] - w[c - sh.ldk + k]) / (rm * 2.0 * dthe);
    const double S11 = dudr, S22 = dvdthe, S33 = dwdphi;
    const double S12 = 0.5 * (dudthe_u + dvdr_u);
    const double S13 = 0.5 * (dudphi_u + dwdr_u);
    const double S23 = 0.5 * (dvdphi_u + dwdthe_u);
    const double S_mag = sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
    const double w_hat = smax(dc, C_lim * S_mag / a.sqrt_bet_star);
    double sig_d = dtkedr * ddisdr + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu;
    sig_d = (sig_d <= 0.0) ? 0.0 : 0.125;
    const double chi_w = fabs(Omega_mag * Omega_mag * S_mag / pow(bet_star * smax(dc, 1.0e-20), 3.0));
    const double f_bet = (1.0 + 85.0 * chi_w) / (1.0 + 100.0 * chi_w);
    const double bet_wc = bet_0 * f_bet;
    SrcOut o;
    o.nue = kc / smax(w_hat, 1.0e-20);
    const double P_k = smin(prod, 20.0 * bet_star * kc * dc);
    const double Y_k = bet_star * kc * dc;
    const double P_w = gam * dc / smax(kc, 1.0e-20) * P_k;
    const double Y_w = bet_wc * dc * dc;
    const double D_w = sig_d / smax(dc, 1.0e-20) * (dtkedr * ddisdr + dtkedthe_neu * ddisdthe_neu + dtkedphi_neu * ddisdphi_neu);
    o.tke_src = P_k - Y_k;
    o.dis_src = P_w - Y_w + D_w;
    return o;
}

// compute_k_epsilon (:885-927).
__device__ __forceinline__ SrcOut keps_cell(const TurbArgs& a, int i, double vs, double kc, double dc, double prod) {
    const double C_eps_1 = 1.35, C_eps_2 = 1.80;
    const double y_phys = a.yphys_ke[i];
    const double y_star = a.ystar_ke[i];
    SrcOut o;
    double nue = kCnue * kc * kc / smax(dc, kDisMin);
    const double d_plus = y_phys * vs / a.nue_air;
    const double f_nue = 1.0 - exp(-0.0115 * d_plus);
    const double Re_T = kc * kc * a.u_0 * a.L_m / smax(dc * a.nue_air, kDisMin * a.nue_air);
    const double f_2 = 1.0 - 0.4 / 1.8 * exp(-Re_T * Re_T / 36.0);
    nue *= f_nue;
    o.nue = nue;
    const double y_star2 = y_star * y_star;
    const double L_k = -2.0 * kc / y_star2;
    const double L_w = -2.0 * dc / y_star2 * exp(-d_plus / 2.0);
    const double P_k = prod;
    const double Y_k = dc;
    const double tke_safe = smax(kc, kDisMin);
    const double P_w = C_eps_1 * dc / tke_safe * P_k;
    const double Y_w = C_eps_2 * f_2 * dc * dc / tke_safe;
    o.tke_src = P_k - Y_k + L_k / a.re_turb;
    o.dis_src = P_w - Y_w + L_w / a.re_turb;
    return o;
}

// The source loop over the interior (i 1..im-2, j 1..jm-2, k 0..km-2 on the periodic ring).
template <int MODEL>
__global__ void __launch_bounds__(256) k_turb_sources(Shape sh, TurbArgs a, int wall_after) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y + 1, i = blockIdx.z + 1;
    const int km = sh.km;
    if (k >= km - 1 || j > sh.jm - 2) return;
    const int kw = (k == 0) ? km - 2 : k - 1, ke = (k == km - 2) ? 0 : k + 1;   // the ring of km-1 meridians
    const size_t c = (size_t)i * sh.plane + (size_t)j * sh.ldk;
    const size_t q = c + k;
    const size_t P = sh.plane, L = sh.ldk;
    const double *u = a.u, *v = a.v, *w = a.w;

    // per-cell geometry (compute_sources, :666-689); exp_rm = 1 (no coordinate stretching)
    const double rm = a.rm[i];
    const double sinthe = a.sin_pair[j];
    const double costhe = a.cos_pair[j];
    const double cotanthe = costhe / sinthe;
    const double inv_rm = a.inv_rm[i];
    const double rmsinthe = rm * sinthe;
    const double inv_rm2dthe = a.inv_2dthe / rm;
    const double inv_rmsinthe2dphi = a.inv_2dphi / rmsinthe;

    const double dudr = (u[q + P] - u[q - P]) * a.inv_2dr;
    const double dvdr = (v[q + P] - v[q - P]) * a.inv_2dr;
    const double dwdr = (w[q + P] - w[q - P]) * a.inv_2dr;
    const double dudthe = (u[q + L] - u[q - L]) * inv_rm2dthe;
    const double dvdthe = (v[q + L] - v[q - L]) * inv_rm2dthe;
    const double dwdthe = (w[q + L] - w[q - L]) * inv_rm2dthe;
    const double dudphi = (u[c + ke] - u[c + kw]) * inv_rmsinthe2dphi;
    const double dvdphi = (v[c + ke] - v[c + kw]) * inv_rmsinthe2dphi;
    const double dwdphi = (w[c + ke] - w[c + kw]) * inv_rmsinthe2dphi;

    // tke and dis as the clamp pass left them: interior neighbours clamped, boundary ones not
    const double* T = a.tke;
    const double* D = a.dis;
    const double kc = tke_in(T[q]), dc = dis_in(D[q]);
    const double tke_im1 = (i - 1 >= 1) ? tke_in(T[q - P]) : T[q - P];
    const double dis_im1 = (i - 1 >= 1) ? dis_in(D[q - P]) : D[q - P];
    const double tke_ip1 = (i + 1 <= sh.im - 2) ? tke_in(T[q + P]) : T[q + P];
    const double dis_ip1 = (i + 1 <= sh.im - 2) ? dis_in(D[q + P]) : D[q + P];
    const double tke_jm1 = (j - 1 >= 1) ? tke_in(T[q - L]) : T[q - L];
    const double dis_jm1 = (j - 1 >= 1) ? dis_in(D[q - L]) : D[q - L];
    const double tke_jp1 = (j + 1 <= sh.jm - 2) ? tke_in(T[q + L]) : T[q + L];
    const double dis_jp1 = (j + 1 <= sh.jm - 2) ? dis_in(D[q + L]) : D[q + L];
    const double tke_km1 = tke_in(T[c + kw]), tke_kp1 = tke_in(T[c + ke]);
    const double dis_km1 = dis_in(D[c + kw]), dis_kp1 = dis_in(D[c + ke]);

    const double dtkedr = (tke_ip1 - tke_im1) * a.inv_2dr;
    const double ddisdr = (dis_ip1 - dis_im1) * a.inv_2dr;

    double cnue;
    {
        const double dis_here = smax(dc, kDisMin);
        const double tke_here = smax(kc, 0.0);
        cnue = (MODEL == ATU_TURB_K_EPSILON) ? kCnue * tke_here * tke_here / dis_here : tke_here / dis_here;
        cnue = smin(cnue, a.nue_max_nd);
    }
    const double u_r = a.curvature ? u[q] * inv_rm : 0.0;
    const double v_r = a.curvature ? v[q] * inv_rm : 0.0;
    const double w_r = a.curvature ? w[q] * inv_rm : 0.0;
    const double v_cot_r = v_r * cotanthe;
    const double w_cot_r = w_r * cotanthe;
    const double g_rr = dudr, g_rt = dudthe - v_r, g_rp = dudphi - w_r;
    const double g_tr = dvdr, g_tt = dvdthe + u_r, g_tp = dvdphi - w_cot_r;
    const double g_pr = dwdr, g_pt = dwdthe, g_pp = dwdphi + u_r + v_cot_r;
    const double der = 0.66667 * (g_rr + g_tt + g_pp);
    const double prod = smax(0.0, (cnue * (2.0 * g_rr - der) - 0.66667 * kc) * g_rr
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

    const double vs = a.vel_star[(size_t)j * sh.ldk + k];
    SrcOut o;
    if (MODEL == ATU_TURB_K_EPSILON)
        o = keps_cell(a, i, vs, kc, dc, prod);
    else if (MODEL == ATU_TURB_K_OMEGA)
        o = komega_cell(a, sh, i, j, k, kw, ke, kc, dc, prod, dtkedr, ddisdr, tke_jm1, tke_jp1, dis_jm1, dis_jp1,
                        tke_km1, tke_kp1, dis_km1, dis_kp1, Omega);
    else
        o = sst_cell(a, i, rm, sinthe, kc, dc, prod, dtkedr, ddisdr, tke_jm1, tke_jp1, dis_jm1, dis_jp1, tke_km1,
                     tke_kp1, dis_km1, dis_kp1, Omega);

    // clamp_out (:636-647): the caps of nue and the sources from the cell's floored tke and dis
    const double k_bg_nd = a.free_flow ? a.k_bg_free_nd : 0.01 * vs * vs / a.kbg_den;
    const double tke_o = smax(k_bg_nd, kc);
    const double dis_o = smax(kDisMin, dc);
    const double nue_o = clamp3(o.nue, 0.0, a.nue_max_nd);
    const double dis_src_max = 20.0 * dis_o * dis_o;
    const double tke_src_max = 20.0 * tke_o * dis_o;
    a.prod[q] = prod;
    a.tke_source[q] = clamp3(o.tke_src, -tke_src_max, tke_src_max);
    a.dis_source[q] = clamp3(o.dis_src, -dis_src_max, dis_src_max);
    // clamp_nue (:564-589) of this cell, and in run() the wall cell below i = 1 (:414-418)
    a.nue[q] = clamp3(nue_o, 0.0, a.nue_max_nd) * a.taper[i];
    if (wall_after && i == 1) {
        const size_t q0 = (size_t)j * sh.ldk + k;
        a.tke[q0] = tke_o;
        a.dis[q0] = dis_o;
        a.nue[q0] = clamp3(nue_o, 0.0, a.nue_max_nd) * a.taper[0];
    }
}

// The clamp-out of tke and dis over the interior (:636-647) and clamp_nue of the interior
// columns' end cells (i = im-1 always, i = 0 in init(), where the wall copy came first).
__global__ void __launch_bounds__(256) k_turb_clamp(Shape sh, TurbArgs a, int wall_after) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y + 1, i = blockIdx.z;
    if (k >= sh.km - 1 || j > sh.jm - 2) return;
    const size_t q = (size_t)i * sh.plane + (size_t)j * sh.ldk + k;
    if (i >= 1 && i <= sh.im - 2) {
        const double vs = a.vel_star[(size_t)j * sh.ldk + k];
        const double k_bg_nd = a.free_flow ? a.k_bg_free_nd : 0.01 * vs * vs / a.kbg_den;
        a.tke[q] = smax(k_bg_nd, tke_in(a.tke[q]));
        a.dis[q] = smax(kDisMin, dis_in(a.dis[q]));
    } else if (i == sh.im - 1 || !wall_after) {
        a.nue[q] = clamp3(a.nue[q], 0.0, a.nue_max_nd) * a.taper[i];
    }
}

// The columns the loop does not reach: the pole rows j = 0, jm-1 (every k) and the seam meridian
// k = km-1 (j = 1..jm-2); column t of the 2 km + jm - 2.
__device__ __forceinline__ size_t edge_column(const Shape& sh, int t) {
    const int km = sh.km, jm = sh.jm;
    int j, k;
    if (t < 2 * km) {
        j = (t < km) ? 0 : jm - 1;
        k = (t < km) ? t : t - km;
    } else {
        j = 1 + (t - 2 * km);
        k = km - 1;
    }
    return (size_t)j * sh.ldk + k;
}

// run(): the wall copy of the edge columns at i = 0 from i = 1 and clamp_nue of the copied cell
// (i = 1 is clamped by k_turb_edge_clamp, launched after this kernel).
__global__ void __launch_bounds__(256) k_turb_edge_wall(Shape sh, TurbArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * sh.km + (sh.jm - 2)) return;
    const size_t col = edge_column(sh, t);
    a.tke[col] = a.tke[sh.plane + col];
    a.dis[col] = a.dis[sh.plane + col];
    a.nue[col] = clamp3(a.nue[sh.plane + col], 0.0, a.nue_max_nd) * a.taper[0];
}

// clamp_nue down the edge columns from level i0 (1 after k_turb_edge_wall, else 0).
__global__ void __launch_bounds__(256) k_turb_edge_clamp(Shape sh, TurbArgs a, int i0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y + i0;
    if (t >= 2 * sh.km + (sh.jm - 2)) return;
    const size_t q = (size_t)i * sh.plane + edge_column(sh, t);
    a.nue[q] = clamp3(a.nue[q], 0.0, a.nue_max_nd) * a.taper[i];
}

// seam_copy (:542-561): k = km-1 takes k = 0 in the closure's eight arrays.
// One thread per (array, row): 8 x im x jm copies.
__global__ void __launch_bounds__(64) k_turb_seam(Shape sh, TurbArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, n = blockIdx.z;
    if (j >= sh.jm) return;
    const size_t r = (size_t)i * sh.plane + (size_t)j * sh.ldk;
    double* f = n == 0 ? a.tke : n == 1 ? a.dis : n == 2 ? a.tken : n == 3 ? a.disn : n == 4 ? a.nue
              : n == 5 ? a.prod : n == 6 ? a.tke_source : a.dis_source;
    f[r + sh.km - 1] = f[r];
}

// fill_lid_corners with the model's own boundaries off (:515-527): the lid i = im-1 (j = 1..jm-2)
// from i = im-2, then the pole rows from j = 1 / jm-2 at every level, so a pole cell of the lid
// takes the lid-filled value, i.e. the cell (im-2, 1 or jm-2). Block row b: b < jm the lid row
// (im-1, b); else the pole rows of level (b - jm) / 2.
__global__ void __launch_bounds__(256) k_turb_lid(Shape sh, TurbArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (k >= sh.km) return;
    const int il = sh.im - 1, jm = sh.jm;
    int i, j;
    if (b < jm) {
        i = il;
        j = b;
    } else {
        i = (b - jm) >> 1;
        j = ((b - jm) & 1) ? jm - 1 : 0;
    }
    const int js = (j == 0) ? 1 : ((j == jm - 1) ? jm - 2 : j);
    const int is = (i == il) ? il - 1 : i;
    const size_t dst = (size_t)i * sh.plane + (size_t)j * sh.ldk + k;
    const size_t src = (size_t)is * sh.plane + (size_t)js * sh.ldk + k;
    double* f[8] = {a.tke, a.dis, a.tken, a.disn, a.nue, a.prod, a.tke_source, a.dis_source};
#pragma unroll
    for (int n = 0; n < 8; n++) f[n][dst] = f[n][src];
}

}  // namespace

hipError_t launch_turbulence(const Shape& sh, const TurbArgs& a, bool init, hipStream_t s) {
    const int bk = 256;
    const dim3 gk((sh.km + bk - 1) / bk);
    hipLaunchKernelGGL(k_turb_vel_star, dim3(gk.x, sh.jm), dim3(bk), 0, s, sh, a);
    const int wall_after = init ? 0 : 1;
    // the interior passes: tiles of 64 meridians x 4 rows (the ring of km-1 meridians in 64-wide waves)
    const dim3 tile(64, 4);
    const int tx = (sh.km - 1 + 63) / 64, ty = (sh.jm - 2 + 3) / 4;
    if (init) {
        hipLaunchKernelGGL(k_turb_init, dim3(tx, ty, sh.im - 2), tile, 0, s, sh, a);
        hipLaunchKernelGGL(k_turb_seam, dim3((sh.jm + 63) / 64, sh.im, 8), dim3(64), 0, s, sh, a);
        hipLaunchKernelGGL(k_turb_wall, dim3(gk.x, sh.jm), dim3(bk), 0, s, sh, a);
    }
    const dim3 gsrc(tx, ty, sh.im - 2);
    switch (a.model) {
        case ATU_TURB_K_OMEGA:
            hipLaunchKernelGGL((k_turb_sources<ATU_TURB_K_OMEGA>), gsrc, tile, 0, s, sh, a, wall_after);
            break;
        case ATU_TURB_K_EPSILON:
            hipLaunchKernelGGL((k_turb_sources<ATU_TURB_K_EPSILON>), gsrc, tile, 0, s, sh, a, wall_after);
            break;
        default:
            hipLaunchKernelGGL((k_turb_sources<ATU_TURB_K_OMEGA_SST>), gsrc, tile, 0, s, sh, a, wall_after);
            break;
    }
    hipLaunchKernelGGL(k_turb_clamp, dim3(tx, ty, sh.im), tile, 0, s, sh, a, wall_after);
    const int ncol = 2 * sh.km + (sh.jm - 2);
    if (wall_after) hipLaunchKernelGGL(k_turb_edge_wall, dim3((ncol + 255) / 256), dim3(256), 0, s, sh, a);
    hipLaunchKernelGGL(k_turb_edge_clamp, dim3((ncol + 255) / 256, sh.im - wall_after), dim3(256), 0, s, sh, a,
                       wall_after);
    hipLaunchKernelGGL(k_turb_seam, dim3((sh.jm + 63) / 64, sh.im, 8), dim3(64), 0, s, sh, a);
    if (a.free_flow)
        hipLaunchKernelGGL(k_turb_lid, dim3(gk.x, sh.jm + 2 * (sh.im - 1)), dim3(bk), 0, s, sh, a);
    return hipGetLastError();
}

}  // namespace atu
