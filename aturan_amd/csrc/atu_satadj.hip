// The reference loop's saturation-adjustment block (planet/cUranusModel.cpp:997-1033) on the
// device: the four SaturationAdjustmentUran calls (H2O, H2S, NH3, CH4), each the shared mixed-phase
// adjustment SaturationAdjustment<cUranusModel>::run (Tao et al. 1988, planet/SaturationAdjustment.h:
// 256-668), then ChemistryUran::EquilibriumNH4SH("loop") (planet/ChemistryUran.h:139-255).
//
// Every cell of the block is independent of every other (the reference loops over cells with no
// neighbour access; the only coupling is the temperature each gas's latent heat changes before the
// next gas reads it), so one thread carries a cell through all five in registers: one read and one
// write of each field instead of the reference's five passes and four clamp passes. The thread
// also writes the stage buffers (FP64 context) or the FP32 mirrors (mixed context) of every field
// it writes, which the context would otherwise copy afterwards.
//
// The expressions keep the reference's operand order (built without FP contraction). The
// saturation vapour pressure is exp(C + (-L0/T + del_alf log T + del_bet T) / (1e-3 R)) and the
// NH4SH equilibrium constant a power of ten: the device's exp, log and pow (the ROCm device
// library's) may round differently from glibc's in the last place, and a cell whose convergence
// test sits at its threshold can then take one pass more or less.

#include "atu_kernels.h"

namespace atu {
namespace {

__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }   // std::max
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }   // std::min

// ATPhys::saturation_vapour_pressure (planet/ATPhys.h:49-52)
__device__ __forceinline__ double svp(double T, double C, double L0, double R, double del_alf, double del_bet) {
    return exp(C + (-L0 / T + del_alf * log(T) + del_bet * T) / (1e-3 * R));
}

// One gas of SaturationAdjustment<cUranusModel>::run for one cell (planet/SaturationAdjustment.h:
// 362-614) and the final clamp of its three fields (:635-647). t_nd is the non-dimensional
// temperature as the model stores it.
__device__ void adjust_gas(const SatAdjArgs& A, const atu_gas& g, double rho_c, double p_u, double& t_nd, double& c,
                           double& cloud, double& ice) {
    const double t_0 = g.t_0, t_00 = g.t_00, ep = g.ep, lv = g.lv, ls = g.ls;
    const double t_range_inv = 1.0 / (t_0 - t_00);
    const bool conserve = A.conserve, newton = A.newton, evap = A.evap, ice_entry = A.ice_entry;
    const int bergeron = A.bergeron;
    const bool freeze_on = A.freeze && conserve && t_00 > 0.0 && t_00 < t_0;
    const double Lf = ls - lv;
    const double q_diff_min = 1.0e-3;
    auto cnd_of = [&](double TT) { return TT <= t_00 ? 0.0 : (TT >= t_0 ? 1.0 : (TT - t_00) * t_range_inv); };
    auto svp_l = [&](double T) { return svp(T, g.C, g.L0, g.R, g.del_alf, g.del_bet); };
    auto svp_i = [&](double T) { return svp(T, g.C_i, g.L0_i, g.R, g.del_alf_i, g.del_bet_i); };

    double t_u = t_nd * A.t_ref;
    if (!(t_u > 0.0) || !(p_u > 0.0)) return;   // no usable state: only the final clamp reaches the cell
    do {
        if (t_u > t_0) {
            if (conserve) {
                if (ice > 0.0) {
                    double dm = ice;
                    if (Lf > 0.0) {
                        const double hc = A.cp_mix * rho_c;
                        const double cap = hc * (t_u - t_0) / Lf;
                        if (dm > cap) dm = cap;
                        t_u -= Lf * dm / hc;
                        t_nd = t_u / A.t_ref;
                    }
                    ice -= dm;
                    cloud += dm;
                }
            } else {
                ice = 0.0;
            }
        }
        if (freeze_on && t_u < t_00 && cloud > 0.0) {
            double dm = cloud;
            if (Lf > 0.0) {
                const double hc = A.cp_mix * rho_c;
                const double cap = hc * (t_00 - t_u) / Lf;
                if (dm > cap) dm = cap;
                t_u += Lf * dm / hc;
                t_nd = t_u / A.t_ref;
            }
            cloud -= dm;
            ice += dm;
        }
        if (c < 0.0) c = 0.0;
        if (cloud < 0.0) cloud = 0.0;
        if (ice < 0.0) ice = 0.0;

        const double q_Rain_0 = rho_c * ep * svp_l(t_u) / p_u;
        if (q_Rain_0 <= 0.0) break;
        double q_entry = q_Rain_0;
        if (ice_entry && cloud <= 0.0 && ice <= 0.0) {
            const double cnd0 = cnd_of(t_u);
            if (cnd0 < 1.0) q_entry = cnd0 * q_Rain_0 + (1.0 - cnd0) * rho_c * ep * svp_i(t_u) / p_u;
        }
        if (bergeron >= 2 && cloud <= 0.0 && ice <= 0.0 && t_u < t_0) q_entry = rho_c * ep * svp_i(t_u) / p_u;
        if (c <= q_entry) {
            const double q_c0 = cloud, q_i0 = ice;
            if (!evap || !conserve || (q_c0 <= 0.0 && q_i0 <= 0.0)) break;
            const double q_Ice_0 = rho_c * ep * svp_i(t_u) / p_u;
            double q_t0;
            if (q_c0 > 0.0 && q_i0 > 0.0) q_t0 = (q_c0 * q_Rain_0 + q_i0 * q_Ice_0) / (q_c0 + q_i0);
            else if (q_i0 <= 0.0) q_t0 = q_Rain_0;
            else q_t0 = q_Ice_0;
            if (fabs(c - q_t0) <= q_diff_min * q_t0) break;
        }

        // the mixed-phase iteration (:432-614)
        double q_v_b = c, q_c_b = cloud, q_i_b = ice;
        double T = t_u;
        double q_v_hyp = q_v_b;
        double q_rain_now = q_Rain_0;
        for (int itr = 1; itr <= 15; itr++) {
            double CND = (T - t_00) * t_range_inv;
            double DEP = (t_0 - T) * t_range_inv;
            if (T <= t_00) { CND = 0.0; DEP = 1.0; }
            if (T >= t_0) { CND = 1.0; DEP = 0.0; }
            const bool berg = bergeron != 0 && T < t_0 && q_v_b <= q_rain_now &&
                              (q_i_b > 0.0 || (bergeron >= 2 && q_c_b <= 0.0));
            if (berg) {
                CND = 0.0;
                DEP = 1.0;
                if (q_c_b > 0.0) {
                    T -= lv * q_c_b / (A.cp_mix * rho_c);
                    q_v_b += q_c_b;
                    q_c_b = 0.0;
                }
            }
            double d_q_v = q_v_hyp - q_v_b;
            double d_q_c = -d_q_v * CND;
            double d_q_i = -d_q_v * DEP;
            if (conserve) {
                if (d_q_v < -q_v_b) {
                    d_q_v = -q_v_b;
                    d_q_c = -d_q_v * CND;
                    d_q_i = -d_q_v * DEP;
                }
                if (q_c_b + d_q_c < 0.0) {
                    const double ex = -(q_c_b + d_q_c);
                    d_q_c = -q_c_b;
                    d_q_i -= ex;
                }
                if (q_i_b + d_q_i < 0.0) {
                    const double ex = -(q_i_b + d_q_i);
                    d_q_i = -q_i_b;
                    d_q_c -= ex;
                    if (q_c_b + d_q_c < 0.0) d_q_c = -q_c_b;
                }
                d_q_v = -(d_q_c + d_q_i);
            }
            T += (lv * d_q_c + ls * d_q_i) / (A.cp_mix * rho_c);
            q_v_b += d_q_v;
            q_c_b += d_q_c;
            q_i_b += d_q_i;
            if (q_v_b < 0.0) q_v_b = 0.0;
            if (q_c_b < 0.0) q_c_b = 0.0;
            if (q_i_b < 0.0) q_i_b = 0.0;

            const double q_Rain = rho_c * ep * svp_l(T) / p_u;
            const double q_Ice = rho_c * ep * svp_i(T) / p_u;
            q_rain_now = q_Rain;
            if (q_c_b > 0.0 && q_i_b > 0.0) q_v_hyp = (q_c_b * q_Rain + q_i_b * q_Ice) / (q_c_b + q_i_b);
            else if (bergeron >= 2 && q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_v_hyp = q_Ice;
            else if (ice_entry && q_c_b <= 0.0 && q_i_b <= 0.0) {
                const double cn = cnd_of(T);
                q_v_hyp = cn * q_Rain + (1.0 - cn) * q_Ice;
            } else if (q_i_b == 0.0) q_v_hyp = q_Rain;
            else q_v_hyp = q_Ice;

            if (T >= t_0) {
                if (conserve) {
                    if (q_i_b > 0.0) {
                        double dm = q_i_b;
                        if (Lf > 0.0) {
                            const double hc = A.cp_mix * rho_c;
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
            if (A.dry_exit && conserve && q_c_b <= 0.0 && q_i_b <= 0.0 && q_v_b < q_v_hyp) break;
            if (newton) {
                const double dT = 0.1;
                const double q_Rain_d = rho_c * ep * svp_l(T + dT) / p_u;
                const double q_Ice_d = rho_c * ep * svp_i(T + dT) / p_u;
                double q_hyp_d;
                if (q_c_b > 0.0 && q_i_b > 0.0) q_hyp_d = (q_c_b * q_Rain_d + q_i_b * q_Ice_d) / (q_c_b + q_i_b);
                else if (bergeron >= 2 && q_c_b <= 0.0 && q_i_b <= 0.0 && T < t_0) q_hyp_d = q_Ice_d;
                else if (ice_entry && q_c_b <= 0.0 && q_i_b <= 0.0) {
                    const double cn = cnd_of(T);
                    q_hyp_d = cn * q_Rain_d + (1.0 - cn) * q_Ice_d;
                } else if (q_i_b == 0.0) q_hyp_d = q_Rain_d;
                else q_hyp_d = q_Ice_d;
                double CNDn = (T - t_00) * t_range_inv;
                if (T <= t_00) CNDn = 0.0;
                if (T >= t_0) CNDn = 1.0;
                if (berg) CNDn = 0.0;
                const double G = (lv * CNDn + ls * (1.0 - CNDn)) / (A.cp_mix * rho_c) * (q_hyp_d - q_v_hyp) / dT;
                double w = 1.0 / (1.0 + smax(G, 0.0));
                if (w < 0.05) w = 0.05;
                q_v_hyp = q_v_b + w * (q_v_hyp - q_v_b);
            } else {
                q_v_hyp = 0.5 * (q_v_hyp + q_v_b);
            }
        }
        c = q_v_b;
        cloud = q_c_b;
        ice = q_i_b;
        t_nd = T / A.t_ref;
    } while (false);
}

// The final clamp of one gas (planet/SaturationAdjustment.h:635-647), every cell.
__device__ __forceinline__ void clamp_gas(double& c, double& cloud, double& ice) {
    if (c <= 0.0) c = 0.0;
    if (cloud <= 0.0) cloud = 0.0;
    if (ice <= 0.0) ice = 0.0;
}

__device__ __forceinline__ double kp_nh4sh(double T) {
    const double atm = 101325.0;
    return pow(10.0, 14.82 - 4705.0 / T) * atm * atm;
}

// ChemistryUran::equilibrateNH4SHCell (planet/ChemistryUran.h:139-180): the NH4SH formed
// (decomposed) to the Lewis equilibrium, by bisection, with its heat.
__device__ void nh4sh_cell(const SatAdjArgs& A, double p_stat, double& t_nd, double& nh3, double& h2s, double& nh4sh) {
    const double R_univ = 8314.46;
    const double T0 = t_nd * A.t_ref;
    if (!(T0 > 0.0)) return;
    double rho = p_stat * 1e5 / (A.R_mix * T0);
    if (!(rho > 0.0) || !isfinite(rho)) rho = A.r_mix;
    const double a = smax(0.0, nh3) / A.m_nh3;
    const double b = smax(0.0, h2s) / A.m_h2s;
    const double s = smax(0.0, nh4sh) / A.m_nh4sh;
    const double heat = A.nh4sh_heat * A.m_nh4sh / (A.cp_mix * rho);
    auto f = [&](double x) {
        const double T = T0 + heat * x, RT = R_univ * T;
        return (a - x) * (b - x) * RT * RT - kp_nh4sh(T);
    };
    const double f0 = f(0.0);
    double lo, hi;
    if (f0 > 0.0) {
        lo = 0.0;
        hi = smin(a, b);
    } else if (s > 0.0) {
        if (f(-s) <= 0.0) {
            lo = hi = -s;
        } else {
            lo = -s;
            hi = 0.0;
        }
    } else {
        return;
    }
    if (hi > lo)
        for (int it = 0; it < 80; it++) {
            const double xm = 0.5 * (lo + hi);
            if (f(xm) > 0.0) lo = xm;
            else hi = xm;
            if (hi - lo <= 1.0e-15 * smax(a, smax(b, s))) break;
        }
    const double x = (f0 > 0.0) ? lo : hi;
    if (x == 0.0) return;
    nh3 = smax(0.0, nh3 - A.m_nh3 * x);
    h2s = smax(0.0, h2s - A.m_h2s * x);
    nh4sh = smax(0.0, nh4sh + A.m_nh4sh * x);
    const double dT = heat * x;
    t_nd = (T0 + dT) / A.t_ref;
}

// Writes a field value to the FP64 master, and to the FP64 stage buffers or the FP32 mirrors.
__device__ __forceinline__ void put(const SatAdjArgs& A, int f, size_t q, double v) {
    const size_t o = (size_t)f * A.fs + q;
    A.y[o] = v;
    if (A.sa) {
        A.sa[o] = v;
        A.sb[o] = v;
    }
    if (A.y32) {
        const float v32 = (float)v;
        A.y32[o] = v32;
        A.sa32[o] = v32;
        A.sb32[o] = v32;
    }
}

__global__ void __launch_bounds__(256) k_satadj(Shape sh, SatAdjArgs A) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= sh.km || j >= sh.jm) return;
    const size_t q = (size_t)i * sh.plane + (size_t)j * sh.ldk + k;
    const double* Y = A.y;
    const size_t fs = A.fs;
    double t_nd = Y[(size_t)ATU_T * fs + q];
    const double p_u = A.p_stat[q];
    double rho_c;
    {
        const double rho = A.rho_mix[q];
        rho_c = (rho > 0.0 && isfinite(rho)) ? rho : A.r_mix;   // cUranusModel::rho_at, LOCAL_RHO 1
    }
    const int vap[4] = {ATU_H2O, ATU_H2S, ATU_NH3, ATU_CH4};
    double sp[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int m = 0; m < 3; m++) sp[n][m] = Y[(size_t)(vap[n] + m) * fs + q];
    double nh4sh = Y[(size_t)ATU_NH4SH * fs + q];
    for (int n = 0; n < 4; n++) {
        if (!(A.gas_mask & (1 << n))) continue;
        adjust_gas(A, A.gas[n], rho_c, p_u, t_nd, sp[n][0], sp[n][1], sp[n][2]);
        clamp_gas(sp[n][0], sp[n][1], sp[n][2]);
    }
    if (A.nh4sh_eq) nh4sh_cell(A, p_u, t_nd, sp[2][0], sp[1][0], nh4sh);
    put(A, ATU_T, q, t_nd);
#pragma unroll
    for (int n = 0; n < 4; n++)
#pragma unroll
        for (int m = 0; m < 3; m++) put(A, vap[n] + m, q, sp[n][m]);
    put(A, ATU_NH4SH, q, nh4sh);
}

}  // namespace

hipError_t launch_satadj(const Shape& sh, const SatAdjArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_satadj, dim3((sh.km + 63) / 64, (sh.jm + 3) / 4, sh.im), dim3(64, 4), 0, s, sh, a);
    return hipGetLastError();
}

}  // namespace atu
