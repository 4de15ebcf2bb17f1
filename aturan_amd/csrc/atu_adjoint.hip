// Discrete adjoint of the step (hydrostatic prep, RK4, domain BC, converging Poisson solve), in FP64,
// for gradients of a zonal-wind misfit. There is no counterpart in the reference; the forward it
// differentiates is the one of atu_rk.hip, atu_hydro_bc.hip, atu_poisson.hip and atu_solve.hip.
//
// Closed subsystem. The tendencies of T, u, v, w read T, u, v, w, p_dyn, p_hydro and the frozen eddy
// viscosity; p_hydro reads T, v, w; the Poisson source reads the stage-3 momentum tendencies. The
// species, k* and omega* tendencies read the winds but no wind or temperature tendency reads them,
// so for a loss of the winds and T their adjoints are exactly zero and are not carried. The adjoint
// state is (lambda_T, lambda_u, lambda_v, lambda_w) of the prognostic fields and lambda_p of p_dyn.
//
// Every kernel is in gather form: each output value is formed by one thread from the values that
// read it in the forward, so there are no atomics and the result does not depend on scheduling.
//
// Field blocks of 4 (T, u, v, w) use the prognostic order of the state (ATU_T..ATU_W = 0..3) with
// the field stride of the padded layout.

#include "atu_kernels.h"

namespace atu {
namespace {

__device__ __forceinline__ bool is_interior(const Shape& sh, int i, int j) {
    return i >= 1 && i <= sh.im - 2 && j >= 1 && j <= sh.jm - 2;
}

__device__ __forceinline__ double endw(int x, int lo, int hi) { return (x == lo || x == hi) ? 1.5 : 1.0; }

// cp_mix / cp_lapse(T_K) and its derivative with respect to T_K (CpH2::Lapse::cp on the H2 table,
// linear inside each 0.5 K cell, constant beyond the table ends).
__device__ __forceinline__ double cp_ratio(const DevGeo& G, const StepConsts& C, double TK, double* dratio) {
    const double* t = G.cp_tab;
    const int n = G.cp_n;
    const bool lo = TK < 5.0;
    const double Tc = lo ? 5.0 : TK;
    const bool hi = Tc > 1500.0;
    const double x = ((hi ? 1500.0 : Tc) - 5.0) / 0.5;
    const int ix = (int)x;
    const int i = (n - 2 < ix) ? n - 2 : ix;
    const double f = x - i;
    const double ch = t[i] + f * (t[i + 1] - t[i]);
    const double cpl = (C.r_h2 * ch + C.cp_a + C.cp_b) / C.cp_den;
    const double dcpl = (lo || hi) ? 0.0 : C.r_h2 * ((t[i + 1] - t[i]) / 0.5) / C.cp_den;
    *dratio = -C.cp_mix / (cpl * cpl) * dcpl;
    return C.cp_mix / cpl;
}

// What the tendencies of one interior cell n contribute, per unit difference, to its neighbours.
struct NbrCoef {
    double lT, lU, lV, lW;        // lambda of the tendencies (u, v, w including the aux seed)
    double lUr, lVr, lWr;         // lambda of rhs_u, rhs_v, rhs_w alone (the pressure gradient terms)
    double u, v, w, cpr;
    double kT, kM;                // diffusivities of T and of momentum
    double irm, irm2, two_irm, irms, irm2s, irm2s2, cirs, cosv;
};

template <bool AUX>
__device__ __forceinline__ void load_nbr(const AdjStageArgs& a, int i, int j, size_t x, NbrCoef& n) {
    const size_t fs = a.sh.fs;
    const DevGeo& G = a.geo;
    const StepConsts& C = a.c;
    const size_t ij = (size_t)i * a.sh.jm + j;
    n.lT = a.lk[x];
    n.lUr = a.lk[fs + x];
    n.lVr = a.lk[2 * fs + x];
    n.lWr = a.lk[3 * fs + x];
    n.lU = n.lUr;
    n.lV = n.lVr;
    n.lW = n.lWr;
    if (AUX) {
        n.lU += a.laux[x];
        n.lV += a.laux[fs + x];
        n.lW += a.laux[2 * fs + x];
    }
    const double T = a.xs[x];
    n.u = a.xs[fs + x];
    n.v = a.xs[2 * fs + x];
    n.w = a.xs[3 * fs + x];
    const double nue = a.nue[x];
    const double nue_t = (C.turb_coupling != 0.0 && isfinite(nue)) ? C.turb_coupling * (0.0 < nue ? nue : 0.0) : 0.0;
    n.kT = 1.0 / C.re_pr + nue_t / C.Pr_t;
    n.kM = 1.0 / C.re + nue_t;
    n.cpr = 1.0;
    if ((C.cp_mode != 0 || C.cp_gas != 0) && T > 0.0) {
        double d;
        n.cpr = cp_ratio(G, C, T * C.t_ref, &d);
    }
    n.irm = G.inv_rm[i];
    n.irm2 = G.inv_rm2[i];
    n.two_irm = G.two_inv_rm[i];
    n.irms = G.inv_rmsinthe[ij];
    n.irm2s = G.inv_rm2sinthe[ij];
    n.irm2s2 = G.inv_rm2sinthe2[ij];
    n.cirs = G.costhe_inv_rm2sinthe[ij];
    n.cosv = G.costhe[j];
}

// Adjoint of one RK stage's tendencies at the stage input X_s, fused with the reverse RK fold:
//   lambda_X = (d R_s / d X_s)^T lambda_R            (R_s = rhs(X_s, p_dyn, p_hydro, nue))
//   ly      += lambda_X                               (X_s = y_n + c_s R_{s-1}, boundary planes = y_n)
//   lk_prev  = ws * ls + cs * lambda_X   (interior)  (the lambda of R_{s-1}), when lk_prev != null
//   lp += (d R_s / d p_dyn)^T lambda_R, lph += (d R_s / d p_hydro)^T lambda_R,
//   lnue += (d R_s / d nue)^T lambda_R
// One thread per cell of the full (i, j) range and k < km-1 (the seam copy is never read).
template <bool AUX>
__global__ void __launch_bounds__(256) k_rhs_adjoint(AdjStageArgs a) {
    const Shape& sh = a.sh;
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= N || j >= sh.jm) return;
    const int ldk = sh.ldk;
    const size_t fs = sh.fs, PL = sh.plane;
    const size_t row = ((size_t)i * sh.jm + j) * ldk;
    const size_t x = row + k;
    const DevGeo& G = a.geo;
    const StepConsts& C = a.c;
    const double sp = C.scale_p, cesp = C.coeff_energy_p * C.scale_p;

    double aT = 0.0, aU = 0.0, aV = 0.0, aW = 0.0, aP = 0.0, aH = 0.0;

    // ---- the transposed difference stencils of the 6 neighbours (and of the cell itself) ----
    // unrolled, so that the neighbours' loads go out together (bit-identical: the accumulations keep
    // their order)
#pragma unroll
    for (int q = 0; q < 7; q++) {
        int axis, sgn;
        int ni = i, nj = j, nk = k;
        if (q < 6) {
            axis = q >> 1;
            sgn = (q & 1) ? 1 : -1;          // n = c + sgn e_axis
            if (axis == 0) ni += sgn;
            else if (axis == 1) nj += sgn;
            else nk = (k + sgn + N) % N;
        } else {
            axis = -1;
            sgn = 0;
        }
        if (!is_interior(sh, ni, nj)) continue;
        const size_t xn = ((size_t)ni * sh.jm + nj) * ldk + nk;
        NbrCoef n;
        load_nbr<AUX>(a, ni, nj, xn, n);
        // d1(n) = (X[n+e] - X[n-e]) i2, d2(n) = (X[n+e] - 2 X[n] + X[n-e]) id2: the cell c is n - sgn e
        for (int ax = 0; ax < 3; ax++) {
            if (q < 6 && ax != axis) continue;
            const double i2 = ax == 0 ? G.inv_2dr : (ax == 1 ? G.inv_2dthe : G.inv_2dphi);
            const double id2 = ax == 0 ? G.inv_dr2 : (ax == 1 ? G.inv_dthe2 : G.inv_dphi2);
            const double w1 = (q < 6) ? (sgn > 0 ? -i2 : i2) : 0.0;
            const double w2 = (q < 6) ? id2 : -2.0 * id2;
            // advection and first-derivative diffusion coefficients along the axis
            double adv, cT, cU, cVW, d2;
            if (ax == 0) {
                adv = -n.u;
                cT = n.two_irm;
                cU = 4.0 * n.irm;
                cVW = n.two_irm;
                d2 = 1.0;
            } else if (ax == 1) {
                adv = -n.v * n.irm;
                cT = cU = cVW = n.cirs;
                d2 = n.irm2;
            } else {
                adv = -n.w * n.irms;
                cT = cU = cVW = 0.0;
                d2 = n.irm2s2;
            }
            double l1T = n.lT * (adv + n.kT * cT), l2T = n.lT * n.kT * d2;
            double l1U = n.lU * (adv + n.kM * cU), l2U = n.lU * n.kM * d2;
            double l1V = n.lV * (adv + n.kM * cVW), l2V = n.lV * n.kM * d2;
            double l1W = n.lW * (adv + n.kM * cVW), l2W = n.lW * n.kM * d2;
            double lP, lH = 0.0;
            if (ax == 0) {
                lP = -n.lUr * sp + n.lT * cesp * n.u * n.cpr;
            } else if (ax == 1) {
                l1U += n.lV * n.kM * 2.0 * n.irm2;                          // 2 d1t(u) / rm^2 in diffusion_v
                lP = -n.lVr * sp * n.irm + n.lT * cesp * (n.v * n.irm) * n.cpr;
                lH = -n.lV * n.irm;
            } else {
                l1U += n.lW * n.kM * 2.0 * n.irm2s;                         // 2 d1p(u) / (rm^2 sin) in diffusion_w
                l1V += n.lW * n.kM * 2.0 * n.cosv * n.irm2s2;               // 2 cos d1p(v) / (rm sin)^2 in diffusion_w
                l1W -= n.lV * n.kM * 2.0 * n.cosv * n.irm2s2;               // -2 cos d1p(w) / (rm sin)^2 in diffusion_v
                lP = -n.lWr * sp * n.irms + n.lT * cesp * (n.w * n.irms) * n.cpr;
                lH = -n.lW * n.irms;
            }
            aT += w1 * l1T + w2 * l2T;
            aU += w1 * l1U + w2 * l2U;
            aV += w1 * l1V + w2 * l2V;
            aW += w1 * l1W + w2 * l2W;
            aP += w1 * lP;
            aH += w1 * lH;
        }
    }

    double ln = 0.0;
    if (is_interior(sh, i, j)) {
        // ---- pointwise derivatives of the cell's own tendencies ----
        NbrCoef n;
        load_nbr<AUX>(a, i, j, x, n);
        const size_t xw = row + (k == 0 ? N - 1 : k - 1), xe = row + (k == N - 1 ? 0 : k + 1);
        const size_t ij = (size_t)i * sh.jm + j;
        double d1r[4], d1t[4], d1p[4], dif[4];
        for (int f = 0; f < 4; f++) {
            const double* X = a.xs + (size_t)f * fs;
            const double x0 = X[x], rp = X[x + PL], rn = X[x - PL], tp = X[x + ldk], tn = X[x - ldk];
            const double pe = X[xe], pw = X[xw];
            d1r[f] = (rp - rn) * G.inv_2dr;
            d1t[f] = (tp - tn) * G.inv_2dthe;
            d1p[f] = (pe - pw) * G.inv_2dphi;
            const double d2r = (rp - 2.0 * x0 + rn) * G.inv_dr2;
            const double d2t = (tp - 2.0 * x0 + tn) * G.inv_dthe2;
            const double d2p = (pe - 2.0 * x0 + pw) * G.inv_dphi2;
            const double c1r = (f == ATU_U) ? 4.0 * n.irm : n.two_irm;
            dif[f] = d2r + d1r[f] * c1r + d2t * n.irm2 + d1t[f] * n.cirs + d2p * n.irm2s2;
        }
        const double vmet = G.v_metric[ij], cot = G.cotanthe[j], sbf = G.sinthe_bf[j], spg = G.sponge[i];
        dif[ATU_U] += 2.0 * n.u * n.irm2;
        dif[ATU_V] += -vmet * n.v + 2.0 * d1t[ATU_U] * n.irm2 - d1p[ATU_W] * 2.0 * n.cosv * n.irm2s2;
        dif[ATU_W] += -vmet * n.w + 2.0 * d1p[ATU_U] * n.irm2s + d1p[ATU_V] * 2.0 * n.cosv * n.irm2s2;
        const double* pd = a.pdyn;
        const double dpdr = (pd[x + PL] - pd[x - PL]) * G.inv_2dr;
        const double dpdthe = (pd[x + ldk] - pd[x - ldk]) * G.inv_2dthe;
        const double dpdphi = (pd[xe] - pd[xw]) * G.inv_2dphi;
        const double corsc = C.Coriolis * C.scale_Cor;
        // transport: -(u d1r + v/rm d1t + w/(rm sin) d1p) of every carried field
        double su = 0.0, sv = 0.0, sw = 0.0;
        const double lam[4] = {n.lT, n.lU, n.lV, n.lW};
        for (int f = 0; f < 4; f++) {
            su -= lam[f] * d1r[f];
            sv -= lam[f] * d1t[f] * n.irm;
            sw -= lam[f] * d1p[f] * n.irms;
        }
        // pressure work of the energy equation
        su += n.lT * cesp * dpdr * n.cpr;
        sv += n.lT * cesp * n.irm * dpdthe * n.cpr;
        sw += n.lT * cesp * n.irms * dpdphi * n.cpr;
        // momentum metric terms, diffusion, sponge and Coriolis
        su += -n.lV * n.v * n.irm - n.lW * n.w * n.irm + n.lU * n.kM * 2.0 * n.irm2 - n.lU * spg
              - n.lW * corsc * C.two_omega * sbf;
        sv += -n.lV * n.u * n.irm - n.lW * n.w * cot * n.irm - n.lV * n.kM * vmet
              - n.lW * corsc * C.two_omega * n.cosv;
        sw += n.lV * 2.0 * n.w * cot * n.irm - n.lW * (n.u + n.v * cot) * n.irm - n.lW * n.kM * vmet
              - n.lV * corsc * C.m_two_omega * n.cosv;
        aU += su;
        aV += sv;
        aW += sw;
        const double T = a.xs[x];
        if ((C.cp_mode != 0 || C.cp_gas != 0) && T > 0.0) {
            double dr;
            cp_ratio(G, C, T * C.t_ref, &dr);
            const double pw0 = cesp * (n.u * dpdr + (n.v * n.irm) * dpdthe + (n.w * n.irms) * dpdphi);
            aT += n.lT * pw0 * dr * C.t_ref;
        }
        // eddy viscosity: nue_t = turb_coupling * max(nue, 0) in the T and momentum diffusivities
        const double nue = a.nue[x];
        if (isfinite(nue) && nue > 0.0 && C.turb_coupling != 0.0)
            ln = C.turb_coupling * (n.lT * dif[ATU_T] / C.Pr_t + n.lU * dif[ATU_U] + n.lV * dif[ATU_V] +
                                    n.lW * dif[ATU_W]);
    }

    // ---- reverse RK fold ----
    const double lx[4] = {aT, aU, aV, aW};
    const bool inner = is_interior(sh, i, j);
    for (int f = 0; f < 4; f++) {
        const size_t q = (size_t)f * fs + x;
        a.ly[q] += lx[f];
        if (a.lk_prev) a.lk_prev[q] = inner ? a.ws * a.ls[q] + a.cs * lx[f] : 0.0;
    }
    a.lp[x] += aP;
    a.lph[x] += aH;
    if (inner) a.lnue[x] += ln;
}

// ---- hydrostatic prep adjoint ----
// Column part: lambda_f(i) = dr/2 (S(i) + S(i+1)) with S the suffix sums of lambda_p_hydro
// (lambda_f(0) = dr/2 S(1)); then the local derivatives of f = buoy + (v^2 + w^2)/rm + cor w, and
// into g the derivative with respect to the level mean, reduced per level afterwards. One thread per
// (j, k) column; the levels go in chunks of kHydChunk whose loads are all issued before the chunk's
// sequential suffix sum, so a column waits for ceil(im / kHydChunk) memory round trips, not im.
constexpr int kHydChunk = 8;
__global__ void __launch_bounds__(256) k_hydro_adj_col(Shape sh, DevGeo g, StepConsts c, const double* t,
                                                       const double* v, const double* w, const double* pstat,
                                                       const double* level, const double* lph, double* ly,
                                                       double* gref) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= sh.km) return;
    const size_t fs = sh.fs;
    const size_t col = (size_t)j * sh.ldk + k;
    const double cor = g.cor_hydro[j];
    double Sp = 0.0;   // S(i + 1)
    for (int top = sh.im - 1; top >= 0; top -= kHydChunk) {
        double vl[kHydChunk], vv[kHydChunk], ww[kHydChunk], tt[kHydChunk], ps[kHydChunk];
        double l0[kHydChunk], l2[kHydChunk], l3[kHydChunk];
#pragma unroll
        for (int q = 0; q < kHydChunk; q++) {
            const int i = top - q;
            if (i < 0) break;
            const size_t x = (size_t)i * sh.plane + col;
            vl[q] = lph[x];
            vv[q] = v[x];
            ww[q] = w[x];
            tt[q] = t[x];
            ps[q] = pstat[x];
            l0[q] = ly[x];
            l2[q] = ly[2 * fs + x];
            l3[q] = ly[3 * fs + x];
        }
#pragma unroll
        for (int q = 0; q < kHydChunk; q++) {
            const int i = top - q;
            if (i < 0) break;
            const size_t x = (size_t)i * sh.plane + col;
            double lf;
            if (i >= 1) {
                const double S = Sp + vl[q];
                lf = 0.5 * (S + Sp) * c.dr;
                Sp = S;
            } else {
                lf = 0.5 * Sp * c.dr;
            }
            ly[2 * fs + x] = l2[q] + lf * (2.0 * vv[q] / g.rm[i]);
            ly[3 * fs + x] = l3[q] + lf * (2.0 * ww[q] / g.rm[i] + cor);
            double gr = 0.0;
            const double T = tt[q];
            if (T > 0.0) {
                const double ref = level[i];
                const double bb = c.g * ps[q] / (c.hyd_bnum * T * c.t_ref);
                const double f = -c.nd_rel * 1.0 * 1.0 * (bb - ref) / ref;
                if (isfinite(f)) {
                    ly[x] = l0[q] + lf * (c.nd_rel * bb / (ref * T));
                    gr = lf * (c.nd_rel * bb / (ref * ref));
                }
            }
            gref[x] = gr;
        }
    }
}

// Level part: lambda_ref(i) = sum over the level of g; then
// d ref / d T(j, k) = -sin_weight(j) bb / (T W), W the level's weight sum over the cells it counts.
// Each level is cut into kHydParts parts of consecutive cells: the first kernel sums g and the
// weights of one part (256 strided partials, then a halving tree), the second sums a level's parts in
// part order (the same order in every block) and updates one part's cells.
constexpr int kHydParts = kHydAdjParts;
__device__ __forceinline__ double hyd_bb(const StepConsts& c, double p, double T) {
    return c.g * p / (c.hyd_bnum * T * c.t_ref);
}
__global__ void __launch_bounds__(256) k_hydro_adj_level_sums(Shape sh, DevGeo g, StepConsts c, const double* t,
                                                              const double* pstat, const double* gref,
                                                              double* parts) {
    const int i = blockIdx.y, part = blockIdx.x;
    const int n = sh.jm * sh.km;
    const int q0 = (int)((long)n * part / kHydParts), q1 = (int)((long)n * (part + 1) / kHydParts);
    __shared__ double red[2][256];
    double sg = 0.0, sw = 0.0;
    for (int q = q0 + (int)threadIdx.x; q < q1; q += 256) {
        const int j = q / sh.km, k = q - j * sh.km;
        const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
        sg += gref[x];
        const double T = t[x];
        if (T > 0.0 && isfinite(hyd_bb(c, pstat[x], T))) sw += g.sin_weight[j];
    }
    red[0][threadIdx.x] = sg;
    red[1][threadIdx.x] = sw;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        parts[2 * ((size_t)i * kHydParts + part)] = red[0][0];
        parts[2 * ((size_t)i * kHydParts + part) + 1] = red[1][0];
    }
}
__global__ void __launch_bounds__(256) k_hydro_adj_level(Shape sh, DevGeo g, StepConsts c, const double* t,
                                                         const double* pstat, const double* parts, double* ly) {
    const int i = blockIdx.y, part = blockIdx.x;
    const int n = sh.jm * sh.km;
    const int q0 = (int)((long)n * part / kHydParts), q1 = (int)((long)n * (part + 1) / kHydParts);
    double lref = 0.0, W = 0.0;
    for (int r = 0; r < kHydParts; r++) {
        lref += parts[2 * ((size_t)i * kHydParts + r)];
        W += parts[2 * ((size_t)i * kHydParts + r) + 1];
    }
    if (!(W > 0.0)) return;
    for (int q = q0 + (int)threadIdx.x; q < q1; q += 256) {
        const int j = q / sh.km, k = q - j * sh.km;
        const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
        const double T = t[x];
        if (T > 0.0) {
            const double bb = hyd_bb(c, pstat[x], T);
            if (isfinite(bb)) ly[x] += lref * (-g.sin_weight[j] * bb / (T * W));
        }
    }
}

// ---- domain BC adjoint (reverse order: phi, theta, radius) on the 4 carried fields ----
__global__ void __launch_bounds__(256) k_bc_phi_adj(Shape sh, double* ly) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= sh.jm) return;
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    for (int f = 0; f < 4; f++) {
        double* L = ly + (size_t)f * sh.fs + row;
        L[0] += L[sh.km - 1];
        L[sh.km - 1] = 0.0;
    }
}

__global__ void __launch_bounds__(256) k_bc_theta_adj(Shape sh, double c43, double c13, double* ly) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (k >= sh.km) return;
    const int ldk = sh.ldk, Jl = sh.jm - 1;
    for (int f = 0; f < 4; f++) {
        double* L = ly + (size_t)f * sh.fs + (size_t)i * sh.plane + k;
        if (f == ATU_V || f == ATU_W) {   // zeroed at both poles
            L[0] = 0.0;
            L[(size_t)Jl * ldk] = 0.0;
            continue;
        }
        const double lo = L[0], hi = L[(size_t)Jl * ldk];
        L[0] = 0.0;
        L[(size_t)Jl * ldk] = 0.0;
        L[(size_t)1 * ldk] += c43 * lo;
        L[(size_t)2 * ldk] -= c13 * lo;
        L[(size_t)(Jl - 1) * ldk] += c43 * hi;
        L[(size_t)(Jl - 2) * ldk] -= c13 * hi;
    }
}

__global__ void __launch_bounds__(256) k_bc_radius_adj(Shape sh, double c43, double c13, int rigid_lid, int lid_vw,
                                                       double* ly) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= sh.km) return;
    const size_t P = sh.plane;
    const int L = sh.im - 1;
    const size_t col = (size_t)j * sh.ldk + k;
    // the lid form of v and w overwrote their top wall last
    if (lid_vw == 2 || lid_vw == 3) {
        for (int f = ATU_V; f <= ATU_W; f++) {
            double* Y = ly + (size_t)f * sh.fs + col;
            const double lt = Y[(size_t)L * P];
            Y[(size_t)L * P] = 0.0;
            if (lid_vw == 2) {
                Y[(size_t)(L - 1) * P] += 2.0 * lt;
                Y[(size_t)(L - 2) * P] -= lt;
            } else {
                Y[(size_t)(L - 3) * P] += lt;
                Y[(size_t)(L - 2) * P] -= 3.0 * lt;
                Y[(size_t)(L - 1) * P] += 3.0 * lt;
            }
        }
    }
    if (rigid_lid) {
        double* Y = ly + (size_t)ATU_U * sh.fs + col;
        Y[0] = 0.0;
        Y[(size_t)L * P] = 0.0;
    }
    for (int f = 0; f < 4; f++) {
        double* Y = ly + (size_t)f * sh.fs + col;
        const double lo = Y[0], hi = Y[(size_t)L * P];
        Y[0] = 0.0;
        Y[(size_t)L * P] = 0.0;
        Y[P] += c43 * lo;
        Y[2 * P] -= c13 * lo;
        Y[(size_t)(L - 1) * P] += c43 * hi;
        Y[(size_t)(L - 2) * P] -= c13 * hi;
    }
}

// ---- the final RK fold: y_{n+1} = clamp_T(y_n + dt/6 (k1 + 2 k2 + 2 k3 + k4)) on the interior ----
// ls = dt/6 m lambda, lyn = m lambda (interior) or lambda (boundary planes, which RK4 leaves as y_n);
// m = 0 where the temperature clamp was active.
__global__ void __launch_bounds__(256) k_final_fold(Shape sh, StepConsts c, const double* tnext, const double* ly,
                                                    double* ls, double* lyn) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.km) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    const bool inner = is_interior(sh, i, j) && k < sh.km - 1;
    for (int f = 0; f < 4; f++) {
        const size_t q = (size_t)f * sh.fs + x;
        double lam = ly[q];
        if (inner && f == ATU_T) {
            const double T = tnext[x];
            if (!(T > c.t_min && T < c.t_max)) lam = 0.0;
        }
        ls[q] = inner ? c.dt * lam / 6.0 : 0.0;
        lyn[q] = lam;
    }
}

// ---- Poisson solve adjoint ----
// lambda of the interior unknowns: the wall values p_wall = c43 p_1 - c13 p_2 (radial walls at
// interior j, polar walls at interior i) folded back onto the rows they are formed from.
__global__ void __launch_bounds__(256) k_pwall_adj(Shape sh, double c43, double c13, const double* lp, double* lpi) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.ldk) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    if (!is_interior(sh, i, j) || k >= sh.km - 1) {
        lpi[x] = 0.0;
        return;
    }
    const size_t P = sh.plane, ldk = sh.ldk;
    const int L = sh.im - 1, J = sh.jm - 1;
    double l = lp[x];
    if (i == 1) l += c43 * lp[x - P];
    if (i == 2) l -= c13 * lp[x - 2 * P];
    if (i == L - 1) l += c43 * lp[x + P];
    if (i == L - 2) l -= c13 * lp[x + 2 * P];
    if (j == 1) l += c43 * lp[x - ldk];
    if (j == 2) l -= c13 * lp[x - 2 * ldk];
    if (j == J - 1) l += c43 * lp[x + ldk];
    if (j == J - 2) l -= c13 * lp[x + 2 * ldk];
    lpi[x] = l;
}

// b = (lambda_int - mean) / y on the interior, 0 elsewhere; the mean is sc[0] / ncell.
// (A^T = Y A Y^-1 with Y = diag(y), y = a_i b_j: A^T mu = g  <=>  mu = Y A^-1 Y^-1 g.)
__global__ void __launch_bounds__(256) k_adj_rhs(Shape sh, const double* lpi, const double* sc, double ncell,
                                                 double* b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.ldk) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    if (!is_interior(sh, i, j) || k >= sh.km - 1) {
        b[x] = 0.0;
        return;
    }
    const double y = endw(i, 1, sh.im - 2) * endw(j, 1, sh.jm - 2);
    b[x] = (lpi[x] - sc[0] / ncell) / y;
}

// lambda_src = mask * (mu - y (1.mu) / (y.1)), mu = y z; sc[0] = sum y z; mask = 0 where the
// source clamp |div| <= den * p_dyn_cap was active.
__global__ void __launch_bounds__(256) k_adj_src(Shape sh, DevGeo g, StepConsts c, const double* z, const double* sc,
                                                 double yw, const double* src, double* lsrc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.ldk) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    if (!is_interior(sh, i, j) || k >= sh.km - 1) {
        lsrc[x] = 0.0;
        return;
    }
    const double y = endw(i, 1, sh.im - 2) * endw(j, 1, sh.jm - 2);
    const double cap = g.den[(size_t)i * sh.jm + j] * c.p_dyn_cap;
    const double s = src[x];
    const bool free = (s < cap) && (s > -cap);
    lsrc[x] = free ? y * (z[x] - sc[0] / yw) : 0.0;
}

// lambda_aux = div^T lambda_src on the interior (the walls the divergence reads are constants:
// aux_u = 0 at the rigid lid and aux_v = 0 at the poles).
__global__ void __launch_bounds__(256) k_adj_div(Shape sh, DevGeo g, const double* ls, double* laux) {
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.ldk) return;
    const size_t fs = sh.fs, P = sh.plane, ldk = sh.ldk;
    const size_t x = ((size_t)i * sh.jm + j) * ldk + k;
    if (!is_interior(sh, i, j) || k >= N) {
        laux[x] = 0.0;
        laux[fs + x] = 0.0;
        laux[2 * fs + x] = 0.0;
        return;
    }
    const size_t row = x - k;
    double lu = 0.0, lv = 0.0, lw = 0.0;
    // du_dr(n) = (aux_u[n + r] - aux_u[n - r]) / (2 dr)
    if (i - 1 >= 1) lu += g.inv_2dr * ls[x - P];
    if (i + 1 <= sh.im - 2) lu -= g.inv_2dr * ls[x + P];
    // dv_dthe(n) * inv_rm[i]
    if (j - 1 >= 1) lv += g.inv_2dthe * g.inv_rm[i] * ls[x - ldk];
    if (j + 1 <= sh.jm - 2) lv -= g.inv_2dthe * g.inv_rm[i] * ls[x + ldk];
    // dw_dphi(n) * inv_rmsinthe_s[ij], periodic over the N meridians
    const double cw = g.inv_2dphi * g.inv_rmsinthe_s[(size_t)i * sh.jm + j];
    lw += cw * ls[row + (k == 0 ? N - 1 : k - 1)];
    lw -= cw * ls[row + (k == N - 1 ? 0 : k + 1)];
    laux[x] = lu;
    laux[fs + x] = lv;
    laux[2 * fs + x] = lw;
}

// ---- adjoint of the red-black sweeps (ATU_SOLVE_REFERENCE, planet/PressureSolver.h:318-746) ----
// The forward sweep is: red pass, black pass (each cell of the colour set from its six neighbours,
// acc = num1 (p_r+ + p_r-) + num2_i (p_t+ + p_t-) + num3_ij (p_f+ + p_f-), p = (1 - w) p + w (acc - src)
// / den), then the radial and polar walls from the interior; after the last sweep the corner rows
// and the seam copy. Its adjoint runs the pieces in reverse: seam, corners, then per sweep (last
// first) the walls folded onto the rows they are formed from, the black pass transposed, the red
// pass transposed. A colour pass transposed: a cell x of the colour passes mu_x = w lambda_x / den_x
// to every neighbour it read, with the weight it read it with (num1, num2 of its own i, num3 of its
// own (i, j): the coefficients transposed), keeps (1 - w) lambda_x, and gives -mu_x to lambda_src.

// the seam copy p(k = km-1) = p(k = 0), folded back; one thread per (i, j)
__global__ void __launch_bounds__(256) k_seam_adj(Shape sh, double* lp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= sh.jm) return;
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    lp[row] += lp[row + sh.km - 1];
    lp[row + sh.km - 1] = 0.0;
}

// the corner rows (i = 0, im-1 at j = 0, jm-1) formed from the radial walls at j = 1, 2 and
// jm-2, jm-3, folded back; one thread per k < km-1
__global__ void __launch_bounds__(256) k_corners_adj(Shape sh, double c43, double c13, double* lp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= sh.km - 1) return;
    const size_t ldk = sh.ldk;
    const int J = sh.jm - 1;
    for (int s = 0; s < 2; s++) {
        const size_t b = (size_t)(s == 0 ? 0 : sh.im - 1) * sh.plane + k;
        const double lo = lp[b], hi = lp[b + (size_t)J * ldk];
        lp[b] = 0.0;
        lp[b + (size_t)J * ldk] = 0.0;
        lp[b + ldk] += c43 * lo;
        lp[b + 2 * ldk] -= c13 * lo;
        lp[b + (size_t)(J - 1) * ldk] += c43 * hi;
        lp[b + (size_t)(J - 2) * ldk] -= c13 * hi;
    }
}

// one colour pass transposed, out of place (lin -> lout); lsrc accumulates
__global__ void __launch_bounds__(256) k_rb_pass_adj(Shape sh, DevGeo g, double omega, int colour, const double* lin,
                                                     double* lout, double* lsrc) {
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.ldk) return;
    const size_t ldk = sh.ldk, P = sh.plane;
    const size_t x = ((size_t)i * sh.jm + j) * ldk + k;
    if (k >= N) {   // the seam and the padding: read by no pass
        lout[x] = lin[x];
        return;
    }
    const int jm = sh.jm;
    if (is_interior(sh, i, j) && ((i + j + k) & 1) == colour) {
        const double mu = omega * lin[x] / g.den[(size_t)i * jm + j];
        lsrc[x] -= mu;
        lout[x] = (1.0 - omega) * lin[x];
        return;
    }
    // mu of a neighbour that is an interior cell of the colour (its update read this cell), else 0
    auto mu = [&](int ii, int jj, int kk, size_t xx) -> double {
        if (!is_interior(sh, ii, jj) || ((ii + jj + kk) & 1) != colour) return 0.0;
        return omega * lin[xx] / g.den[(size_t)ii * jm + jj];
    };
    double l = lin[x];
    double r = 0.0;
    if (i + 1 <= sh.im - 1) r += mu(i + 1, j, k, x + P);
    if (i - 1 >= 0) r += mu(i - 1, j, k, x - P);
    l += g.num1 * r;
    double t = 0.0;
    if (j + 1 <= sh.jm - 1) t += mu(i, j + 1, k, x + ldk);
    if (j - 1 >= 0) t += mu(i, j - 1, k, x - ldk);
    l += g.num2[i] * t;
    if (is_interior(sh, i, j)) {   // the phi neighbours of a non-interior cell are not interior
        const int ke = (k == N - 1) ? 0 : k + 1, kw = (k == 0) ? N - 1 : k - 1;
        const size_t row = x - k;
        l += g.num3[(size_t)i * jm + j] * (mu(i, j, ke, row + ke) + mu(i, j, kw, row + kw));
    }
    lout[x] = l;
}

// lambda_src of the clamped source: zero where the clamp |div| <= den * p_dyn_cap was active
__global__ void __launch_bounds__(256) k_src_mask(Shape sh, DevGeo g, StepConsts c, const double* src, double* lsrc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.ldk) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    if (!is_interior(sh, i, j) || k >= sh.km - 1) {
        lsrc[x] = 0.0;
        return;
    }
    const double cap = g.den[(size_t)i * sh.jm + j] * c.p_dyn_cap;
    const double v = src[x];
    if (!((v < cap) && (v > -cap))) lsrc[x] = 0.0;
}

// p_int += g on the interior: the held interior mean makes p_{n+1} depend on mean(p_n)
__global__ void __launch_bounds__(256) k_add_mean(Shape sh, const double* sc, double ncell, double* lp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.km - 1 || !is_interior(sh, i, j)) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    lp[x] += sc[0] / ncell;
}

// ---- loss: mean over the interior (r, theta) of (zonal mean of w - target(theta))^2 ----
// One block per interior (i, j) row. rows[] receives the squared misfit of the row, lw the seed
// scale * d loss / d w = scale * 2 (zm - target) / (ni nj N) on the row's N meridians (0 elsewhere in the row).
__global__ void __launch_bounds__(512) k_loss_rows(Shape sh, const double* w, const double* target, double scale,
                                                   double* rows, double* lw) {
    const int N = sh.km - 1;
    const int j = 1 + blockIdx.x, i = 1 + blockIdx.y;
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    __shared__ double red[512];
    double s = 0.0;
    for (int k = threadIdx.x; k < N; k += blockDim.x) s += w[row + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    const double d = red[0] / N - target[j];
    const double nrow = (double)(sh.im - 2) * (sh.jm - 2);
    for (int k = threadIdx.x; k < sh.ldk; k += blockDim.x) lw[row + k] = (k < N) ? scale * (2.0 * d / (nrow * N)) : 0.0;
    if (threadIdx.x == 0) rows[(size_t)(i - 1) * (sh.jm - 2) + (j - 1)] = d * d;
}

// sum of n values in a fixed order (one block)
__global__ void __launch_bounds__(1024) k_sum(const double* v, size_t n, double* out) {
    __shared__ double red[1024];
    double s = 0.0;
    for (size_t q = threadIdx.x; q < n; q += blockDim.x) s += v[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// per-block partial sums of a long vector (first pass of a two-pass sum)
__global__ void __launch_bounds__(256) k_partial_sum(const double* v, size_t n, double* part) {
    __shared__ double red[256];
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        s += v[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_zero(double* v, size_t n) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) v[q] = 0.0;
}

__global__ void __launch_bounds__(256) k_extrapolate(double* z, const double* z1, const double* z2, const double* z3,
                                                     size_t n) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) z[q] = z3 ? 3.0 * z1[q] - 3.0 * z2[q] + z3[q] : (z2 ? 2.0 * z1[q] - z2[q] : z1[q]);
}

inline dim3 cells_grid(const Shape& sh, int kext) { return dim3((kext + 255) / 256, sh.jm, sh.im); }

}  // namespace

hipError_t launch_rhs_adjoint(const AdjStageArgs& a, hipStream_t s) {
    const int N = a.sh.km - 1;
    const dim3 grid((N + 63) / 64, (a.sh.jm + 3) / 4, a.sh.im), block(64, 4);
    if (a.laux)
        hipLaunchKernelGGL(k_rhs_adjoint<true>, grid, block, 0, s, a);
    else
        hipLaunchKernelGGL(k_rhs_adjoint<false>, grid, block, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_hydro_adjoint(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* x4,
                                const double* pstat, const double* level, const double* lph, double* gref, double* ly,
                                double* parts, hipStream_t s) {
    const size_t fs = sh.fs;
    hipLaunchKernelGGL(k_hydro_adj_col, dim3((sh.km + 255) / 256, sh.jm), dim3(256), 0, s, sh, g, c, x4,
                       x4 + 2 * fs, x4 + 3 * fs, pstat, level, lph, ly, gref);
    hipLaunchKernelGGL(k_hydro_adj_level_sums, dim3(kHydParts, sh.im), dim3(256), 0, s, sh, g, c, x4, pstat,
                       (const double*)gref, parts);
    hipLaunchKernelGGL(k_hydro_adj_level, dim3(kHydParts, sh.im), dim3(256), 0, s, sh, g, c, x4, pstat,
                       (const double*)parts, ly);
    return hipGetLastError();
}

hipError_t launch_bc_adjoint(const Shape& sh, const atu_params& p, double* ly, hipStream_t s) {
    hipLaunchKernelGGL(k_bc_phi_adj, dim3((sh.jm + 255) / 256, sh.im), dim3(256), 0, s, sh, ly);
    hipLaunchKernelGGL(k_bc_theta_adj, dim3((sh.km + 255) / 256, sh.im), dim3(256), 0, s, sh, 4.0 / 3.0, 1.0 / 3.0, ly);
    hipLaunchKernelGGL(k_bc_radius_adj, dim3((sh.km + 255) / 256, sh.jm), dim3(256), 0, s, sh, 4.0 / 3.0, 1.0 / 3.0,
                       p.rigid_lid, p.lid_vw, ly);
    return hipGetLastError();
}

hipError_t launch_final_fold(const Shape& sh, const StepConsts& c, const double* tnext, const double* ly, double* ls,
                             double* lyn, hipStream_t s) {
    hipLaunchKernelGGL(k_final_fold, cells_grid(sh, sh.km), dim3(256), 0, s, sh, c, tnext, ly, ls, lyn);
    return hipGetLastError();
}

hipError_t launch_pwall_adjoint(const Shape& sh, const double* lp, double* lpi, hipStream_t s) {
    hipLaunchKernelGGL(k_pwall_adj, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, 4.0 / 3.0, 1.0 / 3.0, lp, lpi);
    return hipGetLastError();
}

hipError_t launch_adj_rhs(const Shape& sh, const double* lpi, const double* sc, double ncell, double* b,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_adj_rhs, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, lpi, sc, ncell, b);
    return hipGetLastError();
}

hipError_t launch_adj_src(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* z, const double* sc,
                          double yw, const double* src, double* lsrc, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_src, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, g, c, z, sc, yw, src, lsrc);
    return hipGetLastError();
}

hipError_t launch_adj_div(const Shape& sh, const DevGeo& g, const double* lsrc, double* laux, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_div, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, g, lsrc, laux);
    return hipGetLastError();
}

hipError_t launch_rb_sweeps_adjoint(const Shape& sh, const DevGeo& g, const StepConsts& c, int sweeps, double omega,
                                    const double* src, double* lp, double* t1, double* t2, double* lsrc,
                                    hipStream_t s) {
    hipLaunchKernelGGL(k_seam_adj, dim3((sh.jm + 255) / 256, sh.im), dim3(256), 0, s, sh, lp);
    hipLaunchKernelGGL(k_corners_adj, dim3((sh.km - 1 + 255) / 256), dim3(256), 0, s, sh, 4.0 / 3.0, 1.0 / 3.0, lp);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((sh.fs + 255) / 256)), dim3(256), 0, s, lsrc, sh.fs);
    for (int q = 0; q < sweeps; q++) {
        hipLaunchKernelGGL(k_pwall_adj, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, 4.0 / 3.0, 1.0 / 3.0,
                           (const double*)lp, t1);
        hipLaunchKernelGGL(k_rb_pass_adj, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, g, omega, 1, (const double*)t1,
                           t2, lsrc);
        hipLaunchKernelGGL(k_rb_pass_adj, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, g, omega, 0, (const double*)t2,
                           lp, lsrc);
    }
    hipLaunchKernelGGL(k_src_mask, cells_grid(sh, sh.ldk), dim3(256), 0, s, sh, g, c, src, lsrc);
    return hipGetLastError();
}

hipError_t launch_add_mean(const Shape& sh, const double* sc, double ncell, double* lp, hipStream_t s) {
    hipLaunchKernelGGL(k_add_mean, cells_grid(sh, sh.km), dim3(256), 0, s, sh, sc, ncell, lp);
    return hipGetLastError();
}

hipError_t launch_loss(const Shape& sh, const double* w, const double* target, double scale, double* rows,
                       double* lw, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_rows, dim3(sh.jm - 2, sh.im - 2), dim3(512), 0, s, sh, w, target, scale, rows, lw);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, (const double*)rows, (size_t)(sh.im - 2) * (sh.jm - 2), out);
    return hipGetLastError();
}

hipError_t launch_sum(const double* v, size_t n, double* part, int nparts, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_partial_sum, dim3(nparts), dim3(256), 0, s, v, n, part);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, (const double*)part, (size_t)nparts, out);
    return hipGetLastError();
}

hipError_t launch_extrapolate(double* z, const double* z1, const double* z2, const double* z3, size_t n,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_extrapolate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z, z1, z2, z3, n);
    return hipGetLastError();
}

hipError_t launch_zero(double* v, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, n);
    return hipGetLastError();
}

}  // namespace atu
