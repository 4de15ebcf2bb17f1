// Hydrostatic prep (planet/RungeKutta_Uran_Turb.cpp:18-177) and the domain boundary conditions
// (planet/BoundaryConditions.h:152-518, planet/BC_Uran.h:91-171, field lists planet/BC_Uran.cpp).

#include "atu_kernels.h"

namespace atu {
namespace {

// ---- level means: sum over the full (j, k) plane of sin(theta) * b, skipping T <= 0 and
// non-finite b (computeBuoyancyRefLevel, :18-37).
//
// The reference adds the terms of a level one after another in (j, k) order, and the buoyancy
// anomaly divides by this mean, which magnifies any rounding difference of the sum ~200-fold in
// p_hydro. So the sum is formed in exactly that order: one workgroup per level. Its lanes load and
// form 64 consecutive terms at a time (a skipped term contributes an exact +0.0); a group of
// kLevelGroup such chunks is parked in LDS and added in sequence while the loads of the group three
// ahead are in flight, the LDS reads of each batch of the chain issued while the previous batch adds.
// Two waves per level: wave 0 runs the chain of the weighted terms, wave 1 that of the weights, each
// on its own SIMD. The chain is ~65k dependent adds per level; the step runs this kernel on a side
// stream, concurrently with the previous step's Poisson solve, which it does not depend on.
constexpr int kLevelGroup = 8;
__global__ void __launch_bounds__(128) k_level_serial(Shape sh, DevGeo g, StepConsts c, const double* t,
                                                      const double* pstat, double* level) {
    const int i = blockIdx.x;
    const int n = sh.jm * sh.km;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;   // wave 0: the sum, wave 1: the weights
    constexpr int kG = kLevelGroup * 64;
    __shared__ double tv[2][kG];
    __shared__ double wsum;
    double* mine = tv[wv];
    // T, p_stat and sin(theta) of three groups in flight: the loads of a group go out while the
    // chain of the group three before it runs (the chain shares the HBM with the stage kernels)
    double tA[kLevelGroup], pA[kLevelGroup], wA[kLevelGroup];
    double tB[kLevelGroup], pB[kLevelGroup], wB[kLevelGroup];
    double tC[kLevelGroup], pC[kLevelGroup], wC[kLevelGroup];
    auto load_group = [&](double (&rt)[kLevelGroup], double (&rp)[kLevelGroup], double (&rw)[kLevelGroup], int q0) {
#pragma unroll
        for (int u = 0; u < kLevelGroup; u++) {
            const int q = q0 + u * 64 + lane;
            rt[u] = 0.0;   // a cell past the plane: T = 0, skipped
            rp[u] = 0.0;
            rw[u] = 0.0;
            if (q < n) {
                const int j = q / sh.km, k = q - j * sh.km;
                const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
                rt[u] = t[x];
                rp[u] = pstat[x];
                rw[u] = g.sin_weight[j];
            }
        }
    };
    double acc = 0.0;
    auto step = [&](double (&rt)[kLevelGroup], double (&rp)[kLevelGroup], double (&rw)[kLevelGroup], int q0) {
#pragma unroll
        for (int u = 0; u < kLevelGroup; u++) {
            double term = 0.0, wt = 0.0;
            const double T = rt[u];
            if (T > 0.0) {
                const double b = c.g * rp[u] / (c.hyd_bnum * T * c.t_ref);
                if (isfinite(b)) {
                    term = rw[u] * b;
                    wt = rw[u];
                }
            }
            mine[u * 64 + lane] = (wv == 0) ? term : wt;
        }
        __syncthreads();
        load_group(rt, rp, rw, q0 + 3 * kG);
        // the whole group (terms past the plane are +0.0, which leaves the sum unchanged) in
        // batches held in two register sets: the LDS reads of one batch are in flight while the
        // other batch adds
        constexpr int kBatch = 8;
        double xa[kBatch], xb[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) xa[u] = mine[u];
#pragma unroll
        for (int l0 = 0; l0 < kG; l0 += 2 * kBatch) {
#pragma unroll
            for (int u = 0; u < kBatch; u++) xb[u] = mine[l0 + kBatch + u];
#pragma unroll
            for (int u = 0; u < kBatch; u++) acc += xa[u];
            const int l2 = (l0 + 2 * kBatch < kG) ? l0 + 2 * kBatch : 0;
#pragma unroll
            for (int u = 0; u < kBatch; u++) xa[u] = mine[l2 + u];
#pragma unroll
            for (int u = 0; u < kBatch; u++) acc += xb[u];
        }
        __syncthreads();
    };
    load_group(tA, pA, wA, 0);
    load_group(tB, pB, wB, kG);
    load_group(tC, pC, wC, 2 * kG);
    for (int q0 = 0; q0 < n; q0 += 3 * kG) {
        step(tA, pA, wA, q0);
        if (q0 + kG < n) step(tB, pB, wB, q0 + kG);
        if (q0 + 2 * kG < n) step(tC, pC, wC, q0 + 2 * kG);
    }
    if (wv == 1 && lane == 0) wsum = acc;
    __syncthreads();
    if (wv == 0 && lane == 0) level[i] = (wsum > 0.0) ? acc / wsum : 0.0;
}

// ---- the column scan: p_hydro[0] = 0, p_hydro[i] = p_hydro[i-1] + 0.5 (f(i-1) + f(i)) dr with
// f = buoyancy anomaly relative to the level mean (BUOY_RHO_LEVEL) + (v^2 + w^2)/rm + Coriolis
// (METRIC_SPLIT + CORIOLIS_SPLIT branch, :104-161). One thread per (j, k) column.
__global__ void __launch_bounds__(256) k_hydro_scan(Shape sh, DevGeo g, StepConsts c, const double* t,
                                                    const double* v, const double* w, const double* pstat,
                                                    const double* level, double* phydro) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= sh.km) return;
    const double cor = g.cor_hydro[j];
    auto f_at = [&](int i) -> double {
        const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
        const double T = t[x];
        double buoy = 0.0;
        if (T > 0.0) {
            const double ref = level[i];
            const double b = c.g * pstat[x] / (c.hyd_bnum * T * c.t_ref) - ref;
            const double f = -c.nd_rel * 1.0 * 1.0 * b / ref;
            buoy = isfinite(f) ? f : 0.0;
        }
        const double vv = v[x], ww = w[x];
        return buoy + (vv * vv + ww * ww) / g.rm[i] + cor * ww;
    };
    double p = 0.0;
    double f_prev = f_at(0);
    phydro[(size_t)j * sh.ldk + k] = 0.0;
    for (int i = 1; i < sh.im; i++) {
        const double fi = f_at(i);
        p = p + 0.5 * (f_prev + fi) * c.dr;
        const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
        phydro[x] = p;
        f_prev = fi;
    }
}

template <typename A, typename B>
__global__ void __launch_bounds__(256) k_convert(const A* src, B* dst, size_t n) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = (B)src[q];
}

__global__ void __launch_bounds__(256) k_fill(float* dst, float v, size_t n, const int* done) {
    if (done && *done) return;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = v;
}

// ---- boundary conditions ----
struct BcCtx {
    Shape sh;
    BcTargets t;
    double c43, c13;
    int lid_vw, rigid_lid, radius_positive;
    int f0, f1;     // the prognostic fields [f0, f1) of this pass
    int rest;       // this pass also does the frozen fields and the field-specific forms
};

__device__ __forceinline__ void bc_put(const BcCtx& b, size_t q, double val) {
    b.t.y[q] = val;
    if (b.t.m64a) b.t.m64a[q] = val;
    if (b.t.m64b) b.t.m64b[q] = val;
    if (b.t.m32a) b.t.m32a[q] = (float)val;
    if (b.t.m32b) b.t.m32b[q] = (float)val;
    if (b.t.m32c) b.t.m32c[q] = (float)val;
}

constexpr int kNumBcFields = 17;   // t, u, v, w and the 13 species: the prognostic part of the lists
// the frozen inputs the lists name: the reaction mass fluxes and the NH4SH flux limiter
__device__ __forceinline__ int frozen_bc(int q) {
    return q == 0 ? ATU_MF_H2S - ATU_NUE
                  : (q == 1 ? ATU_MF_NH3 - ATU_NUE : (q == 2 ? ATU_MF_NH4SH - ATU_NUE : ATU_FLUXLIM_NH4SH - ATU_NUE));
}
__device__ __forceinline__ void fz_put(const BcCtx& b, size_t q, double val) {
    b.t.frozen[q] = val;
    if (b.t.frozen32) b.t.frozen32[q] = (float)val;
}

// The BC kernels take one field per blockIdx.z: the prognostic fields f0 + z, then (with `rest`)
// the four frozen fields. The fields do not interact in the BC, and each field's own forms follow
// its NEUMANN form in the same thread, as in the reference's order.

// bcRadius with margin 0 (every j, every k), NEUMANN form, then the rigid lid on u, the species
// floor (floorRadialSpecies) and the lid form of v and w (BC_LID_VW).
__global__ void __launch_bounds__(256) k_bc_radius(BcCtx b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= b.sh.km) return;
    const Shape& sh = b.sh;
    const size_t col = (size_t)j * sh.ldk + k;
    const size_t P = sh.plane;
    const int L = sh.im - 1;
    const int z = blockIdx.z, nf = b.f1 - b.f0;
    if (z >= nf) {   // a frozen field
        const size_t o = (size_t)frozen_bc(z - nf) * sh.fs + col;
        const double* F = b.t.frozen + o;
        const double lo = b.c43 * F[1 * P] - b.c13 * F[2 * P];
        const double hi = b.c43 * F[(size_t)(L - 1) * P] - b.c13 * F[(size_t)(L - 2) * P];
        fz_put(b, o, lo);
        fz_put(b, o + (size_t)L * P, hi);
        return;
    }
    const int f = b.f0 + z;
    const double* F = b.t.y + (size_t)f * sh.fs + col;
    const double lo = b.c43 * F[1 * P] - b.c13 * F[2 * P];
    const double hi = b.c43 * F[(size_t)(L - 1) * P] - b.c13 * F[(size_t)(L - 2) * P];
    bc_put(b, (size_t)f * sh.fs + col, lo);
    bc_put(b, (size_t)f * sh.fs + col + (size_t)L * P, hi);
    if (!b.rest) return;
    if (f == ATU_U && b.rigid_lid) {
        bc_put(b, (size_t)ATU_U * sh.fs + col, 0.0);
        bc_put(b, (size_t)ATU_U * sh.fs + col + (size_t)L * P, 0.0);
    }
    if (b.radius_positive && f >= ATU_CH4 && f <= ATU_NH4SH) {
        const double f0 = F[0], fl = F[(size_t)L * P];
        if (f0 < 0.0) {
            const double f1 = F[P];
            bc_put(b, (size_t)f * sh.fs + col, b.radius_positive == 2 ? (0.0 < f1 ? f1 : 0.0) : 0.0);
        }
        if (fl < 0.0) {
            const double f1 = F[(size_t)(L - 1) * P];
            bc_put(b, (size_t)f * sh.fs + col + (size_t)L * P, b.radius_positive == 2 ? (0.0 < f1 ? f1 : 0.0) : 0.0);
        }
    }
    if ((b.lid_vw == 2 || b.lid_vw == 3) && (f == ATU_V || f == ATU_W)) {
        const double fa = F[(size_t)(L - 1) * P], fb = F[(size_t)(L - 2) * P], fc = F[(size_t)(L - 3) * P];
        const double val = (b.lid_vw == 2) ? 2.0 * fa - fb : fc - 3.0 * fb + 3.0 * fa;
        bc_put(b, (size_t)f * sh.fs + col + (size_t)L * P, val);
    }
}

// bcTheta with margin 0: v, w zeroed at both poles, t, u and the species by the NEUMANN form.
__global__ void __launch_bounds__(256) k_bc_theta(BcCtx b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (k >= b.sh.km) return;
    const Shape& sh = b.sh;
    const int ldk = sh.ldk;
    const int Jl = sh.jm - 1;
    const size_t base = (size_t)i * sh.plane + k;
    const int z = blockIdx.z, nf = b.f1 - b.f0;
    if (z >= nf) {
        const size_t o = (size_t)frozen_bc(z - nf) * sh.fs + base;
        fz_put(b, o, 0.0);
        fz_put(b, o + (size_t)Jl * ldk, 0.0);
        return;
    }
    const int f = b.f0 + z;
    const size_t q0 = (size_t)f * sh.fs + base;
    if (f == ATU_V || f == ATU_W) {
        bc_put(b, q0, 0.0);
        bc_put(b, q0 + (size_t)Jl * ldk, 0.0);
        return;
    }
    const double* F = b.t.y + q0;
    const double lo = b.c43 * F[(size_t)1 * ldk] - b.c13 * F[(size_t)2 * ldk];
    const double hi = b.c43 * F[(size_t)(Jl - 1) * ldk] - b.c13 * F[(size_t)(Jl - 2) * ldk];
    bc_put(b, q0, lo);
    bc_put(b, q0 + (size_t)Jl * ldk, hi);
}

// bcPhi with the periodic seam: k = km-1 is the copy of k = 0.
__global__ void __launch_bounds__(256) k_bc_phi(BcCtx b) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= b.sh.jm) return;
    const Shape& sh = b.sh;
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    const int z = blockIdx.z, nf = b.f1 - b.f0;
    if (z >= nf) {
        const size_t o = (size_t)frozen_bc(z - nf) * sh.fs + row;
        fz_put(b, o + sh.km - 1, b.t.frozen[o]);
        return;
    }
    const size_t q = (size_t)(b.f0 + z) * sh.fs + row;
    bc_put(b, q + sh.km - 1, b.t.y[q]);
}

// clampNegativeSpecies (planet/BC_Uran.h:110-128): every species value that is not >= 0 -> 0.
__global__ void __launch_bounds__(256) k_species_clamp(BcCtx b) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= b.sh.km) return;
    const size_t x = ((size_t)i * b.sh.jm + j) * b.sh.ldk + k;
    for (int f = ATU_CH4; f <= ATU_NH4SH; f++) {
        const size_t q = (size_t)f * b.sh.fs + x;
        if (!(b.t.y[q] >= 0.0)) bc_put(b, q, 0.0);
    }
}

}  // namespace

hipError_t launch_level(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t, const double* pstat,
                        double* level_out, hipStream_t s) {
    hipLaunchKernelGGL(k_level_serial, dim3(sh.im), dim3(128), 0, s, sh, g, c, t, pstat, level_out);
    return hipGetLastError();
}

hipError_t launch_hydro_scan(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t, const double* v,
                             const double* w, const double* pstat, const double* level, double* phydro,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_hydro_scan, dim3((sh.km + 255) / 256, sh.jm), dim3(256), 0, s, sh, g, c, t, v, w, pstat,
                       level, phydro);
    return hipGetLastError();
}

hipError_t launch_hydro(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t, const double* v,
                        const double* w, const double* pstat, double* level_out, double* phydro,
                        hipStream_t s) {
    const hipError_t e = launch_level(sh, g, c, t, pstat, level_out, s);
    if (e != hipSuccess) return e;
    return launch_hydro_scan(sh, g, c, t, v, w, pstat, level_out, phydro, s);
}

hipError_t launch_convert(const double* src, float* dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_convert(const float* src, double* dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_fill(float* dst, float v, size_t n, hipStream_t s, const int* done) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, v, n, done);
    return hipGetLastError();
}

hipError_t launch_bc_part(const Shape& sh, const atu_params& p, const BcTargets& t, int f0, int f1, bool rest,
                          hipStream_t s) {
    BcCtx b;
    b.sh = sh;
    b.t = t;
    b.c43 = 4.0 / 3.0;
    b.c13 = 1.0 / 3.0;
    b.lid_vw = p.lid_vw;
    b.rigid_lid = p.rigid_lid;
    b.radius_positive = p.radius_positive;
    b.f0 = f0 < 0 ? 0 : f0;
    b.f1 = f1 > kNumBcFields ? kNumBcFields : f1;
    b.rest = rest ? 1 : 0;
    const int nz = (b.f1 - b.f0) + (rest ? 4 : 0);
    if (nz <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bc_radius, dim3((sh.km + 255) / 256, sh.jm, nz), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_bc_theta, dim3((sh.km + 255) / 256, sh.im, nz), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_bc_phi, dim3((sh.jm + 255) / 256, sh.im, nz), dim3(256), 0, s, b);
    if (rest && p.species_clamp)
        hipLaunchKernelGGL(k_species_clamp, dim3((sh.km + 255) / 256, sh.jm, sh.im), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_bc(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s) {
    return launch_bc_part(sh, p, t, 0, kNumBcFields, true, s);
}

}  // namespace atu
