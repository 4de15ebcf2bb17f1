// Hydrostatic prep (planet/RungeKutta_Uran_Turb.cpp:18-177) and the domain boundary conditions
// (planet/BoundaryConditions.h:152-518, planet/BC_Uran.h:91-171, field lists planet/BC_Uran.cpp).

#include "atu_kernels.h"

namespace atu {
namespace {

// ---- level means: sum over the full (j, k) plane of sin(theta) * b, skipping T <= 0 and
// non-finite b (computeBuoyancyRefLevel, :18-37).
//
// The reference adds the terms of a level one after another in (j, k) order. Here every level sum
// (and its sum of weights) runs in one fixed parallel order, which the oracle restates term for
// term: per (i, j) row, the km terms in chunks of 64 consecutive k, each chunk reduced by the halving
// tree v[l] += v[l + h], h = 32..1 (a k past the row, or a skipped cell, is +0.0), the chunk results
// of a row added in sequence to 0.0; the jm row sums dealt to 64 partials (partial t adds rows t,
// t + 64, ... in sequence to 0.0), and those reduced by the halving tree. The row sums run one
// wavefront per row (k_level_rows), the partials and the tree one wavefront per level
// (k_level_final).
__device__ __forceinline__ double tree64(double v) {
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) v += __shfl_down(v, h, 64);
    return v;   // the tree's result in lane 0
}

// The row sums: one wavefront per (i, j) row (four rows per workgroup), into rows[i jm + j] (terms) and
// rows[im jm + i jm + j] (weights).
__global__ void __launch_bounds__(256) k_level_rows(Shape sh, DevGeo g, StepConsts c, const double* t,
                                                    const double* pstat, double* rows) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6), i = blockIdx.y;
    if (j >= sh.jm) return;
    const double wgt = g.sin_weight[j];
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    double s = 0.0, ws = 0.0;
    for (int k0 = 0; k0 < sh.km; k0 += 64) {
        const int k = k0 + lane;
        double term = 0.0, wt = 0.0;
        if (k < sh.km) {
            const double T = t[row + k];
            if (T > 0.0) {
                const double b = c.g * pstat[row + k] / (c.hyd_bnum * T * c.t_ref);
                if (isfinite(b)) {
                    term = wgt * b;
                    wt = wgt;
                }
            }
        }
        term = tree64(term);
        wt = tree64(wt);
        s = s + term;
        ws = ws + wt;
    }
    if (lane == 0) {
        const size_t q = (size_t)i * sh.jm + j;
        rows[q] = s;
        rows[(size_t)sh.im * sh.jm + q] = ws;
    }
}

// The level means from the row sums: one wavefront per level.
__global__ void __launch_bounds__(64) k_level_final(Shape sh, const double* rows, double* level) {
    const int i = blockIdx.x, lane = threadIdx.x;
    const double* rs = rows + (size_t)i * sh.jm;
    const double* rw = rows + (size_t)sh.im * sh.jm + (size_t)i * sh.jm;
    double ps = 0.0, pw = 0.0;
    for (int q = lane; q < sh.jm; q += 64) {
        ps = ps + rs[q];
        pw = pw + rw[q];
    }
    ps = tree64(ps);
    pw = tree64(pw);
    if (lane == 0) level[i] = (pw > 0.0) ? ps / pw : 0.0;
}

// ---- the column scan: p_hydro[0] = 0, p_hydro[i] = p_hydro[i-1] + 0.5 (f(i-1) + f(i)) dr with
// f = buoyancy anomaly relative to the level mean (BUOY_RHO_LEVEL) + (v^2 + w^2)/rm + Coriolis
// (METRIC_SPLIT + CORIOLIS_SPLIT branch, :104-161). One thread per (j, k) column.
__global__ void __launch_bounds__(256) k_hydro_scan(Shape sh, DevGeo g, StepConsts c, const double* __restrict__ t,
                                                    const double* __restrict__ v, const double* __restrict__ w,
                                                    const double* __restrict__ pstat, const double* __restrict__ level,
                                                    double* __restrict__ phydro) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= sh.km) return;
    const double cor = g.cor_hydro[j];
    auto f_of = [&](int i, double T, double ps, double vv, double ww) -> double {
        double buoy = 0.0;
        if (T > 0.0) {
            const double ref = level[i];
            const double b = c.g * ps / (c.hyd_bnum * T * c.t_ref) - ref;
            const double f = -c.nd_rel * 1.0 * 1.0 * b / ref;
            buoy = isfinite(f) ? f : 0.0;
        }
        return buoy + (vv * vv + ww * ww) / g.rm[i] + cor * ww;
    };
    // kScanBatch levels at a time: every value they read is loaded first (pstat too, whatever the sign
    // of t), so that the loads are all in flight before the sum's dependent adds consume them
    constexpr int kScanBatch = 8;
    const size_t col = (size_t)j * sh.ldk + k;
    double p = 0.0;
    double f_prev = f_of(0, t[col], pstat[col], v[col], w[col]);
    phydro[col] = 0.0;
    for (int i0 = 1; i0 < sh.im; i0 += kScanBatch) {
        double T[kScanBatch], ps[kScanBatch], vv[kScanBatch], ww[kScanBatch];
#pragma unroll
        for (int u = 0; u < kScanBatch; u++) {
            const size_t x = (size_t)min(i0 + u, sh.im - 1) * sh.plane + col;
            T[u] = t[x];
            ps[u] = pstat[x];
            vv[u] = v[x];
            ww[u] = w[x];
        }
#pragma unroll
        for (int u = 0; u < kScanBatch; u++) {
            const int i = i0 + u;
            if (i >= sh.im) break;
            const double f = f_of(i, T[u], ps[u], vv[u], ww[u]);
            p = p + 0.5 * (f_prev + f) * c.dr;
            phydro[(size_t)i * sh.plane + col] = p;
            f_prev = f;
        }
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

// clampNegativeSpecies (planet/BC_Uran.h:110-128): every species value that is not >= 0 (NaN
// included) -> 0. A flat pass over each species field's storage, blockIdx.y = species, two values
// per thread (the padding past km holds zeros, which the test leaves alone); a value is written only
// where it changes, to the state and to every mirror.
__global__ void __launch_bounds__(256) k_species_clamp(BcCtx b) {
    const size_t q2 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * q2 >= b.sh.fs) return;
    const size_t q = (size_t)(ATU_CH4 + blockIdx.y) * b.sh.fs + 2 * q2;
    const double2 v = *reinterpret_cast<const double2*>(b.t.y + q);
    if (!(v.x >= 0.0)) bc_put(b, q, 0.0);
    if (!(v.y >= 0.0)) bc_put(b, q + 1, 0.0);
}

// The clamp on the shell of rows the BC writes or reads: the planes i = 0, 1, 2, im-3, im-2, im-1 (every
// j), then at the other levels the rows j = 0, 1, 2, jm-3, jm-2, jm-1; blockIdx.x = row of that list,
// blockIdx.y = species, one thread per value of the row's storage (the padding holds zeros).
__global__ void __launch_bounds__(256) k_species_clamp_shell(BcCtx b) {
    const Shape& sh = b.sh;
    const int r = blockIdx.x;
    int i, j;
    if (r < 6 * sh.jm) {
        const int pi = r / sh.jm;
        i = pi < 3 ? pi : sh.im - 6 + pi;
        j = r - pi * sh.jm;
    } else {
        const int t = r - 6 * sh.jm, pj = t % 6;
        i = 3 + t / 6;
        j = pj < 3 ? pj : sh.jm - 6 + pj;
    }
    const size_t q0 = (size_t)(ATU_CH4 + blockIdx.y) * sh.fs + ((size_t)i * sh.jm + j) * sh.ldk;
    for (int k = threadIdx.x; k < sh.ldk; k += blockDim.x) {
        const double v = b.t.y[q0 + k];
        if (!(v >= 0.0)) bc_put(b, q0 + k, 0.0);
    }
}

}  // namespace

hipError_t launch_level(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t, const double* pstat,
                        double* rows, double* level_out, hipStream_t s) {
    hipLaunchKernelGGL(k_level_rows, dim3((sh.jm + 3) / 4, sh.im), dim3(256), 0, s, sh, g, c, t, pstat, rows);
    hipLaunchKernelGGL(k_level_final, dim3(sh.im), dim3(64), 0, s, sh, rows, level_out);
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
                        const double* w, const double* pstat, double* rows, double* level_out, double* phydro,
                        hipStream_t s) {
    const hipError_t e = launch_level(sh, g, c, t, pstat, rows, level_out, s);
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
                          hipStream_t s, bool clamp) {
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
    if (rest && clamp && p.species_clamp)   // fs is a multiple of ldk, itself a multiple of 32: even
        hipLaunchKernelGGL(k_species_clamp, dim3((unsigned)((sh.fs / 2 + 255) / 256), ATU_NH4SH - ATU_CH4 + 1),
                           dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_bc(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s, bool clamp) {
    return launch_bc_part(sh, p, t, 0, kNumBcFields, true, s, clamp);
}

hipError_t launch_species_clamp(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s) {
    if (!p.species_clamp) return hipSuccess;
    BcCtx b;
    b.sh = sh;
    b.t = t;
    hipLaunchKernelGGL(k_species_clamp, dim3((unsigned)((sh.fs / 2 + 255) / 256), ATU_NH4SH - ATU_CH4 + 1), dim3(256),
                       0, s, b);
    return hipGetLastError();
}

hipError_t launch_species_clamp_shell(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s) {
    if (!p.species_clamp) return hipSuccess;
    if (sh.im < 6 || sh.jm < 6) return hipErrorInvalidValue;
    BcCtx b;
    b.sh = sh;
    b.t = t;
    const int rows = 6 * sh.jm + 6 * (sh.im - 6);
    hipLaunchKernelGGL(k_species_clamp_shell, dim3((unsigned)rows, ATU_NH4SH - ATU_CH4 + 1), dim3(256), 0, s, b);
    return hipGetLastError();
}

}  // namespace atu
