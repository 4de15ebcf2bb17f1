// Red-black relaxation of the dynamic-pressure Poisson equation: prepareProjectionBoundaries
// (planet/Pressure_Uran.cpp:30-146), the clamped divergence source and the colour passes of
// PressureSolver<Planet>::run (planet/PressureSolver.h:318-495) with its wall passes (:710-746).
//
// Layout of one colour pass: a wavefront holds 64 cells of one colour in one (i, j) row, i.e. every
// other k; its loads of p touch the full row segment, which the other colour's pass reuses.
//
// The reference forms the radial and polar wall values after each sweep from the interior it has
// just relaxed. Here the black pass forms them: for every wall cell exactly one of its two interior
// sources (i = 1, 2 or j = 1, 2 and their mirrors) is black, so the thread relaxing it has the new
// value of one and reads the other, already final, from memory. No cell reads a wall value during
// the black pass that the same pass writes, so the result equals a separate wall pass.

#include "atu_kernels.h"

namespace atu {
namespace {

// ---- prepareProjectionBoundaries on aux (rigid lid; rhs_* are not kept on this path) ----
template <typename S>
__global__ void __launch_bounds__(256) k_proj_walls(Shape sh, int rigid_lid, S* aux) {
    // index set 1: radial planes, j = 1..jm-2, k = 0..km-2; index set 2: polar rows, i = 1..im-2;
    // index set 3: the seam copy (k = km-1 <- k = 0) of the rows neither wall set writes (the interior
    // rows and the four corner rows); the wall sets write the seam copy of the values they form at k = 0
    const int nk = sh.km - 1;
    const int n1 = (sh.jm - 2) * nk;
    const int n2 = (sh.im - 2) * nk;
    const int n3 = (sh.im - 2) * (sh.jm - 2) + 4;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t fs = sh.fs, P = sh.plane;
    const int ldk = sh.ldk;
    if (t < n1) {
        const int j = 1 + t / nk, k = t % nk;
        const size_t col = (size_t)j * ldk + k;
        const int L = sh.im - 1;
        S* au = aux;
        S* av = aux + fs;
        S* aw = aux + 2 * fs;
        S ulo, uhi;
        if (rigid_lid) {
            ulo = S(0.0);
            uhi = S(0.0);
        } else {
            ulo = au[col + 3 * P] - S(3.0) * au[col + 2 * P] + S(3.0) * au[col + P];
            uhi = au[col + (L - 3) * P] - S(3.0) * au[col + (L - 2) * P] + S(3.0) * au[col + (L - 1) * P];
        }
        au[col] = ulo;
        au[col + L * P] = uhi;
        if (k == 0) {
            au[col + nk] = ulo;
            au[col + L * P + nk] = uhi;
        }
        for (int f = 0; f < 2; f++) {
            S* a = (f == 0) ? av : aw;
            const S lo = a[col + 3 * P] - S(3.0) * a[col + 2 * P] + S(3.0) * a[col + P];
            const S hi = a[col + (L - 3) * P] - S(3.0) * a[col + (L - 2) * P] + S(3.0) * a[col + (L - 1) * P];
            a[col] = lo;
            a[col + L * P] = hi;
            if (k == 0) {
                a[col + nk] = lo;
                a[col + L * P + nk] = hi;
            }
        }
    } else if (t < n1 + n2) {
        const int q = t - n1;
        const int i = 1 + q / nk, k = q % nk;
        const size_t b = (size_t)i * P + k;
        const int J = sh.jm - 1;
        S* au = aux;
        const S lo = au[b + 3 * ldk] - S(3.0) * au[b + 2 * ldk] + S(3.0) * au[b + ldk];
        const S hi = au[b + (size_t)(J - 3) * ldk] - S(3.0) * au[b + (size_t)(J - 2) * ldk]
                     + S(3.0) * au[b + (size_t)(J - 1) * ldk];
        au[b] = lo;
        au[b + (size_t)J * ldk] = hi;
        aux[fs + b] = S(0.0);
        aux[2 * fs + b] = S(0.0);
        aux[fs + b + (size_t)J * ldk] = S(0.0);
        aux[2 * fs + b + (size_t)J * ldk] = S(0.0);
        if (k == 0) {
            au[b + nk] = lo;
            au[b + (size_t)J * ldk + nk] = hi;
            aux[fs + b + nk] = S(0.0);
            aux[2 * fs + b + nk] = S(0.0);
            aux[fs + b + (size_t)J * ldk + nk] = S(0.0);
            aux[2 * fs + b + (size_t)J * ldk + nk] = S(0.0);
        }
    } else if (t < n1 + n2 + n3) {
        const int q = t - n1 - n2, nin = (sh.im - 2) * (sh.jm - 2);
        int i, j;
        if (q < nin) {
            i = 1 + q / (sh.jm - 2);
            j = 1 + q % (sh.jm - 2);
        } else {
            i = (q - nin) & 1 ? sh.im - 1 : 0;
            j = (q - nin) & 2 ? sh.jm - 1 : 0;
        }
        const size_t row = ((size_t)i * sh.jm + j) * ldk;
        const S a0 = aux[row], a1 = aux[fs + row], a2 = aux[2 * fs + row];
        aux[row + nk] = a0;
        aux[fs + row + nk] = a1;
        aux[2 * fs + row + nk] = a2;
    }
}

// ---- the clamped divergence source div(aux) (planet/PressureSolver.h:448-487) ----
template <typename S>
__global__ void __launch_bounds__(256) k_src(Shape sh, DevGeo g, StepConsts c, const S* aux, double* src) {
    const int nk = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 1 + blockIdx.z;
    if (k >= nk || j > sh.jm - 2) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    src[x] = poisson_src_at<S>(sh, g, c, aux, i, j, k, x);
}

// ---- one colour pass; colour 1 also forms the wall values ----
template <typename P>
__global__ void __launch_bounds__(256) k_rb_pass(Shape sh, DevGeo g, double c43, double c13, P* p,
                                                 const double* src, int colour, double omega, int form_walls) {
    const int nk = sh.km - 1;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 1 + blockIdx.z;
    if (j > sh.jm - 2) return;
    const int k = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x) + ((colour + i + j) & 1);
    if (k >= nk) return;
    const int ldk = sh.ldk;
    const size_t PL = sh.plane;
    const size_t row = ((size_t)i * sh.jm + j) * ldk;
    const size_t x = row + k;
    const size_t xw = row + (k == 0 ? nk - 1 : k - 1), xe = row + (k == nk - 1 ? 0 : k + 1);
    const size_t ij = (size_t)i * sh.jm + j;
    const double num1 = g.num1, num2 = g.num2[i], num3 = g.num3[ij], den = g.den[ij];
    const double acc = ((double)p[x + PL] + (double)p[x - PL]) * num1
                       + ((double)p[x + ldk] + (double)p[x - ldk]) * num2
                       + ((double)p[xe] + (double)p[xw]) * num3;
    double pn = (acc - src[x]) / den;
    if (omega != 1.0) pn = (1.0 - omega) * (double)p[x] + omega * pn;
    const P pv = (P)pn;
    p[x] = pv;
    if (k == 0) p[x + nk] = pv;
    if (colour != 1 || !form_walls) return;
    // wall values of this sweep (planet/PressureSolver.h:718-746)
    auto put = [&](size_t q, double val) {
        p[q] = (P)val;
        if (k == 0) p[q + nk] = (P)val;
    };
    const int im = sh.im, jm = sh.jm;
    if (i == 1 || i == 2) {
        const double p1 = (i == 1) ? (double)pv : (double)p[x - PL];
        const double p2 = (i == 2) ? (double)pv : (double)p[x + PL];
        put(x - (size_t)i * PL, c43 * p1 - c13 * p2);
    }
    if (i == im - 2 || i == im - 3) {
        const double pa = (i == im - 2) ? (double)pv : (double)p[x + PL];
        const double pb = (i == im - 3) ? (double)pv : (double)p[x - PL];
        put(x + (size_t)(im - 1 - i) * PL, c43 * pa - c13 * pb);
    }
    if (j == 1 || j == 2) {
        const double p1 = (j == 1) ? (double)pv : (double)p[x - ldk];
        const double p2 = (j == 2) ? (double)pv : (double)p[x + ldk];
        put(x - (size_t)j * ldk, c43 * p1 - c13 * p2);
    }
    if (j == jm - 2 || j == jm - 3) {
        const double pa = (j == jm - 2) ? (double)pv : (double)p[x + ldk];
        const double pb = (j == jm - 3) ? (double)pv : (double)p[x - ldk];
        put(x + (size_t)(jm - 1 - j) * ldk, c43 * pa - c13 * pb);
    }
}

// The corner rows: the polar pass at i = 0 and im-1 reads the radial walls of the same sweep.
template <typename P>
__global__ void k_p_corners(Shape sh, double c43, double c13, P* p, const int* skip) {
    if (skip && *skip) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nk = sh.km - 1;
    if (k >= nk) return;
    const int ldk = sh.ldk;
    for (int s = 0; s < 2; s++) {
        const int i = s == 0 ? 0 : sh.im - 1;
        const size_t b = (size_t)i * sh.plane + k;
        const int J = sh.jm - 1;
        const double lo = c43 * (double)p[b + ldk] - c13 * (double)p[b + 2 * ldk];
        const double hi = c43 * (double)p[b + (size_t)(J - 1) * ldk] - c13 * (double)p[b + (size_t)(J - 2) * ldk];
        p[b] = (P)lo;
        p[b + (size_t)J * ldk] = (P)hi;
        if (k == 0) {
            p[b + nk] = (P)lo;
            p[b + (size_t)J * ldk + nk] = (P)hi;
        }
    }
}

// ---- residual r = src - Lc p, Lc p = acc - den p, with per-block fixed-order partial sums ----
constexpr int kResBlock = 256;

template <typename P>
__global__ void __launch_bounds__(kResBlock) k_residual(Shape sh, DevGeo g, const P* p, const double* src,
                                                        double* r_out, double* partials) {
    const int nk = sh.km - 1;
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = 1 + blockIdx.z;
    double rr = 0.0, ss = 0.0;
    if (k < nk && j <= sh.jm - 2) {
        const int ldk = sh.ldk;
        const size_t row = ((size_t)i * sh.jm + j) * ldk;
        const size_t x = row + k;
        const size_t xw = row + (k == 0 ? nk - 1 : k - 1), xe = row + (k == nk - 1 ? 0 : k + 1);
        const size_t ij = (size_t)i * sh.jm + j;
        const double acc = ((double)p[x + sh.plane] + (double)p[x - sh.plane]) * g.num1
                           + ((double)p[x + ldk] + (double)p[x - ldk]) * g.num2[i]
                           + ((double)p[xe] + (double)p[xw]) * g.num3[ij];
        const double s = src[x];
        const double r = s - (acc - g.den[ij] * (double)p[x]);
        if (r_out) r_out[x] = r;
        rr = r * r;
        ss = s * s;
    }
    __shared__ double sr[kResBlock], sq[kResBlock];
    sr[threadIdx.x] = rr;
    sq[threadIdx.x] = ss;
    __syncthreads();
    for (int h = kResBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sr[threadIdx.x] += sr[threadIdx.x + h];
            sq[threadIdx.x] += sq[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        partials[2 * b] = sr[0];
        partials[2 * b + 1] = sq[0];
    }
}

__global__ void __launch_bounds__(kResBlock) k_sum_partials(int nblocks, const double* partials, double* out) {
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += kResBlock) {
        a += partials[2 * q];
        b += partials[2 * q + 1];
    }
    __shared__ double sa[kResBlock], sb[kResBlock];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int h = kResBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sa[threadIdx.x] += sa[threadIdx.x + h];
            sb[threadIdx.x] += sb[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = sb[0];
    }
}

dim3 residual_grid(const Shape& sh) { return dim3((sh.km - 1 + 63) / 64, (sh.jm - 2 + 3) / 4, sh.im - 2); }

}  // namespace

int residual_partials_count(const Shape& sh) {
    const dim3 gr = residual_grid(sh);
    return (int)(gr.x * gr.y * gr.z);
}

template <typename S>
hipError_t launch_proj_src(const Shape& sh, const DevGeo& g, const StepConsts& c, S* aux, double* src,
                           hipStream_t s, bool form_src) {
    const int nk = sh.km - 1;
    const int n = (sh.jm - 2) * nk + (sh.im - 2) * nk + (sh.im - 2) * (sh.jm - 2) + 4;
    hipLaunchKernelGGL((k_proj_walls<S>), dim3((n + 255) / 256), dim3(256), 0, s, sh, 1, aux);
    if (form_src)
        hipLaunchKernelGGL((k_src<S>), dim3((nk + 63) / 64, (sh.jm - 2 + 3) / 4, sh.im - 2), dim3(64, 4), 0, s, sh, g,
                           c, (const S*)aux, src);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_rb_pass(const Shape& sh, const DevGeo& g, const StepConsts& c, P* p, const double* src, int colour,
                          double omega, hipStream_t s, int form_walls) {
    const int half = (sh.km - 1 + 1) / 2;
    hipLaunchKernelGGL((k_rb_pass<P>), dim3((half + 63) / 64, (sh.jm - 2 + 3) / 4, sh.im - 2), dim3(64, 4), 0, s, sh,
                       g, c.c43, c.c13, p, src, colour, omega, form_walls);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_p_corners(const Shape& sh, const StepConsts& c, P* p, hipStream_t s, const int* skip) {
    hipLaunchKernelGGL((k_p_corners<P>), dim3((sh.km - 1 + 255) / 256), dim3(256), 0, s, sh, c.c43, c.c13, p, skip);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_residual(const Shape& sh, const DevGeo& g, const P* p, const double* src, double* r_out,
                           double* partials, double* out, hipStream_t s) {
    const dim3 gr = residual_grid(sh);
    hipLaunchKernelGGL((k_residual<P>), gr, dim3(kResBlock), 0, s, sh, g, p, src, r_out, partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kResBlock), 0, s, (int)(gr.x * gr.y * gr.z),
                       (const double*)partials, out);
    return hipGetLastError();
}

template hipError_t launch_proj_src<double>(const Shape&, const DevGeo&, const StepConsts&, double*, double*,
                                            hipStream_t, bool);
template hipError_t launch_proj_src<float>(const Shape&, const DevGeo&, const StepConsts&, float*, double*,
                                           hipStream_t, bool);
template hipError_t launch_rb_pass<double>(const Shape&, const DevGeo&, const StepConsts&, double*, const double*, int,
                                           double, hipStream_t, int);
template hipError_t launch_rb_pass<float>(const Shape&, const DevGeo&, const StepConsts&, float*, const double*, int,
                                          double, hipStream_t, int);
template hipError_t launch_p_corners<double>(const Shape&, const StepConsts&, double*, hipStream_t, const int*);
template hipError_t launch_p_corners<float>(const Shape&, const StepConsts&, float*, hipStream_t, const int*);
template hipError_t launch_residual<double>(const Shape&, const DevGeo&, const double*, const double*, double*, double*,
                                            double*, hipStream_t);
template hipError_t launch_residual<float>(const Shape&, const DevGeo&, const float*, const double*, double*, double*,
                                           double*, hipStream_t);

}  // namespace atu
