// The red-black solver's two options, PRESS_LINE and PRESS_BAROTROPIC (planet/PressureSolver.h:
// 284-303, 497-708), as the reference runs them after the pointwise colour passes of every sweep and
// before its wall passes. Every output is formed with the reference's operations in its order, so an
// FP64 sweep equals the reference's at one thread (and the CPU oracle's) bit for bit, including the
// interior mean PRESS_LINE holds (summed in the serial order, applied by the caller).
//
//   line pass (per colour)   one thread per (j, k) column of the colour (j + k) & 1: the right-hand
//                            sides of a group of planes are loaded together, the forward sweep
//                            divides by the row's host-built pivots, its d' values go to a scratch
//                            field, the back substitution writes every finite value into p
//   column residuals         R(j, k) = sum_i (src - Lc p) / A2 over i in order, one thread per column,
//                            the end rows read as the wall pass will form them
//   forward DFT              per (mode, row): the cosine and sine sums over k in order, modes above 0
//                            of round-off size (below 1e-12 of the row's sum |R|) dropped
//   theta solves             one lane per (mode, part) out of LDS: mode 0's consistency shift, its pin
//                            at j = 1 and its zero sin-weighted mean
//   inverse DFT              per (row, k): the modes in order, divided by N, added to every level

#include "atu_kernels.h"

namespace atu {
namespace {

constexpr int kGroup = 8;          // planes whose right-hand sides are loaded together
constexpr int kMaxN = 1024;        // meridians per LDS table
constexpr int kRows = 4;           // rows of a DFT block (one wave each)
constexpr int kMaxJ = 1024;        // rows of a theta system staged in LDS

__global__ void __launch_bounds__(256) k_rline(Shape sh, DevPressRef r, double* p, const double* src, double* dps,
                                               int colour) {
    const int N = sh.km - 1;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    if (j > sh.jm - 2) return;
    const int k = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x) + ((colour + j) & 1);
    if (k >= N) return;
    const int ldk = sh.ldk, jm = sh.jm, ni = r.ni;
    const int kw = (k == 0) ? N - 1 : k - 1, ke = (k == N - 1) ? 0 : k + 1;
    const double* cp = r.line_cp + (size_t)j * ni;
    const double* dn = r.line_dn + (size_t)j * ni;
    const size_t stride = sh.plane;
    const size_t x1 = ((size_t)jm + j) * ldk + k;   // the column's cell at i = 1
    const size_t r1 = x1 - k;
    double dprev = 0.0;
    for (int s0 = 0; s0 < ni; s0 += kGroup) {
        double rhs[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int s = s0 + u;
            if (s < ni) {
                const int i = s + 1;
                const size_t x = x1 + (size_t)s * stride, row = r1 + (size_t)s * stride;
                const double acc_h = r.num2[i] * (p[x + ldk] + p[x - ldk])
                                     + r.num3[(size_t)i * jm + j] * (p[row + ke] + p[row + kw]);
                rhs[u] = acc_h - src[x];
            }
        }
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int s = s0 + u;
            if (s < ni) {
                const double a = (s == ni - 1) ? r.a_hi : r.a_mid;
                const double dp = (s == 0) ? rhs[u] / dn[0] : (rhs[u] - a * dprev) / dn[s];
                dps[x1 + (size_t)s * stride] = dp;
                dprev = dp;
            }
        }
    }
    double xv = dprev;
    if (isfinite(xv)) p[x1 + (size_t)(ni - 1) * stride] = xv;
    for (int s1 = ni - 2; s1 >= 0; s1 -= kGroup) {
        double dv[kGroup], cv[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int s = s1 - u;
            if (s >= 0) {
                dv[u] = dps[x1 + (size_t)s * stride];
                cv[u] = cp[s];
            }
        }
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int s = s1 - u;
            if (s >= 0) {
                xv = dv[u] - cv[u] * xv;
                if (isfinite(xv)) p[x1 + (size_t)s * stride] = xv;
            }
        }
    }
}

// R(jj, k) = sum over i = 1..im-2 of (src - Lc p) / A2
__global__ void __launch_bounds__(256) k_rcol_resid(Shape sh, DevPressRef r, const double* p, const double* src,
                                                    double* R) {
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    if (k >= N || j > sh.jm - 2) return;
    const int im = sh.im, jm = sh.jm, ldk = sh.ldk;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const int kw = (k == 0) ? N - 1 : k - 1, ke = (k == N - 1) ? 0 : k + 1;
    const size_t PL = sh.plane;
    const size_t col = (size_t)j * ldk;
    double s = 0.0;
    for (int i = 1; i < im - 1; i++) {
        const size_t row = i * PL + col;
        const size_t x = row + k;
        const double p0 = p[x];
        const double pi_up = (i + 1 == im - 1) ? c43 * p[x] - c13 * p[x - PL] : p[x + PL];
        const double pi_dn = (i - 1 == 0) ? c43 * p[x] - c13 * p[x + PL] : p[x - PL];
        const double pj_up = (j + 1 == jm - 1) ? c43 * p[x] - c13 * p[x - ldk] : p[x + ldk];
        const double pj_dn = (j - 1 == 0) ? c43 * p[x] - c13 * p[x + ldk] : p[x - ldk];
        const double lc = r.num1 * (pi_up + pi_dn - 2.0 * p0) + r.num2[i] * (pj_up + pj_dn - 2.0 * p0)
                          + r.num3[(size_t)i * jm + j] * (p[row + ke] + p[row + kw] - 2.0 * p0);
        s += src[x] - lc;
    }
    R[(size_t)(j - 1) * N + k] = s / r.A2;
}

// Copy n values src(q) -> dst[q] with a block of nt threads.
template <typename Src>
__device__ __forceinline__ void stage(double* dst, int n, int tid, int nt, Src src) {
    for (int q = tid; q < n; q += nt) dst[q] = src(q);
}

// Ac[m][jj] = sum_k R[jj][k] cos(2 pi m k / N), As likewise with sin; a block is 64 modes x kRows rows
__global__ void __launch_bounds__(64 * kRows) k_rdft(DevSolve t, const double* R, double* Ac, double* As) {
    __shared__ double c1[kMaxN], s1[kMaxN], rr[kRows][kMaxN];
    const int N = t.N, M1 = t.NM + 1, J = t.J;
    const int tid = threadIdx.y * 64 + threadIdx.x, nt = 64 * kRows;
    const int jj0 = blockIdx.y * kRows;
    stage(c1, N, tid, nt, [&](int q) { return t.ct[(size_t)N + q]; });   // cos(2 pi q / N)
    stage(s1, N, tid, nt, [&](int q) { return t.st[(size_t)N + q]; });
    stage(&rr[0][0], kRows * kMaxN, tid, nt, [&](int q) {
        const int w = q / kMaxN, k = q - w * kMaxN;
        return (jj0 + w < J && k < N) ? R[(size_t)(jj0 + w) * N + k] : 0.0;
    });
    __syncthreads();
    const int m = blockIdx.x * 64 + threadIdx.x;
    const int w = threadIdx.y;
    if (m >= M1 || jj0 + w >= J) return;
    double sc = 0.0, ss = 0.0, ra = 0.0;
    int q = 0;   // (m k) mod N
    for (int k = 0; k < N; k++) {
        const double rv = rr[w][k];
        sc += rv * c1[q];
        ss += rv * s1[q];
        ra += fabs(rv);
        q += m;
        if (q >= N) q -= N;
    }
    if (m > 0) {
        const double tol = 1.0e-12 * ra;
        if (fabs(sc) < tol) sc = 0.0;
        if (fabs(ss) < tol) ss = 0.0;
    }
    Ac[(size_t)m * J + jj0 + w] = sc;
    As[(size_t)m * J + jj0 + w] = ss;
}

// One 64-lane block per (mode, part); lane 0 runs the sequential work out of LDS.
__global__ void __launch_bounds__(64) k_rtri(DevSolve t, DevPressRef r, double* Ac, double* As) {
    const int m = blockIdx.x >> 1, part = blockIdx.x & 1;
    const int J = t.J, N = t.N;
    if (part == 1 && (m == 0 || 2 * m == N)) return;   // no sine part (the inverse DFT never reads it)
    double* d = (part == 0 ? Ac : As) + (size_t)m * J;
    __shared__ double sd[kMaxJ], scp[kMaxJ], sdn[kMaxJ];
    const double* cp = t.mode_cp + (size_t)m * J;
    const double* dn = t.mode_dn + (size_t)m * J;
    for (int jj = threadIdx.x; jj < J; jj += blockDim.x) {
        sd[jj] = d[jj];
        scp[jj] = cp[jj];
        sdn[jj] = dn[jj];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double e = t.e, c13 = 1.0 / 3.0;
        if (m == 0) {   // consistency: the part the all-Neumann operator cannot reach, then the pin
            double yr = 0.0;
            for (int jj = 0; jj < J; jj++) {
                const double y = (jj == 0 || jj == J - 1) ? 1.5 : 1.0;
                yr += y * sd[jj];
            }
            const double f = yr / r.ys;
            for (int jj = 0; jj < J; jj++) sd[jj] -= f * r.abs_sin[jj];
            sd[0] = 0.0;
        }
        double prev = sd[0] / sdn[0];
        sd[0] = prev;
        for (int jj = 1; jj < J; jj++) {
            const double a = (jj == J - 1) ? e - c13 * e : e;
            prev = (sd[jj] - a * prev) / sdn[jj];
            sd[jj] = prev;
        }
        double x = sd[J - 1];
        for (int jj = J - 2; jj >= 0; jj--) {
            x = sd[jj] - scp[jj] * x;
            sd[jj] = x;
        }
        if (m == 0) {   // zero sin-weighted mean
            double sq = 0.0;
            for (int jj = 0; jj < J; jj++) sq += r.abs_sin[jj] * sd[jj];
            const double f = sq / r.sw;
            for (int jj = 0; jj < J; jj++) sd[jj] -= f;
        }
    }
    __syncthreads();
    for (int jj = threadIdx.x; jj < J; jj += blockDim.x) d[jj] = sd[jj];
}

// q(jj, k) = sum_m w_m (Ac c + As s) / N, added to p at every level i = 0..im-1 of row jj + 1;
// a block is 64 k x kRows rows with the rows' coefficients in LDS
constexpr int kMaxM = kMaxN / 2 + 1;
__global__ void __launch_bounds__(64 * kRows) k_ridft_add(Shape sh, DevSolve t, const double* Ac, const double* As,
                                                          double* p) {
    __shared__ double c1[kMaxN], s1[kMaxN], ta[kRows][kMaxM], tb[kRows][kMaxM];
    const int N = t.N, NM = t.NM, J = t.J;
    const int tid = threadIdx.y * 64 + threadIdx.x, nt = 64 * kRows;
    const int jj0 = blockIdx.y * kRows;
    stage(c1, N, tid, nt, [&](int q) { return t.ct[(size_t)N + q]; });
    stage(s1, N, tid, nt, [&](int q) { return t.st[(size_t)N + q]; });
    stage(&ta[0][0], kRows * kMaxM, tid, nt, [&](int q) {
        const int w = q / kMaxM, m = q - w * kMaxM;
        return (jj0 + w < J && m <= NM) ? Ac[(size_t)m * J + jj0 + w] : 0.0;
    });
    stage(&tb[0][0], kRows * kMaxM, tid, nt, [&](int q) {
        const int w = q / kMaxM, m = q - w * kMaxM;
        return (jj0 + w < J && m <= NM) ? As[(size_t)m * J + jj0 + w] : 0.0;
    });
    __syncthreads();
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int w = threadIdx.y;
    if (k >= N || jj0 + w >= J) return;
    double q = 0.0;
    int qi = 0;   // (m k) mod N
    for (int m = 0; m <= NM; m++) {
        const bool edge = (m == 0 || 2 * m == N);
        const double wm = edge ? 1.0 : 2.0;
        q += wm * (ta[w][m] * c1[qi] + (edge ? 0.0 : tb[w][m] * s1[qi]));
        qi += k;
        if (qi >= N) qi -= N;
    }
    q /= N;
    const size_t x = (size_t)(jj0 + w + 1) * sh.ldk + k;
    for (int i = 0; i < sh.im; i++) p[x + i * sh.plane] += q;
}

// The interior sum of p in the reference's serial (i, j, k) order, the order its OpenMP reduction
// takes on one thread (planet/PressureSolver.h:309-316). A sum whose rounding is that of one
// sequential chain has no parallel form: one wave loads the rows in coalesced chunks of 64 values,
// kGroupS chunks at a time, into LDS (the next group's loads in flight while the current group is
// added), and every lane adds the group's values in order from LDS (a broadcast read), so the chain
// of dependent FP64 adds is the only serial part. The held mean feeds back into every cell of every
// sweep, and the barotropic solve magnifies a one-ulp change of it to ~1e-12 of p within 20 sweeps,
// which is why the order is the reference's rather than a tree's.
constexpr int kGroupS = 8;
__global__ void __launch_bounds__(64) k_serial_interior_sum(Shape sh, const double* p, double* out) {
    __shared__ double buf[2][kGroupS * 64];
    __shared__ int cnt[2][kGroupS];
    const int lane = threadIdx.x;
    const int N = sh.km - 1, nj = sh.jm - 2;
    const int cpr = (N + 63) / 64;                       // chunks per row
    const long nch = (long)(sh.im - 2) * nj * cpr;
    // chunk c: row c / cpr in (i, j) order, values k0 .. k0 + n - 1 of it; lanes beyond n load nothing
    auto load = [&](long c, int& n) -> double {
        n = 0;
        if (c >= nch) return 0.0;
        const long r = c / cpr;
        const int k0 = (int)(c - r * cpr) * 64;
        const int i = 1 + (int)(r / nj), j = 1 + (int)(r % nj);
        n = min(64, N - k0);
        const double* row = p + (size_t)i * sh.plane + (size_t)j * sh.ldk;
        return lane < n ? row[k0 + lane] : 0.0;
    };
    auto stash = [&](int b, const double* v, const int* n) {
        for (int u = 0; u < kGroupS; u++) {
            buf[b][u * 64 + lane] = v[u];
            if (lane == 0) cnt[b][u] = n[u];
        }
    };
    double v[kGroupS];
    int nv[kGroupS];
#pragma unroll
    for (int u = 0; u < kGroupS; u++) v[u] = load(u, nv[u]);
    stash(0, v, nv);
    __syncthreads();
    double s = 0.0;
    int cur = 0;
    for (long c = 0; c < nch; c += kGroupS) {
#pragma unroll
        for (int u = 0; u < kGroupS; u++) v[u] = load(c + kGroupS + u, nv[u]);   // next group, in flight
        for (int u = 0; u < kGroupS; u++) {
            const double* b = &buf[cur][u * 64];
            const int n = cnt[cur][u];
            if (n == 64) {
#pragma unroll 16
                for (int q = 0; q < 64; q++) s += b[q];
            } else {
                for (int q = 0; q < n; q++) s += b[q];
            }
        }
        stash(cur ^ 1, v, nv);
        __syncthreads();
        cur ^= 1;
    }
    if (lane == 0) *out = s;
}

}  // namespace

hipError_t launch_serial_interior_sum(const Shape& sh, const double* p, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_serial_interior_sum, dim3(1), dim3(64), 0, s, sh, p, out);
    return hipGetLastError();
}

hipError_t launch_rline(const Shape& sh, const DevPressRef& r, double* p, const double* src, double* scratch,
                        int colour, hipStream_t s) {
    if (r.ni < 2) return hipErrorInvalidValue;
    const int half = (sh.km - 1 + 1) / 2;
    hipLaunchKernelGGL(k_rline, dim3((half + 63) / 64, (sh.jm - 2 + 3) / 4), dim3(64, 4), 0, s, sh, r, p, src, scratch,
                       colour);
    return hipGetLastError();
}

hipError_t launch_rbaro(const Shape& sh, const DevPressRef& r, const DevSolve& t, double* p, const double* src,
                        double* R, double* Ac, double* As, hipStream_t s) {
    if (t.N > kMaxN || t.NM + 1 > kMaxM || t.J > kMaxJ || t.J < 2 || t.N < 2) return hipErrorInvalidValue;
    const int N = sh.km - 1;
    hipLaunchKernelGGL(k_rcol_resid, dim3((N + 63) / 64, (sh.jm - 2 + 3) / 4), dim3(64, 4), 0, s, sh, r,
                       (const double*)p, src, R);
    hipLaunchKernelGGL(k_rdft, dim3((t.NM + 1 + 63) / 64, (t.J + kRows - 1) / kRows), dim3(64, kRows), 0, s, t,
                       (const double*)R, Ac, As);
    hipLaunchKernelGGL(k_rtri, dim3(2 * (t.NM + 1)), dim3(64), 0, s, t, r, Ac, As);
    hipLaunchKernelGGL(k_ridft_add, dim3((N + 63) / 64, (t.J + kRows - 1) / kRows), dim3(64, kRows), 0, s, sh, t,
                       (const double*)Ac, (const double*)As, p);
    return hipGetLastError();
}

}  // namespace atu
