// Converging solve of the dynamic-pressure system (ATU_SOLVE_CONVERGE): zebra red-black radial line
// relaxation plus an exact barotropic correction, to a relative residual target.
//
// The system is the red-black sweep's (planet/PressureSolver.h:350-491) with its wall passes
// (:718-732) folded into the equations, so its fixed point is what the reference's sweeps approach.
// The reference's own cures for this radially dominated operator are its options PRESS_LINE (exact
// solve along r, :284-303, 497-570) and PRESS_BAROTROPIC (the column-mean mode by a DFT in phi and a
// tridiagonal solve in theta, :49-67, 572-708); this file runs both every iteration:
//
//   line pass (colour 0, 1)   one thread per (j, k) column of a colour, Thomas along i with
//                             host-built factors; the column's own values are overwritten in place
//                             (the horizontal neighbours it reads belong to the other colour)
//   walls                     radial and polar walls re-formed from the interior
//   column residual           R(j, k) = sum_i a_i (b - A p) / S2, a_i = 3/2 at the end rows
//   DFT / theta solves / iDFT the correction q(j, k) that zeroes R, added to every level
//   residual norm             in the canonical order shared with the oracle
//
// Every sum runs in a fixed sequential order (the same order the CPU oracle uses), so the FP64
// result is bit-identical to the oracle at equal iteration counts.
//
// Templates: P is the storage of the unknown, B of the right-hand side, R the arithmetic of the line
// pass. The FP64 solve is <double, double, double>; the mixed path's inner solve of the refinement
// is <float, float, float> on the correction e with the FP32 residual as its right-hand side.

#include "atu_kernels.h"

namespace atu {
namespace {

__device__ __forceinline__ double endw(int x, int lo, int hi) { return (x == lo || x == hi) ? 1.5 : 1.0; }

template <typename P, typename B, typename R>
__global__ void __launch_bounds__(256) k_line(Shape sh, DevGeo g, DevSolve t, P* p, const B* b, int colour,
                                              const int* done) {
    if (done && *done) return;
    const int N = sh.km - 1;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    if (j > sh.jm - 2) return;
    const int k = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x) + ((colour + j) & 1);
    if (k >= N) return;
    const int ldk = sh.ldk, jm = sh.jm, ni = t.ni;
    const R c13 = R(1.0) / R(3.0);
    const int kw = (k == 0) ? N - 1 : k - 1, ke = (k == N - 1) ? 0 : k + 1;
    const double* cp = t.line_cp + (size_t)j * ni;
    // FP64: divide by the Thomas pivots, as the oracle; FP32 (the refinement's correction):
    // multiply by their reciprocals, off the division's latency
    constexpr bool kDiv = sizeof(R) == sizeof(double);
    const double* dn = (kDiv ? t.line_dn : t.line_inv_dn) + (size_t)j * ni;
    const size_t stride = sh.plane;
    const size_t x1 = ((size_t)jm + j) * ldk + k;   // the column's cell at i = 1
    const size_t r1 = x1 - k;                        // and its row
    const R num1 = (R)g.num1;
    // Forward elimination in groups of LINE_GROUP planes: the right-hand sides of a group (which
    // read only the other colour and b) are loaded and formed first, so their loads are in flight
    // together, then the group's recurrence steps run and store.
    constexpr int LINE_GROUP = 8;
    R dprev = R(0.0);
    for (int s0 = 0; s0 < ni; s0 += LINE_GROUP) {
        R rhs[LINE_GROUP], dns[LINE_GROUP];
#pragma unroll
        for (int u = 0; u < LINE_GROUP; u++) {
            const int s = s0 + u;
            if (s < ni) {
                const int i = s + 1;
                const size_t x = x1 + (size_t)s * stride, row = r1 + (size_t)s * stride;
                const R num2 = (R)g.num2[i], num3 = (R)g.num3[(size_t)i * jm + j];
                R th;
                if (j == 1)
                    th = (num2 - num2 * c13) * (R)p[x + ldk];
                else if (j == jm - 2)
                    th = (num2 - num2 * c13) * (R)p[x - ldk];
                else
                    th = ((R)p[x + ldk] + (R)p[x - ldk]) * num2;
                rhs[u] = (R)b[x] - th - ((R)p[row + ke] + (R)p[row + kw]) * num3;
                dns[u] = (R)dn[s];
            }
        }
#pragma unroll
        for (int u = 0; u < LINE_GROUP; u++) {
            const int s = s0 + u;
            if (s < ni) {
                const R a = (s == 0) ? R(0.0) : ((s == ni - 1) ? num1 - num1 * c13 : num1);
                const R dp = kDiv ? (rhs[u] - a * dprev) / dns[u] : (rhs[u] - a * dprev) * dns[u];
                p[x1 + (size_t)s * stride] = (P)dp;
                dprev = dp;
            }
        }
    }
    // Back substitution x_s = d_s - cp_s x_{s+1}, from the last plane (whose value is dprev), in
    // groups loaded ahead the same way.
    R xv = dprev;
    for (int s1 = ni - 2; s1 >= 0; s1 -= LINE_GROUP) {
        R dv[LINE_GROUP], cv[LINE_GROUP];
#pragma unroll
        for (int u = 0; u < LINE_GROUP; u++) {
            const int s = s1 - u;
            if (s >= 0) {
                dv[u] = (R)p[x1 + (size_t)s * stride];
                cv[u] = (R)cp[s];
            }
        }
#pragma unroll
        for (int u = 0; u < LINE_GROUP; u++) {
            const int s = s1 - u;
            if (s >= 0) {
                xv = dv[u] - cv[u] * xv;
                p[x1 + (size_t)s * stride] = (P)xv;
            }
        }
    }
}

// The same line pass with the column's work split three ways. A block is 64 columns of one colour
// in one row j times kLineWaves waves: the waves form the right-hand sides of all planes at once
// (their loads all in flight) into LDS, wave 0 runs the two recurrences out of LDS with the same
// operations in the same order as k_line, and all waves store the column. The dynamic LDS holds
// ni x 64 values of R.
constexpr int kLineWaves = 8;
template <typename P, typename B, typename R>
__global__ void __launch_bounds__(64 * kLineWaves) k_line_par(Shape sh, DevGeo g, DevSolve t, P* p, const B* b,
                                                              int colour, const int* done) {
    if (done && *done) return;
    extern __shared__ unsigned char line_lds[];
    const int ni = t.ni;
    R* sv = reinterpret_cast<R*>(line_lds);                                        // [ni][64]
    double* scp = reinterpret_cast<double*>(line_lds + (size_t)ni * 64 * sizeof(R));   // the row's factors
    double* sdn = scp + ni;
    const int N = sh.km - 1;
    const int j = 1 + blockIdx.y;
    const int lane = threadIdx.x, w = threadIdx.y;
    const int k = 2 * (int)(blockIdx.x * 64 + lane) + ((colour + j) & 1);
    const bool on = k < N;
    const int ldk = sh.ldk, jm = sh.jm;
    {   // the row's Thomas factors into LDS, off the recurrences' chain
        constexpr bool kDivF = sizeof(R) == sizeof(double);
        const double* cpg = t.line_cp + (size_t)j * ni;
        const double* dng = (kDivF ? t.line_dn : t.line_inv_dn) + (size_t)j * ni;
        for (int q = w * 64 + lane; q < ni; q += 64 * kLineWaves) {
            scp[q] = cpg[q];
            sdn[q] = dng[q];
        }
    }
    const R c13 = R(1.0) / R(3.0);
    const int kw = (k == 0) ? N - 1 : k - 1, ke = (k == N - 1) ? 0 : k + 1;
    const size_t stride = sh.plane;
    const size_t x1 = ((size_t)jm + j) * ldk + k;
    const size_t r1 = x1 - k;
    if (on) {
        for (int s = w; s < ni; s += kLineWaves) {
            const int i = s + 1;
            const size_t x = x1 + (size_t)s * stride, row = r1 + (size_t)s * stride;
            const R num2 = (R)g.num2[i], num3 = (R)g.num3[(size_t)i * jm + j];
            R th;
            if (j == 1)
                th = (num2 - num2 * c13) * (R)p[x + ldk];
            else if (j == jm - 2)
                th = (num2 - num2 * c13) * (R)p[x - ldk];
            else
                th = ((R)p[x + ldk] + (R)p[x - ldk]) * num2;
            sv[s * 64 + lane] = (R)b[x] - th - ((R)p[row + ke] + (R)p[row + kw]) * num3;
        }
    }
    __syncthreads();
    if (w == 0 && on) {
        constexpr bool kDiv = sizeof(R) == sizeof(double);
        const double* cp = scp;
        const double* dn = sdn;
        const R num1 = (R)g.num1;
        R dprev = R(0.0);
#pragma unroll 4
        for (int s = 0; s < ni; s++) {
            const R a = (s == 0) ? R(0.0) : ((s == ni - 1) ? num1 - num1 * c13 : num1);
            const R dp = kDiv ? (sv[s * 64 + lane] - a * dprev) / (R)dn[s] : (sv[s * 64 + lane] - a * dprev) * (R)dn[s];
            sv[s * 64 + lane] = dp;
            dprev = dp;
        }
        R xv = dprev;
#pragma unroll 4
        for (int s = ni - 2; s >= 0; s--) {
            xv = sv[s * 64 + lane] - (R)cp[s] * xv;
            sv[s * 64 + lane] = xv;
        }
    }
    __syncthreads();
    if (on)
        for (int s = w; s < ni; s += kLineWaves) p[x1 + (size_t)s * stride] = (P)sv[s * 64 + lane];
}

// radial walls (j = 1..jm-2) and polar walls (i = 1..im-2), k = 0..N-1
template <typename P>
__global__ void __launch_bounds__(256) k_walls(Shape sh, double c43, double c13, P* p, const int* done) {
    if (done && *done) return;
    const int N = sh.km - 1;
    const int n1 = (sh.jm - 2) * N, n2 = (sh.im - 2) * N;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int ldk = sh.ldk;
    const size_t PL = sh.plane;
    if (q < n1) {
        const int j = 1 + q / N, k = q % N;
        const size_t c = (size_t)j * ldk + k;
        const int L = sh.im - 1;
        p[c] = (P)(c43 * (double)p[c + PL] - c13 * (double)p[c + 2 * PL]);
        p[c + L * PL] = (P)(c43 * (double)p[c + (L - 1) * PL] - c13 * (double)p[c + (L - 2) * PL]);
    } else if (q < n1 + n2) {
        const int r = q - n1;
        const int i = 1 + r / N, k = r % N;
        const size_t c = (size_t)i * PL + k;
        const int J = sh.jm - 1;
        p[c] = (P)(c43 * (double)p[c + ldk] - c13 * (double)p[c + 2 * ldk]);
        p[c + (size_t)J * ldk] = (P)(c43 * (double)p[c + (size_t)(J - 1) * ldk] - c13 * (double)p[c + (size_t)(J - 2) * ldk]);
    }
}

// k = km-1 is the copy of k = 0 for every (i, j)
template <typename P>
__global__ void __launch_bounds__(256) k_seam(Shape sh, P* p, const int* skip) {
    if (skip && *skip) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= sh.jm) return;
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    p[row + sh.km - 1] = p[row];
}

// b - A p at an interior cell. The radial and polar wall values it reads are formed from the
// interior as the wall pass forms them (4/3 p_1 - 1/3 p_2, rounded to P), so no wall pass is
// needed between an update of the interior and a residual: the iterations form the walls only once,
// after the last one.
template <typename P, typename B>
__device__ __forceinline__ double resid_at(const Shape& sh, const DevGeo& g, const P* p, const B* b, int i, int j,
                                           int k) {
    const int N = sh.km - 1, ldk = sh.ldk;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const size_t row = ((size_t)i * sh.jm + j) * ldk;
    const size_t x = row + k;
    const size_t xw = row + (k == 0 ? N - 1 : k - 1), xe = row + (k == N - 1 ? 0 : k + 1);
    const size_t ij = (size_t)i * sh.jm + j;
    const double pc = (double)p[x];
    const double rp = (double)p[x + sh.plane], rn = (double)p[x - sh.plane];
    const double tp = (double)p[x + ldk], tn = (double)p[x - ldk];
    const double r_up = (i + 1 == sh.im - 1) ? (double)(P)(c43 * pc - c13 * rn) : rp;
    const double r_dn = (i - 1 == 0) ? (double)(P)(c43 * pc - c13 * rp) : rn;
    const double t_up = (j + 1 == sh.jm - 1) ? (double)(P)(c43 * pc - c13 * tn) : tp;
    const double t_dn = (j - 1 == 0) ? (double)(P)(c43 * pc - c13 * tp) : tn;
    const double acc = (r_up + r_dn) * g.num1 + (t_up + t_dn) * g.num2[i] + ((double)p[xe] + (double)p[xw]) * g.num3[ij];
    return (double)b[x] - (acc - g.den[ij] * pc);
}

// R(jj, k) = sum_i a_i r_i / S2 over each interior column
template <typename P, typename B>
__global__ void __launch_bounds__(256) k_col_resid(Shape sh, DevGeo g, DevSolve t, const P* p, const B* b,
                                                   double* R, const int* done) {
    if (done && *done) return;
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    if (k >= N || j > sh.jm - 2) return;
    double s = 0.0;
    for (int i = 1; i < sh.im - 1; i++) s += endw(i, 1, sh.im - 2) * resid_at(sh, g, p, b, i, j, k);
    R[(size_t)(j - 1) * N + k] = s / t.S2;
}

// The same column residual with the levels' terms formed by all waves at once into LDS and summed
// in order by wave 0 (dynamic LDS: ni x 64 doubles).
template <typename P, typename B>
__global__ void __launch_bounds__(64 * kLineWaves) k_col_resid_par(Shape sh, DevGeo g, DevSolve t, const P* p,
                                                                   const B* b, double* R, const int* done) {
    if (done && *done) return;
    extern __shared__ unsigned char resid_lds[];
    double* sv = reinterpret_cast<double*>(resid_lds);   // [ni][64]
    const int N = sh.km - 1;
    const int lane = threadIdx.x, w = threadIdx.y;
    const int k = blockIdx.x * 64 + lane;
    const int j = 1 + blockIdx.y;
    const int ni = sh.im - 2;
    const bool on = k < N;
    if (on)
        for (int s = w; s < ni; s += kLineWaves)
            sv[s * 64 + lane] = endw(s + 1, 1, sh.im - 2) * resid_at(sh, g, p, b, s + 1, j, k);
    __syncthreads();
    if (w == 0 && on) {
        double acc = 0.0;
#pragma unroll 8
        for (int s = 0; s < ni; s++) acc += sv[s * 64 + lane];
        R[(size_t)(j - 1) * N + k] = acc / t.S2;
    }
}

// The DFT tables depend on (m k mod N) only: ct[m][k] = ct[1][(m k) mod N], the same bits (the host
// builds every entry from the reduced angle). The DFT kernels stage that one row of N cosines and
// sines in LDS with the rows of data they need, and index it with the running (m k) mod N.
constexpr int kDftMaxN = 1024;

// Copy n values src(q) -> dst[q] with a block of nt threads, kStageBatch loads in flight per thread
// (the loads of a batch are issued before any of its stores).
constexpr int kStageBatch = 8;
template <typename Src>
__device__ __forceinline__ void stage_lds(double* dst, int n, int tid, int nt, Src src) {
    for (int q0 = tid; q0 < n; q0 += kStageBatch * nt) {
        double v[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; u++) {
            const int q = q0 + u * nt;
            v[u] = q < n ? src(q) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; u++) {
            const int q = q0 + u * nt;
            if (q < n) dst[q] = v[u];
        }
    }
}

// forward real DFT: Ac[m][jj] = sum_k R[jj][k] cos(2 pi m k / N), As likewise with sin. Each sum in
// the order shared with the oracle (dft_sum16): 16 partials, partial l adding the terms k = l, l + 16,
// ... in sequence from 0, then a halving tree over the 16. A group of 16 lanes forms one (m, jj);
// the groups of a block are consecutive modes of one row, which share the row's loads; the twiddles
// cos / sin(2 pi q / N), q = m k mod N, come from one table row in LDS.
constexpr int kDftLanes = 16, kDftGroups = 16;   // groups of a 256-thread block
__device__ __forceinline__ double tree16(double v) {
    for (int h = kDftLanes / 2; h > 0; h >>= 1) v += __shfl_down(v, h, kDftLanes);
    return v;
}
__global__ void __launch_bounds__(kDftLanes * kDftGroups) k_dft_fwd(DevSolve t, const double* R, double* Ac,
                                                                     double* As, const int* done) {
    if (done && *done) return;
    __shared__ double c1[kDftMaxN], s1[kDftMaxN], r[kDftMaxN];
    const int N = t.N, M1 = t.NM + 1, J = t.J;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int mblocks = (M1 + kDftGroups - 1) / kDftGroups;
    const int jj = blockIdx.x / mblocks;
    stage_lds(c1, N, tid, nt, [&](int q) { return t.ct[(size_t)N + q]; });
    stage_lds(s1, N, tid, nt, [&](int q) { return t.st[(size_t)N + q]; });
    stage_lds(r, N, tid, nt, [&](int q) { return R[(size_t)jj * N + q]; });   // the block's row
    __syncthreads();
    const int l = tid & (kDftLanes - 1), g = tid / kDftLanes;
    const int m = (blockIdx.x - jj * mblocks) * kDftGroups + g;
    const bool on = m < M1 && jj < J;
    double sc = 0.0, ss = 0.0;
    if (on) {
        const int step = (m * kDftLanes) % N;
        int q = (m * l) % N;   // (m k) mod N
#pragma unroll 4
        for (int k = l; k < N; k += kDftLanes) {
            const double rv = r[k];
            sc += rv * c1[q];
            ss += rv * s1[q];
            q += step;
            if (q >= N) q -= N;
        }
    }
    sc = tree16(sc);
    ss = tree16(ss);
    if (on && l == 0) {
        Ac[(size_t)m * J + jj] = sc;
        As[(size_t)m * J + jj] = ss;
    }
}

// Per mode and part: the folded theta system, mode 0 pinned at j = 1. One 64-lane block per
// (mode, part): the right-hand side, the mode's factors and the reciprocals of its pivots are
// staged in LDS with coalesced loads, one lane runs the two sequential sweeps out of LDS (the
// forward sweep multiplies by the reciprocals, off the division's latency), and the block writes
// the result back.
constexpr int kBaroMaxJ = 1024;
__global__ void __launch_bounds__(64) k_baro_tri(DevSolve t, double c13, double* Ac, double* As, const int* done) {
    if (done && *done) return;
    const int m = blockIdx.x >> 1, part = blockIdx.x & 1;
    const int J = t.J;
    double* d = (part == 0 ? Ac : As) + (size_t)m * J;
    if (part == 1 && (m == 0 || 2 * m == t.N)) {
        for (int jj = threadIdx.x; jj < J; jj += blockDim.x) d[jj] = 0.0;
        return;
    }
    __shared__ double sd[kBaroMaxJ], scp[kBaroMaxJ], sdn[kBaroMaxJ];
    const double* cp = t.mode_cp + (size_t)m * J;
    const double* dn = t.mode_dn + (size_t)m * J;
    for (int jj = threadIdx.x; jj < J; jj += blockDim.x) {
        sd[jj] = d[jj];
        scp[jj] = cp[jj];
        sdn[jj] = 1.0 / dn[jj];   // the sweep multiplies by the pivots' reciprocals
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double e = t.e;
        double prev = 0.0;
#pragma unroll 8
        for (int jj = 0; jj < J; jj++) {
            double a = (jj == 0) ? 0.0 : e;
            if (jj == J - 1) a = e - c13 * e;
            double rhs = sd[jj];
            if (m == 0 && jj == 0) rhs = 0.0;
            const double dp = (rhs - a * prev) * sdn[jj];
            sd[jj] = dp;
            prev = dp;
        }
        double x = sd[J - 1];
#pragma unroll 8
        for (int jj = J - 2; jj >= 0; jj--) {
            x = sd[jj] - scp[jj] * x;
            sd[jj] = x;
        }
    }
    __syncthreads();
    for (int jj = threadIdx.x; jj < J; jj += blockDim.x) d[jj] = sd[jj];
}

// inverse DFT of the correction: Q[jj][k] = (sum over m = 0..NM of t_m) / N with t_0 = Ac[0] c_0,
// t_m = 2 (Ac[m] c_m + As[m] s_m) for 0 < m < NM and t_NM = Ac[NM] c_NM (N even), summed in the
// order shared with the oracle: 16 partials over the modes m = l, l + 16, ..., then a halving tree.
// A group of 16 lanes forms one (jj, k); the groups of a block are consecutive k of one row. Q may
// overwrite the column residual.
__global__ void __launch_bounds__(kDftLanes * kDftGroups) k_dft_inv(DevSolve t, const double* Ac, const double* As,
                                                                     double* Q, const int* done) {
    if (done && *done) return;
    __shared__ double c1[kDftMaxN], s1[kDftMaxN], ta[kDftMaxN / 2 + 1], tb[kDftMaxN / 2 + 1];
    const int N = t.N, NM = t.NM, J = t.J;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int kblocks = (N + kDftGroups - 1) / kDftGroups;
    const int jj = blockIdx.x / kblocks;
    stage_lds(c1, N, tid, nt, [&](int q) { return t.ct[(size_t)N + q]; });
    stage_lds(s1, N, tid, nt, [&](int q) { return t.st[(size_t)N + q]; });
    stage_lds(ta, NM + 1, tid, nt, [&](int m) { return Ac[(size_t)m * J + jj]; });   // the block's row's modes
    stage_lds(tb, NM + 1, tid, nt, [&](int m) { return As[(size_t)m * J + jj]; });
    __syncthreads();
    const int l = tid & (kDftLanes - 1), g = tid / kDftLanes;
    const int k = (blockIdx.x - jj * kblocks) * kDftGroups + g;
    const bool on = k < N && jj < J;
    double s = 0.0;
    if (on) {
        const int step = (k * kDftLanes) % N;
        int q = (l * k) % N;   // (m k) mod N
#pragma unroll 4
        for (int m = l; m <= NM; m += kDftLanes) {
            const double a = ta[m];
            double tm;
            if (m == 0 || m == NM)
                tm = a * c1[q];
            else
                tm = 2.0 * (a * c1[q] + tb[m] * s1[q]);
            s += tm;
            q += step;
            if (q >= N) q -= N;
        }
    }
    s = tree16(s);
    if (on && l == 0) Q[(size_t)jj * N + k] = s / N;
}

// the correction added to every interior level: p += Q on the interior
template <typename P>
__global__ void __launch_bounds__(256) k_add_q(Shape sh, const double* Q, P* p, const int* done) {
    if (done && *done) return;
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 1 + blockIdx.z;
    if (k >= N || j > sh.jm - 2) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    p[x] = (P)((double)p[x] + Q[(size_t)(j - 1) * N + k]);
}

// ---- canonical interior sums: per row, 64-wide trees, chunks, rows and levels in sequence ----
// One wave per interior (i, j) row (kRowWaves rows per block, no barriers): the row's terms in
// chunks of 64 k (lane = k mod 64, zero past km-2), each chunk reduced by the 64-wide halving tree,
// the chunk sums added in order from 0. A wave forms kRowBatch chunks' terms before it reduces them,
// so that their loads go out together. The row index r = (i-1)(jm-2) + (j-1).
constexpr int kRowWaves = 4;
constexpr int kRowBatch = 3;

__device__ __forceinline__ double tree64(double t) {
    for (int h = 32; h > 0; h >>= 1) t += __shfl_down(t, h, 64);
    return t;   // the chunk's sum in lane 0
}

struct RowWave {
    int r, i, j, lane;
};
__device__ __forceinline__ bool row_wave(const Shape& sh, RowWave& w) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    w.r = blockIdx.x * kRowWaves + wave;
    if (w.r >= (sh.im - 2) * (sh.jm - 2)) return false;
    w.i = 1 + w.r / (sh.jm - 2);
    w.j = 1 + w.r % (sh.jm - 2);
    w.lane = threadIdx.x & 63;
    return true;
}

template <int MODE, typename P>
__global__ void __launch_bounds__(64 * kRowWaves) k_dsum_rows(Shape sh, DevGeo g, const P* __restrict__ p,
                                                               const double* __restrict__ v, double* __restrict__ rows,
                                                               float* __restrict__ r_out, const int* done) {
    if (done && *done) return;
    RowWave w;
    if (!row_wave(sh, w)) return;
    const int N = sh.km - 1;
    const size_t row = ((size_t)w.i * sh.jm + w.j) * sh.ldk;
    const double wy = endw(w.i, 1, sh.im - 2) * endw(w.j, 1, sh.jm - 2);
    double sum = 0.0;
    for (int c0 = 0; c0 < N; c0 += 64 * kRowBatch) {
        double t[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; u++) {
            const int k = c0 + 64 * u + w.lane;
            t[u] = 0.0;
            if (k < N) {
                const size_t x = row + k;
                if (MODE == DSUM_P) t[u] = (double)p[x];
                if (MODE == DSUM_YSRC) t[u] = wy * v[x];
                if (MODE == DSUM_SQ) t[u] = v[x] * v[x];
                if (MODE == DSUM_RESID2) {
                    const double r = resid_at(sh, g, p, v, w.i, w.j, k);
                    if (r_out) r_out[x] = (float)r;
                    t[u] = r * r;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; u++)
            if (c0 + 64 * u < N) sum += tree64(t[u]);
    }
    if (w.lane == 0) rows[w.r] = sum;
}

// The end of a converging solve in one pass, one wave per (i, j) row of the whole grid: the radial
// and polar wall values of the final interior (as k_walls forms them), the corner rows from the radial
// wall values they read (formed again here, the same bits), the seam copy k = km-1 <- k = 0 of every
// row, and each interior row's sum of p in k_dsum_rows<DSUM_P>'s order. No row reads a value this pass
// writes: the walls and corners read interior cells below km-1, the sums too, and the interior rows'
// only store is their seam cell.
template <typename P>
__global__ void __launch_bounds__(64 * kRowWaves) k_post_rows(Shape sh, double c43, double c13, P* __restrict__ p,
                                                               double* __restrict__ rows, const int* skip) {
    if (skip && *skip) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int q = blockIdx.x * kRowWaves + wave;
    if (q >= sh.im * sh.jm) return;
    const int i = q / sh.jm, j = q % sh.jm, lane = threadIdx.x & 63;
    const int N = sh.km - 1, L = sh.im - 1, J = sh.jm - 1, ldk = sh.ldk;
    const size_t PL = sh.plane;
    const size_t row = ((size_t)i * sh.jm + j) * ldk;
    const bool ri = (i == 0 || i == L), rj = (j == 0 || j == J);
    if (!ri && !rj) {
        double sum = 0.0;
        for (int c0 = 0; c0 < N; c0 += 64 * kRowBatch) {
            double t[kRowBatch];
#pragma unroll
            for (int u = 0; u < kRowBatch; u++) t[u] = (double)p[row + min(c0 + 64 * u + lane, N - 1)];
#pragma unroll
            for (int u = 0; u < kRowBatch; u++)
                if (c0 + 64 * u < N) sum += tree64(c0 + 64 * u + lane < N ? t[u] : 0.0);
        }
        if (lane == 0) {
            p[row + N] = p[row];
            rows[(size_t)(i - 1) * (sh.jm - 2) + (j - 1)] = sum;
        }
        return;
    }
    // radial (i) and polar (j) steps towards the interior
    const ptrdiff_t di = i == 0 ? (ptrdiff_t)PL : -(ptrdiff_t)PL, dj = j == 0 ? (ptrdiff_t)ldk : -(ptrdiff_t)ldk;
    for (int k = lane; k < N; k += 64) {
        const size_t x = row + k;
        P v;
        if (ri && !rj) {
            v = (P)(c43 * (double)p[x + di] - c13 * (double)p[x + 2 * di]);
        } else if (rj && !ri) {
            v = (P)(c43 * (double)p[x + dj] - c13 * (double)p[x + 2 * dj]);
        } else {   // a corner row: the polar form over the radial walls at j +- 1, j +- 2
            const P w1 = (P)(c43 * (double)p[x + dj + di] - c13 * (double)p[x + dj + 2 * di]);
            const P w2 = (P)(c43 * (double)p[x + 2 * dj + di] - c13 * (double)p[x + 2 * dj + 2 * di]);
            v = (P)(c43 * (double)w1 - c13 * (double)w2);
        }
        p[x] = v;
        if (k == 0) p[row + N] = v;
    }
}

// The p part of the solve's first pass at one cell, from p, pprev and pprev2 as read there (the row
// passes below load them ahead): p is the value whose interior sum the solve holds; with pprev, the
// warm start. With pprev2 (quadratic start) the history rotates by buffer: p goes
// into pprev2's buffer (its p_(n-2) read first) and the caller swaps the two pointers, so pprev
// holds p_n and pprev2 p_(n-1) after the pass.
__device__ __forceinline__ void setup_p_store(double* p, double* pprev, double* pprev2, int extrap, size_t x, double tp,
                                              double p1, double p2) {
    if (pprev) {
        if (pprev2) {
            if (extrap == 2) p[x] = 3.0 * tp - 3.0 * p1 + p2;
            else if (extrap) p[x] = 2.0 * tp - p1;
            pprev2[x] = tp;
        } else {
            if (extrap) p[x] = 2.0 * tp - p1;
            pprev[x] = tp;
        }
    }
}

// The two sums the solve starts with, in one pass: y . src and the interior sum of p, with the warm
// start of setup_p_store. Each batch of chunks loads every value first (lanes past km-2 load the row's last
// cell and discard it), then forms and stores.
__global__ void __launch_bounds__(64 * kRowWaves) k_dsum_rows_ysrc_p(Shape sh, double* __restrict__ p,
                                                                      const double* __restrict__ src,
                                                                      double* __restrict__ rows_y,
                                                                      double* __restrict__ rows_p,
                                                                      double* __restrict__ pprev,
                                                                      double* __restrict__ pprev2, int extrap) {
    RowWave w;
    if (!row_wave(sh, w)) return;
    const int N = sh.km - 1;
    const size_t row = ((size_t)w.i * sh.jm + w.j) * sh.ldk;
    const double wy = endw(w.i, 1, sh.im - 2) * endw(w.j, 1, sh.jm - 2);
    const double* q1 = pprev ? pprev : p;
    const double* q2 = pprev2 ? pprev2 : p;
    double sy = 0.0, sp = 0.0;
    for (int c0 = 0; c0 < N; c0 += 64 * kRowBatch) {
        double sv[kRowBatch], tp[kRowBatch], p1[kRowBatch], p2[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; u++) {
            const size_t x = row + min(c0 + 64 * u + w.lane, N - 1);
            sv[u] = src[x];
            tp[u] = p[x];
            p1[u] = q1[x];
            p2[u] = q2[x];
        }
        double ty[kRowBatch], tq[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; u++) {
            const int k = c0 + 64 * u + w.lane;
            ty[u] = 0.0;
            tq[u] = 0.0;
            if (k < N) {
                ty[u] = wy * sv[u];
                tq[u] = tp[u];
                setup_p_store(p, pprev, pprev2, extrap, row + k, tp[u], p1[u], p2[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; u++)
            if (c0 + 64 * u < N) {
                sy += tree64(ty[u]);
                sp += tree64(tq[u]);
            }
    }
    if (w.lane == 0) {
        rows_y[w.r] = sy;
        rows_p[w.r] = sp;
    }
}

// The same pass forming the source from aux first (the projection's clamped divergence, written to
// src): k_src and k_dsum_rows_ysrc_p in one launch, the same values.
template <typename S>
__global__ void __launch_bounds__(64 * kRowWaves) k_src_dsum_rows(Shape sh, DevGeo g, StepConsts c,
                                                                   const S* __restrict__ aux, double* __restrict__ src,
                                                                   double* __restrict__ p, double* __restrict__ rows_y,
                                                                   double* __restrict__ rows_p,
                                                                   double* __restrict__ pprev,
                                                                   double* __restrict__ pprev2, int extrap) {
    RowWave w;
    if (!row_wave(sh, w)) return;
    const int N = sh.km - 1;
    const size_t row = ((size_t)w.i * sh.jm + w.j) * sh.ldk;
    const size_t ij = (size_t)w.i * sh.jm + w.j;
    const double wy = endw(w.i, 1, sh.im - 2) * endw(w.j, 1, sh.jm - 2);
    const double* q1 = pprev ? pprev : p;
    const double* q2 = pprev2 ? pprev2 : p;
    const S* au = aux;
    const S* av = aux + sh.fs;
    const S* aw = aux + 2 * sh.fs;
    double sy = 0.0, sp = 0.0;
    for (int c0 = 0; c0 < N; c0 += 64 * kRowBatch) {
        double up[kRowBatch], un[kRowBatch], vp[kRowBatch], vn[kRowBatch], we[kRowBatch], ww[kRowBatch];
        double tp[kRowBatch], p1[kRowBatch], p2[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; u++) {
            const int k = min(c0 + 64 * u + w.lane, N - 1);
            const size_t x = row + k;
            up[u] = (double)au[x + sh.plane];
            un[u] = (double)au[x - sh.plane];
            vp[u] = (double)av[x + sh.ldk];
            vn[u] = (double)av[x - sh.ldk];
            we[u] = (double)aw[row + (k == N - 1 ? 0 : k + 1)];
            ww[u] = (double)aw[row + (k == 0 ? N - 1 : k - 1)];
            tp[u] = p[x];
            p1[u] = q1[x];
            p2[u] = q2[x];
        }
        double ty[kRowBatch], tq[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; u++) {
            const int k = c0 + 64 * u + w.lane;
            ty[u] = 0.0;
            tq[u] = 0.0;
            if (k < N) {
                const size_t x = row + k;
                const double sv = poisson_src_of(g, c, w.i, ij, up[u], un[u], vp[u], vn[u], we[u], ww[u]);
                src[x] = sv;
                ty[u] = wy * sv;
                tq[u] = tp[u];
                setup_p_store(p, pprev, pprev2, extrap, x, tp[u], p1[u], p2[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; u++)
            if (c0 + 64 * u < N) {
                sy += tree64(ty[u]);
                sp += tree64(tq[u]);
            }
    }
    if (w.lane == 0) {
        rows_y[w.r] = sy;
        rows_p[w.r] = sp;
    }
}

// The consistent source b = src - (y . src) / (y . 1) and the rows of |b|^2, in one pass.
__global__ void __launch_bounds__(64 * kRowWaves) k_make_b_sq_rows(Shape sh, const double* __restrict__ src,
                                                                    const double* __restrict__ sc, double yw,
                                                                    double* __restrict__ b, double* __restrict__ rows) {
    RowWave w;
    if (!row_wave(sh, w)) return;
    const int N = sh.km - 1;
    const size_t row = ((size_t)w.i * sh.jm + w.j) * sh.ldk;
    const double shift = sc[0] / yw;
    double sum = 0.0;
    for (int c0 = 0; c0 < N; c0 += 64 * kRowBatch) {
        double t[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; u++) {
            const int k = c0 + 64 * u + w.lane;
            t[u] = 0.0;
            if (k < N) {
                const double bv = src[row + k] - shift;
                b[row + k] = bv;
                t[u] = bv * bv;
            }
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; u++)
            if (c0 + 64 * u < N) sum += tree64(t[u]);
    }
    if (w.lane == 0) rows[w.r] = sum;
}

// The total of the row sums: partial t (one thread of 256) adds rows t, t + 256, ... in sequence,
// each wave reduces its 64 partials by the halving tree, and the 4 wave results are reduced by the
// halving tree. With a control block (the residual norm of the converging solve) it also decides
// convergence on the device: the counters take the work just done, then rel = |r| / |b| (|b|^2 in
// *bb) <= tol sets done.
constexpr int kDsumParts = 256;
// Partial t of the 256: rows t, t + 256, t + 512, ... added in sequence; the loads go out in batches
// of eight ahead of their adds (which keep the order).
__device__ __forceinline__ double partial_256(int nrows, const double* rows, int t) {
    double s = 0.0;
    int q = t;
    for (; q + 7 * kDsumParts < nrows; q += 8 * kDsumParts) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = rows[q + u * kDsumParts];
#pragma unroll
        for (int u = 0; u < 8; u++) s += x[u];
    }
    for (; q < nrows; q += kDsumParts) s += rows[q];
    return s;
}

__global__ void __launch_bounds__(kDsumParts) k_dsum_final(int nrows, const double* rows, double* out, SolveCtl ctl) {
    if (ctl.flags && ctl.flags[CTL_DONE]) return;
    if (ctl.skip && *ctl.skip) return;
    __shared__ double wsum[kDsumParts / 64];
    const int t = threadIdx.x;
    double s = partial_256(nrows, rows, t);
    for (int h = 32; h > 0; h >>= 1) s += __shfl_down(s, h, 64);
    if ((t & 63) == 0) wsum[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        const double total = (wsum[0] + wsum[2]) + (wsum[1] + wsum[3]);
        *out = total;
        if (ctl.flags) {
            int* f = ctl.flags;
            f[CTL_ITERS] += ctl.iters;
            f[CTL_OUTER] += ctl.outer;
            const double b2 = *ctl.bb;
            const double rel = b2 > 0.0 ? sqrt(total) / sqrt(b2) : sqrt(total);
            if (rel <= ctl.tol) f[CTL_DONE] = 1;
            f[CTL_NOTDONE] = f[CTL_DONE] ? 0 : 1;
            if (ctl.mbox) {
                for (int q = 0; q < CTL_N; q++) ctl.mbox->flags[q] = f[q];
                ctl.mbox->rr = total;
                ctl.mbox->bb = b2;
                __threadfence_system();
            }
        }
    }
}

// Two totals in the order of k_dsum_final (each array on its own), and, when given, the solve's
// control flags cleared for the iterations that follow.
__device__ __forceinline__ double total_256(int nrows, const double* rows, double* wsum) {
    const int t = threadIdx.x;
    double s = partial_256(nrows, rows, t);
    for (int h = 32; h > 0; h >>= 1) s += __shfl_down(s, h, 64);
    if ((t & 63) == 0) wsum[t >> 6] = s;
    __syncthreads();
    return (wsum[0] + wsum[2]) + (wsum[1] + wsum[3]);
}
__global__ void __launch_bounds__(kDsumParts) k_dsum_final2(int nrows, const double* r1, const double* r2, double* out1,
                                                            double* out2, int* reset_flags) {
    __shared__ double w1[kDsumParts / 64], w2[kDsumParts / 64];
    const double a = total_256(nrows, r1, w1);
    const double b = total_256(nrows, r2, w2);
    if (threadIdx.x == 0) {
        *out1 = a;
        *out2 = b;
        if (reset_flags) {
            for (int q = 0; q < CTL_N; q++) reset_flags[q] = 0;
            reset_flags[CTL_NOTDONE] = 1;
        }
    }
}

__global__ void __launch_bounds__(256) k_make_b(Shape sh, const double* src, const double* sc, double yw, double* b) {
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 1 + blockIdx.z;
    if (k >= N || j > sh.jm - 2) return;
    const double shift = sc[0] / yw;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    b[x] = src[x] - shift;
}

// hold the interior mean: p -= (mean_now - mean_before) on every cell
template <typename P>
__global__ void __launch_bounds__(256) k_shift_all(Shape sh, const double* sc, int slot_before, int slot_now,
                                                   double ncell, P* p, const int* skip) {
    if (skip && *skip) return;
    const size_t n = sh.fs;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double mean0 = sc[slot_before] / ncell;
    const double dm = sc[slot_now] / ncell - mean0;
    p[q] = (P)((double)p[q] - dm);
}

// refinement: p += e on the interior (FP64 accumulation of the FP32 correction)
__global__ void __launch_bounds__(256) k_add_correction(Shape sh, const float* e, double* p, const int* done) {
    if (done && *done) return;
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const int i = 1 + blockIdx.z;
    if (k >= N || j > sh.jm - 2) return;
    const size_t x = ((size_t)i * sh.jm + j) * sh.ldk + k;
    p[x] = p[x] + (double)e[x];
}

}  // namespace

// ---- launchers ----
template <typename P, typename B, typename R>
hipError_t launch_line(const Shape& sh, const DevGeo& g, const DevSolve& t, P* p, const B* b, int colour,
                       hipStream_t s, const int* done) {
    const int half = (sh.km - 1 + 1) / 2;
    const size_t lds = (size_t)t.ni * 64 * sizeof(R) + 2 * (size_t)t.ni * sizeof(double);
    if (lds <= 48 * 1024)
        hipLaunchKernelGGL((k_line_par<P, B, R>), dim3((half + 63) / 64, sh.jm - 2), dim3(64, kLineWaves), lds, s, sh,
                           g, t, p, b, colour, done);
    else   // a column too deep for the block's LDS: one thread per column
        hipLaunchKernelGGL((k_line<P, B, R>), dim3((half + 63) / 64, sh.jm - 2), dim3(64), 0, s, sh, g, t, p, b, colour,
                           done);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_walls(const Shape& sh, P* p, hipStream_t s, const int* done) {
    const int n = (sh.jm - 2 + sh.im - 2) * (sh.km - 1);
    hipLaunchKernelGGL((k_walls<P>), dim3((n + 255) / 256), dim3(256), 0, s, sh, 4.0 / 3.0, 1.0 / 3.0, p, done);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_seam(const Shape& sh, P* p, hipStream_t s, const int* skip) {
    hipLaunchKernelGGL((k_seam<P>), dim3((sh.jm + 255) / 256, sh.im), dim3(256), 0, s, sh, p, skip);
    return hipGetLastError();
}

template <typename P, typename B>
hipError_t launch_baro(const Shape& sh, const DevGeo& g, const DevSolve& t, P* p, const B* b, double* R, double* Ac,
                       double* As, hipStream_t s, const int* done) {
    if (t.J > kBaroMaxJ) return hipErrorInvalidValue;   // the theta solves stage a mode in LDS
    if (t.N > kDftMaxN || t.NM < 1 || 2 * t.NM != t.N) return hipErrorInvalidValue;   // LDS tables; N even
    const int N = sh.km - 1;
    const size_t lds = (size_t)(sh.im - 2) * 64 * sizeof(double);
    if (lds <= 48 * 1024)
        hipLaunchKernelGGL((k_col_resid_par<P, B>), dim3((N + 63) / 64, sh.jm - 2), dim3(64, kLineWaves), lds, s, sh, g,
                           t, (const P*)p, b, R, done);
    else
        hipLaunchKernelGGL((k_col_resid<P, B>), dim3((N + 63) / 64, sh.jm - 2), dim3(64), 0, s, sh, g, t, (const P*)p,
                           b, R, done);
    hipLaunchKernelGGL(k_dft_fwd, dim3(((t.NM + 1 + kDftGroups - 1) / kDftGroups) * t.J), dim3(kDftLanes * kDftGroups),
                       0, s, t, (const double*)R, Ac, As, done);
    hipLaunchKernelGGL(k_baro_tri, dim3(2 * (t.NM + 1)), dim3(64), 0, s, t, 1.0 / 3.0, Ac, As, done);
    hipLaunchKernelGGL(k_dft_inv, dim3(((N + kDftGroups - 1) / kDftGroups) * t.J), dim3(kDftLanes * kDftGroups), 0, s,
                       t, (const double*)Ac, (const double*)As, R, done);
    hipLaunchKernelGGL((k_add_q<P>), dim3((N + 63) / 64, (sh.jm - 2 + 3) / 4, sh.im - 2), dim3(64, 4), 0, s, sh,
                       (const double*)R, p, done);
    return hipGetLastError();
}

// blocks of kRowWaves rows over the interior rows
inline unsigned row_blocks(const Shape& sh) {
    return (unsigned)(((sh.im - 2) * (sh.jm - 2) + kRowWaves - 1) / kRowWaves);
}

template <typename P>
hipError_t launch_dsum(int mode, const Shape& sh, const DevGeo& g, const P* p, const double* v, double* rows,
                       double* out, hipStream_t s, float* r_out, const SolveCtl& ctl) {
    const int* done = ctl.skip ? ctl.skip : (ctl.flags ? ctl.flags + CTL_DONE : nullptr);
    const dim3 grid(row_blocks(sh)), threads(64 * kRowWaves);
    switch (mode) {
        case DSUM_P:
            hipLaunchKernelGGL((k_dsum_rows<DSUM_P, P>), grid, threads, 0, s, sh, g, p, v, rows, nullptr, done);
            break;
        case DSUM_YSRC:
            hipLaunchKernelGGL((k_dsum_rows<DSUM_YSRC, P>), grid, threads, 0, s, sh, g, p, v, rows, nullptr, done);
            break;
        case DSUM_SQ:
            hipLaunchKernelGGL((k_dsum_rows<DSUM_SQ, P>), grid, threads, 0, s, sh, g, p, v, rows, nullptr, done);
            break;
        default:
            hipLaunchKernelGGL((k_dsum_rows<DSUM_RESID2, P>), grid, threads, 0, s, sh, g, p, v, rows, r_out, done);
            break;
    }
    hipLaunchKernelGGL(k_dsum_final, dim3(1), dim3(kDsumParts), 0, s, (sh.im - 2) * (sh.jm - 2), (const double*)rows,
                       out, ctl);
    return hipGetLastError();
}

hipError_t launch_solve_setup(const Shape& sh, double* p, const double* src, double* rows, double* rows2,
                              double* scal, double yw, double* b, int* flags, double* pprev, double* pprev2,
                              int extrap, hipStream_t s, const void* aux, bool aux_f32, const DevGeo* g,
                              const StepConsts* c) {
    const dim3 grid(row_blocks(sh)), threads(64 * kRowWaves);
    const int nrows = (sh.im - 2) * (sh.jm - 2);
    if (!aux)
        hipLaunchKernelGGL(k_dsum_rows_ysrc_p, grid, threads, 0, s, sh, p, src, rows, rows2, pprev, pprev2, extrap);
    else if (aux_f32)
        hipLaunchKernelGGL((k_src_dsum_rows<float>), grid, threads, 0, s, sh, *g, *c, (const float*)aux,
                           (double*)src, p, rows, rows2, pprev, pprev2, extrap);
    else
        hipLaunchKernelGGL((k_src_dsum_rows<double>), grid, threads, 0, s, sh, *g, *c, (const double*)aux,
                           (double*)src, p, rows, rows2, pprev, pprev2, extrap);
    hipLaunchKernelGGL(k_dsum_final2, dim3(1), dim3(kDsumParts), 0, s, nrows, (const double*)rows,
                       (const double*)rows2, scal + 0, scal + 1, flags);
    hipLaunchKernelGGL(k_make_b_sq_rows, grid, threads, 0, s, sh, src, (const double*)scal, yw, b, rows);
    hipLaunchKernelGGL(k_dsum_final, dim3(1), dim3(kDsumParts), 0, s, nrows, (const double*)rows, scal + 2, SolveCtl());
    return hipGetLastError();
}

hipError_t launch_make_b(const Shape& sh, const double* src, const double* sc, double yw, double* b, hipStream_t s) {
    hipLaunchKernelGGL(k_make_b, dim3((sh.km - 1 + 63) / 64, (sh.jm - 2 + 3) / 4, sh.im - 2), dim3(64, 4), 0, s, sh,
                       src, sc, yw, b);
    return hipGetLastError();
}

hipError_t launch_add_correction(const Shape& sh, const float* e, double* p, hipStream_t s, const int* done) {
    hipLaunchKernelGGL(k_add_correction, dim3((sh.km - 1 + 63) / 64, (sh.jm - 2 + 3) / 4, sh.im - 2), dim3(64, 4), 0,
                       s, sh, e, p, done);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_post_dsum(const Shape& sh, P* p, double* rows, double* out, hipStream_t s, const int* skip) {
    const unsigned nb = (unsigned)((sh.im * sh.jm + kRowWaves - 1) / kRowWaves);
    hipLaunchKernelGGL((k_post_rows<P>), dim3(nb), dim3(64 * kRowWaves), 0, s, sh, 4.0 / 3.0, 1.0 / 3.0, p, rows, skip);
    SolveCtl k;
    k.skip = skip;
    hipLaunchKernelGGL(k_dsum_final, dim3(1), dim3(kDsumParts), 0, s, (sh.im - 2) * (sh.jm - 2), (const double*)rows,
                       out, k);
    return hipGetLastError();
}

template <typename P>
hipError_t launch_shift_all(const Shape& sh, const double* sc, int slot_before, int slot_now, double ncell, P* p,
                            hipStream_t s, const int* skip) {
    hipLaunchKernelGGL((k_shift_all<P>), dim3((unsigned)((sh.fs + 255) / 256)), dim3(256), 0, s, sh, sc, slot_before,
                       slot_now, ncell, p, skip);
    return hipGetLastError();
}

template hipError_t launch_line<double, double, double>(const Shape&, const DevGeo&, const DevSolve&, double*,
                                                        const double*, int, hipStream_t, const int*);
template hipError_t launch_line<float, float, float>(const Shape&, const DevGeo&, const DevSolve&, float*,
                                                     const float*, int, hipStream_t, const int*);
template hipError_t launch_baro<double, double>(const Shape&, const DevGeo&, const DevSolve&, double*, const double*,
                                                double*, double*, double*, hipStream_t, const int*);
template hipError_t launch_baro<float, float>(const Shape&, const DevGeo&, const DevSolve&, float*, const float*,
                                              double*, double*, double*, hipStream_t, const int*);
template hipError_t launch_post_dsum<double>(const Shape&, double*, double*, double*, hipStream_t, const int*);
template hipError_t launch_walls<double>(const Shape&, double*, hipStream_t, const int*);
template hipError_t launch_walls<float>(const Shape&, float*, hipStream_t, const int*);
template hipError_t launch_seam<double>(const Shape&, double*, hipStream_t, const int*);
template hipError_t launch_seam<float>(const Shape&, float*, hipStream_t, const int*);
template hipError_t launch_dsum<double>(int, const Shape&, const DevGeo&, const double*, const double*, double*,
                                        double*, hipStream_t, float*, const SolveCtl&);
template hipError_t launch_shift_all<double>(const Shape&, const double*, int, int, double, double*, hipStream_t,
                                              const int*);

}  // namespace atu
