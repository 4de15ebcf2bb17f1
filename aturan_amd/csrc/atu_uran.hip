// PressureSolverUran::run (planet/PressureSolverUran.h:33-250), the reference's default pressure
// solver, at its defaults: periodic phi, PRESS_CONSISTENT (metric radius, src_scale = r_mix u0^2 /
// 1e5) and PRESS_INCR (the defect-correction source src_scale div(rhs) + Lc(p_old)).
//
// The reference relaxes p_dyn by one serial lexicographic Gauss-Seidel sweep over (i, j, k). A cell
// reads its i-1, j-1 and k-1 neighbours already updated and its i+1, j+1, k+1 neighbours not yet;
// on the periodic ring the west neighbour of k = 0 is k = N-1 (not yet updated) and the east
// neighbour of k = N-1 is k = 0 (updated). Every one of these is a cell of a lower (updated) or
// higher (not yet updated) hyperplane s = i + j + k, so sweeping the hyperplanes in order, all
// cells of one hyperplane at once, gives the lexicographic sweep's result exactly. One workgroup
// runs the hyperplanes with a barrier between them (no grid-wide synchronisation).
//
// The weights are host-built with the reference's two expression sets (the source's and the
// sweep's, which round differently) and every expression keeps its operand order.

#include "atu_kernels.h"

namespace atu {
namespace {

// radial walls of aux_u, aux_v, aux_w at j = 1..jm-2, k = 0..km-2 (:73-84)
__global__ void __launch_bounds__(256) k_uran_aux_walls(Shape sh, double* aux) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y;
    if (k >= sh.km - 1) return;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const size_t P = sh.plane, L = sh.im - 1;
    for (int q = 0; q < 3; q++) {
        double* A = aux + (size_t)q * sh.fs + (size_t)j * sh.ldk + k;
        A[0] = c43 * A[P] - c13 * A[2 * P];
        A[L * P] = c43 * A[(L - 1) * P] - c13 * A[(L - 2) * P];
    }
}

// incr = src_scale * div(rhs) + Lc(p_old) on the interior (:139-175); rhs_u's radial walls are its
// 2-point extrapolations, rhs_v's polar walls are whatever the array holds (0: the stage kernel
// writes the interior only)
__global__ void __launch_bounds__(256) k_uran_src(Shape sh, UranTabs t, const double* rhs3, const double* p,
                                                 double* incr) {
    const int N = sh.km - 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y, i = 1 + blockIdx.z;
    if (k >= N) return;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const int L = sh.im - 1, ldk = sh.ldk;
    const size_t P = sh.plane, fs = sh.fs;
    const size_t row = ((size_t)i * sh.jm + j) * ldk;
    const size_t x = row + k;
    const size_t xw = row + (k == 0 ? N - 1 : k - 1), xe = row + (k == N - 1 ? 0 : k + 1);
    const size_t ij = (size_t)i * sh.jm + j;
    const double* ru = rhs3;
    const double* rv = rhs3 + fs;
    const double* rw = rhs3 + 2 * fs;
    const size_t col = (size_t)j * ldk + k;
    const double up = (i + 1 == L) ? c43 * ru[col + (size_t)(L - 1) * P] - c13 * ru[col + (size_t)(L - 2) * P]
                                   : ru[x + P];
    const double um = (i - 1 == 0) ? c43 * ru[col + P] - c13 * ru[col + 2 * P] : ru[x - P];
    const double div_rhs = (up - um) / t.two_dr + (rv[x + ldk] - rv[x - ldk]) / t.two_dthe * t.a_inv_rm[i]
                           + (rw[xe] - rw[xw]) / t.two_dphi * t.a_inv_rmsinthe[ij];
    const double lc_old = (p[x + P] + p[x - P]) * t.a_num1 + (p[x + ldk] + p[x - ldk]) * t.a_num2[i]
                          + (p[xe] + p[xw]) * t.a_num3[ij] - t.a_denom[ij] * p[x];
    incr[x] = t.src_scale * div_rhs + lc_old;
}

// The lexicographic sweep by hyperplanes s = (i-1) + (j-1) + k. Each thread takes (i, j) pairs
// q, q + 1024, ...; for a hyperplane it loads the inputs of up to URAN_GROUP of its cells before
// storing any of them (the cells of one hyperplane do not read each other).
constexpr int URAN_GROUP = 8;
__global__ void __launch_bounds__(1024) k_uran_sweep(Shape sh, UranTabs t, double* p, const double* incr) {
    const int ni = sh.im - 2, nj = sh.jm - 2, N = sh.km - 1;
    const int npairs = ni * nj;
    const int ldk = sh.ldk;
    const size_t P = sh.plane;
    const int smax = (ni - 1) + (nj - 1) + (N - 1);
    for (int s = 0; s <= smax; s++) {
        for (int q0 = threadIdx.x; q0 < npairs; q0 += URAN_GROUP * blockDim.x) {
            size_t xs[URAN_GROUP], xws[URAN_GROUP], xes[URAN_GROUP], ijs[URAN_GROUP];
            double v[URAN_GROUP];
            bool on[URAN_GROUP];
#pragma unroll
            for (int u = 0; u < URAN_GROUP; u++) {
                const int q = q0 + u * blockDim.x;
                const int ii = q / nj, jj = q - (q / nj) * nj;
                const int k = s - ii - jj;
                on[u] = q < npairs && k >= 0 && k < N;
                if (on[u]) {
                    const int i = ii + 1, j = jj + 1;
                    const size_t row = ((size_t)i * sh.jm + j) * ldk;
                    xs[u] = row + k;
                    xws[u] = row + (k == 0 ? N - 1 : k - 1);
                    xes[u] = row + (k == N - 1 ? 0 : k + 1);
                    ijs[u] = (size_t)i * sh.jm + j;
                    const size_t x = xs[u];
                    v[u] = ((p[x + P] + p[x - P]) * t.m_num1 + (p[x + ldk] + p[x - ldk]) * t.m_num2[i]
                            + (p[xes[u]] + p[xws[u]]) * t.m_num3[ijs[u]] - incr[x]) * t.m_inv_denom[ijs[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < URAN_GROUP; u++)
                if (on[u]) p[xs[u]] = v[u];
        }
        __syncthreads();
    }
}

// p_dyn walls as the reference forms them after its sweep (:213-244): radial planes at every j and
// k, then the polar rows at every i, then the seam copy
__global__ void __launch_bounds__(256) k_uran_walls_r(Shape sh, double* p) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (k >= sh.km) return;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const size_t P = sh.plane, L = sh.im - 1;
    double* c = p + (size_t)j * sh.ldk + k;
    c[0] = c43 * c[P] - c13 * c[2 * P];
    c[L * P] = c43 * c[(L - 1) * P] - c13 * c[(L - 2) * P];
}

__global__ void __launch_bounds__(256) k_uran_walls_t(Shape sh, double* p) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (k >= sh.km) return;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const size_t ldk = sh.ldk, J = sh.jm - 1;
    double* c = p + (size_t)i * sh.plane + k;
    c[0] = c43 * c[ldk] - c13 * c[2 * ldk];
    c[J * ldk] = c43 * c[(J - 1) * ldk] - c13 * c[(J - 2) * ldk];
}

__global__ void __launch_bounds__(256) k_uran_seam(Shape sh, double* p) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= sh.jm) return;
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    p[row + sh.km - 1] = p[row];
}

}  // namespace

hipError_t launch_uran_call(const Shape& sh, const UranTabs& t, double* aux, const double* rhs3, double* p,
                            double* incr, hipStream_t s) {
    const int N = sh.km - 1;
    hipLaunchKernelGGL(k_uran_aux_walls, dim3((N + 255) / 256, sh.jm - 2), dim3(256), 0, s, sh, aux);
    hipLaunchKernelGGL(k_uran_src, dim3((N + 255) / 256, sh.jm - 2, sh.im - 2), dim3(256), 0, s, sh, t, rhs3,
                       (const double*)p, incr);
    hipLaunchKernelGGL(k_uran_sweep, dim3(1), dim3(1024), 0, s, sh, t, p, (const double*)incr);
    hipLaunchKernelGGL(k_uran_walls_r, dim3((sh.km + 255) / 256, sh.jm), dim3(256), 0, s, sh, p);
    hipLaunchKernelGGL(k_uran_walls_t, dim3((sh.km + 255) / 256, sh.im), dim3(256), 0, s, sh, p);
    hipLaunchKernelGGL(k_uran_seam, dim3((sh.jm + 255) / 256, sh.im), dim3(256), 0, s, sh, p);
    return hipGetLastError();
}

}  // namespace atu
