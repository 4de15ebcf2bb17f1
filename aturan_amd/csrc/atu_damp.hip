// Shapiro (1-2-1) wiggle damping, AtomUtils::damp_wiggles (lib/Utils.cpp:224-306): per pass along
// k, along j, along i, each f + strength/4 (f[n-1] - 2 f + f[n+1]) from the field as the previous
// direction left it. No fluid floor (i_surface = nullptr, as every call in the reference's loop
// passes, planet/cUranusModel.cpp:1218-1240); the phi seam periodic as the model sets it
// (planet/cUranusModel.cpp:459): k = 0..km-2 over the ring of km-1 meridians, k = km-1 the copy of
// the new k = 0. j is clamped at the poles, i at the bottom and the top.
//
// One thread per cell; each direction reads one buffer and writes the other (the reference's copy
// into tmp), so a pass needs no synchronisation inside a launch.

#include <utility>

#include "atu_kernels.h"

namespace atu {
namespace {

__global__ void __launch_bounds__(256) k_damp_k(Shape sh, double coeff, const double* src, double* dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    const int km = sh.km;
    if (k >= km) return;
    const size_t row = ((size_t)i * sh.jm + j) * sh.ldk;
    const int kk = (k == km - 1) ? 0 : k;   // the seam column takes the new k = 0
    const int km1 = (kk == 0) ? km - 2 : kk - 1, kp1 = (kk == km - 2) ? 0 : kk + 1;
    const double c = src[row + kk];
    dst[row + k] = c + coeff * (src[row + km1] - 2.0 * c + src[row + kp1]);
}

__global__ void __launch_bounds__(256) k_damp_j(Shape sh, double coeff, const double* src, double* dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.km) return;
    const int jm1 = j > 0 ? j - 1 : 0, jp1 = j < sh.jm - 1 ? j + 1 : sh.jm - 1;
    const size_t lev = (size_t)i * sh.plane;
    const double c = src[lev + (size_t)j * sh.ldk + k];
    dst[lev + (size_t)j * sh.ldk + k] =
        c + coeff * (src[lev + (size_t)jm1 * sh.ldk + k] - 2.0 * c + src[lev + (size_t)jp1 * sh.ldk + k]);
}

__global__ void __launch_bounds__(256) k_damp_i(Shape sh, double coeff, const double* src, double* dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= sh.km) return;
    const int im1 = i > 0 ? i - 1 : 0, ip1 = i < sh.im - 1 ? i + 1 : sh.im - 1;
    const size_t col = (size_t)j * sh.ldk + k;
    const double c = src[(size_t)i * sh.plane + col];
    dst[(size_t)i * sh.plane + col] =
        c + coeff * (src[(size_t)im1 * sh.plane + col] - 2.0 * c + src[(size_t)ip1 * sh.plane + col]);
}

}  // namespace

hipError_t launch_damp_wiggles(const Shape& sh, double* field, double* scratch, bool along_i, bool along_j,
                               bool along_k, double strength, int passes, hipStream_t s) {
    const double coeff = strength * 0.25;
    const dim3 grid((sh.km + 255) / 256, sh.jm, sh.im);
    double* a = field;
    double* b = scratch;
    for (int p = 0; p < passes; p++) {
        if (along_k) {
            hipLaunchKernelGGL(k_damp_k, grid, dim3(256), 0, s, sh, coeff, (const double*)a, b);
            std::swap(a, b);
        }
        if (along_j) {
            hipLaunchKernelGGL(k_damp_j, grid, dim3(256), 0, s, sh, coeff, (const double*)a, b);
            std::swap(a, b);
        }
        if (along_i) {
            hipLaunchKernelGGL(k_damp_i, grid, dim3(256), 0, s, sh, coeff, (const double*)a, b);
            std::swap(a, b);
        }
    }
    if (a != field) {
        const hipError_t e = hipMemcpyAsync(field, a, sh.fs * sizeof(double), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace atu
