// atu_pressure_residual_static: the relative residual of the red-black Poisson system, as
// atu_pressure_residual forms it for a context, of caller-held device arrays, with nothing allocated.
// The sweep's weights and the residual's block partials live in one file-scope device array,
// `scratch`, of kScratchBytes per GPU; its device address is looked up once per device (with that
// device current) and kept.

#include "atu_kernels.h"

#include <cmath>
#include <cstdio>
#include <mutex>

namespace atu {
namespace {

constexpr size_t kScratchBytes = size_t(4) << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
constexpr int kMaxDevices = 64;

// num2 [im], num3 [im][jm], den [im][jm], the residual partials (2 per block) and the two totals
__device__ __align__(256) double scratch[kScratchDoubles];

// Host side: the array's address on each device, and the staging of the weights. Every call holds
// the lock from its first HIP call until its result is on the host, so two calls never share them.
double* g_scratch_dev[kMaxDevices];
double g_stage[kScratchDoubles];
std::mutex g_lock;

}  // namespace
}  // namespace atu

using namespace atu;

extern "C" int atu_pressure_residual_static(const atu_grid* g, const atu_params* p, const double* p_dyn,
                                            const double* src, int ld_k, void* hip_stream, double* rel,
                                            double* srcn) {
    if (!g || !p || !p_dyn || !src) {
        set_free_error("atu_pressure_residual_static: null grid, params or array");
        return ATU_EINVAL;
    }
    if (g->im < 3 || g->jm < 3 || g->km < 3 || ld_k < g->km) {
        set_free_error("atu_pressure_residual_static: need im, jm, km >= 3 and ld_k >= km");
        return ATU_EINVAL;
    }
    Shape sh{};
    sh.im = g->im;
    sh.jm = g->jm;
    sh.km = g->km;
    sh.ldk = ld_k;
    sh.plane = (size_t)g->jm * ld_k;
    sh.fs = (size_t)g->im * sh.plane;
    const size_t nij = (size_t)g->im * g->jm;
    const size_t n_weights = (size_t)g->im + 2 * nij;
    const size_t n_part = 2 * (size_t)residual_partials_count(sh);
    const size_t off_part = (n_weights + 31) / 32 * 32;            // partials on a 256-byte boundary
    const size_t off_out = (off_part + n_part + 31) / 32 * 32;
    if (off_out + 2 > kScratchDoubles) {
        char m[256];
        std::snprintf(m, sizeof m,
                      "atu_pressure_residual_static: a %dx%dx%d grid needs %zu bytes of scratch, more than the "
                      "%zu this entry point holds",
                      g->im, g->jm, g->km, (off_out + 2) * sizeof(double), kScratchBytes);
        set_free_error(m);
        return ATU_EINVAL;
    }
    std::lock_guard<std::mutex> hold(g_lock);
    auto fail = [](hipError_t e, const char* what) {
        char m[256];
        std::snprintf(m, sizeof m, "atu_pressure_residual_static: %s: %s", what, hipGetErrorString(e));
        set_free_error(m);
        return ATU_EHIP;
    };
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) {
        set_free_error("atu_pressure_residual_static: device ordinal beyond the 64 this entry point tracks");
        return ATU_EINVAL;
    }
    if (!g_scratch_dev[dev]) {   // the device's copy of the array (the host-side name is no address)
        void* a = nullptr;
        e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_dev[dev] = static_cast<double*>(a);
    }
    double* base = g_scratch_dev[dev];
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    // the weights, as a context's geometry holds them (every value is written: nothing to clear)
    DevGeo geo{};
    geo.num1 = solver_weights(*g, *p, g_stage, g_stage + g->im, g_stage + g->im + nij);
    e = hipMemcpyAsync(base, g_stage, n_weights * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(e, "hipMemcpyAsync(weights)");
    geo.num2 = base;
    geo.num3 = base + g->im;
    geo.den = base + g->im + nij;
    // every block writes its partial and the final sum writes both totals: nothing to clear
    e = launch_residual<double>(sh, geo, p_dyn, src, nullptr, base + off_part, base + off_out, s);
    if (e != hipSuccess) return fail(e, "residual");
    double h[2] = {0.0, 0.0};
    e = hipMemcpyAsync(h, base + off_out, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return fail(e, "hipMemcpyAsync(result)");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(e, "hipStreamSynchronize");
    const double rn = std::sqrt(h[0]), sn = std::sqrt(h[1]);
    if (srcn) *srcn = sn;
    if (rel) *rel = sn > 0.0 ? rn / sn : rn;
    return ATU_OK;
}
