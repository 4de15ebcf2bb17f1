// Device-side argument blocks and launcher declarations of libaturan_amd.
#pragma once

#include <hip/hip_runtime.h>

#include "atu_internal.h"

namespace atu {

// Geometry tables in device memory (FP64), built by build_geometry().
struct DevGeo {
    const double *rm, *inv_rm, *inv_rm2, *two_inv_rm, *sponge, *num2;                 // [im]
    const double *costhe, *cotanthe, *sinthe_bf, *cor_hydro, *sin_weight;             // [jm]
    const double *inv_rmsinthe, *inv_rm2sinthe, *inv_rm2sinthe2, *costhe_inv_rm2sinthe,
        *v_metric, *num3, *den, *inv_rmsinthe_s;                                       // [im][jm]
    const double* cp_tab;                                                              // [cp_n]
    int cp_n;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2, num1;
};

// Shape of every 3-D device field: [im][jm][ldk], k fastest; fields of one buffer are `fs` apart.
struct Shape {
    int im, jm, km, ldk;
    size_t plane;   // jm * ldk
    size_t fs;      // im * jm * ldk
};

// One RK stage. `xs` is the state the stencils read; `yn` the FP64 start-of-step state; `acc` the
// FP64 running sum k1 + 2k2 + 2k3 + k4; `xnext` the next stage's input (stages 0-2); stage 3 writes
// y_{n+1} into `yn` (own cell only) and, when non-null, its FP32 mirror `ymirror`.
template <typename S>
struct StageArgs {
    const S* xs;
    const double* yn;
    double* yout;
    S* xnext;
    S* ymirror;
    double* acc;
    const S* frozen;     // 21 fields
    const S* pdyn;
    const S* phydro;
    S* aux;              // 3 fields, written at stage 3
    Shape sh;
    DevGeo geo;
    StepConsts c;
};

// Launchers (stream-ordered, asynchronous).
hipError_t launch_rk_stage_f64(int stage, const StageArgs<double>& a, hipStream_t s);
hipError_t launch_rk_stage_mixed(int stage, const StageArgs<float>& a, hipStream_t s);

// Hydrostatic prep: level means (in the reference's serial summation order) into `level_out` [im],
// then the column scan into `phydro`.
hipError_t launch_hydro(const Shape& sh, const DevGeo& g, const StepConsts& c, const double* t,
                        const double* v, const double* w, const double* pstat, double* level_out,
                        double* phydro, hipStream_t s);

// Domain BC on the fields of the path that the reference's lists name: the 17 prognostic fields of
// `y` (FP64; every value written also goes to each non-null mirror, FP64 `m64a`, `m64b`, FP32
// `m32a..c`) and the three reaction mass fluxes and the NH4SH flux limiter of `frozen` (with the
// optional FP32 mirror `frozen32`).
struct BcTargets {
    double* y;
    double* m64a;
    double* m64b;
    float* m32a;
    float* m32b;
    float* m32c;
    double* frozen;
    float* frozen32;
};
hipError_t launch_bc(const Shape& sh, const atu_params& p, const BcTargets& t, hipStream_t s);

// Projection prep on aux (prepareProjectionBoundaries) and the clamped divergence source.
template <typename S>
hipError_t launch_proj_src(const Shape& sh, const DevGeo& g, const StepConsts& c, S* aux, double* src,
                           hipStream_t s);

// One colour pass of the red-black relaxation on p (FP64 or FP32 storage, FP64 source). The black
// pass (colour 1) also writes the radial and polar wall values and the seam copy that the
// reference's boundary passes form after each sweep.
template <typename P>
hipError_t launch_rb_pass(const Shape& sh, const DevGeo& g, const StepConsts& c, P* p, const double* src,
                          int colour, double omega, hipStream_t s);
// The four corner rows (i = 0, im-1 at j = 0, jm-1) and their seam copies, after the last sweep.
template <typename P>
hipError_t launch_p_corners(const Shape& sh, const StepConsts& c, P* p, hipStream_t s);

// Residual r = src - Lc p over the interior: writes per-block partial sums of r^2 and src^2 into
// `partials` (2 * nblocks doubles) and the fixed-order total into out[0] (r^2), out[1] (src^2).
int residual_partials_count(const Shape& sh);
template <typename P>
hipError_t launch_residual(const Shape& sh, const DevGeo& g, const P* p, const double* src,
                           double* r_out, double* partials, double* out, hipStream_t s);

}  // namespace atu
