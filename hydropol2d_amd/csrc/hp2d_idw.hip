// hp2d_idw.hip - inverse-distance-weighted interpolation of rain gauges onto
// the grid (Rainfall_Interpolator.m:40-52, idw_function.m:18-63), the spatial
// rainfall maps of update_spatial_BC.m:469-490.
//
// One thread per grid cell, FP64; the gauges (coordinates and values) are
// staged in LDS in chunks, so every thread reads them from LDS.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <string>

#include "hp2d.h"

namespace hp2d {
void set_global_error(const std::string &m);

namespace {

constexpr int kIdwBlock = 256;
constexpr int kIdwChunk = 512;  // gauges per LDS chunk

__global__ __launch_bounds__(kIdwBlock) void idw_kernel(int n_obs, const double *__restrict__ xo,
                                                        const double *__restrict__ yo,
                                                        const double *__restrict__ fo, int nxg,
                                                        const double *__restrict__ xg, int nyg,
                                                        const double *__restrict__ yg, double p,
                                                        double rad, double *__restrict__ out) {
    __shared__ double sx[kIdwChunk], sy[kIdwChunk], sf[kIdwChunk];
    const long long n = (long long)nxg * nyg;
    const long long c = (long long)blockIdx.x * kIdwBlock + threadIdx.x;
    const bool live = c < n;
    const int i = live ? (int)(c % nyg) : 0, j = live ? (int)(c / nyg) : 0;
    const double x = live ? xg[j] : 0.0, y = live ? yg[i] : 0.0;
    double num = 0.0, den = 0.0;
    for (int g0 = 0; g0 < n_obs; g0 += kIdwChunk) {
        const int m = min(kIdwChunk, n_obs - g0);
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += kIdwBlock) {
            sx[k] = xo[g0 + k];
            sy[k] = yo[g0 + k];
            sf[k] = fo[g0 + k];
        }
        __syncthreads();
        for (int k = 0; k < m; ++k) {
            const double dx = fabs(x - sx[k]), dy = fabs(y - sy[k]);
            // (|dx|^2 + |dy|^2)^(1/2); D = 0 -> eps; beyond the radius -> Inf
            double D = sqrt(dx * dx + dy * dy);
            if (D == 0) D = DBL_EPSILON;
            if (D > rad) D = INFINITY;
            const double W = p == 2.0 ? 1.0 / (D * D) : 1.0 / pow(D, p);
            num += W * sf[k];
            den += W;
        }
    }
    if (live) out[c] = num / den;
}

}  // namespace
}  // namespace hp2d

extern "C" {

hp2d_status hp2d_idw_interpolate(int32_t n_obs, const double *x_obs, const double *y_obs,
                                 const double *f_obs, int32_t nx_grid, const double *x_grid,
                                 int32_t ny_grid, const double *y_grid, double p, double rad,
                                 double *out) {
    if (n_obs < 1 || nx_grid < 1 || ny_grid < 1 || !x_obs || !y_obs || !f_obs || !x_grid ||
        !y_grid || !out || !(p > 0) || !(rad > 0)) {
        hp2d::set_global_error("hp2d_idw_interpolate: bad arguments");
        return HP2D_E_ARG;
    }
    const long long n = (long long)nx_grid * ny_grid;
    if (n > (1LL << 31)) {
        hp2d::set_global_error("hp2d_idw_interpolate: grid beyond 2^31 cells");
        return HP2D_E_ARG;
    }
    double *d = nullptr;
    const size_t obs = (size_t)n_obs * sizeof(double);
    const size_t bytes = 3 * obs + ((size_t)nx_grid + ny_grid + n) * sizeof(double);
    hipError_t e = hipMalloc((void **)&d, bytes);
    if (e == hipSuccess) {
        double *xo = d, *yo = d + n_obs, *fo = d + 2 * n_obs;
        double *xg = d + 3 * n_obs, *yg = xg + nx_grid, *o = yg + ny_grid;
        e = hipMemcpy(xo, x_obs, obs, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(yo, y_obs, obs, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(fo, f_obs, obs, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(xg, x_grid, nx_grid * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(yg, y_grid, ny_grid * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(hp2d::idw_kernel, dim3((unsigned)((n + 255) / 256)),
                               dim3(hp2d::kIdwBlock), 0, 0, n_obs, xo, yo, fo, nx_grid, xg,
                               ny_grid, yg, p, rad, o);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(out, o, n * sizeof(double), hipMemcpyDeviceToHost);
        hipFree(d);
    }
    if (e != hipSuccess) {
        hp2d::set_global_error(std::string("hp2d_idw_interpolate: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? HP2D_E_NOMEM : HP2D_E_HIP;
    }
    return HP2D_OK;
}

}  // extern "C"
