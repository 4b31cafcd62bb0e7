// hp2d_bates.hip - elementwise Inertial_Solver (Inertial_Solver.m:19-26), the
// stand-alone form of the face update that the fused step inlines.
#include "hp2d_internal.hpp"

namespace hp2d {

__global__ __launch_bounds__(256) void bates_kernel(const double *q_p, double dt, const double *Hf,
                                                    const double *S, const double *n_sq,
                                                    long long n_cells, int n_pages, double *q) {
    const long long total = n_cells * n_pages;
    for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < total;
         o += (long long)gridDim.x * blockDim.x) {
        const long long c = o % n_cells;
        const double g = 9.81;
        const double num = q_p[o] - g * Hf[o] * dt * S[o];
        const double den = 1.0 + g * dt * n_sq[c] * fabs(q_p[o]) / pow(Hf[o], 7.0 / 3.0);
        q[o] = num / den;
    }
}

hipError_t launch_bates(const double *q_p, double dt, const double *Hf, const double *S,
                        const double *n_sq, long long n_cells, int n_pages, double *q,
                        hipStream_t s) {
    const long long total = n_cells * n_pages;
    int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(bates_kernel, dim3(blocks), dim3(256), 0, s, q_p, dt, Hf, S, n_sq, n_cells,
                       n_pages, q);
    return hipGetLastError();
}

}  // namespace hp2d
