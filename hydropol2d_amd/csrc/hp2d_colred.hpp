// hp2d_colred.hpp - per-column reductions of column-major grids (nansum /
// NaN-omitting max of MATLAB over the first dimension).
//
// One 256-thread block per column: thread t folds rows t, t + 256, ... (the
// wave reads 64 consecutive rows of the column, one coalesced access), then
// the block combines its waves in a fixed order, so results are deterministic.
#pragma once

#include "hp2d_internal.hpp"

namespace hp2d {

constexpr int kColThreads = 256;

// load(c, v) writes the F values of cell c into v; is_max[f] selects a
// NaN-omitting maximum (NaN when every value is NaN) instead of a NaN-skipping
// sum.  out[F * j + f] receives field f of column j.
template <int F, typename Load>
__device__ __forceinline__ void column_reduce(int ny, int j, const bool (&is_max)[F], Load load,
                                              double *out) {
    double acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = is_max[f] ? NAN : 0.0;
    const long long base = (long long)j * ny;
    for (int i = threadIdx.x; i < ny; i += kColThreads) {
        double v[F];
        load(base + i, v);
#pragma unroll
        for (int f = 0; f < F; ++f) {
            if (is_max[f]) acc[f] = fmax(acc[f], v[f]);
            else if (!isnan(v[f])) acc[f] += v[f];
        }
    }
    __shared__ double sh[F][kColThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int f = 0; f < F; ++f) {
        const double r = is_max[f] ? wave_fmax(acc[f]) : wave_sum(acc[f]);
        if (lane == 0) sh[f][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < F) {
        const int f = threadIdx.x;
        double r = sh[f][0];
        for (int k = 1; k < kColThreads / kWave; ++k) r = is_max[f] ? fmax(r, sh[f][k]) : r + sh[f][k];
        out[(long long)F * j + f] = r;
    }
}

}  // namespace hp2d
