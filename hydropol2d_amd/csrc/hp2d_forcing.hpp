// hp2d_forcing.hpp - exact step integral of an end-stamped piecewise-constant
// forcing series, usable on host and device.
//
// Restates localAverageRateOverStep / localAdvanceEndStampedCursor /
// localEstimateConstantDt of update_spatial_BC.m:1095-1185,1240-1250:
// times[i] is the END of interval i (interval 0 starts at times[0] minus the
// series spacing), the cursor (1-based, as in the reference) only moves
// forward, and the result is the time-averaged rate over [t0, t1].
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hp2d {

// The series read through an accessor: times(k), values(k) for 0-based k.
template <typename Series>
__host__ __device__ inline double rain_average(const Series &ser, int n, double t0, double t1,
                                               int *cursor) {
    if (n <= 0 || !__builtin_isfinite(t0) || !__builtin_isfinite(t1) || t1 <= t0) return 0.0;
    const double span = t1 - t0;
    double acc = 0.0;
    int cur = *cursor;
    if (cur < 1) cur = 1;
    if (cur > n) cur = n;
    while (cur < n && t0 > ser.time(cur - 1)) ++cur;
    const double spacing = (n >= 2) ? ser.time(1) - ser.time(0) : 0.0;
    for (int i = cur; i <= n; ++i) {
        const double start = (i == 1) ? ser.time(0) - spacing : ser.time(i - 2);
        const double end = ser.time(i - 1);
        const double v = ser.value(i - 1);
        const double overlap = fmax(0.0, fmin(t1, end) - fmax(t0, start));
        if (overlap > 0 && __builtin_isfinite(v)) acc = acc + v * overlap;
        if (end >= t1) {
            cur = i;
            break;
        }
    }
    *cursor = cur;
    return acc / span;
}

// the series as two arrays
struct SeriesArrays {
    const double *t, *v;
    __host__ __device__ double time(int k) const { return t[k]; }
    __host__ __device__ double value(int k) const { return v[k]; }
};

// delta_p_agg [mm] of update_spatial_BC.m:422-448 for the step [t_prev, t].
template <typename Series>
__host__ __device__ inline double rain_depth(const Series &ser, int n, double t_prev, double t,
                                             int *cursor) {
    const double dt_step = fmax(t - t_prev, 0.0);
    double depth = 0.0;
    if (dt_step > 0) {
        double avg = rain_average(ser, n, t_prev, t, cursor);
        if (__builtin_isnan(avg)) avg = 0.0;
        depth = avg * dt_step / 60.0;
    }
    return depth;
}
__host__ __device__ inline double rain_depth(const double *times, const double *values, int n,
                                             double t_prev, double t, int *cursor) {
    return rain_depth(SeriesArrays{times, values}, n, t_prev, t, cursor);
}

// ---------------------------------------------------------------------------
// Rainfall maps (update_spatial_BC.m:469-518 with localAverageSpatialGaugeRainOverStep
// and localAverageInputRainMapOverStep, :1287-1408): the step's depth map is
//   delta_p_agg = (sum_i overlap_i * map_i) / (t1 - t0) * ((t1 - t0) / 60)
// over the maps whose end-stamped intervals overlap (t0, t1].  The terms
// (map index, overlap) are found once per step; the per-cell sum runs in the
// step kernel, in the same order.
// ---------------------------------------------------------------------------
constexpr int kMaxMapTerms = 4;

struct MapTerms {
    int n;                     // maps overlapping the step (0: no rain this step)
    int idx[kMaxMapTerms];     // 0-based map index
    double w[kMaxMapTerms];    // overlap [min]
    double dt_total, k;        // t1 - t0 [min] and dt / 60 [h]
};

// Returns false when more than kMaxMapTerms maps overlap one step.
__host__ __device__ inline bool map_terms(const double *times, int K, double t0, double t1,
                                          int *cursor, MapTerms *mt) {
    mt->n = 0;
    mt->dt_total = 0.0;
    mt->k = 0.0;
    if (K <= 0 || !(t1 > t0)) return true;
    mt->dt_total = t1 - t0;
    mt->k = fmax(t1 - t0, 0.0) / 60.0;
    int cur = *cursor;
    if (cur < 1) cur = 1;
    if (cur > K) cur = K;
    while (cur < K && t0 > times[cur - 1]) ++cur;
    const double spacing = (K >= 2 && times[1] - times[0] > 0) ? times[1] - times[0] : 0.0;
    for (int i = cur; i <= K; ++i) {
        const double start = (i == 1) ? times[0] - spacing : times[i - 2];
        const double end = times[i - 1];
        const double overlap = fmax(0.0, fmin(t1, end) - fmax(t0, start));
        if (overlap > 0) {
            if (mt->n == kMaxMapTerms) return false;
            mt->idx[mt->n] = i - 1;
            mt->w[mt->n] = overlap;
            mt->n += 1;
        }
        if (end >= t1) {
            cur = i;
            break;
        }
    }
    *cursor = cur;
    return true;
}

// ---------------------------------------------------------------------------
// Inflow hydrographs (update_spatial_BC.m:362-378): BC_States.delta_inflow_agg
// of each gauge over (t0, t1] [mm], the gauges' shared cursor carried from one
// to the next as in the reference's loop.  rates: [n_times x n_gauges].
// ---------------------------------------------------------------------------
constexpr int kMaxInflowGauges = 16;

__host__ __device__ inline void inflow_depths(const double *times, const double *rates,
                                              int n_times, int n_gauges, double t0, double t1,
                                              int *cursor, double *agg) {
    for (int g = 0; g < n_gauges; ++g) agg[g] = 0.0;
    const double dt_step = fmax(t1 - t0, 0.0);
    if (!(dt_step > 0)) return;
    for (int g = 0; g < n_gauges; ++g) {
        const double avg = rain_average(SeriesArrays{times, rates + (long long)g * n_times},
                                        n_times, t0, t1, cursor);
        agg[g] = avg * dt_step / 60.0;
    }
}

}  // namespace hp2d
