// hp2d_maps.hpp - map-stack snapshots of the device loop (save_output_maps.m:148-258).
#pragma once

#include "hp2d.h"
#include "hp2d_internal.hpp"

namespace hp2d {

template <typename T> struct SnapArgs {
    const SimState *st;
    long long n;
    int ny;
    int slots;                       // records the stack holds between host drains
    T *stack[HP2D_MAP_FIELDS];       // field f, record r at stack[f] + r * n (or NULL)
    const T *d[2], *I[2];            // ping-pong state (current = (k + 1) & 1)
    const T *vras, *vtot;            // velocity_raster and total_velocity of the event step
    const T *n2s;                    // NaN outside the domain (FP32), outlet flag in the sign
    const double *elev;              // FP64 elevation (FP64 runs; NaN outside the domain)
    // Human_Instability_Module mode 1
    int hi;
    double hi_weight, hi_w1, hi_w2, hi_mu, hi_Cd, hi_ro, hi_g;
    // groundwater depth h - (z - soil depth), FP64 module state
    const double *gw_h, *gw_elev, *gw_sd;
    double gw_sd_u;
    // water quality: B_t ping-pong (current = wq_n & 1), P_conc, the outflow pages
    const T *wqB[2], *wqPc;
    const T *qx[2], *qy[2];
    const long long *out_idx;
    const T *out_q;
    int n_outlet;
    double R;
};

template <typename T> hipError_t launch_snapshot(const SnapArgs<T> &a, hipStream_t s);

}  // namespace hp2d
