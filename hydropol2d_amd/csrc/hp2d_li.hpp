// hp2d_li.hpp - launch arguments of the fused local-inertial step kernel.
#pragma once

#include "hp2d_internal.hpp"

namespace hp2d {

enum { kModeDropIn = 0, kModeSim = 1 };

template <typename T> struct LiArgs {
    int ny, nx;
    int tj;               // columns per wave task
    int rows;             // rows per lane (1, or 2 for an even ny)
    int n_slabs;          // li_n_slabs(ny, rows)
    int waves_per_block;  // 1..4
    // physics constants
    T R, R2, cell_area, C_a;
    T inv_R, k_mmh_q, k_q_mmh;  // 1/R, R/3.6e6 (mm/h -> m2/s), 3.6e6/A (m3/s -> mm/h)
    int flag_critical, outlet_type, adaptive;
    T sqrtS_type1;        // sqrt(|slope_outlet|)
    T depth_tol_mm, alfa_min;
    T dt_s, dt_h;         // drop-in mode only
    // static fields
    const T *z;           // elevation [m] (FP64 kernels)
    const T *dzx, *dzy;   // FP32 kernels: z(i,j+1) - z(i,j) and z(i-1,j) - z(i,j) [m],
                          // taken in FP64 and rounded once (NaN outside the domain)
    const T *n2s;         // Manning n^2, sign bit set on outlet cells
    // drop-in mode state
    const T *d_src, *qx_src, *qy_src;
    T *d_dst, *qx_dst, *qy_dst;
    const T *I_src, *T_src;
    T *I_dst, *T_dst;
    // simulation mode ping-pong buffers (index = step parity)
    T *d_buf[2], *qx_buf[2], *qy_buf[2], *I_buf[2], *T_buf[2];
    const SimState *st;
    double routing_time;  // decides on the device whether a mass-balance report can follow
    // hydrology (sim mode)
    const T *rain_maps;           // rainfall maps [mm/h], map k at k * ny * nx (or NULL)
    const uint8_t *inflow_gauge;  // the cell's inflow gauge 1..n, 0 = none (or NULL)
    const T *rain_pattern;
    const uint8_t *soil_class;
    const GaClass *ga;    // NULL = no infiltration
    const float *ga_f32;  // the same table as 8 floats per class (FP32 path)
    // van Genuchten infiltration (Infiltration_Module) in place of Green-Ampt:
    // class tables (NULL = off), the depth to the water table [m] per cell
    // (sign bit set on impervious cells), Ltop, dh_max, Ksurf_weight, l_vg
    const VgClass *vg;
    const float *vg_f32;
    const T *vg_zwt;
    T vg_Ltop, vg_dhmax, vg_wK, vg_l;
    int n_soil_classes;
    T I_min;
    int track_dmax;       // reduce the depth maxima every step (a late rollback can follow)
    // Human_Instability_Module mode 1: max_risk map (or NULL) and the person's parameters
    T *max_risk;
    T hi_weight, hi_w1, hi_w2, hi_mu, hi_Cd, hi_ro, hi_g;
    // map stacks: the event step's velocity_raster and total_velocity (or NULL)
    T *snap_vras, *snap_vtot;
    double rec_maps;
    T *dmax_map;
    T *tpeak_map;         // optional: model time [min] at which dmax was last raised
    // outlets (sorted linear indices)
    const long long *out_idx;
    const T *out_n;
    int n_outlet;
    T *out_q, *out_d;
    // drop-in optional outputs
    T *q3, *hfx, *hfy, *hf3, *outlet_flow, *qleft, *qright, *qup, *qdown, *itot;
    // sim partials (structure of arrays, kPartCount fields of part_cap entries)
    double *partials;
    long long part_cap;
};

template <typename T> hipError_t launch_li_step(const LiArgs<T> &a, int mode, hipStream_t s);

// n_tiles tiles of several grids in one launch (simulation mode); all grids
// share waves_per_block and the critical-velocity flag
template <typename T>
hipError_t launch_li_step_batch(const LiArgs<T> *args, const int *tile_grid,
                                const int *grid_first_tile, int n_tiles, int waves_per_block,
                                int critical, hipStream_t s);

}  // namespace hp2d
