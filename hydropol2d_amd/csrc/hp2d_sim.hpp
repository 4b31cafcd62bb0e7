// hp2d_sim.hpp - arguments of the per-step finalize kernel.
#pragma once

#include "hp2d_internal.hpp"

namespace hp2d {

template <typename T> struct FinArgs {
    SimState *st;
    const double *partials;
    long long part_cap;
    int n_partials;
    // running_control / flags
    double routing_time, tsc, inc, alfa_min, alfa_max, R, A, rec, report_pct;
    double factor_reduction, drainage_area, sum_pattern, n_valid;
    int flag_timestep, adaptive, ga, flag_wq, flag_gw, has_pattern, n_rec, vmax_squared;
    int ca;  // CA routing: keep the I_tot_end_cell flag of the step's routing
    // lumped rainfall series
    const double *rain_t, *rain_v;
    int n_rain;
    // rainfall maps: end stamps [min] and each map's sum over the valid cells [mm/h]
    const double *map_t, *map_sum;
    int n_maps;
    // inflow hydrographs: shared end stamps, rates [n_times x n_gauges], valid cells per gauge
    const double *inflow_t, *inflow_v, *gauge_cells;
    int n_inflow_times, n_gauges;
    // map stacks: record spacing [min] (0 = off), device slots, record times
    double rec_maps;
    int snap_slots;
    double *snap_t;
    // outlets
    const T *out_q, *out_d;
    int n_outlet;
    // records
    double *hq, *ht, *hd;
    int n_hydro_max;
    double *dt_hist;
    int n_dt_max;
    double *vol_err_hist;
    int n_rep_max;
    int dry_run;           // 1: reduce the partials only (autotuner timing)
    double *partials_out;  // sink of the dry-run reduction
};

template <typename T> hipError_t launch_finalize(const FinArgs<T> &a, hipStream_t s);
// one block per grid of an ensemble; block b finalises grid ids[b]
template <typename T>
hipError_t launch_finalize_batch(const FinArgs<T> *args, const int *ids, int n_grids, hipStream_t s);
template <typename T>
hipError_t launch_storage(const T *d, const T *I, long long n, double *part, int blocks,
                          hipStream_t s);
// The depth to the water table of the van Genuchten closure after a change of
// the head h [m]: zw = max(sd - (h - (elev - sd)), 0), keeping the sign bit of
// the old zw (the impervious flag); sd from soil_depth, or sd_u when NULL
template <typename T>
hipError_t launch_vg_water_table(const double *h, const double *elev, const double *soil_depth,
                                 double sd_u, long long n, T *zw, hipStream_t s);

}  // namespace hp2d
