/*
 * hp2d_oracle.h - FP64 CPU restatement of the HydroPol2D step (TEST INFRASTRUCTURE).
 *
 * This is the parity oracle.  It is compiled into oracle/_build/libhp2d_oracle.so
 * and is only loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, as the checker.  It is never linked into the product
 * library (hydropol2d_amd/libhp2d.so).
 *
 * Every function restates a MATLAB routine of the reference, statement by
 * statement, on full-grid column-major arrays (element (i,j) at i + j*ny,
 * 0-based), with MATLAB's NaN-omitting max/min (C fmax/fmin) and the same
 * operation order.  The file:line each function follows is given at its
 * definition in hp2d_oracle.c.
 */
#ifndef HP2D_ORACLE_H
#define HP2D_ORACLE_H

#include <stdint.h>
#include "../include/hp2d.h"

/* Element-wise loops (no accumulation across cells) may run on several
 * threads when the oracle is built with OpenMP (the all-core CPU baseline);
 * results are bit-identical either way, since no summation order changes. */
#ifdef _OPENMP
#define ORACLE_PAR _Pragma("omp parallel for schedule(static)")
#else
#define ORACLE_PAR
#endif

#ifdef __cplusplus
extern "C" {
#endif

void oracle_inertial_solver(int32_t scheme, int64_t n_cells, int32_t n_pages,
                            const double *q_p, double dt, const double *Hf,
                            const double *S, const double *n_sq, double *q);

int32_t oracle_li_d4(int32_t ny, int32_t nx, int32_t scheme, const double *z,
                     const double *d_tot, const double *roughness_cell,
                     const double *roughness_squared, double cell_area,
                     double time_step_min, double resolution, int32_t outlet_type,
                     double slope_outlet, const int64_t *outlet_idx, int32_t n_outlet,
                     const double *outflow_prev, int32_t flag_critical,
                     double *qout_left, double *qout_right, double *qout_up,
                     double *qout_down, double *outlet_flow, double *d_t,
                     double *I_tot_end_cell, double *outflow, double *Hf);

typedef struct oracle_dt_params {
    double resolution, alfa_min, alfa_max, max_time_step, min_time_step;
    double time_step_increments, depth_tolerance_mm;
    int32_t flag_timestep, flag_adaptive_timestepping, flag_waterquality;
    double tmin_wq;
    double nt_raw;  /* out: the new time step before the snap to the increments */
    double *total_velocity; /* optional out: velocities.total_velocity per cell [m/s] */
} oracle_dt_params;

/* returns time_calculation_routing [s]; writes max_velocity and optionally
 * velocity_raster (ny*nx). */
double oracle_refresh_dt(int32_t ny, int32_t nx, const double *qout_right,
                         const double *qout_up, const double *Hf,
                         oracle_dt_params *p, double *max_velocity,
                         double *velocity_raster);

/* the same with flags.flag_CA = 1: q4 = qout_left, right, up, down; d_t the
 * depth after the gather */
double oracle_refresh_dt_ca(int32_t ny, int32_t nx, const double *q4, const double *d_t,
                            oracle_dt_params *p, double *max_velocity, double *velocity_raster);

double oracle_rain_avg(const double *times, const double *values, int32_t n, double t0,
                       double t1, int32_t *cursor);

/* Complete Main_While run (FP64); arrays in/out as in hp2d_run_inputs.
 * Outputs are written to `out` (caller allocates the pointed arrays). */
typedef struct oracle_run_out {
    double *d_t;            /* ny*nx final depth [mm] */
    double *outflow;        /* ny*nx*3 */
    double *I_t;            /* ny*nx (GA) or NULL */
    double *dmax;           /* ny*nx */
    double *B_t;            /* ny*nx or NULL */
    double *gw_h;           /* ny*nx or NULL */
    double *hydro_q;        /* n_hydro_max */
    double *hydro_t;
    double *hydro_depth;
    int32_t n_hydro_max, n_hydro;
    double *dt_hist;        /* n_dt_max accepted-step dt [s] */
    int32_t n_dt_max;
    int64_t max_steps;      /* stop after this many steps (<=0: run to the end) */
    int64_t inject_fail_step; /* reject step k == this once (<=0: never) */
    hp2d_run_report rep;
    double *volume_error_hist; /* n_report_max */
    int32_t n_report_max;
    double *P_conc, *Tot_Washed;  /* ny*nx or NULL (water quality) */
    /* Optional accepted-step dt sequence of another implementation.  Where the
     * oracle's own pre-snap dt lies within 1e-9 (relative) of an increment
     * boundary and the given dt is one increment away, the floor() of
     * refreshing_timestep.m:344 is a floating-point tie and the given dt is
     * adopted (n_dt_ties counts them); any other difference is counted in
     * n_dt_mismatch and the oracle keeps its own dt. */
    const double *dt_follow;
    int64_t n_follow, n_dt_ties, n_dt_mismatch;
    /* Stop right after the first catch block (the rollback of
     * HydroPol2D_Main_While.m:926-993) has run, so that the restored state and
     * the new dt can be read; rollback_max_dt_s is the max_dt it leaves. */
    int32_t stop_after_rollback;
    double rollback_dt_s, rollback_max_dt_s;
    /* With dt_force set, the oracle takes every refreshed dt from dt_follow
     * (n_dt_mismatch still counts the steps where its own differed): the FP32
     * device path is then compared with it on the same time levels. */
    int32_t dt_force;
    /* Human_Instability_Module max_risk (ny*nx, optional) and the map-stack
     * records (save_output_maps.m:148-258): n_rec_max records of
     * HP2D_MAP_FIELDS maps each, record r field f at rec + (r * F + f) * ny*nx
     * (optional), with the model time of each record */
    double *max_risk;
    double *rec, *rec_t;
    int32_t n_rec_max, n_rec;
} oracle_run_out;

/* Green-Ampt derived parameters of one soil class (HydroPol2D_preprocessing.m:
 * 2249-2253): out6 = {ksat [mm/h], psi [mm], dtheta, Lu [mm], k_out [mm/h], Tr [h]};
 * lu_m NaN derives Lu from ksat. */
void oracle_ga_params(double ksat_mmh, double psi_mm, double dtheta, double theta_sat_minus_r,
                      double lu_m, double *out6);

/* The legacy explicit Green-Ampt closure on n cells of one class, the same
 * function the loop applies (HydroPol2D_Main_While_Ensemble.m:243-298). */
void oracle_ga_cells(int64_t n, int32_t first_step, double dt_h, double I_min_mm,
                     const double *params6, double *d, const double *I_p, double *I_t,
                     double *T, double *f_out);

int32_t oracle_run(const hp2d_run_config *cfg, const hp2d_run_inputs *in,
                   oracle_run_out *out);

int32_t oracle_ca_routing(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                          const double *roughness, double cell_area, double time_step_min,
                          const double *h_0, double resolution,
                          const double *I_tot_end_cell_prev, int32_t outlet_type,
                          double slope_outlet, const int64_t *outlet_idx,
                          int32_t n_outlet, int32_t flag_critical, double *qout_left,
                          double *qout_right, double *qout_up, double *qout_down,
                          double *outlet_flow, double *d_t, double *I_tot_end_cell);

/* returns 1 when the step was split internally, 0 when not, -1 when the split
 * would exceed 1e7 sub-steps (tmin_wq = 0 loops for ever in the reference) */
int32_t oracle_washoff(int32_t ny, int32_t nx, const double *C3, const double *C4,
                       const double *ql, const double *qr, const double *qu, const double *qd,
                       const double *qo, double *B_t, double time_step_min, double cell_area,
                       const uint8_t *nan_mask, int32_t model, double *mass_lost,
                       double *Tot_Washed, double Bmin, double Bmax, double min_Bt,
                       double *P_conc, double *Out_Conc, double *tmin_wq, double *tot_W_out);

/* D8 routines (hp2d_d8.c): 9-page arrays in the reference page order
 * left, right, up, down, outlet, NE, SE, SW, NW; outflow pages in mm/h. */
int32_t oracle_bates_8d(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                        const double *d_p, const double *rough, double cell_area,
                        double time_step_min, double R, int32_t outlet_type,
                        double slope_outlet, const int64_t *outlet_idx, int32_t n_outlet,
                        const double *outflow_prev, int32_t flag_critical, double *outflow,
                        double *d_t, double *I_tot_end_cell, double *Hf_out);
int32_t oracle_ca_8d(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                     const double *rough, double cell_area, double time_step_min,
                     const double *h_0, double R, const double *I_prev_in, int32_t outlet_type,
                     double slope_outlet, const int64_t *outlet_idx, int32_t n_outlet,
                     int32_t flag_critical, double d_tolerance, double *outflow, double *d_t,
                     double *I_tot_end_cell);
/* outflows become inflows (n_pages 5: four directions, 9: eight); d_t, when
 * not NULL, gets qin * time_step / 60 added */
void oracle_gather(int32_t ny, int32_t nx, const double *outflow, int32_t n_pages,
                   double time_step_min, double *qin, double *d_t);
/* returns as oracle_washoff */
int32_t oracle_washoff_8d(int32_t ny, int32_t nx, const double *C3, const double *C4,
                          const double *q9, double *B_t, double time_step_min, double cell_area,
                          const uint8_t *nan_mask, int32_t model, double *mass_lost,
                          double *Tot_Washed, double Bmin, double Bmax, double min_Bt,
                          double *P_conc, double *Out_Conc, double *tmin_wq, double *tot_W_out);

/* lookup-table subgrid (hp2d_subgrid.c) */
int32_t oracle_subgrid_build(const double *dem, int32_t fny, int32_t fnx, double cellsize,
                             const double *rough, int32_t rough_is_fine, double coarse_res,
                             int32_t ny, int32_t nx, int32_t nl, hp2d_subgrid_tables *T);
void oracle_subgrid_queries(const hp2d_subgrid_tables *T, int32_t ny, int32_t nx, const double *zs,
                            double *V, double *zs_back, double *A);
void oracle_subgrid_face_state(const hp2d_subgrid_faces *F, int64_t nf, int32_t nl,
                               const double *zu, double *HG, double *nr, double *phi);
void oracle_subgrid_volume(const hp2d_subgrid_tables *T, int32_t ny, int32_t nx,
                           const double *drep_old, const double *delta_mm, double cell_area,
                           double *drep_new, double *V_old, double *V_new, double *A_new,
                           double *residual);
int32_t oracle_li_subgrid(int32_t ny, int32_t nx, const hp2d_subgrid_tables *T,
                          const double *d_tot, double time_step_min, double R,
                          int32_t outlet_type, double slope_outlet, const int64_t *outlet_idx,
                          const int32_t *outlet_side, int32_t n_outlet,
                          const double *outflow_prev, int32_t n_prev_pages,
                          const uint8_t *idx_nan, double *outflow, double *d_t,
                          double *I_tot_end_cell, double *Hf_out, double *C_a_out,
                          double *eta_out, double *V_out, int32_t *n_sub);

/* Full_Momentum_Model_D4 (hp2d_fm.c): face_masks NULL or {right, left, up,
 * down} (entries may be NULL); outflow 5 pages, Hf 3 pages */
int32_t oracle_full_momentum(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                             const double *rough, const double *n2_in, const double *cell_area,
                             double time_step_min, double R, const uint8_t *outlet_index,
                             int32_t outlet_type, double slope_outlet, const int64_t *outlet_idx,
                             int32_t n_outlet, double d_tolerance, const double *outflow_prev,
                             int32_t n_prev_pages, const uint8_t *idx_nan, int32_t flag_HR,
                             const uint8_t *const *face_masks, double *outflow, double *d_t,
                             double *I_tot, double *Hf, double *C_a_out);

int32_t oracle_boussinesq_2d(int32_t ny, int32_t nx, double dt, double dx, double dy,
                             const double *h, const double *z0, const double *Sy, const double *R,
                             const double *K, double courant, const double *h_soil,
                             const uint8_t *catchment, const uint8_t *dirichlet,
                             const double *h_dirichlet, double *h_t, double *u_x, double *u_y,
                             double *q_exf, double *error, int32_t *n_sub);

int32_t oracle_gw_recharge(int64_t n, const double *infiltration_rate, const double *S0,
                           double dt, const double *min_S, const double *max_S,
                           const double *theta_s, const double *theta_r, const double *n_vg,
                           const double *alpha_vg, double l_vg, const double *ksat_mmh,
                           double *recharge_rate, double *S_out, double *cumulative_recharge);

void oracle_conv3x3_fwd(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                        const float *x, const float *w, const float *b, int32_t relu, float *y);
void oracle_conv3x3_bwd(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                        const float *x, const float *w, const float *gy, float *gx, float *gw,
                        float *gb);

double oracle_vg_cell(double dt_h, double A, double C_a, double zwt_in, double ts, double tr,
                      double nv, double a, double ksat_mmh, int imp, double l_vg, double Ltop,
                      double dh_max, double wK, double *d, double *I);
int32_t oracle_infiltration(int32_t ny, int32_t nx, double time_step_min, double resolution,
                            double C_a, const double *elevation, const double *h_t,
                            const double *soil_depth, const double *theta_sat,
                            const double *theta_r, const double *n_vg, const double *alpha_vg,
                            const double *ksat_mmh, const uint8_t *idx_imp, double l_vg,
                            double Ltop_in, double dh_max, double Ksurf_weight, double *d_t,
                            double *I_t, double *I_p, double *f_out, double *cum_inf,
                            double *inf_volume, double *dS_inf);

/* Groundwater_Module step (no layered soil) */
typedef struct oracle_gwm_params {
    int32_t ny, nx;
    double time_step_min, resolution, t_min, routing_time_min;
    int32_t flag_baseflow, flag_async;
    double target_dt_min, min_dt_min, max_head_change_m, courant, l_vg;
    const double *elevation, *soil_depth, *theta_sat, *theta_r, *n_vg, *alpha_vg;
    const double *ksat_mmh, *ksat_gw_mmh, *Sy;
    const uint8_t *domain;   /* catchment mask for the Boussinesq update, NULL = all */
} oracle_gwm_params;

typedef struct oracle_gwm_state {
    double *pending_net, *pending_recharge, *pending_capillary;  /* ny*nx [m] */
    double elapsed_s, total_recharge_m3, total_capillary_m3, total_exfiltration_m3, last_dt_s;
    int32_t n_updates;
} oracle_gwm_state;

int32_t oracle_groundwater_module(const oracle_gwm_params *p, oracle_gwm_state *g, double *h_t,
                                  double *I_t, double *d_t, double *cumulative_recharge,
                                  double *max_gw_depth, double *q_exf, double *error,
                                  int32_t *ran, double *dt_gw_out);

/* Rainfall_Interpolator / idw_function (inverse-distance weighting) */
void oracle_idw(int32_t n_obs, const double *x_obs, const double *y_obs, const double *f_obs,
                int32_t nx_grid, const double *x_grid, int32_t ny_grid, const double *y_grid,
                double p, double rad, double *out);

#ifdef __cplusplus
}
#endif

#endif
