/*
 * hp2d.h - C ABI of the MI355X (gfx950) explicit 2D surface-water step.
 *
 * The library replaces the per-step physics of HydroPol2D
 * (HydroPol2D_Functions/HydroPol2D_Main_While.m:403-994) behind the
 * reference's own function boundaries:
 *
 *   Inertial_Solver            HydroPol2D_Functions/Inertial_Solver.m:1-26
 *   Local_Inertial_Model_D4    HydroPol2D_Functions/Local_Inertial_Model_D4.m:1-516
 *   CA_Routing                 HydroPol2D_Functions/CA_Routing.m:1-171
 *   build_up_wash_off          HydroPol2D_Functions/build_up_wash_off.m:1-183
 *   Bates_Inertial_8D          HydroPol2D_Functions/Bates_Inertial_8D.m:1-274
 *   CA_Routing_8D              HydroPol2D_Functions/CA_Routing_8D.m:1-188
 *   build_up_wash_off_8D       HydroPol2D_Functions/build_up_wash_off_8D.m:1-143
 *   Local_Inertial_Model_D4_Subgrid  HydroPol2D_Functions/Local_Inertial_Model_D4_Subgrid.m:1-317
 *   Full_Momentum_Model_D4     HydroPol2D_Functions/Full_Momentum_Model_D4.m:1-724
 *   Subgrid_Properties_Lookup  HydroPol2D_Functions/Subgrid_Properties_Lookup.m:1-510
 *   Boussinesq_2D_explicit     HydroPol2D_Functions/Boussinesq_2D_explicit.m:1-195
 *   refreshing_timestep        HydroPol2D_Functions/refreshing_timestep.m:17-368
 *   the Main_While step loop   HydroPol2D_Functions/HydroPol2D_Main_While.m:403-994
 *
 * Conventions (identical to the MATLAB state layout):
 *   - every grid is column-major ny x nx: element (i,j) (1-based in MATLAB)
 *     lives at offset (i-1) + (j-1)*ny; 3-page arrays (outflow, Hf) stack the
 *     pages: page p at offset p*ny*nx;
 *   - depths in mm, face fluxes (outflow pages) in mm/h of cell area, Hf in m,
 *     time_step in minutes, every *_s field in seconds, model time in minutes;
 *   - NaN elevation marks cells outside the catchment (idx_nan);
 *   - outlet rows/cols are 1-based, as MATLAB's find() returns them.
 *
 * Every function returns hp2d_status (0 = OK). On failure, hp2d_last_error()
 * of the handle (or hp2d_global_last_error() when no handle exists) gives the
 * text. A non-OK status from a step means the step was rejected; the caller's
 * state arrays are left untouched, so the MATLAB catch/rollback of
 * HydroPol2D_Main_While.m:926-993 can run unchanged.
 *
 * Threading: one handle is bound to one HIP stream and is not re-entrant.
 * Different handles may be used from different threads.
 */
#ifndef HP2D_H
#define HP2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hp2d_status;
enum {
    HP2D_OK = 0,
    HP2D_E_ARG = 1,          /* bad argument / shape */
    HP2D_E_HIP = 2,          /* HIP runtime error */
    HP2D_E_NOMEM = 3,        /* device allocation failed */
    HP2D_E_STATE = 4,        /* call out of order (e.g. run before terrain) */
    HP2D_E_UNSUPPORTED = 5,  /* flag combination not implemented */
    HP2D_E_INSTABILITY = 6,  /* step rejected (NaN state, WQ instability, injected failure) */
    HP2D_E_NO_DEVICE = 7     /* no HIP device visible */
};

enum { HP2D_FP64 = 0, HP2D_FP32 = 1 };

enum { HP2D_ROUTING_INERTIAL = 1, HP2D_ROUTING_CA = 2 };

/* ------------------------------------------------------------------------ */
/* Library / handle                                                          */
/* ------------------------------------------------------------------------ */

typedef struct hp2d_handle hp2d_handle;

/* Version string of the library build (never NULL). */
const char *hp2d_version(void);

/* Number of visible HIP devices (0 when none). */
int32_t hp2d_device_count(void);

/* Text of the last error raised on a thread without a handle. */
const char *hp2d_global_last_error(void);

/* Create a handle for grids of ny x nx cells computing in `precision`
 * (HP2D_FP64 = MATLAB double, HP2D_FP32 = MATLAB single / flag_single).
 * ny * nx must stay below 2^32 - 2^16 (HP2D_E_ARG otherwise).
 * The handle owns device buffers sized for the grid and one HIP stream.
 * Replaces the per-call allocations of Local_Inertial_Model_D4.m:73-84,145-149. */
hp2d_status hp2d_create(int32_t ny, int32_t nx, int32_t precision, hp2d_handle **out);
void hp2d_destroy(hp2d_handle *h);
const char *hp2d_last_error(const hp2d_handle *h);

/* ------------------------------------------------------------------------ */
/* Drop-in per-call boundaries (host arrays in, host arrays out)             */
/* ------------------------------------------------------------------------ */

/* q = Inertial_Solver(flag_numerical_scheme, q_p, dt, Hf, S, n_sq, dx, idx_nan)
 * Inertial_Solver.m:1-26 (scheme 1, Bates 2010).  q_p, Hf, S, q_out hold
 * n_pages pages of ny*nx; n_sq is one page, broadcast over pages as MATLAB's
 * implicit expansion does.  dt in seconds.  Non-finite results are returned
 * as computed (the caller zeroes them, Local_Inertial_Model_D4.m:264). */
hp2d_status hp2d_inertial_solver(hp2d_handle *h, int32_t flag_numerical_scheme,
                                 const double *q_p, double dt, const double *Hf,
                                 const double *S, const double *n_sq, int32_t n_pages,
                                 double *q_out);

/* Local_Inertial_Model_D4.m:1-5 with flag_subgrid = 0, flag_reservoir = 0
 * (the arguments that are unused on that branch are omitted).
 * Inputs:  z, d_tot [mm], roughness_cell, roughness_squared: ny*nx;
 *          outflow_prev: ny*nx*3 [mm/h] (the `outflow` argument);
 *          row_outlet/col_outlet: n_outlet 1-based indices.
 * Outputs (caller-allocated; any may be NULL to skip it):
 *          qout_left, qout_right, qout_up, qout_down, outlet_flow [mm/h],
 *          d_t [mm], I_tot_end_cell [m3]: ny*nx;
 *          outflow [mm/h], Hf [m]: ny*nx*3.
 * d_p, d_tolerance, idx_nan and the subgrid/reservoir arguments do not enter
 * the computation on this branch (Local_Inertial_Model_D4.m:9,13).
 * Returns the scalar outputs Qc = Qf = Qci = Qfi = 0 and C_a = resolution^2
 * implicitly (Local_Inertial_Model_D4.m:257-266). */
hp2d_status hp2d_local_inertial_d4(hp2d_handle *h, int32_t flag_numerical_scheme,
                                   const double *z, const double *d_tot,
                                   const double *roughness_cell,
                                   const double *roughness_squared, double cell_area,
                                   double time_step_min, double resolution,
                                   int32_t outlet_type, double slope_outlet,
                                   const int32_t *row_outlet, const int32_t *col_outlet,
                                   int32_t n_outlet, const double *outflow_prev,
                                   int32_t flag_critical,
                                   double *qout_left, double *qout_right, double *qout_up,
                                   double *qout_down, double *outlet_flow, double *d_t,
                                   double *I_tot_end_cell, double *outflow, double *Hf);

/* Same boundary on single-precision arrays (MATLAB flag_single = 1,
 * HydroPol2D_preprocessing.m:3287-3569).  Handle must be HP2D_FP32. */
hp2d_status hp2d_local_inertial_d4_f32(hp2d_handle *h, int32_t flag_numerical_scheme,
                                       const float *z, const float *d_tot,
                                       const float *roughness_cell,
                                       const float *roughness_squared, float cell_area,
                                       float time_step_min, float resolution,
                                       int32_t outlet_type, float slope_outlet,
                                       const int32_t *row_outlet, const int32_t *col_outlet,
                                       int32_t n_outlet, const float *outflow_prev,
                                       int32_t flag_critical,
                                       float *qout_left, float *qout_right, float *qout_up,
                                       float *qout_down, float *outlet_flow, float *d_t,
                                       float *I_tot_end_cell, float *outflow, float *Hf);

/* [qout_left,qout_right,qout_up,qout_down,outlet_flow,d_t,I_tot_end_cell] =
 *   CA_Routing(..., z, d_tot, roughness, cell_area, time_step, h_0, Resolution,
 *              I_tot_end_cell, outlet_index, outlet_type, slope_outlet,
 *              row_outlet, col_outlet, idx_nan, flag_critical)
 * CA_Routing.m:1-171 (flag_reservoir = 0).  I_tot_end_cell_prev in, the new
 * I_tot_end_cell out (may alias).  h_0 [mm] may be NULL (zeros). */
hp2d_status hp2d_ca_routing(hp2d_handle *h, const double *z, const double *d_tot,
                            const double *roughness, double cell_area, double time_step_min,
                            const double *h_0, double resolution,
                            const double *I_tot_end_cell_prev, int32_t outlet_type,
                            double slope_outlet, const int32_t *row_outlet,
                            const int32_t *col_outlet, int32_t n_outlet,
                            int32_t flag_critical,
                            double *qout_left, double *qout_right, double *qout_up,
                            double *qout_down, double *outlet_flow, double *d_t,
                            double *I_tot_end_cell);

/* [B_t,P_conc,Out_Conc,tmin_wq,tot_W_out,mass_lost,Tot_Washed] =
 *   build_up_wash_off(C_3,C_4,qout_left,qout_right,qout_up,qout_down,
 *                     outlet_flow,B_t,time_step,nx,ny,cell_area,outlet_index,
 *                     idx_nan_5,flag_wq_model,mass_lost,Tot_Washed,Bmin,Bmax,min_Bt)
 * replaces build_up_wash_off.m:1-183 (called at HydroPol2D_Main_While.m:685).
 * Host arrays, column-major ny*nx doubles: C_3, C_4 (per-cell wash-off
 * coefficient and exponent), the five flow pages [mm/h], B_t [kg],
 * Tot_Washed [kg] (NULL = zeros).  nan_mask = idx_nan_5(:,:,1) as 0/1 bytes
 * (NULL = none); outlet_index is not read by the routine and is not taken.
 * time_step_min [min], cell_area [m2], Bmin/Bmax [g/m2] (flag_wq_model = 1
 * only), min_Bt [g/m2].  Outputs ny*nx: B_t_out [kg], P_conc [mg/L],
 * tot_W_out [kg/h] (may be NULL), Tot_Washed_out [kg]; scalars Out_Conc
 * [mg/L], tmin_wq [s]; *mass_lost [kg] is read and updated.  More than 1e7
 * internal sub-steps is refused with HP2D_E_HIP. */
hp2d_status hp2d_build_up_wash_off(hp2d_handle *h, const double *C_3, const double *C_4,
                                   const double *qout_left, const double *qout_right,
                                   const double *qout_up, const double *qout_down,
                                   const double *outlet_flow, const double *B_t,
                                   double time_step_min, double cell_area,
                                   const uint8_t *nan_mask, int32_t flag_wq_model,
                                   const double *Tot_Washed, double Bmin, double Bmax,
                                   double min_Bt,
                                   double *B_t_out, double *P_conc, double *Out_Conc,
                                   double *tmin_wq, double *tot_W_out, double *mass_lost,
                                   double *Tot_Washed_out);

/* ------------------------------------------------------------------------ */
/* Eight-direction (D8) routines (HydroPol2D_Main_While.m:444-460, 684-688,   */
/* flags.flag_D8 = 1).  The nine flow pages are one ny*nx*9 array in the      */
/* reference's page order: qout_left, qout_right, qout_up, qout_down,         */
/* outlet_flow, qout_ne, qout_se, qout_sw, qout_nw [mm/h] (page p at offset   */
/* p*ny*nx).  Host arrays, column-major doubles, computed in FP64 on a handle */
/* of the grid's shape.                                                      */
/* idx_nan is the set of cells with NaN elevation.                           */
/* ------------------------------------------------------------------------ */

/* [qout_left,...,qout_nw,d_t,I_tot_end_cell,outflow,Hf] = Bates_Inertial_8D(
 *     ..., flag_reservoir = 0, z, d_tot, d_p, roughness_cell, cell_area,
 *     time_step, Resolution, outlet_index, outlet_type, slope_outlet,
 *     row_outlet, col_outlet, d_tolerance, outflow, idx_nan, flag_critical)
 * replaces Bates_Inertial_8D.m:1-274 (called at HydroPol2D_Main_While.m:454-456).
 * Inputs ny*nx: z [m], d_tot [mm] (the depth the outflow is taken from),
 * d_p [mm] (the depth of the water surface and face depths), roughness_cell;
 * outflow_prev: the nine pages of the previous call [mm/h].  Outputs (NULL
 * skips): outflow, the nine new pages [mm/h] (the nine qout arrays and the
 * `outflow` output are the same values); d_t [mm], I_tot_end_cell [m3]: ny*nx;
 * Hf [m]: nine pages.  d_tolerance and outlet_index do not enter the result
 * (Bates_Inertial_8D.m:29, 84-98 read the outlet rows/cols). */
hp2d_status hp2d_bates_inertial_8d(hp2d_handle *h, const double *z, const double *d_tot,
                                   const double *d_p, const double *roughness_cell,
                                   double cell_area, double time_step_min, double resolution,
                                   int32_t outlet_type, double slope_outlet,
                                   const int32_t *row_outlet, const int32_t *col_outlet,
                                   int32_t n_outlet, const double *outflow_prev,
                                   int32_t flag_critical, double *outflow, double *d_t,
                                   double *I_tot_end_cell, double *Hf);

/* [qout_left,...,qout_nw,d_t,I_tot_end_cell] = CA_Routing_8D(..., flag_reservoir = 0,
 *     elevation_cell, d_tot, roughness_cell, cell_area, time_step, h_0_cell,
 *     Resolution, I_tot_end_cell, outlet_index, outlet_type, slope_outlet,
 *     row_outlet, col_outlet, idx_nan, flag_critical, d_tolerance)
 * replaces CA_Routing_8D.m:1-188 (called at HydroPol2D_Main_While.m:446-450).
 * I_tot_end_cell_prev in [m3], the new I_tot_end_cell out (may alias);
 * h_0 [mm] may be NULL (zeros); d_tolerance [mm].  Outputs: outflow (nine
 * pages [mm/h]), d_t [mm] (outflow taken, inflow not yet added). */
hp2d_status hp2d_ca_routing_8d(hp2d_handle *h, const double *z, const double *d_tot,
                               const double *roughness, double cell_area, double time_step_min,
                               const double *h_0, double resolution,
                               const double *I_tot_end_cell_prev, int32_t outlet_type,
                               double slope_outlet, const int32_t *row_outlet,
                               const int32_t *col_outlet, int32_t n_outlet, int32_t flag_critical,
                               double d_tolerance, double *outflow, double *d_t,
                               double *I_tot_end_cell);

/* The caller's gather of the routing outflows into inflows and the depth
 * update of the cellular-automata branch, HydroPol2D_Main_While.m:631-660:
 * qin_t from the four (n_pages = 5: left, right, up, down, outlet) or eight
 * (n_pages = 9) neighbour pages of `outflow` [mm/h], summed in the order of
 * :634-645; d_t [mm] (in/out, NULL skips) gets qin_t * time_step / 60.
 * qin_t [mm/h] may be NULL. */
hp2d_status hp2d_inflow_gather(hp2d_handle *h, const double *outflow, int32_t n_pages,
                               double time_step_min, double *qin_t, double *d_t);

/* Device time [ms] of the compute kernels of the last hp2d_bates_inertial_8d,
 * hp2d_ca_routing_8d or hp2d_inflow_gather call on the handle, without its
 * host<->device copies (HIP events on the handle's stream).  No counterpart in
 * the reference: for callers that separate the kernel from the PCIe cost.
 * HP2D_E_STATE before the first such call. */
hp2d_status hp2d_last_kernel_time(hp2d_handle *h, double *ms);

/* [B_t,P_conc,Out_Conc,tmin_wq,tot_W_out,mass_lost,Tot_Washed] =
 *   build_up_wash_off_8D(C_3,C_4,qout_left_t,qout_right_t,qout_up_t,qout_down_t,
 *     outlet_flow,qout_ne_t,qout_se_t,qout_sw_t,qout_nw_t,B_t,time_step,nx,ny,
 *     cell_area,outlet_index,idx_nan_5,flag_wq_model,mass_lost,Tot_Washed,
 *     Bmin,Bmax,min_Bt)
 * replaces build_up_wash_off_8D.m:1-143 (called at HydroPol2D_Main_While.m:687).
 * As hp2d_build_up_wash_off, with the nine flow pages in `outflow`; B_t_out
 * may alias B_t and Tot_Washed_out may alias Tot_Washed. */
hp2d_status hp2d_build_up_wash_off_8d(hp2d_handle *h, const double *C_3, const double *C_4,
                                      const double *outflow, const double *B_t,
                                      double time_step_min, double cell_area,
                                      const uint8_t *nan_mask, int32_t flag_wq_model,
                                      const double *Tot_Washed, double Bmin, double Bmax,
                                      double min_Bt, double *B_t_out, double *P_conc,
                                      double *Out_Conc, double *tmin_wq, double *tot_W_out,
                                      double *mass_lost, double *Tot_Washed_out);

/* [qout_left,qout_right,qout_up,qout_down,outlet_flow,d_t,I_tot_end_cell,outflow,
 *  Hf,Qc,Qf,Qci,Qfi,C_a] = Full_Momentum_Model_D4(..., flag_reservoir = 0, z, d_tot, d_p,
 *  roughness_cell, roughness_squared, cell_area, time_step, Resolution, outlet_index,
 *  outlet_type, slope_outlet, row_outlet, col_outlet, d_tolerance, outflow, idx_nan, ...,
 *  flag_subgrid = 0, ..., Outlet_Properties)
 * replaces Full_Momentum_Model_D4.m:1-724 on the ordinary grid (the lookup-subgrid
 * storage branch and reservoirs are not taken).  Inputs ny*nx: z [m], d_tot [mm],
 * roughness_cell; roughness_squared (NULL: roughness_cell^2); cell_area_grid (NULL:
 * the scalar cell_area everywhere); outlet_index, idx_nan: 0/1 bytes or NULL;
 * outflow_prev: NULL or n_prev_pages pages [mm/h, and m2/s in pages 4-5]: five or
 * more pages carry the momentum memory (pages 4, 5), fewer the face fluxes of pages
 * 1-2 (a MATLAB 2-D matrix is 0 pages); slope_outlet: scalar.  Outlet_Properties:
 * flag_HR_full_momentum (hydrostatic reconstruction) and the face masks
 * face_right, face_left, face_up, face_down (0/1 bytes or NULL); when a face mask
 * marks an active cell the outlets are face-based, else the row/col cells.
 * Outputs (NULL skips): outflow, 5 pages (x face, y face, outlet [mm/h]; hu, hv
 * [m2/s]); d_t [mm]; I_tot_end_cell [m3]; Hf [m], 3 pages; C_a [m2].  The qout
 * pages are the reference's views of outflow (qout_right = page 1, qout_up =
 * page 2, qout_left / qout_down their negated shifts, outlet_flow = page 3), and
 * Qc = Qf = Qci = Qfi = 0. */
hp2d_status hp2d_full_momentum_d4(hp2d_handle *h, const double *z, const double *d_tot,
                                  const double *roughness_cell, const double *roughness_squared,
                                  double cell_area, const double *cell_area_grid,
                                  double time_step_min, double resolution,
                                  const uint8_t *outlet_index, int32_t outlet_type,
                                  double slope_outlet, const int32_t *row_outlet,
                                  const int32_t *col_outlet, int32_t n_outlet, double d_tolerance,
                                  const double *outflow_prev, int32_t n_prev_pages,
                                  const uint8_t *idx_nan, int32_t flag_HR_full_momentum,
                                  const uint8_t *face_right, const uint8_t *face_left,
                                  const uint8_t *face_up, const uint8_t *face_down,
                                  double *outflow, double *d_t, double *I_tot_end_cell,
                                  double *Hf, double *C_a);

/* ------------------------------------------------------------------------ */
/* Lookup-table subgrid local inertial (flags.flag_subgrid = 1,              */
/* HydroPol2D_Main_While.m:564-582) and its SFINCS-style tables.             */
/* ------------------------------------------------------------------------ */

/* One set of velocity-point tables (Subgrid_Properties_Lookup.m:296-345):
 * zmin, zmax, navg, ffit: one value per face; havg, nrep, pwet: nr_levels
 * pages of one value per face (level k at offset k * n_faces). */
typedef struct hp2d_subgrid_faces {
    double *zmin, *zmax, *havg, *nrep, *pwet, *navg, *ffit;
} hp2d_subgrid_faces;

/* SubgridTables (Subgrid_Properties_Lookup.m:56-100) of a coarse ny x nx grid,
 * column-major as MATLAB holds the struct's fields:
 *   z_zmin, z_zmax, z_volmax          ny*nx
 *   z_level, z_volume, z_area         ny*nx*nr_levels
 *   u: faces between columns j, j+1   ny*max(nx-1,1)   [*nr_levels]
 *   v: faces between rows i, i+1      max(ny-1,1)*nx   [*nr_levels]
 *   side[0..3]: north, south, west, east boundary faces of each cell
 *                                     ny*nx            [*nr_levels]
 * cell_area = coarse_res^2.  The caller owns the arrays. */
typedef struct hp2d_subgrid_tables {
    int32_t nr_levels;
    double cell_area;
    double *z_zmin, *z_zmax, *z_volmax, *z_level, *z_volume, *z_area;
    hp2d_subgrid_faces u, v, side[4];
} hp2d_subgrid_tables;

/* [SubgridTables, invert_el] = Subgrid_Properties_Lookup(DEM_raster,
 *     Roughness_raster, Reference_raster, coarse_res, 'nr_levels', nr_levels)
 * replaces Subgrid_Properties_Lookup.m:1-510 (called at
 * HydroPol2D_preprocessing.m:3174, 3217).  dem: the fine DEM, fny x fnx,
 * cell size `cellsize` [m]; roughness: Manning n on the fine grid
 * (rough_is_fine = 1) or on the coarse grid (0).  The coarse grid is the
 * handle's ny x nx at coarse_res [m]; coarse_res / cellsize must round to an
 * even integer (HP2D_E_HIP with a message otherwise, where the reference
 * errors).  Built on the GPU; the tables stay resident on the handle for
 * hp2d_local_inertial_d4_subgrid, and are copied into `out` when it is not
 * NULL (every array allocated by the caller with the sizes above;
 * invert_el = out->z_zmin).  'dz', 'maxDepth', 'tol_rel' and 'verbose' do
 * not change the tables and are not taken. */
hp2d_status hp2d_subgrid_tables_build(hp2d_handle *h, const double *dem, int32_t fny,
                                      int32_t fnx, double cellsize, const double *roughness,
                                      int32_t rough_is_fine, double coarse_res,
                                      int32_t nr_levels, hp2d_subgrid_tables *out);

/* Upload SubgridTables (e.g. read back from a MAT-file) to the handle, where
 * they stay resident for hp2d_local_inertial_d4_subgrid and
 * hp2d_subgrid_apply_volume_change.  No counterpart in the reference (whose
 * tables live in the MATLAB workspace): the device copy of the struct. */
hp2d_status hp2d_subgrid_set_tables(hp2d_handle *h, const hp2d_subgrid_tables *tables);

/* [qout_left,qout_right,qout_up,qout_down,outlet_flow,d_t,I_tot_end_cell,outflow,
 *  Hf,Qc,Qf,Qci,Qfi,C_a,eta_t,V_t] = Local_Inertial_Model_D4_Subgrid(...,
 *  flag_reservoir = 0, z, d_tot, ..., time_step, Resolution, outlet_index,
 *  outlet_type, slope_outlet, row_outlet, col_outlet, ..., outflow, idx_nan, ...,
 *  SubgridTables)
 * replaces Local_Inertial_Model_D4_Subgrid.m:1-317 (called at
 * HydroPol2D_Main_While.m:566-582) on the handle's resident tables.
 * d_tot [mm]: depth above z_zmin; outflow_prev: NULL or n_prev_pages pages of
 * ny*nx [mm/h] (pages 1-2: the x and y face fluxes the previous call
 * returned); idx_nan: 0/1 bytes or NULL; slope_outlet: scalar.  Outputs
 * (NULL skips): outflow [mm/h] 3 pages (x faces, y faces, outlet), d_t [mm],
 * I_tot_end_cell [m3], Hf [m] 3 pages, C_a [m2], eta_t [m], V_t [m3];
 * n_substeps: the accepted internal sub-steps.  The qout pages are the
 * reference's views of `outflow` (qout_right = page 1, qout_up = page 2,
 * qout_left / qout_down their negated shifts, outlet_flow = page 3), and
 * Qc = Qf = Qci = Qfi = 0. */
hp2d_status hp2d_local_inertial_d4_subgrid(hp2d_handle *h, const double *d_tot,
                                           double time_step_min, double resolution,
                                           int32_t outlet_type, double slope_outlet,
                                           const int32_t *row_outlet, const int32_t *col_outlet,
                                           int32_t n_outlet, const double *outflow_prev,
                                           int32_t n_prev_pages, const uint8_t *idx_nan,
                                           double *outflow, double *d_t, double *I_tot_end_cell,
                                           double *Hf, double *C_a, double *eta_t, double *V_t,
                                           int32_t *n_substeps);

/* [drep_new,V_old,V_new,A_new,residual_m3] = hp2d_subgrid_apply_volume_change(
 *     drep_old, delta_depth_mm, SubgridTables, cell_area)
 * replaces hp2d_subgrid_apply_volume_change.m (SFINCS-exact branch) on the
 * handle's resident tables: drep [m], delta_depth_mm [mm] per cell, cell_area
 * [m2].  Outputs ny*nx (NULL skips all but drep_new). */
hp2d_status hp2d_subgrid_apply_volume_change(hp2d_handle *h, const double *drep_old,
                                             const double *delta_depth_mm, double cell_area,
                                             double *drep_new, double *V_old, double *V_new,
                                             double *A_new, double *residual_m3);

/* [h_t,u_x,u_y,q_exf,q_river,error] = Boussinesq_2D_explicit(dt,dx,dy,h,z0,Sy,R,K,
 *     river_mask,K_river,h_river,z_river,Courant,h_soil,catchment_mask,
 *     dirichlet_mask,h_dirichlet,perimeter)
 * replaces Boussinesq_2D_explicit.m:1-195 (called at Groundwater_Module.m:229-247 and by
 * Validation/Groundwater/Hillslope_Analytical/run_hillslope_analytical_validation.m:81,159).
 * Host arrays, column-major ny*nx doubles (ny, nx >= 2): h [m], z0 (aquifer base) [m],
 * Sy [-], R [m/s], K [m/s], h_soil [m] (surface = z0 + h_soil).  dt [s] is the span
 * integrated with the reference's Courant-limited sub-steps.  catchment_mask /
 * dirichlet_mask: 0/1 bytes or NULL (all inside / none); h_dirichlet required with a
 * Dirichlet mask.  The river arguments and the perimeter are not read by the reference
 * (q_river is zeros) and are not taken.  Outputs ny*nx: h_t [m] (NaN outside the
 * catchment), u_x, u_y [m/s] (may be NULL), q_exf [m/s], q_river (zeros, may be NULL);
 * error [m3] as the reference computes it; n_substeps (may be NULL).  On an FP32
 * handle the sub-steps take the FP32 runs' two-pass form (face fluxes recomputed,
 * reciprocal multiplies); FP64 keeps the reference's operation order. */
hp2d_status hp2d_boussinesq_2d(hp2d_handle *h, double dt, double dx, double dy,
                               const double *h_in, const double *z0, const double *Sy,
                               const double *R, const double *K, double courant,
                               const double *h_soil, const uint8_t *catchment_mask,
                               const uint8_t *dirichlet_mask, const double *h_dirichlet,
                               double *h_t, double *u_x, double *u_y, double *q_exf,
                               double *q_river, double *error, int32_t *n_substeps);

/* Infiltration_Module.m:101-583 (flag_infiltration = 1, no subgrid, no
 * layered soil; called from Hydrological_Model.m): van Genuchten / Mualem
 * surface-interface capacity with a Darcy driving head, limited by the supply
 * d_t/dt and the remaining unsaturated-zone storage.  Host arrays ny*nx:
 * elevation [m], h_t (groundwater head) [m], soil_depth [m], theta_sat,
 * theta_r, n_vg, alpha_vg [1/m], ksat_mmh [mm/h], idx_imp (0/1 bytes or NULL),
 * d_t [mm], I_t [mm] (read-only).  Scalars: time_step [min], resolution [m],
 * C_a [m2] (continuity area), l_vg, Ltop [m], dh_max [m], Ksurf_weight
 * (the reference's defaults are 0.5, 0.05, 1.0, 0.5).  Outputs ny*nx: d_t_out,
 * I_t_out [mm], I_p [mm] (= I_t in, may be NULL), f [mm/h]; cum_inf [m] is
 * read and updated when given; inf_volume [m3] and dS_inf (errors(4)) [m3]. */
hp2d_status hp2d_infiltration_module(hp2d_handle *h, double time_step_min, double resolution,
                                     double C_a, const double *elevation, const double *h_t,
                                     const double *soil_depth, const double *theta_sat,
                                     const double *theta_r, const double *n_vg,
                                     const double *alpha_vg, const double *ksat_mmh,
                                     const uint8_t *idx_imp, double l_vg, double Ltop,
                                     double dh_max, double Ksurf_weight, const double *d_t,
                                     const double *I_t, double *d_t_out, double *I_t_out,
                                     double *I_p, double *f, double *cum_inf, double *inf_volume,
                                     double *dS_inf);

/* Groundwater_Module.m:37-404 (flag_groundwater_modeling = 1, no layered
 * soil, so no capillary rise) with initialize_groundwater_scheduler.m,
 * simulate_groundwater_recharge.m:55-215 and
 * estimate_groundwater_stable_timestep.m: one call per Main_While step.  The
 * unsaturated-zone bucket drains into the pending exchange; when the
 * scheduler's update is due (flag_groundwater_async = 0, or the elapsed time
 * reaches min(target, stable dt), or the pending head change reaches the cap,
 * or the final step) the exchange drives Boussinesq_2D_explicit
 * (flag_baseflow = 1) or a local head change, and the water above the ground
 * and the unsaturated-zone excess return to the surface. */
typedef struct hp2d_gw_params {
    int32_t ny, nx;
    double time_step_min, resolution, t_min, routing_time_min;
    int32_t flag_baseflow, flag_async;
    double target_dt_min, min_dt_min, max_head_change_m, courant, l_vg; /* 1440, 1, 0.25, 0.25, 0.5 */
    const double *elevation, *soil_depth, *theta_sat, *theta_r, *n_vg, *alpha_vg;
    const double *ksat_mmh, *ksat_gw_mmh, *Sy;   /* ny*nx; NaN elevation = outside */
    const uint8_t *domain;                        /* catchment mask (NULL = all) */
} hp2d_gw_params;

/* GW_States: pending exchange maps [m] (ny*nx, NaN outside at first) and scalars */
typedef struct hp2d_gw_state {
    double *pending_net, *pending_recharge, *pending_capillary;
    double elapsed_s, total_recharge_m3, total_capillary_m3, total_exfiltration_m3, last_dt_s;
    int32_t n_updates;
} hp2d_gw_state;

/* Host arrays ny*nx updated in place: h_t [m], I_t [mm], d_t [mm],
 * cumulative_recharge [mm], max_gw_depth [m]; state updated; q_exf [m/s] out;
 * *error = errors(5) [m3]; *ran = 1 when the update ran. */
/* On an FP32 handle the Boussinesq sub-steps take the FP32 runs' two-pass form
 * (face fluxes recomputed, reciprocal multiplies); FP64 keeps the reference's
 * operation order. */
hp2d_status hp2d_groundwater_module(hp2d_handle *h, const hp2d_gw_params *p, hp2d_gw_state *state,
                                    double *h_t, double *I_t, double *d_t,
                                    double *cumulative_recharge, double *max_gw_depth,
                                    double *q_exf, double *error, int32_t *ran);

/* Depression filling of a DEM (host preprocessing; the reference fills with
 * TopoToolbox): Priority-Flood+epsilon (Barnes et al., Computers & Geosciences
 * 62, 2014) seeded at the seed_mask cells (NULL = the grid perimeter and the
 * edge of the NaN mask), D8 neighbours.  z, out: ny*nx [m] (may alias). */
hp2d_status hp2d_fill_sinks(int32_t ny, int32_t nx, const double *z, const uint8_t *seed_mask,
                            double epsilon, double *out);

/* D8 upslope contributing cells of every cell (host preprocessing; the
 * reference derives flow accumulation with TopoToolbox FLOWobj): each valid
 * cell passes its count, itself included, to its steepest-descent neighbour
 * (drop / distance, diagonals at sqrt(2)), cells visited from high to low;
 * cells without a lower neighbour keep theirs.  NaN cells get 0.  out: ny*nx. */
hp2d_status hp2d_flow_accumulation(int32_t ny, int32_t nx, const double *z, double *out);

/* ------------------------------------------------------------------------ */
/* Device-resident step loop (the performance path)                         */
/* ------------------------------------------------------------------------ */

/* Flags and running_control fields of one run (Config/input_data_bypass_script.m
 * :160-170,314-381; input_data_script.m:84-160). */
typedef struct hp2d_run_config {
    int32_t ny, nx;
    int32_t precision;                 /* HP2D_FP64 / HP2D_FP32 (state + fluxes) */
    int32_t routing;                   /* HP2D_ROUTING_INERTIAL, or HP2D_ROUTING_CA: CA_Routing
                                          (HydroPol2D_Main_While.m:461-465) with the inflow gather and
                                          mass balance of :631-660 and the CA time step of
                                          refreshing_timestep.m:31-53; no wash-off, one launch shape,
                                          no ensembles */
    int32_t flag_numerical_scheme;     /* 1 = Bates */
    int32_t flag_critical;
    int32_t outlet_type;               /* 1 normal slope, 2 critical depth */
    double slope_outlet;
    double resolution;                 /* dx [m] */
    double routing_time_min;           /* running_control.routing_time */
    double min_time_step_s, max_time_step_s;
    double time_step_increments_s, time_step_change_s;
    double alfa_min, alfa_max;
    int32_t flag_timestep;             /* 1: min(dt_adv, dt_wave) */
    int32_t flag_adaptive_timestepping;/* 1: dt_wave from celerity only */
    double depth_tolerance_mm;         /* CA_States.depth_tolerance */
    double record_time_hydrographs_min;
    double report_every_percent;       /* forced to 5 by Main_While.m:106 */
    double factor_reduction;           /* rollback dt shrink, default 2 */
    int32_t flag_infiltration;         /* 1 = Green-Ampt (legacy explicit closure);
                                          2 = the van Genuchten Infiltration_Module
                                          (Infiltration_Module.m:182-473, Hydrological_Model.m:24) */
    int32_t flag_waterquality;         /* 1 = build-up / wash-off */
    int32_t flag_wq_model;             /* 0 mass model, 1 rating model */
    int32_t flag_groundwater;          /* 1 = Groundwater_Module (Boussinesq baseflow) every
                                          gw_every_steps accepted steps; needs flag_infiltration
                                          1 or 2 */
    int32_t gw_every_steps;
    double gw_courant;                 /* flags.groundwater_courant */
    int32_t n_soil_classes;            /* GA class tables below */
    int32_t n_lulc_classes;            /* WQ class tables below */
    int32_t flag_inflow;               /* 1 = inflow hydrographs on gauge cells
                                          (update_spatial_BC.m:362-398, Hydrological_Model.m:30) */
    int32_t n_inflow_gauges;           /* 1..HP2D_MAX_INFLOW_GAUGES */
    int32_t n_rain_maps;               /* > 0: rainfall as a series of rate maps instead of the
                                          lumped series (update_spatial_BC.m:469-518) */
    int32_t flag_human_instability;    /* 1 = Human_Instability_Module.m:4-26 (risk, max_risk) */
    double hi_weight_kg, hi_width1_m, hi_width2_m, hi_mu, hi_Cd, hi_ro_water, hi_gravity;
    double record_time_maps_min;       /* > 0: map stacks at running_control.time_records
                                          (save_output_maps.m:24-27, 148-258) */
    int32_t gw_exact;                  /* FP32 runs: 1 = the Boussinesq sub-steps in the
                                          reference's four passes and operation order (as FP64
                                          runs take them); 0 = two passes that recompute the face
                                          fluxes, with reciprocal multiplies */
    /* van Genuchten infiltration (flag_infiltration = 2): Soil_Properties.Ltop [m],
     * dh_max [m], Ksurf_weight [-] (the reference's defaults 0.05, 1.0, 0.5;
     * Infiltration_Module.m:149-163) */
    double vg_Ltop, vg_dh_max, vg_Ksurf_weight;
} hp2d_run_config;

/* Fields of a map-stack record (save_output_maps.m:148-258: Maps.Hydro.d, velocity,
 * hazard_dv, risk, I_t, GWdepth_save; Maps.WQ_States.Pol_Conc_Map, Pol_Mass_Map,
 * Pol_Load_Map) */
enum {
    HP2D_MAP_DEPTH = 0, HP2D_MAP_VELOCITY, HP2D_MAP_HAZARD, HP2D_MAP_RISK, HP2D_MAP_INFILTRATION,
    HP2D_MAP_GW_DEPTH, HP2D_MAP_POL_CONC, HP2D_MAP_POL_MASS, HP2D_MAP_POL_LOAD, HP2D_MAP_FIELDS
};

#define HP2D_MAX_INFLOW_GAUGES 16

/* Host inputs of a run; all grids ny*nx column-major.  Pointers may be NULL
 * where marked optional. */
typedef struct hp2d_run_inputs {
    const double *elevation;           /* [m], NaN outside the domain */
    const double *roughness;           /* Manning n */
    const double *d0;                  /* initial depth [mm] (optional: zeros) */
    const uint8_t *outlet_mask;        /* 1 = outlet cell (required unless no outlet) */
    /* lumped, end-stamped piecewise-constant rainfall (update_spatial_BC.m:1095-1185) */
    const double *rain_time_min;       /* end times [min] */
    const double *rain_intensity_mmh;  /* rates [mm/h] */
    int32_t n_rain;
    const double *rain_pattern;        /* optional per-cell multiplier of the lumped rate */
    /* Green-Ampt, class-indexed (HydroPol2D_preprocessing.m:2249-2253) */
    const uint8_t *soil_class;         /* optional (all class 0) */
    const double *ksat_mmh;            /* [n_soil_classes] */
    const double *psi_mm;
    const double *dtheta;              /* theta_sat - theta_i */
    const double *theta_sat_minus_r;   /* for k_out */
    const double *lu_m;                /* optional top-layer depth override (NaN = derived) */
    const double *I0_mm;               /* optional initial infiltrated depth per cell */
    double I_min_mm;                   /* min_soil_moisture (HydroPol2D_preprocessing.m:2386-2389) */
    /* Water quality, class-indexed by LULC (HydroPol2D_preprocessing.m:2159-2174) */
    const uint8_t *lulc_class;
    const double *wq_C1, *wq_C2, *wq_C3, *wq_C4; /* [n_lulc_classes] */
    double wq_ADD_days;
    double wq_Bmin, wq_Bmax, wq_min_Bt;
    /* Groundwater (Groundwater_Module.m with simulate_groundwater_recharge.m and
     * Boussinesq_2D_explicit.m): the vadose bucket is I_t, the soil classes
     * carry the van Genuchten parameters, the soil Ks is ksat_mmh above */
    const double *gw_h0;               /* initial water-table head [m] */
    const double *gw_soil_depth;       /* soil depth above the aquifer base [m] */
    double gw_ksat_mmh, gw_Sy;         /* aquifer conductivity [mm/h], specific yield */
    const double *gw_theta_sat, *gw_theta_r, *gw_n_vg, *gw_alpha_vg; /* [n_soil_classes] */
    double gw_l_vg;                    /* Mualem pore-connectivity (0.5) */
    /* Inflow hydrographs (flag_inflow): end-stamped series of rates [mm/h over
     * the gauge's cells] shared time stamps, one column per gauge */
    const double *inflow_time_min;     /* [n_inflow_times] end stamps [min] */
    const double *inflow_rate_mmh;     /* [n_inflow_times x n_inflow_gauges], times fastest */
    int32_t n_inflow_times;
    const uint8_t *inflow_gauge;       /* ny*nx: the cell's gauge 1..n_inflow_gauges, 0 = none
                                          (Wshed_Properties.inflow_cells as one index map) */
    /* Rainfall maps (n_rain_maps > 0): piecewise-constant rate maps [mm/h],
     * end-stamped, from interpolated gauges (hp2d_idw_interpolate) or input
     * rasters; they replace rain_time_min / rain_intensity_mmh / rain_pattern */
    const double *rain_map_time_min;   /* [n_rain_maps] end stamps [min] */
    const double *rain_maps_mmh;       /* [ny*nx x n_rain_maps], NaN outside the domain */
    /* van Genuchten infiltration (flag_infiltration = 2): the soil classes'
     * gw_theta_sat, gw_theta_r, gw_n_vg (> 1), gw_alpha_vg and ksat_mmh, gw_l_vg,
     * the head gw_h0 and gw_soil_depth above give the unsaturated zone; with the
     * groundwater module the head follows its updates.  I0_mm / I_min_mm set
     * the initial storage I_t [mm]. */
    const uint8_t *impervious;         /* optional idx_imp (1 = impervious: no infiltration) */
    const double *ca_h0_mm;            /* CA routing: optional LULC_Properties.h_0 [mm] per cell
                                          (CA_Routing.m:110-113; NULL = 0) */
} hp2d_run_inputs;

/* Rainfall_Interpolator.m:40-52 / idw_function.m:18-63: inverse-distance
 * weighting of n_obs gauges (x_obs, y_obs [m], values f_obs) onto the grid
 * meshgrid(x_grid, y_grid), out ny_grid x nx_grid column-major:
 * out = sum W f / sum W with W = 1 / D^p, D the Euclidean distance, D = 0
 * taken as eps and D > rad as Inf (the reference's defaults: p = 2, rad = Inf).
 * Builds the spatial rainfall maps of update_spatial_BC.m:469-490. */
hp2d_status hp2d_idw_interpolate(int32_t n_obs, const double *x_obs, const double *y_obs,
                                 const double *f_obs, int32_t nx_grid, const double *x_grid,
                                 int32_t ny_grid, const double *y_grid, double p, double rad,
                                 double *out);

/* ------------------------------------------------------------------------ */
/* On-disk formats (host code)                                              */
/* ------------------------------------------------------------------------ */

/* GeoTIFF rasters as GRIDobj reads them (HydroPol2D_preprocessing.m:192-194):
 * one band, strips or tiles, uncompressed / PackBits / LZW / Deflate,
 * predictor 1-3, 8-64-bit integers or floats, either byte order.  info gives
 * the size and geo6 = {x of the upper-left corner, y of it, dx, dy, nodata,
 * EPSG code}; read writes Z (ny*nx column-major, row 1 = the first image row,
 * GDAL_NODATA cells NaN). */
hp2d_status hp2d_geotiff_info(const char *path, int32_t *ny, int32_t *nx, double *geo6);
hp2d_status hp2d_geotiff_read(const char *path, double *Z);

/* MATLAB Level 5 MAT-file writer for the temporary outputs of
 * save_output_maps.m:575-580, 620-830 (save_map_hydro_N.mat, which holds the
 * Maps struct, and current_time_series_outputs.mat, which holds TempTS):
 * each variable is named by a dotted path ("Maps.Hydro.d") and the paths are
 * written as nested 1 x 1 structs; data are column-major. */
enum { HP2D_MAT_DOUBLE = 0, HP2D_MAT_SINGLE = 1, HP2D_MAT_INT32 = 2, HP2D_MAT_UINT8 = 3,
       HP2D_MAT_CHAR = 4 };
typedef struct hp2d_mat_var {
    const char *name;
    int32_t mclass;                    /* HP2D_MAT_* (CHAR: ASCII bytes) */
    int32_t ndims;
    const int64_t *dims;
    const void *data;
} hp2d_mat_var;
hp2d_status hp2d_mat_write(const char *path, int32_t n_vars, const hp2d_mat_var *vars);

/* Results of hp2d_sim_run, accumulated on the device in FP64. */
typedef struct hp2d_run_report {
    int64_t steps;                     /* accepted steps (k - 1) */
    int64_t rejected_steps;            /* rollbacks */
    double t_min;                      /* model time at exit [min] */
    double time_step_s;                /* dt of the next step [s] */
    double max_velocity;               /* velocities.max_velocity */
    double volume_error_m3;            /* last mass_balance_check volume_error */
    int32_t n_reports;                 /* mass-balance reports so far */
    double outflow_volume_m3;          /* BC_States.outflow_volume */
    double inflow_volume_m3;           /* BC_States.inflow_volume */
    double outlet_runoff_volume_mm;    /* save_output_maps.m:611-617 */
    double cum_rain_m3, cum_infiltration_m3, cum_outflow_m3; /* every-step FP64 ledgers */
    double storage_m3;                 /* surface storage at exit */
    int32_t catch_index;
    /* water quality (WQ_States, build_up_wash_off outputs of the last step) */
    double wq_mass_kg;                 /* sum of B_t over the domain */
    double wq_mass_lost_kg;            /* WQ_States.mass_lost */
    double wq_out_conc_mgl;            /* Out_Conc */
    double wq_tmin_s;                  /* tmin_wq */
    double wq_mass_outlet_mg;          /* WQ_States.mass_outlet (Main_While.m:728) */
    double wq_vol_outlet;              /* WQ_States.vol_outlet (Main_While.m:729) */
    int64_t wq_substep_steps;          /* steps whose wash-off was split internally */
    /* groundwater (GW_States) */
    double gw_storage_m3;              /* sum A*Sy*(h_t - z_bed) */
    double gw_total_recharge_m3, gw_total_exfiltration_m3;
    double gw_error_m3;                /* errors(5) of the last update */
    int32_t gw_updates;
    double cum_inflow_m3;              /* every-step FP64 ledger of the inflow hydrographs */
    int32_t n_map_records;             /* map-stack records taken so far */
} hp2d_run_report;

typedef struct hp2d_sim hp2d_sim;

/* Upload a run; allocates every device buffer once.  ny * nx must stay below
 * 2^32 - 2^16 (HP2D_E_ARG otherwise). */
hp2d_status hp2d_sim_create(const hp2d_run_config *cfg, const hp2d_run_inputs *in,
                            hp2d_sim **out);
void hp2d_sim_destroy(hp2d_sim *s);
const char *hp2d_sim_last_error(const hp2d_sim *s);

/* Advance up to max_steps steps of the Main_While loop (or until t passes
 * routing_time + min_dt, HydroPol2D_Main_While.m:403).  No host round-trip
 * per step: dt, rainfall, hydrograph and mass ledgers live on the device.
 * *done = 1 once the end of the simulation is reached. */
hp2d_status hp2d_sim_run(hp2d_sim *s, int64_t max_steps, int32_t *done);

/* Advance exactly n_steps LI steps with a fixed step pattern and no time
 * control bookkeeping beyond the device dt (timing / throughput harness).
 * Uses a captured hipGraph of `graph_steps` steps when graph_steps > 0. */
hp2d_status hp2d_sim_run_fixed(hp2d_sim *s, int64_t n_steps, int32_t graph_steps);

hp2d_status hp2d_sim_report(hp2d_sim *s, hp2d_run_report *rep);

/* Copy state back to the host (arrays of ny*nx unless noted; NULL skips).
 * outflow: ny*nx*3 [mm/h]; Hf: ny*nx*3 [m] (of the last step). */
hp2d_status hp2d_sim_get_state(hp2d_sim *s, double *d_t, double *outflow, double *Hf,
                               double *I_t, double *dmax, double *B_t, double *gw_h);

/* Water-quality maps of the last step: P_conc [mg/L] and Tot_Washed [kg]
 * (build_up_wash_off.m:177,182); ny*nx each, NULL skips. */
hp2d_status hp2d_sim_get_wq_maps(hp2d_sim *s, double *P_conc, double *Tot_Washed);

/* Hydrograph records (save_output_maps.m:266-269,409,481-490):
 * q_m3s[n], time_min[n], depth_outlet_mm[n]; n_max capacity, *n_out filled. */
hp2d_status hp2d_sim_get_hydrograph(hp2d_sim *s, double *q_m3s, double *time_min,
                                    double *depth_outlet_mm, int32_t n_max, int32_t *n_out);

/* Per-step dt history [s] (last n_max accepted steps, oldest first). */
hp2d_status hp2d_sim_get_dt_history(hp2d_sim *s, double *dt_s, int32_t n_max,
                                    int32_t *n_out);

/* Force the step with counter k == step_k to be rejected (kind 1), to test the
 * host rollback path of HydroPol2D_Main_While.m:926-993 without a real fault. */
/* Human_Instability_Module.m:22 Human_Instability.max_risk (ny*nx) */
hp2d_status hp2d_sim_get_max_risk(hp2d_sim *s, double *max_risk);
/* The map-stack records taken so far (save_output_maps.m:148-258, Maps.Hydro and
 * Maps.WQ_States), field one of HP2D_MAP_*: records first .. first + count - 1 into
 * out (count * ny*nx, optional) with their model times t_min (optional);
 * *n_records receives the number taken.  The records are held in host memory
 * as the reference's Maps structure is. */
hp2d_status hp2d_sim_get_map_records(hp2d_sim *s, int32_t field, int32_t first, int32_t count,
                                     double *out, double *t_min, int32_t *n_records);
hp2d_status hp2d_debug_inject_failure(hp2d_sim *s, int64_t step_k, int32_t kind);

/* Project autotuner of the fused step's launch shape: times the candidate
 * (columns per wave task, waves per block) pairs on this sim's own data
 * without advancing it, keeps the fastest, and caches the choice per
 * (device, grid shape, precision) for the process.  Outputs may be NULL. */
hp2d_status hp2d_sim_autotune(hp2d_sim *s, int32_t *best_tj, int32_t *best_wpb, double *best_ms);

/* The fused step's launch shape (no reference counterpart): columns per wave
 * task tj (16..4096), waves per block wpb (1, 2 or 4) and grid rows per lane
 * (1, or 2 when ny is even).  Every shape gives bit-identical results; the
 * setter refuses a shape outside these limits.  Outputs may be NULL. */
hp2d_status hp2d_sim_launch_shape(hp2d_sim *s, int32_t *tj, int32_t *wpb, int32_t *rows);
hp2d_status hp2d_sim_set_launch_shape(hp2d_sim *s, int32_t tj, int32_t wpb, int32_t rows);

/* Average duration [ms] of the fused local-inertial kernel over the last
 * hp2d_sim_run_fixed call, measured with HIP events on the sim's stream. */
hp2d_status hp2d_sim_kernel_time(hp2d_sim *s, double *li_ms, double *total_ms,
                                 int64_t *launches);

/* Time to peak: once enabled, every accepted step whose depth raises a cell's
 * dmax also records the model time [min] at the end of that step.  Derived
 * output with no counterpart in the reference (save_output_maps.m:588-600
 * keeps only the peak).  Enabling it costs one conditional 4/8-byte write
 * per raised cell. */
hp2d_status hp2d_sim_record_time_to_peak(hp2d_sim *s);
hp2d_status hp2d_sim_get_time_to_peak(hp2d_sim *s, double *tpeak_min);

/* ------------------------------------------------------------------------ */
/* Ensemble: independent runs advanced together                              */
/* ------------------------------------------------------------------------ */

/* Members of an ensemble / calibration / probabilistic study (the member loops
 * of HydroPol2D_Main_While_Ensemble.m and Calibration_Water_Quality.m, each a
 * full HydroPol2D_Main_While run) with their own grid sizes, inputs, clocks
 * and adaptive dt.  One batched launch of the fused step covers the tiles of
 * every unfinished member and one batched launch runs every member's dt /
 * bookkeeping tail; finished members drop out of the tile table.  Members
 * share precision and flag_critical.  cols_per_wave: 16, 32, 64, 128 for all
 * members, or 0 for each member's own default. */
typedef struct hp2d_ensemble hp2d_ensemble;
hp2d_status hp2d_ensemble_create(int32_t n_members, const hp2d_run_config *cfgs,
                                 const hp2d_run_inputs *inputs, int32_t cols_per_wave,
                                 hp2d_ensemble **out);
void hp2d_ensemble_destroy(hp2d_ensemble *e);
const char *hp2d_ensemble_last_error(const hp2d_ensemble *e);
/* Advance every unfinished member by up to max_steps steps (hipGraphs of
 * graph_steps batched steps when graph_steps > 0).  *n_done = members finished. */
hp2d_status hp2d_ensemble_run(hp2d_ensemble *e, int64_t max_steps, int32_t graph_steps,
                              int32_t *n_done);
/* Borrow member i's simulation for the hp2d_sim_* getters (owned by the ensemble). */
hp2d_status hp2d_ensemble_member(hp2d_ensemble *e, int32_t i, hp2d_sim **sim);
hp2d_status hp2d_ensemble_record_time_to_peak(hp2d_ensemble *e);
/* Device time [ms] of the last hp2d_ensemble_run call. */
hp2d_status hp2d_ensemble_elapsed(hp2d_ensemble *e, double *ms);

/* ------------------------------------------------------------------------ */
/* Flood surrogate kernels (no counterpart in the reference)                 */
/* ------------------------------------------------------------------------ */
/* A convolutional encoder-decoder maps per-member input maps to peak-depth
 * and time-to-peak maps; its layers run on these kernels.  All pointers are
 * device pointers; `stream` is a hipStream_t (NULL = default stream).
 * Activations are NHWC bf16 stored as uint16 (channels a multiple of 32);
 * 3x3 weights are [Cout][3][3][Cin] bf16; biases, weight gradients, optimiser
 * state and the head are FP32.  Every 3x3 convolution is an implicit GEMM on
 * gfx950 BF16 MFMA with FP32 accumulation. */

/* y = act(conv3x3(x, w) + bias) (act = ReLU when relu != 0); with skip != NULL
 * also y_skip = y + skip.  bias may be NULL. */
hp2d_status hp2d_nn_conv3x3_fwd(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                                const uint16_t *x, const uint16_t *w, const float *bias,
                                int32_t relu, const uint16_t *skip, uint16_t *y, uint16_t *y_skip,
                                void *stream);
/* gx = d/dx of conv3x3 for the output gradient gy (before the activation);
 * w_flip is a 9*Cin*Cout bf16 workspace. */
hp2d_status hp2d_nn_conv3x3_dgrad(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                                  const uint16_t *gy, const uint16_t *w, uint16_t *w_flip,
                                  uint16_t *gx, void *stream);
/* hp2d_nn_conv3x3_dgrad without the w_flip workspace: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops).  It allocates nothing: the flipped weights
 * live in a fixed per-GPU device array of 9*512*256 bf16 (2.25 MiB, held
 * until the process ends), so 9*Cin*Cout must not exceed 1179648
 * (HP2D_E_ARG otherwise).  Calls that overlap on one GPU share that array
 * and race; the results equal hp2d_nn_conv3x3_dgrad's bit for bit. */
hp2d_status hp2d_nn_conv3x3_dgrad_noalloc(int32_t N, int32_t H, int32_t W, int32_t Cin,
                                          int32_t Cout, const uint16_t *gy, const uint16_t *w,
                                          uint16_t *gx, void *stream);
/* gw [Cout][9*Cin] += d/dw, gb [Cout] += d/db (gb may be NULL), FP32. */
hp2d_status hp2d_nn_conv3x3_wgrad(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                                  const uint16_t *x, const uint16_t *gy, float *gw, float *gb,
                                  void *stream);
/* g[i] = 0 where y[i] <= 0 (ReLU backward on the stored activation); g and y
 * 16-byte aligned (8 elements per thread). */
hp2d_status hp2d_nn_relu_mask(uint16_t *g, const uint16_t *y, int64_t n, void *stream);
/* out = a + b (bf16); a, b, out 16-byte aligned. */
hp2d_status hp2d_nn_add(const uint16_t *a, const uint16_t *b, uint16_t *out, int64_t n,
                        void *stream);
/* 2x2 average pooling of x [N,H,W,C] into y [N,H/2,W/2,C]; with backward != 0,
 * x is the gradient of the pooled map and y the gradient of the input [N,H,W,C].
 * C a multiple of 8, tensors 16-byte aligned. */
hp2d_status hp2d_nn_avgpool2(const uint16_t *x, int32_t N, int32_t H, int32_t W, int32_t C,
                             uint16_t *y, int32_t backward, void *stream);
/* nearest 2x upsampling of x [N,H/2,W/2,C] into y [N,H,W,C]; with backward != 0,
 * x is the gradient [N,H,W,C] and y receives [N,H/2,W/2,C].  C a multiple of 8,
 * tensors 16-byte aligned. */
hp2d_status hp2d_nn_upsample2(const uint16_t *x, int32_t N, int32_t H, int32_t W, int32_t C,
                              uint16_t *y, int32_t backward, void *stream);
/* 1x1 head: out [P][O] = b + x [P][C] w^T (w [O][C]); with gy != NULL the
 * backward instead: gx (bf16), gw, gb accumulated.  C a multiple of 8 and
 * O <= 4 (the backward: C = 32, O = 2); x 16-byte aligned. */
hp2d_status hp2d_nn_head(const uint16_t *x, int64_t P, int32_t C, const float *w, const float *b,
                         int32_t O, float *out, const float *gy, uint16_t *gx, float *gw, float *gb,
                         void *stream);
/* masked MSE: *loss += sum(mask*(out-target)^2)*inv_count, gy = 2(out-target)*mask*inv_count */
hp2d_status hp2d_nn_mse(const float *out, const float *target, const uint8_t *mask, int64_t P,
                        int32_t O, float inv_count, float *gy, float *loss, void *stream);
/* one Adam step (Kingma & Ba 2015) of n FP32 weights; w_bf16 (may be NULL) gets the bf16 copy */
hp2d_status hp2d_nn_adam(float *w, float *m, float *v, const float *g, int64_t n, float lr,
                         float beta1, float beta2, float eps, int32_t step, uint16_t *w_bf16,
                         void *stream);
hp2d_status hp2d_nn_to_bf16(const float *x, int64_t n, uint16_t *y, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HP2D_H */
