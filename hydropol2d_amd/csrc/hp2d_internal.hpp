// hp2d_internal.hpp - shared device/host definitions of the gfx950 HydroPol2D step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

namespace hp2d {

constexpr int kWave = 64;                  // CDNA wavefront
constexpr int kHalo = 2;                   // radius of the step stencil (donor limiter)
constexpr int kRowsOut = kWave - 2 * kHalo; // rows finalised per wave slab (one row per lane)

// The step kernel's slabs when each lane holds `rows` consecutive rows (1 or
// 2): 64*rows rows, of which 64*rows - 4 are finalised.  Two rows per lane
// need an even row count (a lane's rows are then both inside or both outside
// the grid, and its loads are aligned 2-element vector loads).
inline int li_rows_out(int rows) { return kWave * rows - 2 * kHalo; }
inline int li_n_slabs(int ny, int rows) { return (ny + li_rows_out(rows) - 1) / li_rows_out(rows); }
inline bool li_rows_ok(int ny, int rows) { return rows == 1 || (rows == 2 && ny % 2 == 0); }
constexpr int kXcds = 8;                    // accelerator complex dies of an MI355X
constexpr int kMaxWavesPerBlock = 4;

// XCD-aware workgroup order: the hardware deals workgroups round-robin to the
// 8 XCDs, so workgroup w runs on XCD w % 8.  Give each XCD a contiguous run of
// logical tiles so that the halo shared by neighbouring tiles is fetched into
// one L2.
__device__ __forceinline__ int xcd_logical_tile() {
    const int n_wg = gridDim.x;
    const int xcd = blockIdx.x % kXcds, k_in = blockIdx.x / kXcds;
    const int per = n_wg / kXcds, rem = n_wg % kXcds;
    return xcd * per + min(xcd, rem) + k_in;
}
constexpr int kMaxSoilClasses = 256;        // class index is one byte per cell

// ----------------------------------------------------------------------------
// Per-wave partial results of one step (structure of arrays, FP64), reduced
// by the finalize kernel in a fixed order (deterministic).  Field f of wave w
// lives at partials[f * cap + w].
// ----------------------------------------------------------------------------
enum PartField {
    kPartStorage = 0,  // sum A*d_t/1000 over valid cells      [m3]
    kPartUz,           // sum A*I_t/1000 (Green-Ampt)          [m3]
    kPartInfil,        // infiltrated volume this step         [m3]
    kPartVmax,         // max total cell velocity              [m/s]  (NaN = none)
    kPartDtw,          // min Courant wave candidate           [s]    (NaN = none)
    kPartRaster,       // max velocity_raster                  [m/s]
    kPartDtotMax,      // max d_tot/1000 (rollback celerity)   [m]
    kPartDprevMax,     // max d_p/1000                         [m]
    kPartNeg,          // 1 if a depth < -1e-8 mm was seen before clean-up
    kPartDnewMax,      // max new d_t/1000 (late rollback celerity)  [m]
    kPartCaItot,       // CA routing: 1 if any I_tot_end_cell of the step is not 0
    kPartCount
};

// ----------------------------------------------------------------------------
// Device-resident time-control state of the Main_While loop (FP64).
// ----------------------------------------------------------------------------
struct SimState {
    double t, t_previous, time_step_min, time_save_previous, delta_p_agg;
    double progress, next_report, S_prev;
    double max_dt, min_dt;
    double outflow_volume, inflow_volume, volume_error, runoff_mm;
    double cum_rain, cum_inf, cum_out, storage, uz;
    double max_velocity, max_vel_raster, tcr;
    double tmin_wq;
    // water quality (WQ_States and the per-step outcome of build_up_wash_off)
    double wq_initial_mass, wq_mass_lost, wq_sum_wout, wq_sum_b, wq_out_conc;
    double wq_mass_outlet, wq_vol_outlet;
    // groundwater: saturated storage + pending exchange [m3], model time since the last call
    double gw_storage, gw_elapsed_min;
    long long k, accepted, rejected, inject_k, stop_accepted;
    long long wq_n, wq_split_steps;  // wash-off calls (B_t buffer parity), split calls
    int cursor, catch_index, done, last_record, n_hydro, n_reports, neg_events, paused;
    int hold;    // a wash-off split longer than the captured sub-steps waits for the host
    int wq_err;  // the wash-off refused a split beyond 1e7 sub-steps
    // rainfall maps of the next step (update_spatial_BC.m:469-518): the maps
    // overlapping it, their overlaps [min], (t - t_previous) and dt/60
    int map_n, map_cursor, map_err, inflow_cursor;
    int map_idx[4];
    double map_w[4], map_dtot, map_k;
    // inflow hydrographs of the next step: BC_States.delta_inflow_agg [mm]
    double inflow_agg[16];
    double cum_inflow;
    // map stacks (save_output_maps.m:24-27): the last record state, whether this
    // step took a record, the records waiting in the device stack, the records
    // taken so far, and the dt the step used
    int map_last, snap_event, snap_n, n_map_records;
    double dt_used_min;
    // CA routing: 1 when the last routing left some I_tot_end_cell other than 0
    // (CA_Routing.m:121 restarts the limiter from 0.9 A d when they sum to 0)
    double ca_itot_any;
};
static_assert(sizeof(SimState) % 4 == 0, "SimState is copied as 32-bit words");

struct GaClass {
    double ksat, psi, dtheta, Lu_mm, k_out, Tr;
};

// van Genuchten / Mualem soil class of the Infiltration_Module closure:
// theta_sat, theta_r, n_vg, alpha_vg [1/m], ksat [mm/h].  The FP32 table holds
// per class {theta_r, theta_sat, 1/n, m, 1/m, -1/m, 1/alpha, Ks [m/s]} with
// m = 1 - 1/n.
struct VgClass {
    double ts, tr, nv, alpha, ksat;
};

// ----------------------------------------------------------------------------
// Small math helpers with MATLAB semantics (max/min omit NaN -> fmax/fmin).
// ----------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T dmax(T a, T b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T dmin(T a, T b) { return fmin(a, b); }

__device__ __forceinline__ double pow73(double h) { return pow(h, 7.0 / 3.0); }
__device__ __forceinline__ float pow73(float h) { return h * h * cbrtf(h); }
__device__ __forceinline__ double pow53(double h) { return pow(h, 5.0 / 3.0); }
__device__ __forceinline__ float pow53(float h) { float c = cbrtf(h); return h * c * c; }
__device__ __forceinline__ double powm13(double h) { return pow(h, -1.0 / 3.0); }
__device__ __forceinline__ float powm13(float h) { return 1.0f / cbrtf(h); }

template <typename T> __device__ __forceinline__ T eps_of();
template <> __device__ __forceinline__ double eps_of<double>() { return 2.220446049250313e-16; }
template <> __device__ __forceinline__ float eps_of<float>() { return 1.1920929e-07f; }

template <typename T> __device__ __forceinline__ T wave_fmax(T v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}
template <typename T> __device__ __forceinline__ T wave_fmin(T v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

}  // namespace hp2d

// Error helpers shared by the C-ABI implementation.
#define HP2D_HIP_TRY(expr)                                                          \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return HP2D_E_HIP;                                                       \
        }                                                                            \
    } while (0)
