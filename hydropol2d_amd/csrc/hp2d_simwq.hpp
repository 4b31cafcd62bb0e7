// hp2d_simwq.hpp - pollutant build-up / wash-off inside the device step loop.
#pragma once

#include "hp2d_internal.hpp"

namespace hp2d {

// per-cell flag word of the wash-off: class in the low byte, then two flags
constexpr unsigned kWqNanBit = 1u << 8;     // idx_nan_5 (NaN elevation)
constexpr unsigned kWqOutletBit = 1u << 9;  // an outlet cell (outlet_flow page)

// block partial fields (FP64), laid out as [field][block]
enum WqField {
    kWqMinDW = 0,   // min dW of the pass                        [kg/h]
    kWqMinMid,      // min B_t_mid (first pass)                  [kg]
    kWqCand0,       // min m_calc/(|dW-| + 1e-16) of the entry B [h]
    kWqCandEnd,     // the same of the closing B                 [h]
    kWqSumWout,     // sum of the outlet page W (first pass)     [kg/h]
    kWqSumB,        // sum of the closing B_t over ~isinf        [kg]
    kWqSumNeg,      // sum of the negative B_t before the clamp  [kg]
    kWqOutMass,     // outlet mass over the sub-steps            [kg]
    kWqFieldCount
};

// internal time-step split of one wash-off call (build_up_wash_off.m:80-88)
struct WqCtl {
    double n_dec, n_floor, tmin0, dt_s;
    long long steps;
    int split;
};

template <typename T> struct WqSimArgs {
    int ny, nx;
    long long n;
    double A, min_Bt, Bmin, Bmax;
    int model;       // flag_wq_model
    int n_blocks;    // grid of the per-cell kernels (grid-stride)
    int s_cap;       // sub-steps captured per step; more raise the hold flag
    const uint16_t *cell;       // class | flags
    const long long *out_idx;   // sorted outlet cells
    const T *out_q;             // outlet_flow per outlet [mm/h]
    int n_outlet;
    T *qx_buf[2], *qy_buf[2];   // face fluxes [mm/h] by step parity
    const T *C3, *C4;           // class tables
    int n_classes;
    T *B[2];                    // B_t [kg] by wash-off call parity
    T *Bs;                      // second buffer of the sub-step ping-pong
    T *Bx[2];                   // B_t_extra of the rating model (sub-steps)
    T *omass;                   // outflow_mass of the sub-steps [kg]
    T *TW[2];                   // Tot_Washed [kg] by wash-off call parity
    T *Pc;                      // P_conc [mg/L]
    double *part;               // [kWqFieldCount][n_blocks]
    WqCtl *ctl;
    SimState *st;
};

// First pass, split decision, up to s_cap guarded sub-steps and their fold:
// the launches of one step (graph-capturable, no host synchronisation).
template <typename T> hipError_t launch_wq_step(const WqSimArgs<T> &a, hipStream_t s);
// Finish a step whose split exceeded s_cap (the hold flag is set): every
// sub-step, the fold, and the hold cleared.
template <typename T>
hipError_t launch_wq_finish(const WqSimArgs<T> &a, long long steps, hipStream_t s);

}  // namespace hp2d
