// hp2d_simca.hpp - the device loop's step with cellular-automata routing
// (flags.flag_CA = 1: HydroPol2D_Main_While.m:461-465 then the inflow gather
// and mass balance of :631-660).
#pragma once

#include "hp2d_li.hpp"

namespace hp2d {

template <typename T> struct CaSimArgs {
    // hydrology inputs, state buffers (d, I, T ping-pong by step parity; the
    // right / up outflows in the qx / qy buffers), partials and outlets: the
    // fields the local-inertial step reads
    LiArgs<T> li;
    T *dtot;             // depth after the hydrology, the routing's d_tot [mm]
    T *ql, *qd, *qo;     // left / down / outlet outflows of the step [mm/h]
    T *itot;             // CA_States.I_tot_end_cell [m3], updated in place (a rejected
                         // step keeps its values, as the reference's catch block does)
    const T *h0;         // LULC_Properties.h_0 [mm] per cell, or NULL (0)
    int blocks;          // workgroups of every pass = partials of the step
    T depth_tol_mm;      // CA_States.depth_tolerance (velocity of dry cells)
    T slope_outlet;      // outlet_type 1: the outlet's normal slope
};

// Three passes over the grid (hydrology; routing; gather + the per-cell terms of
// the time step), each writing its fields of the step's partials.
template <typename T> hipError_t launch_ca_sim(const CaSimArgs<T> &a, hipStream_t s);

}  // namespace hp2d
