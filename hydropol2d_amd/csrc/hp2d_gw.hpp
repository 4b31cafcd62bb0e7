// hp2d_gw.hpp - the explicit Boussinesq solver's device state and its
// device-pointer entry, shared by the per-call boundary and the groundwater
// module.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "hp2d.h"

namespace hp2d {

constexpr int kGwBlock = 256;
constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

struct GwCtl {
    double t, T, dt, dt_used, n_steps;
    unsigned long long dtgw_bits;  // min positive finite Courant candidate, as bits
    int done, nsub;
    int last;  // the running sub-step is the call's last one
};

struct GwArgs {
    int ny, nx;
    long long n;
    double dx, dy, courant;
    const double *z0, *Sy, *R, *K, *h_soil, *h_dir;
    const uint8_t *catchment, *dirichlet;  // NULL = all inside / none
    GwCtl *ctl;
    // a uniform specific yield / conductivity [m/s] / soil depth [m] in place
    // of the Sy / K / h_soil maps when those are NULL
    double Sy_u = 0.0, K_u = 0.0, h_soil_u = 0.0;
    // FP32 runs: the sub-step as two passes that recompute the face fluxes,
    // with reciprocal multiplies (gw_run then returns the final head in
    // w.ht or w.Fx, by the parity of the sub-step count)
    int fast = 0;
};

// Sy and K of cell c: the map, or the uniform value
__device__ __forceinline__ double gw_Sy(const GwArgs &a, long long c) { return a.Sy ? a.Sy[c] : a.Sy_u; }
__device__ __forceinline__ double gw_K(const GwArgs &a, long long c) { return a.K ? a.K[c] : a.K_u; }

// device buffers of one solve (carved from a workspace)
struct GwScratch {
    double *ht, *hp, *tmp, *Fx, *Fy, *ux, *uy, *q, *col;
    GwCtl *ctl;
};

size_t gw_scratch_bytes(int ny, int nx);
GwScratch gw_scratch(char *base, int ny, int nx);

// Boussinesq_2D_explicit on device arrays: a holds the device fields (z0, Sy,
// R, K, h_soil, Dirichlet, masks), dh0 the entry head; the final head, q_exf
// and velocities are left in w.ht, w.q, w.ux, w.uy; *error as the reference.
hp2d_status gw_run(hp2d_handle *h, const GwArgs &a, double dt, const double *dh0,
                   const GwScratch &w, double *error, int32_t *n_substeps,
                   const double **h_final = nullptr);

// Device fields of the groundwater module step (Groundwater_Module.m:37-404):
// soil / aquifer maps, the workspace state it updates in place (head h [m],
// vadose storage I [mm], surface depth d [mm], cumulative recharge [mm],
// maximum water-table depth, pending exchange maps [m]) and temporaries.
struct GwmArgs {
    int ny, nx;
    long long n;
    double dt_s, A, l_vg;
    const double *elev, *soil_depth, *ts, *tr, *nv, *alpha, *ksat, *ksat_gw, *Sy;
    // uniform aquifer conductivity [mm/h] and specific yield when ksat_gw / Sy
    // are NULL (Kgw is then NULL too and Kgw_u = ksat_gw_u / 3.6e6)
    double ksat_gw_u = 0.0, Sy_u = 0.0, Kgw_u = 0.0;
    double soil_depth_u = 0.0;  // the soil depth [m] when soil_depth is NULL
    // soil classes in place of the ts / tr / nv / alpha / ksat maps when
    // soil_class is set: a class byte per cell and per-class tables
    const uint8_t *soil_class = nullptr;
    const double *cls_ts = nullptr, *cls_tr = nullptr, *cls_nv = nullptr, *cls_alpha = nullptr,
                 *cls_ksat = nullptr;
    double *h, *I, *d, *cum, *maxd, *pn, *pr, *pc;
    double *z_bed, *Kgw, *rfg, *q;
    double *t2;  // unsaturated-zone excess of the last update [m]
    // FP32 runs: d and I in the run's own buffers (NULL: the FP64 arrays above)
    float *d32 = nullptr, *I32 = nullptr;
    // z_bed and Kgw already hold elev - soil_depth and ksat_gw / 3.6e6 (the
    // recharge pass then leaves them alone)
    int statics_ready = 0;
    // the Boussinesq sub-steps as two passes with reciprocal multiplies
    int fast_boussinesq = 0;
};

// One module call on device arrays (the scalars of p, the state g updated);
// col: 3*nx doubles of scratch; dom: catchment mask or NULL.  Returns after
// the stream has drained; *s_gw (optional) receives the saturated storage
// plus the pending exchange [m3] after the call.
hp2d_status gwm_run(hp2d_handle *h, const hp2d_gw_params *p, hp2d_gw_state *g, GwmArgs a,
                    double *col, const uint8_t *dom, double *error, int32_t *ran, double *s_gw);

}  // namespace hp2d
