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
};

struct GwArgs {
    int ny, nx;
    long long n;
    double dx, dy, courant;
    const double *z0, *Sy, *R, *K, *h_soil, *h_dir;
    const uint8_t *catchment, *dirichlet;  // NULL = all inside / none
    GwCtl *ctl;
};

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
                   const GwScratch &w, double *error, int32_t *n_substeps);

}  // namespace hp2d
