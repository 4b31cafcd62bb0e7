/*
 * hp2d_ca.c - FP64 CPU restatement of CA_Routing (TEST INFRASTRUCTURE ONLY).
 *
 * Follows HydroPol2D_Functions/CA_Routing.m:19-171 (flag_reservoir = 0)
 * statement by statement on full-grid column-major arrays.  idx_nan is the
 * set of cells with NaN elevation (the reference passes it separately; its
 * preprocessing sets both together).
 */
#include <math.h>
#include <stdlib.h>

#include "hp2d_oracle.h"

#define IX(i, j) ((int64_t)(i) + (int64_t)(j) * ny)

int32_t oracle_ca_routing(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                          const double *roughness, double cell_area, double time_step_min,
                          const double *h_0, double resolution,
                          const double *I_tot_end_cell_prev, int32_t outlet_type,
                          double slope_outlet, const int64_t *outlet_idx, int32_t n_outlet,
                          int32_t flag_critical, double *qout_left, double *qout_right,
                          double *qout_up, double *qout_down, double *outlet_flow, double *d_t,
                          double *I_tot_end_cell) {
    const int64_t n = (int64_t)ny * nx;
    const double R = resolution, h_min = 1e-6; /* :25, :32 */
    double *depth = calloc((size_t)n, sizeof(double)), *wse = calloc((size_t)n, sizeof(double));
    double *dh = calloc((size_t)(5 * n), sizeof(double));
    double *S0 = calloc((size_t)n, sizeof(double));
    double *Iprev = calloc((size_t)n, sizeof(double));
    for (int64_t c = 0; c < n; ++c) {
        depth[c] = d_tot[c] / 1000.0; /* :39 */
        wse[c] = z[c] + depth[c];     /* :40 */
        Iprev[c] = I_tot_end_cell_prev[c];
    }
    /* :51-54 depth differences to the 4 neighbours (0 outside the grid) */
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            int64_t c = IX(i, j);
            double zl = j > 0 ? z[IX(i, j - 1)] : 0.0, dl = j > 0 ? depth[IX(i, j - 1)] : 0.0;
            double zr = j < nx - 1 ? z[IX(i, j + 1)] : 0.0, dr = j < nx - 1 ? depth[IX(i, j + 1)] : 0.0;
            double zu = i > 0 ? z[IX(i - 1, j)] : 0.0, du = i > 0 ? depth[IX(i - 1, j)] : 0.0;
            double zd = i < ny - 1 ? z[IX(i + 1, j)] : 0.0, dd = i < ny - 1 ? depth[IX(i + 1, j)] : 0.0;
            dh[c] = wse[c] - zl - dl;
            dh[n + c] = wse[c] - zr - dr;
            dh[2 * n + c] = wse[c] - zu - du;
            dh[3 * n + c] = wse[c] - zd - dd;
        }
    /* :58-72 outlet pseudo-neighbour */
    uint8_t *is_out = calloc((size_t)n, 1);
    for (int32_t k = 0; k < n_outlet; ++k) is_out[outlet_idx[k]] = 1;
    if (outlet_type == 1) {
        for (int64_t c = 0; c < n; ++c) dh[4 * n + c] = slope_outlet * R * (double)is_out[c];
    } else {
        for (int32_t k = 0; k < n_outlet; ++k) {
            int64_t c = outlet_idx[k];
            S0[c] = (pow(depth[c], -1.0 / 6.0)) * sqrt(9.81) * roughness[c];
        }
        for (int64_t c = 0; c < n; ++c) dh[4 * n + c] = S0[c] * R;
    }
    /* :80-84 */
    for (int32_t i = 0; i < ny; ++i) {
        dh[IX(i, 0)] = 0.0;
        dh[n + IX(i, nx - 1)] = 0.0;
    }
    for (int32_t j = 0; j < nx; ++j) {
        dh[2 * n + IX(0, j)] = 0.0;
        dh[3 * n + IX(ny - 1, j)] = 0.0;
    }
    /* :87 */
    for (int64_t c = 0; c < 5 * n; ++c)
        if (dh[c] < h_min || isnan(dh[c])) dh[c] = 0.0;
    /* :90-106 */
    const double cbig = R * R * 1000.0;
    double *weight = calloc((size_t)(5 * n), sizeof(double));
    double *wmax = calloc((size_t)n, sizeof(double)), *vmin = calloc((size_t)n, sizeof(double));
    double *maxh = calloc((size_t)n, sizeof(double));
    for (int64_t c = 0; c < n; ++c) {
        double vt = 0.0, vm = NAN, mh = NAN;
        for (int p = 0; p < 5; ++p) {
            double d = dh[p * n + c];
            double v = cell_area * d;
            vt += v;
            double vn = (d < h_min || isnan(d)) ? cbig : v;
            vm = fmin(vm, vn);
            mh = fmax(mh, d);
        }
        vmin[c] = vm;
        maxh[c] = mh;
        double wm = NAN;
        for (int p = 0; p < 5; ++p) {
            double w = (cell_area * dh[p * n + c]) / (vt + vm);
            if (isnan(w)) w = 0.0;
            weight[p * n + c] = w;
            wm = fmax(wm, w);
        }
        wmax[c] = wm;
    }
    /* :121-123 (global condition) */
    double sprev = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double col = 0.0;
        for (int32_t i = 0; i < ny; ++i) col += Iprev[IX(i, j)];
        sprev += col;
    }
    if (sprev == 0.0)
        for (int64_t c = 0; c < n; ++c) Iprev[c] = 0.9 * cell_area * depth[c];
    /* :110-129 */
    for (int64_t c = 0; c < n; ++c) {
        double h0 = h_0 ? h_0[c] : 0.0;
        double vm;
        double man = 1.0 / roughness[c] * pow(fmax(depth[c] - h0 / 1000.0, 0.0), 2.0 / 3.0) *
                     sqrt(maxh[c] / R);
        if (flag_critical == 1) vm = fmin(sqrt(9.81 * depth[c]), man);
        else vm = man;
        double I = fmin(depth[c] * cell_area, vm * depth[c] / wmax[c] * (R * time_step_min * 60.0));
        I = fmin(I, Iprev[c] + vmin[c]);
        if (d_tot[c] < 1e-6 * 1000.0 || isnan(z[c])) I = 0.0;
        I_tot_end_cell[c] = I;
    }
    /* :132-170 */
    const double dts = time_step_min * 60.0;
    for (int64_t c = 0; c < n; ++c) {
        double tot = 0.0;
        double q[5];
        for (int p = 0; p < 5; ++p) {
            double Ip = weight[p * n + c] * I_tot_end_cell[c];
            tot += Ip;
            q[p] = Ip / dts / cell_area * 1000.0 * 3600.0;
        }
        if (qout_left) qout_left[c] = q[0];
        if (qout_right) qout_right[c] = q[1];
        if (qout_up) qout_up[c] = q[2];
        if (qout_down) qout_down[c] = q[3];
        if (outlet_flow) outlet_flow[c] = q[4];
        I_tot_end_cell[c] = tot;
        d_t[c] = d_tot[c] - tot / cell_area * 1000.0;
    }
    free(depth); free(wse); free(dh); free(S0); free(Iprev); free(is_out);
    free(weight); free(wmax); free(vmin); free(maxh);
    return 0;
}
