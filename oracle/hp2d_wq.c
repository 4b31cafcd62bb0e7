/*
 * hp2d_wq.c - FP64 CPU restatement of build_up_wash_off (TEST INFRASTRUCTURE ONLY).
 *
 * Follows HydroPol2D_Functions/build_up_wash_off.m:11-183 statement by
 * statement.  Face rates below zero (inflow faces of the signed
 * local-inertial outputs) carry no wash-off: q is taken as max(q, 0) before
 * the power law (for the non-negative CA / validation inputs this is the
 * reference expression unchanged); tot_q_out keeps the signed sum of :37.
 * nan_mask is page 1 of idx_nan_5 (all five pages are the same mask).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hp2d_oracle.h"

#define IX(i, j) ((int64_t)(i) + (int64_t)(j) * ny)

static double nanmin_all(const double *x, int64_t n) {
    double m = NAN;
    for (int64_t c = 0; c < n; ++c) m = fmin(m, x[c]);
    return m;
}

/* W_out pages (5) and their sum for the current B_t */
static void washoff_rates(int32_t ny, int32_t nx, const double *C3, const double *C4,
                          const double *q5, const double *B_t, double *B_extra,
                          double cell_area, int32_t model, double Bmin, double Bmax,
                          double min_Bt, double *W, double *totW) {
    const int64_t n = (int64_t)ny * nx;
    const double Br_kg = Bmin / 1000.0 * cell_area, Bm_kg = Bmax / 1000.0 * cell_area;
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        double f_bt = B_t[c];
        if (model == 1) {
            if (B_t[c] * 1000.0 / cell_area <= Bmin) B_extra[c] = 0.0;
            if (B_t[c] * 1000.0 / cell_area >= Bmax) B_extra[c] = Bm_kg;
            f_bt = 1.0 + fmax(B_extra[c] - Br_kg, 0.0);
        }
        const int low = (B_t[c] * 1000.0 / cell_area < min_Bt) || isnan(B_t[c]) || isinf(B_t[c]);
        double tot = 0.0;
        for (int p = 0; p < 5; ++p) {
            double q = q5[p * n + c] < 0.0 ? 0.0 : q5[p * n + c]; /* NaN stays NaN */
            double w = C3[c] * pow(q / 1000.0 / 3600.0 * cell_area, C4[c]) * f_bt;
            if (low) w = 0.0;
            W[p * n + c] = w;
            tot += w;
        }
        totW[c] = tot;
    }
}

/* dW = tot_W_out - sum of the four inflows (:51-59) */
static void net_rate(int32_t ny, int32_t nx, const double *W, const double *totW, double *dW) {
    const int64_t n = (int64_t)ny * nx;
    ORACLE_PAR
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            int64_t c = IX(i, j);
            double in_l = j > 0 ? W[1 * n + IX(i, j - 1)] : 0.0;       /* right outflow of the left cell */
            double in_r = j < nx - 1 ? W[0 * n + IX(i, j + 1)] : 0.0;  /* left outflow of the right cell */
            double in_u = i > 0 ? W[3 * n + IX(i - 1, j)] : 0.0;       /* down outflow of the upper cell */
            double in_d = i < ny - 1 ? W[2 * n + IX(i + 1, j)] : 0.0;  /* up outflow of the lower cell */
            dW[c] = totW[c] - (in_l + in_r + in_u + in_d);
        }
}

static double tmin_of(const double *B_t, const double *dW, int64_t n, double cell_area,
                      double min_Bt, double large_timestep) {
    if (nanmin_all(dW, n) >= 0) return large_timestep;
    double m = NAN;
    for (int64_t c = 0; c < n; ++c) {
        double mc = B_t[c];
        if (mc * 1000.0 / cell_area < min_Bt) mc = INFINITY;
        double dc = dW[c] > 0 ? 0.0 : dW[c];
        m = fmin(m, mc / (fabs(dc) + 1e-16));
    }
    return 3600.0 * m;
}

int32_t oracle_washoff(int32_t ny, int32_t nx, const double *C3, const double *C4,
                       const double *ql, const double *qr, const double *qu, const double *qd,
                       const double *qo, double *B_t, double time_step_min, double cell_area,
                       const uint8_t *nan_mask, int32_t model, double *mass_lost,
                       double *Tot_Washed, double Bmin, double Bmax, double min_Bt,
                       double *P_conc, double *Out_Conc, double *tmin_wq, double *tot_W_out) {
    const int64_t n = (int64_t)ny * nx;
    double *q5 = malloc(sizeof(double) * 5 * n), *W = malloc(sizeof(double) * 5 * n);
    double *totW = malloc(sizeof(double) * n), *totQ = malloc(sizeof(double) * n);
    double *dW = malloc(sizeof(double) * n), *Bx = malloc(sizeof(double) * n);
    double *Wout5 = malloc(sizeof(double) * n);
    memcpy(q5, ql, sizeof(double) * n);
    memcpy(q5 + n, qr, sizeof(double) * n);
    memcpy(q5 + 2 * n, qu, sizeof(double) * n);
    memcpy(q5 + 3 * n, qd, sizeof(double) * n);
    memcpy(q5 + 4 * n, qo, sizeof(double) * n);
    const double large_timestep = time_step_min * 60.0;
    memcpy(Bx, B_t, sizeof(double) * n); /* B_t_extra = B_t (:19) */
    washoff_rates(ny, nx, C3, C4, q5, B_t, Bx, cell_area, model, Bmin, Bmax, min_Bt, W, totW);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        double tq = 0.0;
        for (int p = 0; p < 5; ++p) tq += q5[p * n + c]; /* sum(q_out_t,3), signed */
        totQ[c] = tq;
        if (nan_mask && nan_mask[c]) { /* :41-43 */
            totW[c] = 0.0;
            totQ[c] = 0.0;
            for (int p = 0; p < 5; ++p) W[p * n + c] = 0.0;
        }
    }
    net_rate(ny, nx, W, totW, dW);
    double tmin = tmin_of(B_t, dW, n, cell_area, min_Bt, large_timestep); /* :61-76 */
    double *Bmid = malloc(sizeof(double) * n);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) Bmid[c] = B_t[c] - dW[c] * time_step_min / 60.0;
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) Wout5[c] = W[4 * n + c];
    double lost = 0.0;
    const int32_t split = nanmin_all(Bmid, n) < 0 || tmin < time_step_min * 60.0;
    if (split) { /* :80-135 */
        double n_dec = time_step_min * 60.0 / tmin;
        double n_floor = floor(n_dec);
        if (!(n_floor < 1e7)) { /* the reference would loop for ever (tmin_wq = 0) */
            free(q5); free(W); free(totW); free(totQ); free(dW); free(Bx); free(Wout5); free(Bmid);
            return -1;
        }
        int64_t steps = (int64_t)n_floor + 1;
        double *outflow_mass = calloc((size_t)n, sizeof(double));
        double *outlet_mass = calloc((size_t)n, sizeof(double));
        for (int64_t s = 1; s <= steps; ++s) {
            double dt = (s == steps) ? (n_dec - n_floor) * tmin : time_step_min * 60.0 / (double)steps;
            washoff_rates(ny, nx, C3, C4, q5, B_t, Bx, cell_area, model, Bmin, Bmax, min_Bt, W, totW);
            net_rate(ny, nx, W, totW, dW);
            for (int64_t c = 0; c < n; ++c) {
                outflow_mass[c] = outflow_mass[c] + totW[c] * dt / 3600.0;
                outlet_mass[c] = outlet_mass[c] + W[4 * n + c] * dt / 3600.0;
                B_t[c] = B_t[c] - dW[c] * dt / 3600.0;
            }
        }
        for (int64_t c = 0; c < n; ++c) {
            totW[c] = outflow_mass[c] / (time_step_min / 60.0);
            Wout5[c] = outlet_mass[c] / (time_step_min / 60.0);
        }
        free(outflow_mass);
        free(outlet_mass);
    } else {
        memcpy(B_t, Bmid, sizeof(double) * n);
    }
    for (int64_t c = 0; c < n; ++c) /* sum(sum(B_t(B_t<0))): one column-major list */
        if (B_t[c] < 0) lost += B_t[c];
    *mass_lost = lost + *mass_lost;
    for (int64_t c = 0; c < n; ++c)
        if (B_t[c] < 0) B_t[c] = 0.0;
    tmin = tmin_of(B_t, dW, n, cell_area, min_Bt, large_timestep); /* :144-159 */
    for (int64_t c = 0; c < n; ++c)
        if (fabs(B_t[c]) < 1e-12) B_t[c] = 0.0; /* :171 */
    double sum_wo = 0.0, sum_qo = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double cw = 0.0, cq = 0.0;
        for (int32_t i = 0; i < ny; ++i) {
            cw += Wout5[IX(i, j)];
            cq += qo[IX(i, j)];
        }
        sum_wo += cw;
        sum_qo += cq;
    }
    for (int64_t c = 0; c < n; ++c) {
        double tq = fmax(totQ[c], 10.0); /* :175 */
        P_conc[c] = fmax(1e6 * (totW[c]) / (tq * cell_area), 0.0);
        Tot_Washed[c] = totW[c] * time_step_min / 60.0 + Tot_Washed[c];
        if (tot_W_out) tot_W_out[c] = totW[c];
    }
    *Out_Conc = fmax(1000.0 * sum_wo / (sum_qo / 1000.0 * cell_area), 0.0);
    *tmin_wq = tmin;
    free(q5); free(W); free(totW); free(totQ); free(dW); free(Bx); free(Wout5); free(Bmid);
    return split;
}
