/*
 * hp2d_d8.c - FP64 CPU restatement of the D8 (eight-direction) routines
 * (TEST INFRASTRUCTURE ONLY).
 *
 *   oracle_bates_8d    Bates_Inertial_8D.m:2-274      (flag_reservoir = 0)
 *   oracle_ca_8d       CA_Routing_8D.m:2-188          (flag_reservoir = 0)
 *   oracle_washoff_8d  build_up_wash_off_8D.m:1-143
 *   oracle_gather      HydroPol2D_Main_While.m:631-660 (outflows become inflows,
 *                      d_t += qin * dt / 60 on the cellular-automata branch)
 *
 * Each routine is restated statement by statement on full-grid page arrays
 * (page p of a 9-page array at offset p*ny*nx, pages in the reference order
 * left, right, up, down, outlet, NE, SE, SW, NW), including the edge rows and
 * columns of the matrix_store / Hf concatenations.  MATLAB's max and min
 * omit NaN, as C's fmax and fmin do.  idx_nan is the set of cells with NaN
 * elevation (the reference's preprocessing sets both together).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hp2d_oracle.h"

#define IX(i, j) ((int64_t)(i) + (int64_t)(j) * ny)
#define PG(p, i, j) ((int64_t)(p) * n + IX(i, j))

/* matrix_store (:54-64 of Bates_Inertial_8D.m, :45-55 of CA_Routing_8D.m):
 * the eight water-surface differences, with the raw edge entries the
 * concatenations leave before they are zeroed; page 5 is left untouched */
static void d8_differences(int32_t ny, int32_t nx, const double *y, double *ms) {
    const int64_t n = (int64_t)ny * nx;
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const double yc = y[IX(i, j)];
            ms[PG(0, i, j)] = j == 0 ? yc : yc - y[IX(i, j - 1)];                /* left  */
            ms[PG(1, i, j)] = j == nx - 1 ? yc : yc - y[IX(i, j + 1)];           /* right */
            ms[PG(2, i, j)] = i == 0 ? yc : yc - y[IX(i - 1, j)];                /* up    */
            ms[PG(3, i, j)] = i == ny - 1 ? yc : yc - y[IX(i + 1, j)];           /* down  */
            /* NE: first row y; rows 2..ny: [y(i,j) - y(i-1,j+1), y(i-1,end)] */
            if (i == 0) ms[PG(5, i, j)] = yc;
            else ms[PG(5, i, j)] = j == nx - 1 ? y[IX(i - 1, j)] : yc - y[IX(i - 1, j + 1)];
            /* SE: rows 1..ny-1: [y(i,j) - y(i+1,j+1), y(i+1,end)]; last row y */
            if (i == ny - 1) ms[PG(6, i, j)] = yc;
            else ms[PG(6, i, j)] = j == nx - 1 ? y[IX(i + 1, j)] : yc - y[IX(i + 1, j + 1)];
            /* SW: rows 1..ny-1: [y(i+1,1), y(i,j) - y(i+1,j-1)]; last row y */
            if (i == ny - 1) ms[PG(7, i, j)] = yc;
            else ms[PG(7, i, j)] = j == 0 ? y[IX(i + 1, j)] : yc - y[IX(i + 1, j - 1)];
            /* NW: first row y; rows 2..ny: [y(i-1,1), y(i,j) - y(i-1,j-1)] */
            if (i == 0) ms[PG(8, i, j)] = yc;
            else ms[PG(8, i, j)] = j == 0 ? y[IX(i - 1, j)] : yc - y[IX(i - 1, j - 1)];
        }
}

/* the outlet page (:84-98 / :59-73): normal slope or critical slope */
static void d8_outlet_page(int32_t ny, int32_t nx, const double *depth, const double *rough,
                           double R, int32_t outlet_type, double slope_outlet,
                           const int64_t *outlet_idx, int32_t n_outlet, double *ms) {
    const int64_t n = (int64_t)ny * nx;
    uint8_t *is_out = calloc((size_t)n, 1);
    for (int32_t k = 0; k < n_outlet; ++k) is_out[outlet_idx[k]] = 1;
    if (outlet_type == 1) {
        for (int64_t c = 0; c < n; ++c) ms[4 * n + c] = slope_outlet * R * (double)is_out[c];
    } else {
        double *S0 = calloc((size_t)n, sizeof(double));
        for (int32_t k = 0; k < n_outlet; ++k) {
            const int64_t c = outlet_idx[k];
            S0[c] = (pow(depth[c], -1.0 / 6.0)) * sqrt(9.81) * rough[c];
        }
        for (int64_t c = 0; c < n; ++c) ms[4 * n + c] = S0[c] * R;
        free(S0);
    }
    free(is_out);
}

/* the edge zeros of :106-108 / :81-83 */
static void d8_zero_edges(int32_t ny, int32_t nx, double *ms) {
    const int64_t n = (int64_t)ny * nx;
    for (int32_t i = 0; i < ny; ++i) {
        ms[PG(0, i, 0)] = 0.0;
        ms[PG(1, i, nx - 1)] = 0.0;
        ms[PG(5, i, nx - 1)] = 0.0;
        ms[PG(6, i, nx - 1)] = 0.0;
        ms[PG(8, i, 0)] = 0.0;
        ms[PG(7, i, 0)] = 0.0;
    }
    for (int32_t j = 0; j < nx; ++j) {
        ms[PG(2, 0, j)] = 0.0;
        ms[PG(3, ny - 1, j)] = 0.0;
        ms[PG(7, ny - 1, j)] = 0.0;
        ms[PG(5, 0, j)] = 0.0;
        ms[PG(6, ny - 1, j)] = 0.0;
        ms[PG(8, 0, j)] = 0.0;
    }
}

int32_t oracle_bates_8d(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                        const double *d_p, const double *rough, double cell_area,
                        double time_step_min, double R, int32_t outlet_type,
                        double slope_outlet, const int64_t *outlet_idx, int32_t n_outlet,
                        const double *outflow_prev, int32_t flag_critical, double *outflow,
                        double *d_t, double *I_tot_end_cell, double *Hf_out) {
    const int64_t n = (int64_t)ny * nx;
    const double h_min = 0.0; /* :35 */
    double *depth = calloc((size_t)n, sizeof(double)), *y = calloc((size_t)n, sizeof(double));
    uint8_t *mask = malloc((size_t)n);
    double *ms = malloc(sizeof(double) * 9 * n), *Hf = malloc(sizeof(double) * 9 * n);
    for (int64_t c = 0; c < n; ++c) {
        depth[c] = fmax(d_p[c] / 1000.0, 0.0); /* :40 */
        mask[c] = depth[c] <= h_min;           /* :43 */
        if (mask[c]) depth[c] = 0.0;           /* :44 */
        y[c] = z[c] + depth[c];                /* :51 */
    }
    d8_differences(ny, nx, y, ms);
    /* Hf (:68-79): max(y) - max(z) of the two cells of a face; the edge
     * entries take one cell's depth (for NE/SE/SW/NW on the first or last
     * column, the depth of the cell one row up or down, as concatenated) */
#define HF2(a, b) (fmax(y[a], y[b]) - fmax(z[a], z[b]))
#define HF1(a) (y[a] - z[a])
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t c = IX(i, j);
            Hf[PG(0, i, j)] = j == 0 ? HF1(c) : HF2(c, IX(i, j - 1));
            Hf[PG(1, i, j)] = j == nx - 1 ? HF1(c) : HF2(c, IX(i, j + 1));
            Hf[PG(2, i, j)] = i == 0 ? HF1(c) : HF2(c, IX(i - 1, j));
            Hf[PG(3, i, j)] = i == ny - 1 ? HF1(c) : HF2(c, IX(i + 1, j));
            Hf[PG(4, i, j)] = HF1(c);
            if (i == 0) Hf[PG(5, i, j)] = HF1(c);
            else Hf[PG(5, i, j)] = j == nx - 1 ? HF1(IX(i - 1, j)) : HF2(c, IX(i - 1, j + 1));
            if (i == ny - 1) Hf[PG(6, i, j)] = HF1(c);
            else Hf[PG(6, i, j)] = j == nx - 1 ? HF1(IX(i + 1, j)) : HF2(c, IX(i + 1, j + 1));
            if (i == ny - 1) Hf[PG(7, i, j)] = HF1(c);
            else Hf[PG(7, i, j)] = j == 0 ? HF1(IX(i + 1, j)) : HF2(c, IX(i + 1, j - 1));
            if (i == 0) Hf[PG(8, i, j)] = HF1(c);
            else Hf[PG(8, i, j)] = j == 0 ? HF1(IX(i - 1, j)) : HF2(c, IX(i - 1, j - 1));
        }
#undef HF1
#undef HF2
    d8_outlet_page(ny, nx, depth, rough, R, outlet_type, slope_outlet, outlet_idx, n_outlet, ms);
    d8_zero_edges(ny, nx, ms);
    for (int64_t k = 0; k < 9 * n; ++k) /* :116 */
        if (ms[k] < h_min || isnan(ms[k])) ms[k] = 0.0;
    const double L = sqrt(2.0) * R;
    for (int64_t k = 0; k < 4 * n; ++k) ms[k] = ms[k] / R; /* :119-121 */
    for (int64_t k = 4 * n; k < 5 * n; ++k) ms[k] = ms[k] / R;
    for (int64_t k = 5 * n; k < 9 * n; ++k) ms[k] = ms[k] / L;
    for (int64_t k = 0; k < 9 * n; ++k) {
        if (ms[k] < 0) ms[k] = 0.0; /* :124 */
        ms[k] = -ms[k];             /* :129 */
    }
    const double dt = time_step_min * 60.0, g = 9.81;
    for (int32_t p = 0; p < 9; ++p)
        for (int64_t c = 0; c < n; ++c) {
            const int64_t k = (int64_t)p * n + c;
            const double qp = outflow_prev[k] / 1000.0 / 3600.0 * (R * R) / R; /* :131 */
            const double h = Hf[k];
            double q = (qp - g * h * dt * ms[k]) /
                       (1.0 + g * h * dt * (rough[c] * rough[c]) * qp / pow(h, 10.0 / 3.0)); /* :137 */
            if (flag_critical == 1) q = fmin(q, h * sqrt(g * h)); /* :209-212 */
            q = R * q / (R * R) * 1000.0 * 3600.0;                /* :215 */
            if (q < 0 || isnan(q) || isinf(q)) q = 0.0;           /* :216-217 */
            if (mask[c]) q = 0.0;                                 /* :218 */
            if (isnan(z[c])) q = NAN;                             /* :220 */
            outflow[k] = q;
            if (Hf_out) Hf_out[k] = h;
        }
    const double kq = 1.0 / 1000.0 * 1.0 / 3600.0;
    for (int64_t c = 0; c < n; ++c) {
        double s = 0.0;
        for (int32_t p = 0; p < 9; ++p) s += kq * outflow[(int64_t)p * n + c];
        I_tot_end_cell[c] = dt * s * (R * R);                    /* :223 */
        d_t[c] = d_tot[c] - I_tot_end_cell[c] / cell_area * 1000.0; /* :269 */
    }
    free(depth); free(y); free(mask); free(ms); free(Hf);
    return 0;
}

int32_t oracle_ca_8d(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                     const double *rough, double cell_area, double time_step_min,
                     const double *h_0, double R, const double *I_prev_in, int32_t outlet_type,
                     double slope_outlet, const int64_t *outlet_idx, int32_t n_outlet,
                     int32_t flag_critical, double d_tolerance, double *outflow, double *d_t,
                     double *I_tot_end_cell) {
    const int64_t n = (int64_t)ny * nx;
    const double d_t_min = d_tolerance / 1000.0, h_min = d_t_min; /* :29, :34 */
    double *depth = calloc((size_t)n, sizeof(double)), *y = calloc((size_t)n, sizeof(double));
    double *ms = malloc(sizeof(double) * 9 * n);
    for (int64_t c = 0; c < n; ++c) {
        depth[c] = fmax(d_tot[c] / 1000.0, 0.0); /* :40 */
        y[c] = z[c] + depth[c];                  /* :42 */
    }
    d8_differences(ny, nx, y, ms);
    d8_outlet_page(ny, nx, depth, rough, R, outlet_type, slope_outlet, outlet_idx, n_outlet, ms);
    d8_zero_edges(ny, nx, ms);
    for (int64_t k = 0; k < 9 * n; ++k) /* :85 */
        if (ms[k] < h_min || isnan(ms[k])) ms[k] = 0.0;
    const double L = sqrt(2.0) * R;
    for (int64_t k = 0; k < 5 * n; ++k) ms[k] = ms[k] / R; /* :88-90 */
    for (int64_t k = 5 * n; k < 9 * n; ++k) ms[k] = ms[k] / L;
    /* :91-96 steepest page (first of equal maxima) and its length */
    double *sf = malloc(sizeof(double) * n), *Dist = malloc(sizeof(double) * n);
    for (int64_t c = 0; c < n; ++c) {
        double best = ms[c];
        int32_t idx = 1;
        for (int32_t p = 1; p < 9; ++p)
            if (ms[(int64_t)p * n + c] > best) {
                best = ms[(int64_t)p * n + c];
                idx = p + 1;
            }
        sf[c] = best;
        Dist[c] = idx <= 5 ? R : sqrt(2.0) * R;
    }
    for (int64_t k = 0; k < 5 * n; ++k) ms[k] = ms[k] * R; /* :98-100 */
    for (int64_t k = 5 * n; k < 9 * n; ++k) ms[k] = ms[k] * L;
    for (int64_t k = 0; k < 9 * n; ++k) ms[k] = cell_area * ms[k]; /* :103 */
    const double cbig = R * R * 1000.0;                              /* :108 */
    double *vmin = malloc(sizeof(double) * n), *wmax = malloc(sizeof(double) * n);
    for (int64_t c = 0; c < n; ++c) {
        double m = NAN, tot = 0.0;
        for (int32_t p = 0; p < 9; ++p) {
            const double v = ms[(int64_t)p * n + c];
            m = fmin(m, v + (double)(v == 0.0) * cbig); /* :109 */
            tot += v;                                   /* :112 */
        }
        vmin[c] = m;
        double w_max = NAN;
        for (int32_t p = 0; p < 9; ++p) {
            double w = ms[(int64_t)p * n + c] / (tot + m); /* :115 */
            if (isnan(w)) w = 0.0;                         /* :116 */
            ms[(int64_t)p * n + c] = w;
            w_max = fmax(w_max, w);                        /* :117 */
        }
        wmax[c] = w_max;
    }
    /* :132-134 the reset of the previous outflow volume (sum(sum()) order) */
    double sprev = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double col = 0.0;
        for (int32_t i = 0; i < ny; ++i) col += I_prev_in[IX(i, j)];
        sprev += col;
    }
    for (int64_t c = 0; c < n; ++c) {
        const double h0 = h_0 ? h_0[c] : 0.0;
        const double man = 1.0 / rough[c] * pow(fmax(depth[c] - h0 / 1000.0, 0.0), 2.0 / 3.0) *
                           sqrt(sf[c]);
        const double vm = flag_critical == 1 ? fmin(sqrt(9.81 * depth[c]), man) : man; /* :121-125 */
        const double Ip = sprev == 0.0 ? 0.9 * cell_area * depth[c] : I_prev_in[c];
        double I = fmin(depth[c] * cell_area,
                        vm * depth[c] / wmax[c] * (Dist[c] * time_step_min * 60.0)); /* :136 */
        I = fmin(I, Ip + vmin[c]);                                                   /* :137 */
        if (d_tot[c] < d_t_min * 1000.0 || isnan(z[c])) I = 0.0;                     /* :138 */
        double tot = 0.0;
        for (int32_t p = 0; p < 9; ++p) {
            const double v = ms[(int64_t)p * n + c] * I; /* :141 */
            tot += v;                                    /* :142 */
            outflow[(int64_t)p * n + c] = v / (time_step_min * 60.0) / cell_area * 1000.0 * 3600.0;
        }
        I_tot_end_cell[c] = tot;
        d_t[c] = d_tot[c] - tot / cell_area * 1000.0; /* :185 */
    }
    free(depth); free(y); free(ms); free(sf); free(Dist); free(vmin); free(wmax);
    return 0;
}

void oracle_gather(int32_t ny, int32_t nx, const double *outflow, int32_t n_pages,
                   double time_step_min, double *qin, double *d_t) {
    const int64_t n = (int64_t)ny * nx;
    const double *L_ = outflow, *Rt = outflow + n, *U = outflow + 2 * n, *D = outflow + 3 * n;
    memset(qin, 0, sizeof(double) * n);
    /* HydroPol2D_Main_While.m:634-637, in that order */
    for (int32_t j = 1; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) qin[IX(i, j)] += Rt[IX(i, j - 1)];
    for (int32_t j = 0; j < nx - 1; ++j)
        for (int32_t i = 0; i < ny; ++i) qin[IX(i, j)] += L_[IX(i, j + 1)];
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 1; i < ny; ++i) qin[IX(i, j)] += D[IX(i - 1, j)];
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny - 1; ++i) qin[IX(i, j)] += U[IX(i + 1, j)];
    if (n_pages == 9) { /* :640-645 */
        const double *NE = outflow + 5 * n, *SE = outflow + 6 * n, *SW = outflow + 7 * n,
                     *NW = outflow + 8 * n;
        for (int32_t j = 0; j < nx - 1; ++j)
            for (int32_t i = 1; i < ny; ++i) qin[IX(i, j)] += SW[IX(i - 1, j + 1)];
        for (int32_t j = 0; j < nx - 1; ++j)
            for (int32_t i = 0; i < ny - 1; ++i) qin[IX(i, j)] += NW[IX(i + 1, j + 1)];
        for (int32_t j = 1; j < nx; ++j)
            for (int32_t i = 0; i < ny - 1; ++i) qin[IX(i, j)] += NE[IX(i + 1, j - 1)];
        for (int32_t j = 1; j < nx; ++j)
            for (int32_t i = 1; i < ny; ++i) qin[IX(i, j)] += SE[IX(i - 1, j - 1)];
    }
    if (d_t) /* :657-658 (flag_CA = 1) */
        for (int64_t c = 0; c < n; ++c) d_t[c] = d_t[c] + qin[c] * time_step_min / 60.0;
}

/* W_out pages (9) for the current B_t (:8-19, :74-90) */
static void rates_8d(int64_t n, const double *C3, const double *C4, const double *q9,
                     const double *B_t, double *B_extra, double cell_area, int32_t model,
                     double Bmin, double Bmax, double min_Bt, double *W, double *totW) {
    const double Br_kg = Bmin / 1000.0 * cell_area, Bm_kg = Bmax / 1000.0 * cell_area;
    for (int64_t c = 0; c < n; ++c) {
        double f = B_t[c];
        if (model == 1) {
            if (B_t[c] * 1000.0 / cell_area <= Bmin) B_extra[c] = 0.0;
            if (B_t[c] * 1000.0 / cell_area >= Bmax) B_extra[c] = Bm_kg;
            f = 1.0 + fmax(B_extra[c] - Br_kg, 0.0);
        }
        const int low = (B_t[c] * 1000.0 / cell_area < min_Bt) || isnan(B_t[c]) || isinf(B_t[c]);
        double tot = 0.0;
        for (int32_t p = 0; p < 9; ++p) {
            double w = C3[c] * pow(q9[(int64_t)p * n + c] / 1000.0 / 3600.0 * cell_area, C4[c]) * f;
            if (low) w = 0.0;
            W[(int64_t)p * n + c] = w;
            tot += w;
        }
        totW[c] = tot;
    }
}

/* dW = sum of the eight inflows - tot_W_out (:34-47, :96-106) */
static void net_8d(int32_t ny, int32_t nx, const double *W, const double *totW, double *dW) {
    const int64_t n = (int64_t)ny * nx;
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const double in_l = j > 0 ? W[PG(1, i, j - 1)] : 0.0;
            const double in_r = j < nx - 1 ? W[PG(0, i, j + 1)] : 0.0;
            const double in_u = i > 0 ? W[PG(3, i - 1, j)] : 0.0;
            const double in_d = i < ny - 1 ? W[PG(2, i + 1, j)] : 0.0;
            const double in_ne = (i > 0 && j < nx - 1) ? W[PG(7, i - 1, j + 1)] : 0.0;
            const double in_se = (i < ny - 1 && j < nx - 1) ? W[PG(8, i + 1, j + 1)] : 0.0;
            const double in_sw = (i < ny - 1 && j > 0) ? W[PG(5, i + 1, j - 1)] : 0.0;
            const double in_nw = (i > 0 && j > 0) ? W[PG(6, i - 1, j - 1)] : 0.0;
            dW[IX(i, j)] = (in_l + in_r + in_u + in_d + in_ne + in_se + in_sw + in_nw) - totW[IX(i, j)];
        }
}

static double tmin_8d(const double *B_t, const double *dW, int64_t n, double cell_area,
                      double min_Bt) {
    double mx = NAN;
    for (int64_t c = 0; c < n; ++c) mx = fmax(mx, dW[c]);
    if (mx <= 0) return 10.0; /* large_timestep (:3) */
    double m = NAN;
    for (int64_t c = 0; c < n; ++c) {
        double mc = B_t[c];
        if (mc * 1000.0 / cell_area < min_Bt) mc = INFINITY;
        const double dc = dW[c] < 0 ? 0.0 : dW[c];
        m = fmin(m, mc / (fabs(dc) + 1e-16));
    }
    return 3600.0 * m;
}

int32_t oracle_washoff_8d(int32_t ny, int32_t nx, const double *C3, const double *C4,
                          const double *q9, double *B_t, double time_step_min, double cell_area,
                          const uint8_t *nan_mask, int32_t model, double *mass_lost,
                          double *Tot_Washed, double Bmin, double Bmax, double min_Bt,
                          double *P_conc, double *Out_Conc, double *tmin_wq, double *tot_W_out) {
    const int64_t n = (int64_t)ny * nx;
    double *W = malloc(sizeof(double) * 9 * n), *totW = malloc(sizeof(double) * n);
    double *totQ = malloc(sizeof(double) * n), *dW = malloc(sizeof(double) * n);
    double *Bx = malloc(sizeof(double) * n), *Bbeg = malloc(sizeof(double) * n);
    double *Wo = malloc(sizeof(double) * n), *Bmid = malloc(sizeof(double) * n);
    memcpy(Bx, B_t, sizeof(double) * n);   /* B_t_extra (:6) */
    memcpy(Bbeg, B_t, sizeof(double) * n); /* B_begin (:7) */
    rates_8d(n, C3, C4, q9, B_t, Bx, cell_area, model, Bmin, Bmax, min_Bt, W, totW);
    for (int64_t c = 0; c < n; ++c) {
        double tq = 0.0;
        for (int32_t p = 0; p < 9; ++p) tq += q9[(int64_t)p * n + c]; /* :21 */
        totQ[c] = tq;
        if (nan_mask && nan_mask[c]) { /* :23-25 */
            totW[c] = 0.0;
            totQ[c] = 0.0;
            for (int32_t p = 0; p < 9; ++p) W[(int64_t)p * n + c] = 0.0;
        }
    }
    net_8d(ny, nx, W, totW, dW);
    double tmin = tmin_8d(B_t, dW, n, cell_area, min_Bt); /* :48-60 */
    double bmid_min = NAN;
    for (int64_t c = 0; c < n; ++c) {
        Bmid[c] = B_t[c] + dW[c] * time_step_min / 60.0; /* :61 */
        bmid_min = fmin(bmid_min, Bmid[c]);
        Wo[c] = W[4 * n + c];
    }
    const int32_t split = bmid_min < 0 || tmin < time_step_min * 60.0;
    if (split) { /* :62-115 */
        const double n_dec = time_step_min * 60.0 / tmin;
        const double n_floor = floor(n_dec);
        if (!(n_floor < 1e7)) {
            free(W); free(totW); free(totQ); free(dW); free(Bx); free(Bbeg); free(Wo); free(Bmid);
            return -1;
        }
        const int64_t steps = (int64_t)n_floor + 1;
        double *om = calloc((size_t)n, sizeof(double)), *ol = calloc((size_t)n, sizeof(double));
        for (int64_t s = 1; s <= steps; ++s) {
            const double dt = s == steps ? (n_dec - n_floor) * tmin : time_step_min * 60.0 / (double)steps;
            rates_8d(n, C3, C4, q9, B_t, Bx, cell_area, model, Bmin, Bmax, min_Bt, W, totW);
            net_8d(ny, nx, W, totW, dW);
            for (int64_t c = 0; c < n; ++c) {
                om[c] = om[c] + totW[c] * dt / 3600.0;
                ol[c] = ol[c] + W[4 * n + c] * dt / 3600.0;
                B_t[c] = B_t[c] + dW[c] * dt / 3600.0;
            }
        }
        for (int64_t c = 0; c < n; ++c) {
            totW[c] = om[c] / (time_step_min / 60.0);
            Wo[c] = ol[c] / (time_step_min / 60.0);
        }
        free(om);
        free(ol);
    } else {
        memcpy(B_t, Bmid, sizeof(double) * n);
    }
    double lost = 0.0;
    for (int64_t c = 0; c < n; ++c)
        if (B_t[c] < 0) lost += B_t[c];
    *mass_lost = lost + *mass_lost;
    for (int64_t c = 0; c < n; ++c)
        if (B_t[c] < 0) B_t[c] = 0.0;
    tmin = tmin_8d(B_t, dW, n, cell_area, min_Bt); /* :120-135 */
    for (int64_t c = 0; c < n; ++c)
        if (fabs(B_t[c]) < 1e-12) B_t[c] = 0.0; /* :136 */
    double sum_wo = 0.0, sum_qo = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double cw = 0.0, cq = 0.0;
        for (int32_t i = 0; i < ny; ++i) {
            cw += Wo[IX(i, j)];
            cq += q9[4 * n + IX(i, j)];
        }
        sum_wo += cw;
        sum_qo += cq;
    }
    for (int64_t c = 0; c < n; ++c) {
        const double tq = fmax(totQ[c], 10.0); /* :137 */
        P_conc[c] = fmax(1e6 * (totW[c]) / (tq * cell_area), 0.0);
        Tot_Washed[c] = (fmax(Bbeg[c] - B_t[c], 0.0)) + Tot_Washed[c]; /* :139 */
        if (tot_W_out) tot_W_out[c] = totW[c];
    }
    *Out_Conc = fmax(1000.0 * sum_wo / (sum_qo / 1000.0 * cell_area), 0.0);
    *tmin_wq = tmin;
    free(W); free(totW); free(totQ); free(dW); free(Bx); free(Bbeg); free(Wo); free(Bmid);
    return split;
}
