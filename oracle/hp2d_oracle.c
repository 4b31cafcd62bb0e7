/*
 * hp2d_oracle.c - FP64 CPU restatement of the HydroPol2D explicit step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle of the HIP library.  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never
 * linked into, called by, or shipped with hydropol2d_amd/libhp2d.so.
 *
 * Each function names the reference file:line it restates.  MATLAB max/min
 * omit NaN (fmax/fmin here), nansum skips NaN, and column-major arrays are
 * indexed (i,j) -> i + j*ny with 0-based i, j.
 */
#include "hp2d_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define G 9.81
#define IX(i, j) ((int64_t)(i) + (int64_t)(j) * ny)

static double *zalloc(int64_t n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* MATLAB nansum over a full array, column-major order (nansum(nansum(X))). */
static double nansum2(int32_t ny, int32_t nx, const double *x) {
    double total = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double col = 0.0;
        for (int32_t i = 0; i < ny; ++i) {
            double v = x[IX(i, j)];
            if (!isnan(v)) col += v;
        }
        total += col;
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* A1  Inertial_Solver.m:19-26 (flag_numerical_scheme == 1, Bates 2010)      */
/* ------------------------------------------------------------------------ */
void oracle_inertial_solver(int32_t scheme, int64_t n_cells, int32_t n_pages,
                            const double *q_p, double dt, const double *Hf,
                            const double *S, const double *n_sq, double *q) {
    (void)scheme;
    const double e73 = 7.0 / 3.0;
    for (int32_t p = 0; p < n_pages; ++p) {
        ORACLE_PAR
        for (int64_t c = 0; c < n_cells; ++c) {
            int64_t o = c + (int64_t)p * n_cells;
            double num = q_p[o] - G * Hf[o] * dt * S[o];
            double den = 1.0 + G * dt * n_sq[c] * fabs(q_p[o]) / pow(Hf[o], e73);
            q[o] = num / den;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* A4  apply_coarse_draining_limiter, Local_Inertial_Model_D4.m:799-858      */
/* Qx, Qy: discharge [m3/s] pages, modified in place.                        */
/* ------------------------------------------------------------------------ */
static void coarse_draining_limiter(int32_t ny, int32_t nx, double *Qx, double *Qy,
                                    double *Q3, const double *d_tot, double dt,
                                    double cell_area) {
    int64_t n = (int64_t)ny * nx;
    double *avail = zalloc(n), *out = zalloc(n), *scale = zalloc(n);
    double dtm = fmax(dt, DBL_EPSILON);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) avail[c] = fmax(d_tot[c] / 1000.0, 0.0) * cell_area / dtm;
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) out[c] = 0.0 + fmax(Qx[c], 0.0);
    ORACLE_PAR
    for (int32_t j = 1; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) out[IX(i, j)] += fmax(-Qx[IX(i, j - 1)], 0.0);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) out[c] += fmax(Qy[c], 0.0);
    ORACLE_PAR
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny - 1; ++i) out[IX(i, j)] += fmax(-Qy[IX(i + 1, j)], 0.0);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        double s = 1.0;
        if (out[c] > avail[c] && out[c] > 0.0) s = avail[c] / out[c];
        if (!isfinite(s)) s = 0.0;
        s = fmax(fmin(s, 1.0), 0.0);
        scale[c] = s;
    }
    if (nx > 1) {
        ORACLE_PAR
        for (int32_t j = 0; j < nx - 1; ++j)
            for (int32_t i = 0; i < ny; ++i) {
                double q = Qx[IX(i, j)];
                if (q > 0) q = q * scale[IX(i, j)];
                else if (q < 0) q = q * scale[IX(i, j + 1)];
                Qx[IX(i, j)] = q;
            }
        ORACLE_PAR
        for (int32_t i = 0; i < ny; ++i) Qx[IX(i, nx - 1)] = 0.0;
    }
    if (ny > 1) {
        ORACLE_PAR
        for (int32_t j = 0; j < nx; ++j)
            for (int32_t i = 1; i < ny; ++i) {
                double q = Qy[IX(i, j)];
                if (q > 0) q = q * scale[IX(i, j)];
                else if (q < 0) q = q * scale[IX(i - 1, j)];
                Qy[IX(i, j)] = q;
            }
        ORACLE_PAR
        for (int32_t j = 0; j < nx; ++j) Qy[IX(0, j)] = 0.0;
    }
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        if (!isfinite(Qx[c])) Qx[c] = 0.0;
        if (!isfinite(Qy[c])) Qy[c] = 0.0;
        if (!isfinite(Q3[c])) Q3[c] = 0.0;
    }
    free(avail);
    free(out);
    free(scale);
}

/* ------------------------------------------------------------------------ */
/* A2-A7  Local_Inertial_Model_D4.m:7-516, branch flag_subgrid = 0,          */
/*        flag_reservoir = 0, h_min = 0, flag_HC = 0.                        */
/* outlet_idx: 0-based linear indices (sub2ind - 1).                         */
/* ------------------------------------------------------------------------ */
int32_t oracle_li_d4(int32_t ny, int32_t nx, int32_t scheme, const double *z,
                     const double *d_tot, const double *roughness_cell,
                     const double *roughness_squared, double cell_area,
                     double time_step_min, double resolution, int32_t outlet_type,
                     double slope_outlet, const int64_t *outlet_idx, int32_t n_outlet,
                     const double *outflow_prev, int32_t flag_critical,
                     double *qout_left, double *qout_right, double *qout_up,
                     double *qout_down, double *outlet_flow, double *d_t,
                     double *I_tot_end_cell, double *outflow, double *Hf) {
    const int64_t n = (int64_t)ny * nx;
    const double R = resolution;
    double *y = zalloc(n), *depth = zalloc(n);
    double *S = zalloc(3 * n);     /* matrix_store */
    double *qprev = zalloc(3 * n); /* outflow_prev in m2/s */
    double *q = zalloc(3 * n);

    /* :31 depth_cell = max(d_tot/1000,0); :43 y = z_dem + depth_cell */
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        depth[c] = fmax(d_tot[c] / 1000.0, 0.0);
        y[c] = z[c] + depth[c];
    }
    /* :145-154 matrix_store = 0*outflow; Hf = 0*outflow (page 3 keeps 0*outflow) */
    ORACLE_PAR
    for (int64_t c = 0; c < 3 * n; ++c) {
        S[c] = 0.0 * outflow_prev[c];
        Hf[c] = 0.0 * outflow_prev[c];
    }
    double *Sx = S, *Sy = S + n, *Hfx = Hf, *Hfy = Hf + n;
    /* :91-92, :123-128, :133-142 */
    ORACLE_PAR
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            int64_t c = IX(i, j);
            if (j < nx - 1) {
                int64_t e = IX(i, j + 1);
                Sx[c] = (y[e] - y[c]) / R;
                double zf = fmax(z[c], z[e]);
                Hfx[c] = fmax(fmax(y[c], y[e]) - zf, 0.0);
            } else {
                Sx[c] = 0.0;
                Hfx[c] = 0.0;
            }
            if (i > 0) {
                int64_t u = IX(i - 1, j);
                Sy[c] = (y[u] - y[c]) / R;
                double zf = fmax(z[u], z[c]);
                Hfy[c] = fmax(fmax(y[u], y[c]) - zf, 0.0);
            } else {
                Sy[c] = 0.0;
                Hfy[c] = 0.0;
            }
        }
    /* :184 Hf(Hf<0)=0 */
    ORACLE_PAR
    for (int64_t c = 0; c < 3 * n; ++c)
        if (Hf[c] < 0) Hf[c] = 0.0;
    /* :198 outflow = outflow/1000/3600*Resolution^2/Resolution  [m2/s] */
    const double R2 = R * R;
    ORACLE_PAR
    for (int64_t c = 0; c < 3 * n; ++c) qprev[c] = outflow_prev[c] / 1000.0 / 3600.0 * R2 / R;
    const double dt = time_step_min * 60.0; /* :201 */
    /* :263-264 */
    oracle_inertial_solver(scheme, n, 3, qprev, dt, Hf, S, roughness_squared, q);
    ORACLE_PAR
    for (int64_t c = 0; c < 3 * n; ++c)
        if (!isfinite(q[c])) q[c] = 0.0;
    /* :272-275 critical velocity cap */
    if (flag_critical == 1) {
        for (int64_t c = 0; c < 3 * n; ++c) {
            double cv = Hf[c] * sqrt(9.81 * Hf[c]);
            q[c] = fmin(fmax(q[c], -cv), cv);
        }
    }
    /* :292-294 velocity cap: q = min(max(q,-10 Hf),10 Hf) */
    ORACLE_PAR
    for (int64_t c = 0; c < 3 * n; ++c) {
        double thr = Hf[c] * 10.0;
        q[c] = fmin(fmax(q[c], -thr), thr);
    }
    /* :304-306 outflow_rate = q.*cell_width; limiter */
    double *Q = q; /* reuse: Q = q*R in place */
    ORACLE_PAR
    for (int64_t c = 0; c < 3 * n; ++c) Q[c] = q[c] * R;
    coarse_draining_limiter(ny, nx, Q, Q + n, Q + 2 * n, d_tot, dt, cell_area);
    /* :309 outflow = outflow_rate./cell_area*1000*3600 */
    ORACLE_PAR
    for (int64_t c = 0; c < 3 * n; ++c) outflow[c] = Q[c] / cell_area * 1000.0 * 3600.0;
    const double *Qx = Q, *Qy = Q + n;
    /* :313-315 Vol_Flux; :390 d_t = d_tot + 1000*Vol_Flux/C_a; :415 */
    const double C_a = R * R; /* :266 */
    ORACLE_PAR
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            int64_t c = IX(i, j);
            double west = (j > 0) ? Qx[IX(i, j - 1)] : 0.0;
            double south = (i < ny - 1) ? Qy[IX(i + 1, j)] : 0.0;
            double vol = dt * (west - Qx[c] - Qy[c] + south);
            double dn = d_tot[c] + 1000.0 * vol / C_a;
            if (dn < 0) dn = 0.0;
            d_t[c] = dn;
        }
    /* :347-350 output fluxes (from matrix_store = outflow) */
    ORACLE_PAR
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            int64_t c = IX(i, j);
            if (qout_left) qout_left[c] = -((j > 0) ? outflow[IX(i, j - 1)] : 0.0);
            if (qout_right) qout_right[c] = outflow[c];
            if (qout_up) qout_up[c] = outflow[n + c];
            if (qout_down) qout_down[c] = -((i < ny - 1) ? outflow[n + IX(i + 1, j)] : 0.0);
        }
    /* :418-498 outlet */
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        if (outlet_flow) outlet_flow[c] = 0.0;
        outflow[2 * n + c] = 0.0;
        Hf[2 * n + c] = 0.0;
    }
    /* vectorised MATLAB semantics: every outlet value is computed from the
     * same pre-outlet d_t, then assigned (duplicates collapse). */
    const double dt_h = time_step_min / 60.0;
    double *qo = zalloc(n_outlet > 0 ? n_outlet : 1);
    for (int32_t k = 0; k < n_outlet; ++k) {
        int64_t c = outlet_idx[k];
        double h_out = fmax(d_t[c], 0.0) / 1000.0;
        double sqrtS;
        if (outlet_type == 1) {
            sqrtS = sqrt(fabs(slope_outlet));
        } else {
            double h_safe = fmax(h_out, 1e-12);
            sqrtS = sqrt(9.81 * roughness_squared[c] * pow(h_safe, -1.0 / 3.0));
            if (!isfinite(sqrtS)) sqrtS = 0.0;
        }
        Hf[2 * n + c] = h_out;
        double q_out = (1.0 / roughness_cell[c]) * R * pow(h_out, 5.0 / 3.0) * sqrtS / cell_area *
                       1000.0 * 3600.0;
        double qmax = fmax(d_t[c] - 1000.0 * 0.0, 0.0) / dt_h;
        q_out = fmin(q_out, qmax);
        if (!isfinite(q_out)) q_out = 0.0;
        qo[k] = q_out;
    }
    for (int32_t k = 0; k < n_outlet; ++k) {
        int64_t c = outlet_idx[k];
        if (outlet_flow) outlet_flow[c] = qo[k];
        outflow[2 * n + c] = qo[k];
    }
    double *dnew = zalloc(n_outlet > 0 ? n_outlet : 1);
    for (int32_t k = 0; k < n_outlet; ++k) dnew[k] = d_t[outlet_idx[k]] - qo[k] * dt_h;
    for (int32_t k = 0; k < n_outlet; ++k) d_t[outlet_idx[k]] = dnew[k];
    free(qo);
    free(dnew);
    /* :513-514 I_tot_end_cell = abs(sum(outflow,3))*dt/1000*1/3600*Resolution^2 */
    if (I_tot_end_cell)
        for (int64_t c = 0; c < n; ++c)
            I_tot_end_cell[c] =
                fabs(outflow[c] + outflow[n + c] + outflow[2 * n + c]) * dt / 1000.0 * 1.0 / 3600.0 * R2;
    free(y);
    free(depth);
    free(S);
    free(qprev);
    free(q);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* A8  refreshing_timestep.m:17-355, inertial D4 branch (flag_CA = 0,        */
/*     flag_D8 = 0, flag_full_momentum = 0).                                 */
/* ------------------------------------------------------------------------ */
double oracle_refresh_dt(int32_t ny, int32_t nx, const double *qout_right,
                         const double *qout_up, const double *Hf,
                         oracle_dt_params *p, double *max_velocity,
                         double *velocity_raster) {
    const int64_t n = (int64_t)ny * nx;
    const double R = p->resolution;
    double *vr = zalloc(n), *vu = zalloc(n), *dts = zalloc(n);
    /* :63-77 */
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        double f1 = Hf[c] * 1000.0, f2 = Hf[n + c] * 1000.0, f3 = Hf[2 * n + c] * 1000.0;
        dts[c] = fmax(fmax(f1 / 1000.0, f2 / 1000.0), f3 / 1000.0);
        double s1 = (f1 < p->depth_tolerance_mm) ? 1e12 : f1;
        double s2 = (f2 < p->depth_tolerance_mm) ? 1e12 : f2;
        vr[c] = (qout_right[c] / 1000.0 / 3600.0) * R / (s1 / 1000.0);
        vu[c] = (qout_up[c] / 1000.0 / 3600.0) * R / (s2 / 1000.0);
    }
    double vmax = NAN;
    double dtw = NAN;
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            int64_t c = IX(i, j);
            double vl = -((j > 0) ? vr[c] : 0.0); /* :73 unshifted quirk */
            double vd = -((i < ny - 1) ? vu[IX(i + 1, j)] : 0.0);
            double raster = fmax(fmax(fmax(fabs(vl), fabs(vr[c])), fabs(vu[c])), fabs(vd));
            if (velocity_raster) velocity_raster[c] = raster;
            double rc = 0.5 * (vr[c] - vl), uc = 0.5 * (vu[c] - vd);
            double tot = hypot(rc, uc);
            if (p->total_velocity) p->total_velocity[c] = tot;
            int stagnant = tot < 0.01;
            vmax = fmax(vmax, tot);
            /* :227-238 */
            double cw = sqrt(9.81 * fmax(dts[c], 0.0));
            cw = fmax(cw, 1e-10);
            if (isinf(cw)) cw = NAN;
            if (stagnant) cw = NAN;
            double cand = (p->flag_adaptive_timestepping == 1) ? p->alfa_min * R / cw
                                                                : p->alfa_min * R / (raster + cw);
            dtw = fmin(dtw, cand);
        }
    *max_velocity = vmax;
    /* :205-221 */
    double dt_adv;
    if (vmax > 0) {
        dt_adv = 1.0 * R / fmax(vmax, 1e-10);
        double tsf = fmax(p->alfa_max - 2.0 * fmax(vmax - 5.0, 0.0), p->alfa_min);
        dt_adv = dt_adv * tsf;
        dt_adv = fmin(dt_adv, p->max_time_step);
    } else {
        dt_adv = p->max_time_step;
    }
    if (isnan(dtw)) dtw = p->max_time_step; /* :240-242 */
    double nt = (p->flag_timestep == 1) ? fmin(dt_adv, dtw) : dtw;
    if (p->flag_waterquality == 1) nt = fmin(1.0 * p->tmin_wq, nt);
    if (isnan(nt) || isinf(nt)) nt = p->max_time_step;
    p->nt_raw = nt;
    /* :344-353 */
    double tcr = fmax(p->time_step_increments * floor(nt / p->time_step_increments), p->min_time_step);
    tcr = fmin(tcr, p->max_time_step);
    free(vr);
    free(vu);
    free(dts);
    return tcr;
}

/* refreshing_timestep.m:31-53, 93-108, 146-150, 180-353 with flags.flag_CA = 1
 * (D4): face velocities are the outflow rates over the cell depth d_t, the
 * cell components the plain differences, the wave depth d_t / 1000.
 * q4: qout_left, qout_right, qout_up, qout_down (ny*nx each). */
double oracle_refresh_dt_ca(int32_t ny, int32_t nx, const double *q4, const double *d_t,
                            oracle_dt_params *p, double *max_velocity, double *velocity_raster) {
    const int64_t n = (int64_t)ny * nx;
    const double R = p->resolution;
    double vmax = NAN, dtw = NAN;
    for (int64_t c = 0; c < n; ++c) {
        const double fd = d_t[c];
        const double safe = (fd < p->depth_tolerance_mm) ? 1e12 : fd;
        const double vl = (q4[c] / 1000.0 / 3600.0) * R / (safe / 1000.0);
        const double vr = (q4[n + c] / 1000.0 / 3600.0) * R / (safe / 1000.0);
        const double vu = (q4[2 * n + c] / 1000.0 / 3600.0) * R / (safe / 1000.0);
        const double vd = (q4[3 * n + c] / 1000.0 / 3600.0) * R / (safe / 1000.0);
        const double raster = fmax(fmax(fmax(fabs(vl), fabs(vr)), fabs(vu)), fabs(vd));
        if (velocity_raster) velocity_raster[c] = raster;
        const double rc = vr - vl, uc = vu - vd; /* :146-150, no 0.5 for CA */
        const double tot = hypot(rc, uc);
        if (p->total_velocity) p->total_velocity[c] = tot;
        const int stagnant = tot < 0.01;
        vmax = fmax(vmax, tot);
        double cw = sqrt(9.81 * fmax(fd / 1000.0, 0.0)); /* depth_timestep = d_t / 1000 */
        cw = fmax(cw, 1e-10);
        if (isinf(cw)) cw = NAN;
        if (stagnant) cw = NAN;
        const double cand = (p->flag_adaptive_timestepping == 1) ? p->alfa_min * R / cw
                                                                 : p->alfa_min * R / (raster + cw);
        dtw = fmin(dtw, cand);
    }
    *max_velocity = vmax;
    double dt_adv;
    if (vmax > 0) {
        dt_adv = 1.0 * R / fmax(vmax, 1e-10);
        const double tsf = fmax(p->alfa_max - 2.0 * fmax(vmax - 5.0, 0.0), p->alfa_min);
        dt_adv = dt_adv * tsf;
        dt_adv = fmin(dt_adv, p->max_time_step);
    } else {
        dt_adv = p->max_time_step;
    }
    if (isnan(dtw)) dtw = p->max_time_step;
    double nt = (p->flag_timestep == 1) ? fmin(dt_adv, dtw) : dtw;
    if (p->flag_waterquality == 1) nt = fmin(1.0 * p->tmin_wq, nt);
    if (isnan(nt) || isinf(nt)) nt = p->max_time_step;
    p->nt_raw = nt;
    double tcr = fmax(p->time_step_increments * floor(nt / p->time_step_increments), p->min_time_step);
    return fmin(tcr, p->max_time_step);
}

/* ------------------------------------------------------------------------ */
/* A18  localAverageRateOverStep + localAdvanceEndStampedCursor +            */
/*      localEstimateConstantDt, update_spatial_BC.m:1095-1185,1240-1250.    */
/* cursor is 1-based as in MATLAB.                                           */
/* ------------------------------------------------------------------------ */
double oracle_rain_avg(const double *times, const double *values, int32_t n, double t0,
                       double t1, int32_t *cursor) {
    if (n <= 0 || !isfinite(t0) || !isfinite(t1) || t1 <= t0) return 0.0;
    double dt_total = t1 - t0, accum = 0.0;
    int32_t cur = *cursor;
    if (cur < 1) cur = 1;
    if (cur > n) cur = n;
    while (cur < n && t0 > times[cur - 1]) cur++;
    double dt_force = (n >= 2) ? times[1] - times[0] : 0.0;
    for (int32_t i = cur; i <= n; ++i) {
        double ts = (i == 1) ? times[0] - dt_force : times[i - 2];
        double te = times[i - 1];
        double overlap = fmax(0.0, fmin(t1, te) - fmax(t0, ts));
        if (overlap > 0 && isfinite(values[i - 1])) accum = accum + values[i - 1] * overlap;
        if (te >= t1) {
            cur = i;
            break;
        }
    }
    *cursor = cur;
    return accum / dt_total;
}

/* ------------------------------------------------------------------------ */
/* A20  HydroPol2D_Main_While.m:106-127, 211-260, 403-994 (LI + GA + WQ).    */
/* ------------------------------------------------------------------------ */
typedef struct ga_tables {
    double ksat, psi, dtheta, Lu_mm, k_out, Tr;
} ga_tables;

/* GA derived parameters, HydroPol2D_preprocessing.m:2249-2253 (lu_m NaN: derived) */
static ga_tables ga_derive(double ksat, double psi, double dtheta, double theta_sat_minus_r,
                           double lu_m) {
    ga_tables g;
    g.ksat = ksat;
    g.psi = psi;
    g.dtheta = dtheta;
    double kr = (1.0 / 75.0 * (sqrt(g.ksat / 25.4)));
    g.Tr = 4.5 / sqrt(g.ksat / 25.4);
    double Lu = 4.0 * sqrt(g.ksat / 25.4);
    Lu = Lu * 2.54 / 100.0;
    if (!isnan(lu_m)) Lu = lu_m;
    g.k_out = theta_sat_minus_r * kr * Lu * 1000.0;
    g.Lu_mm = Lu * 1000.0;
    return g;
}

static ga_tables ga_class(const hp2d_run_inputs *in, int32_t k) {
    return ga_derive(in->ksat_mmh[k], in->psi_mm[k], in->dtheta[k], in->theta_sat_minus_r[k],
                     in->lu_m ? in->lu_m[k] : NAN);
}

void oracle_ga_params(double ksat_mmh, double psi_mm, double dtheta, double theta_sat_minus_r,
                      double lu_m, double *out6) {
    const ga_tables g = ga_derive(ksat_mmh, psi_mm, dtheta, theta_sat_minus_r, lu_m);
    out6[0] = g.ksat; out6[1] = g.psi; out6[2] = g.dtheta;
    out6[3] = g.Lu_mm; out6[4] = g.k_out; out6[5] = g.Tr;
}

/* The legacy explicit Green-Ampt closure of one cell, in the
 * Infiltration_Module slot (HydroPol2D_Main_While_Ensemble.m:243-266 on the
 * first step, :274-298 afterwards; eff_depth as in Infiltration_Module.m:128-172):
 * eff is the post-rain depth [mm], Ip the infiltrated depth before the step
 * [mm], *Tc the recovery clock [h].  Writes *It and returns the infiltration
 * rate f [mm/h]; the caller takes f*dt_h from the depth. */
static double ga_cell(const ga_tables *g, int first, double dt_h, double I_min, double eff,
                      double Ip, double *Tc, double *It) {
    double i_a = eff / fmax(dt_h, 1e-12);
    double f;
    if (first) {
        double C = g->ksat * (1.0 + ((eff + g->psi) * g->dtheta) / Ip);
        f = fmin(C, i_a);
        *It = fmax(Ip + f * dt_h, I_min);
        *Tc = g->Tr;
    } else {
        double C = g->ksat * (1.0 + ((eff + g->psi) * g->dtheta) / Ip);
        if (Ip > g->Lu_mm) C = g->ksat;
        int idx_C = (i_a <= C);
        *Tc = *Tc - dt_h;
        if (i_a > C) *Tc = g->Tr;
        f = fmin(C, i_a);
        *It = fmax(Ip + f * dt_h - g->k_out * (double)idx_C * dt_h, I_min);
        if (*Tc < 0) *It = I_min;
    }
    return f;
}

/* n cells of one soil class (parameters as returned by oracle_ga_params):
 * d [mm] in = post-rain depth, out = depth after infiltration; f_out [mm/h]. */
void oracle_ga_cells(int64_t n, int32_t first_step, double dt_h, double I_min_mm,
                     const double *params6, double *d, const double *I_p, double *I_t,
                     double *T, double *f_out) {
    ga_tables g;
    g.ksat = params6[0]; g.psi = params6[1]; g.dtheta = params6[2];
    g.Lu_mm = params6[3]; g.k_out = params6[4]; g.Tr = params6[5];
    for (int64_t c = 0; c < n; ++c) {
        const double f = ga_cell(&g, first_step, dt_h, I_min_mm, d[c], I_p[c], &T[c], &I_t[c]);
        d[c] = d[c] - f * dt_h;
        if (f_out) f_out[c] = f;
    }
}

/* ------------------------------------------------------------------------ */
/* Human_Instability_Module.m:4-26 (mode 1): the risk of one cell at depth h */
/* [m] and total velocity v [m/s].                                           */
/* ------------------------------------------------------------------------ */
static double hi_risk(const hp2d_run_config *cfg, double h, double v) {
    const double F_person = cfg->hi_weight_kg * cfg->hi_gravity;
    const double F_buoy = (cfg->hi_width1_m * cfg->hi_width2_m) * h * cfg->hi_ro_water * cfg->hi_gravity;
    const double avail = cfg->hi_mu * (fmax(F_person - F_buoy, 0.0));
    const double hydro = fmax(1.0 / 2.0 * (cfg->hi_Cd * cfg->hi_ro_water * cfg->hi_width1_m * h * (v * v)), 0.0);
    double r = 0.0;
    if (avail > 0) r = hydro / avail;
    else if (h > 0) r = 1.0;
    return fmin(fmax(r, 0.0), 1.0);
}

/* ------------------------------------------------------------------------ */
/* update_spatial_BC.m:362-398, 469-518: the inflow hydrographs and the      */
/* rainfall maps of the next step over (t_prev, t).                          */
/* ------------------------------------------------------------------------ */
typedef struct forcing_state {
    int32_t n_maps, n_gauges, cursor_maps, cursor_inflow;
    double *dp_map;              /* ny*nx rain depth of the next step [mm] (maps) */
    double agg[HP2D_MAX_INFLOW_GAUGES]; /* BC_States.delta_inflow_agg [mm] */
} forcing_state;

/* localAverageSpatialGaugeRainOverStep / localAverageInputRainMapOverStep's
 * slow path: accum = sum overlap_i * map_i, avg = accum / dt_total, then
 * delta_p_agg = avg * dt / 60 (:486, :514) */
static void forcing_rain_maps(const hp2d_run_inputs *in, int64_t n, const uint8_t *is_nan,
                              double t0, double t1, forcing_state *fs) {
    const double dt_step = fmax(t1 - t0, 0.0);
    const int32_t K = fs->n_maps;
    const double *times = in->rain_map_time_min;
    for (int64_t c = 0; c < n; ++c) fs->dp_map[c] = is_nan[c] ? NAN : 0.0;
    if (K <= 0 || !(t1 > t0)) return;
    const double dt_total = t1 - t0;
    int32_t cur = fs->cursor_maps;
    if (cur < 1) cur = 1;
    if (cur > K) cur = K;
    while (cur < K && t0 > times[cur - 1]) cur++;
    const double dt_force = (K >= 2 && times[1] - times[0] > 0) ? times[1] - times[0] : 0.0;
    for (int64_t c = 0; c < n; ++c) fs->dp_map[c] = 0.0;
    for (int32_t i = cur; i <= K; ++i) {
        const double ts = (i == 1) ? times[0] - dt_force : times[i - 2];
        const double te = times[i - 1];
        const double overlap = fmax(0.0, fmin(t1, te) - fmax(t0, ts));
        if (overlap > 0) {
            const double *m = in->rain_maps_mmh + (int64_t)(i - 1) * n;
            for (int64_t c = 0; c < n; ++c) fs->dp_map[c] = fs->dp_map[c] + m[c] * overlap;
        }
        if (te >= t1) {
            cur = i;
            break;
        }
    }
    fs->cursor_maps = cur;
    for (int64_t c = 0; c < n; ++c)
        fs->dp_map[c] = is_nan[c] ? NAN : fs->dp_map[c] / dt_total * (dt_step / 60.0);
}

/* :362-378: one shared cursor over the gauges' common time stamps */
static void forcing_inflow(const hp2d_run_inputs *in, double t0, double t1, forcing_state *fs) {
    for (int32_t g = 0; g < fs->n_gauges; ++g) fs->agg[g] = 0.0;
    const double dt_step = fmax(t1 - t0, 0.0);
    if (!(dt_step > 0)) return;
    for (int32_t g = 0; g < fs->n_gauges; ++g) {
        const double avg = oracle_rain_avg(in->inflow_time_min,
                                           in->inflow_rate_mmh + (int64_t)g * in->n_inflow_times,
                                           in->n_inflow_times, t0, t1, &fs->cursor_inflow);
        fs->agg[g] = avg * dt_step / 60.0;
    }
}

int32_t oracle_run(const hp2d_run_config *cfg, const hp2d_run_inputs *in,
                   oracle_run_out *out) {
    const int32_t ny = cfg->ny, nx = cfg->nx;
    const int64_t n = (int64_t)ny * nx;
    const double R = cfg->resolution, A = R * R;
    /* CA routing (flags.flag_CA = 1) runs without the wash-off, as the device
     * loop does */
    const int do_ca = cfg->routing == HP2D_ROUTING_CA;
    if (!do_ca && cfg->routing != HP2D_ROUTING_INERTIAL) return HP2D_E_UNSUPPORTED;
    if (do_ca && cfg->flag_waterquality == 1) return HP2D_E_UNSUPPORTED;
    if (cfg->flag_groundwater == 1 && cfg->flag_infiltration != 1 && cfg->flag_infiltration != 2)
        return HP2D_E_UNSUPPORTED;

    double *z = zalloc(n), *nman = zalloc(n), *n2 = zalloc(n);
    double *d_t = zalloc(n), *d_p = zalloc(n), *d_tot = zalloc(n);
    double *outflow = zalloc(3 * n), *outflow_prev = zalloc(3 * n), *Hf = zalloc(3 * n);
    double *qr = zalloc(n), *qu = zalloc(n), *qof = zalloc(n), *dmax = zalloc(n);
    double *I_t = zalloc(n), *I_p = zalloc(n), *Tclk = zalloc(n), *vraster = zalloc(n);
    double *area_nan = zalloc(n), *ga_f = zalloc(n);
    double *vtot = zalloc(n), *max_risk = zalloc(n), *risk = zalloc(n);
    /* CA: the outflow pages (left, right, up, down, outlet), the gathered
     * inflows and CA_States.I_tot_end_cell (kept through rejected steps, as the
     * reference's catch block keeps it) */
    double *ca_q = do_ca ? zalloc(5 * n) : NULL, *ca_qin = do_ca ? zalloc(n) : NULL;
    double *ca_itot = do_ca ? zalloc(n) : NULL;
    const int do_hi = cfg->flag_human_instability == 1;
    const int do_maps = cfg->record_time_maps_min > 0;
    int32_t last_record_maps = 1; /* recording_parameters.last_record_maps (preprocessing.m:2396) */
    uint8_t *is_nan = (uint8_t *)calloc((size_t)n, 1);
    int64_t *oidx = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    int32_t n_out = 0;
    int32_t n_classes = cfg->n_soil_classes > 0 ? cfg->n_soil_classes : 1;
    ga_tables *ga = (ga_tables *)calloc((size_t)n_classes, sizeof(ga_tables));
    const int do_ga = cfg->flag_infiltration == 1;
    /* the van Genuchten Infiltration_Module in the same slot (flag_infiltration = 2) */
    const int do_vg = cfg->flag_infiltration == 2;
    const int do_inf = do_ga || do_vg;
    if (do_ga)
        for (int32_t k = 0; k < n_classes; ++k) ga[k] = ga_class(in, k);

    for (int64_t c = 0; c < n; ++c) {
        z[c] = in->elevation[c];
        nman[c] = in->roughness[c];
        n2[c] = nman[c] * nman[c]; /* Main_While.m:187 roughness.^2 */
        is_nan[c] = isnan(z[c]);
        d_t[c] = in->d0 ? in->d0[c] : 0.0;
        if (is_nan[c]) d_t[c] = NAN;
        area_nan[c] = is_nan[c] ? NAN : A; /* mass_balance_check.m:12-13 */
        if (in->outlet_mask && in->outlet_mask[c]) oidx[n_out++] = c; /* find() order */
        if (do_inf) {
            I_t[c] = in->I0_mm ? in->I0_mm[c] : in->I_min_mm;
            if (is_nan[c]) I_t[c] = NAN;
        }
    }
    memcpy(d_p, d_t, (size_t)n * sizeof(double));
    memcpy(I_p, I_t, (size_t)n * sizeof(double));

    /* Water quality: class-indexed C1..C4 and the exponential build-up of
     * HydroPol2D_preprocessing.m:2170-2174, initial_mass of :2319 */
    const int do_wq = cfg->flag_waterquality == 1;
    double *B = NULL, *C3m = NULL, *C4m = NULL, *Pc = NULL, *TW = NULL, *ql = NULL, *qd = NULL;
    double initial_mass = 0.0, mass_lost = 0.0, out_conc = 0.0, tmin_wq = 0.0;
    double mass_outlet = 0.0, vol_outlet = 0.0;
    int64_t wq_split = 0;
    if (do_wq) {
        B = zalloc(n); C3m = zalloc(n); C4m = zalloc(n); Pc = zalloc(n); TW = zalloc(n);
        ql = zalloc(n); qd = zalloc(n);
        for (int64_t c = 0; c < n; ++c) {
            const int32_t cls = in->lulc_class ? in->lulc_class[c] : 0;
            const double B0 = in->wq_C1[cls] * (1.0 - pow(2.718281828459045, -in->wq_C2[cls] * in->wq_ADD_days));
            B[c] = B0 * A / 1e4 * (is_nan[c] ? 0.0 : 1.0);
            C3m[c] = in->wq_C3[cls];
            C4m[c] = in->wq_C4[cls];
        }
        for (int64_t c = 0; c < n; ++c) /* sum(sum(B_t(~isinf(B_t)))), column-major */
            if (!isinf(B[c])) initial_mass += B[c];
    }

    /* Groundwater: Groundwater_Module (synchronous, Boussinesq baseflow) called
     * every gw_every_steps accepted steps over the interval since its last call,
     * before the next step's hydrology; per-cell van Genuchten parameters from
     * the soil classes (initialize_groundwater_scheduler.m:38-66 state) */
    const int do_gw = cfg->flag_groundwater == 1 && do_inf;
    double *gw_h = NULL, *gw_sd = NULL, *gw_ts = NULL, *gw_tr = NULL, *gw_nv = NULL;
    double *gw_al = NULL, *gw_ks = NULL, *gw_kg = NULL, *gw_Sy = NULL, *gw_q = NULL;
    double *gw_pn = NULL, *gw_pr = NULL, *gw_pc = NULL, *gw_cum = NULL;
    uint8_t *gw_dom = NULL;
    oracle_gwm_params gp;
    oracle_gwm_state gs;
    memset(&gp, 0, sizeof(gp));
    memset(&gs, 0, sizeof(gs));
    int64_t gw_last = 0;
    double gw_elapsed_min = 0.0, gw_err = 0.0;
    if (do_gw) {
        gw_h = zalloc(n); gw_sd = zalloc(n); gw_ts = zalloc(n); gw_tr = zalloc(n);
        gw_nv = zalloc(n); gw_al = zalloc(n); gw_ks = zalloc(n); gw_kg = zalloc(n);
        gw_Sy = zalloc(n); gw_q = zalloc(n); gw_pn = zalloc(n); gw_pr = zalloc(n);
        gw_pc = zalloc(n); gw_cum = zalloc(n);
        gw_dom = (uint8_t *)calloc((size_t)n, 1);
        for (int64_t c = 0; c < n; ++c) {
            const int32_t cls = in->soil_class ? in->soil_class[c] : 0;
            gw_h[c] = in->gw_h0[c];
            gw_sd[c] = in->gw_soil_depth[c];
            gw_ts[c] = in->gw_theta_sat[cls];
            gw_tr[c] = in->gw_theta_r[cls];
            gw_nv[c] = in->gw_n_vg[cls];
            gw_al[c] = in->gw_alpha_vg[cls];
            gw_ks[c] = in->ksat_mmh[cls];
            gw_kg[c] = in->gw_ksat_mmh;
            gw_Sy[c] = in->gw_Sy;
            gw_dom[c] = is_nan[c] ? 0 : 1;
            const double p0 = is_nan[c] ? NAN : 0.0;
            gw_pn[c] = p0;
            gw_pr[c] = p0;
            gw_pc[c] = p0;
        }
        gp.ny = ny;
        gp.nx = nx;
        gp.resolution = R;
        gp.routing_time_min = cfg->routing_time_min;
        gp.flag_baseflow = 1;
        gp.flag_async = 0;
        gp.target_dt_min = 1440.0;
        gp.min_dt_min = 1.0;
        gp.max_head_change_m = 0.25;
        gp.courant = cfg->gw_courant;
        gp.l_vg = in->gw_l_vg;
        gp.elevation = z; gp.soil_depth = gw_sd; gp.theta_sat = gw_ts; gp.theta_r = gw_tr;
        gp.n_vg = gw_nv; gp.alpha_vg = gw_al; gp.ksat_mmh = gw_ks; gp.ksat_gw_mmh = gw_kg;
        gp.Sy = gw_Sy; gp.domain = gw_dom;
        gs.pending_net = gw_pn;
        gs.pending_recharge = gw_pr;
        gs.pending_capillary = gw_pc;
    }

    /* running_control */
    double min_dt0 = cfg->min_time_step_s, max_dt = cfg->max_time_step_s;
    double min_dt = min_dt0;
    double rec = cfg->record_time_hydrographs_min;
    double routing_time = cfg->routing_time_min;
    oracle_dt_params dp = {R, cfg->alfa_min, cfg->alfa_max, 0, 0, cfg->time_step_increments_s,
                           cfg->depth_tolerance_mm, cfg->flag_timestep,
                           cfg->flag_adaptive_timestepping, 0, 0.0, 0.0, NULL};

    /* Main_While.m:109-127 + preprocessing.m:2469,2615,2618 */
    int64_t k = 1;
    double time_step = min_dt0 / 60.0;
    double t = min_dt0 / 60.0;
    double progress = (t / routing_time) * 100.0;
    double t_previous = 0.0;
    int32_t cursor = 1;
    double delta_p_agg = 0.0;
    forcing_state fs;
    memset(&fs, 0, sizeof(fs));
    fs.n_maps = cfg->n_rain_maps > 0 ? cfg->n_rain_maps : 0;
    fs.n_gauges = cfg->flag_inflow == 1 ? cfg->n_inflow_gauges : 0;
    fs.cursor_maps = fs.cursor_inflow = 1;
    if (fs.n_gauges > HP2D_MAX_INFLOW_GAUGES || (fs.n_gauges > 0 && !in->inflow_gauge) ||
        (fs.n_maps > 0 && (!in->rain_maps_mmh || !in->rain_map_time_min)))
        return HP2D_E_ARG;
    if (fs.n_maps > 0) fs.dp_map = zalloc(n);
    {
        double dt_step = fmax(t - t_previous, 0.0);
        double avg = 0.0;
        if (dt_step > 0) {
            avg = oracle_rain_avg(in->rain_time_min, in->rain_intensity_mmh, in->n_rain, t_previous,
                                  t, &cursor);
            if (isnan(avg)) avg = 0.0;
            delta_p_agg = avg * dt_step / 60.0;
        }
        if (fs.n_maps > 0) forcing_rain_maps(in, n, is_nan, t_previous, t, &fs);
        if (fs.n_gauges > 0) forcing_inflow(in, t_previous, t, &fs);
    }
    double cum_inflow = 0.0;
    t_previous = 0.0;
    max_dt = fmin(rec * 60.0, max_dt);
    min_dt = fmin(rec * 60.0, min_dt);
    double time_save_previous = 0.0;
    double tcr = 0.0;

    /* Main_While.m:248-260 initial storage */
    double S_prev;
    {
        double *tmp = zalloc(n);
        for (int64_t c = 0; c < n; ++c) tmp[c] = A * d_t[c] / 1000.0;
        double S_p = nansum2(ny, nx, tmp);
        double S_UZ = 0.0;
        if (do_inf) {
            for (int64_t c = 0; c < n; ++c) tmp[c] = A * I_t[c] / 1000.0;
            S_UZ = nansum2(ny, nx, tmp);
        }
        double S_GW = 0.0;
        if (do_gw) { /* Main_While.m:258 */
            for (int64_t c = 0; c < n; ++c) tmp[c] = A * gw_Sy[c] * (gw_h[c] - (z[c] - gw_sd[c]));
            S_GW = nansum2(ny, nx, tmp);
        }
        S_prev = 0.0 + S_p + S_UZ + S_GW;
        free(tmp);
    }
    int32_t catch_index = 1;
    double next_report = cfg->report_every_percent;
    double outflow_volume = 0.0, inflow_volume = 0.0, volume_error = 0.0, runoff_mm = 0.0;
    int32_t n_reports = 0;
    int32_t last_record = 1;
    double drainage_area = 0.0;
    for (int64_t c = 0; c < n; ++c)
        if (!is_nan[c]) drainage_area += A;
    double cum_rain = 0.0, cum_inf = 0.0, cum_out = 0.0;
    int64_t rejected = 0, accepted = 0;
    double max_velocity = 0.0;
    int64_t inject = out->inject_fail_step;
    int32_t n_hydro = 0;
    int32_t n_dt = 0;
    double *pat = (double *)in->rain_pattern;

    while (t <= (routing_time + min_dt / 60.0)) {
        if (out->max_steps > 0 && accepted >= out->max_steps) break;
        int failed = 0, late = 0;
        /* ------- groundwater update at the end of every gw_every_steps accepted steps ------- */
        if (do_gw && cfg->gw_every_steps > 0 && accepted - gw_last >= cfg->gw_every_steps) {
            int32_t ran = 0;
            double dtg = 0.0;
            gp.time_step_min = gw_elapsed_min;
            gp.t_min = t;
            int32_t rc = oracle_groundwater_module(&gp, &gs, gw_h, I_t, d_t, gw_cum, NULL, gw_q,
                                                   &gw_err, &ran, &dtg);
            if (rc != 0) return rc;
            memcpy(d_p, d_t, (size_t)n * sizeof(double));
            memcpy(I_p, I_t, (size_t)n * sizeof(double));
            gw_last = accepted;
            gw_elapsed_min = 0.0;
        }
        /* ---------------- Hydrological_Model.m:15-55 ---------------- */
        /* interception_module.m:59-71, Snow_Module.m:97-99: d_t += Eff_Rainfall */
        double rain_vol_step = 0.0, inf_vol_step = 0.0;
        const double dt_used_min = time_step;
        for (int64_t c = 0; c < n; ++c) {
            double eff = fs.n_maps > 0 ? fs.dp_map[c] : (pat ? pat[c] * delta_p_agg : delta_p_agg);
            if (is_nan[c]) eff = NAN;
            else rain_vol_step += eff / 1000.0 * A;
            d_t[c] = d_t[c] + eff;
        }
        /* Green-Ampt in the Infiltration_Module slot (legacy closure,
         * HydroPol2D_Main_While_Ensemble.m:243-298; eff_depth as in
         * Infiltration_Module.m:128-172) */
        if (do_ga) {
            double dt_h = time_step / 60.0;
            ORACLE_PAR
            for (int64_t c = 0; c < n; ++c) {
                int32_t cls = in->soil_class ? in->soil_class[c] : 0;
                double f = ga_cell(&ga[cls], k == 1, dt_h, in->I_min_mm, d_t[c], I_p[c], &Tclk[c],
                                   &I_t[c]);
                if (is_nan[c]) {
                    I_t[c] = NAN;
                    f = NAN;
                }
                d_t[c] = d_t[c] - f * dt_h;
                ga_f[c] = f;
            }
            for (int64_t c = 0; c < n; ++c)
                if (!isnan(ga_f[c])) inf_vol_step += ga_f[c] * dt_h / 1000.0 * A;
        }
        /* the van Genuchten closure (Infiltration_Module.m:182-473) on the head of
         * the last groundwater update (gw_h0 without the module) */
        if (do_vg) {
            const double dt_h = time_step / 60.0;
            const double Ltop = fmax(cfg->vg_Ltop, 1e-6);
            const double wK = fmin(fmax(cfg->vg_Ksurf_weight, 0.0), 1.0);
            ORACLE_PAR
            for (int64_t c = 0; c < n; ++c) {
                if (is_nan[c]) {
                    I_t[c] = NAN;
                    ga_f[c] = NAN;
                    continue;
                }
                const int32_t cls = in->soil_class ? in->soil_class[c] : 0;
                const double h = do_gw ? gw_h[c] : in->gw_h0[c];
                const double sd = in->gw_soil_depth[c];
                const double gw_depth = h - (z[c] - sd);
                ga_f[c] = oracle_vg_cell(dt_h, A, A, sd - gw_depth, in->gw_theta_sat[cls],
                                         in->gw_theta_r[cls], in->gw_n_vg[cls],
                                         in->gw_alpha_vg[cls], in->ksat_mmh[cls],
                                         in->impervious && in->impervious[c], in->gw_l_vg, Ltop,
                                         cfg->vg_dh_max, wK, &d_t[c], &I_t[c]);
            }
            for (int64_t c = 0; c < n; ++c)
                if (!isnan(ga_f[c])) inf_vol_step += ga_f[c] * dt_h / 1000.0 * A;
        }
        /* Hydrological_Model.m:30: d_t += BC_States.inflow (NaN outside the domain) */
        double inflow_vol_step = 0.0;
        if (fs.n_gauges > 0) {
            for (int64_t c = 0; c < n; ++c) {
                const int32_t g = in->inflow_gauge[c];
                double q = (g >= 1 && g <= fs.n_gauges) ? fs.agg[g - 1] : 0.0;
                if (is_nan[c]) q = NAN;
                else inflow_vol_step += q / 1000.0 * A;
                d_t[c] = d_t[c] + q;
            }
        }
        /* :45-55 */
        double dmin = NAN;
        for (int64_t c = 0; c < n; ++c) dmin = fmin(dmin, d_t[c]);
        if (dmin < -1e-8) {
            catch_index = catch_index + 1;
        } else {
            for (int64_t c = 0; c < n; ++c)
                if (d_t[c] < 1e-6) d_t[c] = 0.0;
        }
        for (int64_t c = 0; c < n; ++c) {
            if (d_t[c] < 0) d_t[c] = 0.0;
            d_tot[c] = d_t[c];
        }
        if (do_ca) {
            /* ---------------- CA_Routing (:461-465), gather + mass balance (:631-660) ---------------- */
            oracle_ca_routing(ny, nx, z, d_tot, nman, A, time_step, in->ca_h0_mm, R, ca_itot,
                              cfg->outlet_type, cfg->slope_outlet, oidx, n_out, cfg->flag_critical,
                              ca_q, ca_q + n, ca_q + 2 * n, ca_q + 3 * n, ca_q + 4 * n, d_t, ca_itot);
            oracle_gather(ny, nx, ca_q, 5, time_step, ca_qin, d_t);
            /* the outflow pages kept as the local-inertial ones are: right, up, outlet */
            memcpy(outflow, ca_q + n, (size_t)n * sizeof(double));
            memcpy(outflow + n, ca_q + 2 * n, (size_t)n * sizeof(double));
            memcpy(outflow + 2 * n, ca_q + 4 * n, (size_t)n * sizeof(double));
            memcpy(qof, ca_q + 4 * n, (size_t)n * sizeof(double));
        } else {
            /* ---------------- Local_Inertial_Model_D4 (:596-615) ---------------- */
            oracle_li_d4(ny, nx, cfg->flag_numerical_scheme, z, d_tot, nman, n2, A, time_step, R,
                         cfg->outlet_type, cfg->slope_outlet, oidx, n_out, outflow_prev,
                         cfg->flag_critical, ql, qr, qu, qd, qof, d_t, NULL, outflow, Hf);
        }
        if (inject > 0 && k == inject) {
            failed = 1;
            inject = 0;
        }
        /* ---------------- build_up_wash_off (:683-695) ---------------- */
        if (do_wq && !failed) {
            int32_t rc = oracle_washoff(ny, nx, C3m, C4m, ql, qr, qu, qd, qof, B, time_step, A,
                                        is_nan, cfg->flag_wq_model, &mass_lost, TW,
                                        in->wq_Bmin, in->wq_Bmax, in->wq_min_Bt, Pc, &out_conc,
                                        &tmin_wq, NULL);
            if (rc < 0) return HP2D_E_ARG;
            wq_split += rc;
            double sB = 0.0; /* sum(sum(B_t(~isinf(B_t)))) > 1.2*initial_mass: error (:691-695) */
            for (int64_t c = 0; c < n; ++c)
                if (!isinf(B[c])) sB += B[c];
            if (sB > 1.2 * initial_mass) failed = 1;
        }
        if (!failed) {
            /* ---------------- refreshing_timestep (:693-705) ---------------- */
            int64_t pos_save = (int64_t)ceil((t * 60.0) / cfg->time_step_change_s);
            double time_save = (double)(pos_save - 1) * cfg->time_step_change_s / 60.0;
            double delta_time_save = time_save - time_save_previous;
            time_save_previous = time_save;
            if (delta_time_save > 0 || k == 1) {
                dp.max_time_step = max_dt;
                dp.min_time_step = min_dt;
                dp.flag_waterquality = do_wq;  /* refreshing_timestep.m:329-332 */
                dp.tmin_wq = tmin_wq;
                dp.total_velocity = vtot;
                tcr = do_ca ? oracle_refresh_dt_ca(ny, nx, ca_q, d_t, &dp, &max_velocity, vraster)
                            : oracle_refresh_dt(ny, nx, qr, qu, Hf, &dp, &max_velocity, vraster);
                if (out->dt_follow && n_dt < out->n_follow && out->dt_follow[n_dt] != tcr) {
                    const double inc = cfg->time_step_increments_s;
                    const double u = dp.nt_raw / inc;
                    const int tie = fabs(u - nearbyint(u)) <= 1e-9 * fmax(1.0, fabs(u)) &&
                                    fabs(out->dt_follow[n_dt] - tcr) <= inc * (1.0 + 1e-9);
                    if (tie) {
                        tcr = out->dt_follow[n_dt];
                        out->n_dt_ties++;
                    } else {
                        out->n_dt_mismatch++;
                        if (out->dt_force) tcr = out->dt_follow[n_dt];
                    }
                }
                time_step = tcr / 60.0;
            } else {
                tcr = time_step * 60.0;
            }
            /* :711-715 first report check (consumes thresholds) */
            if (progress >= next_report || progress >= 100.0) next_report += cfg->report_every_percent;
            /* :718 Human_Instability_Module (mode 1) on the new depths and velocities */
            if (do_hi) {
                for (int64_t c = 0; c < n; ++c) {
                    risk[c] = is_nan[c] ? NAN : hi_risk(cfg, d_t[c] / 1000.0, vtot[c]);
                    max_risk[c] = fmax(risk[c], k == 1 ? (is_nan[c] ? NAN : 0.0) : max_risk[c]);
                }
            }
            /* update_spatial_BC (:721) */
            {
                double dt_step = fmax(t - t_previous, 0.0);
                delta_p_agg = 0.0;
                if (dt_step > 0) {
                    double avg = oracle_rain_avg(in->rain_time_min, in->rain_intensity_mmh,
                                                 in->n_rain, t_previous, t, &cursor);
                    if (isnan(avg)) avg = 0.0;
                    delta_p_agg = avg * dt_step / 60.0;
                }
                if (fs.n_maps > 0) forcing_rain_maps(in, n, is_nan, t_previous, t, &fs);
                if (fs.n_gauges > 0) forcing_inflow(in, t_previous, t, &fs);
            }
            /* EMC (:724-731), with the time_step refreshing_timestep just set */
            if (do_wq) {
                double q_cols = 0.0, q_all = 0.0;
                for (int32_t j = 0; j < nx; ++j) {
                    double col = 0.0;
                    for (int32_t i = 0; i < ny; ++i)
                        if (!isnan(qof[IX(i, j)])) col += qof[IX(i, j)];
                    q_cols += col / 1000.0 / 3600.0 * 1000.0;
                    q_all += col;
                }
                mass_outlet = fmax(mass_outlet + out_conc * q_cols * (time_step * 60.0), 0.0);
                vol_outlet = fmax(q_all / 1000.0 / 3600.0 * 1000.0 * (time_step * 60.0) + vol_outlet, 0.0);
            }
            /* :734-740 */
            if (k == 1) t_previous = tcr / 60.0;
            else t_previous = t;
            double t_save = t + tcr / 60.0;
            /* save_output_maps.m:266-269, 409-411, 481-490 */
            int32_t actual = 0;
            {
                int32_t n_rec = (int32_t)floor(routing_time / rec + 1e-12) + 1;
                for (int32_t r = 1; r <= n_rec; ++r)
                    if ((double)(r - 1) * rec <= t_save) actual = r;
            }
            double qsum = 0.0;
            for (int32_t kk = 0; kk < n_out; ++kk) qsum += qof[oidx[kk]] * A; /* nansum(outlet_flow.*C_a) */
            double dsum = 0.0;
            for (int32_t kk = 0; kk < n_out; ++kk) dsum += d_t[oidx[kk]];
            double dmean = n_out > 0 ? dsum / n_out : NAN;
            if (k == 1) {
                if (n_hydro < out->n_hydro_max) {
                    out->hydro_q[0] = qsum / 1000.0 / 3600.0;
                    out->hydro_t[0] = tcr / 60.0;
                    out->hydro_depth[0] = dmean;
                    if (n_hydro < 1) n_hydro = 1;
                }
            } else if (actual - last_record > 0) {
                int32_t ts = actual;
                if (ts - 1 < out->n_hydro_max) {
                    out->hydro_q[ts - 1] = qsum / 1000.0 / 3600.0;
                    out->hydro_t[ts - 1] = t;
                    out->hydro_depth[ts - 1] = dmean;
                    if (n_hydro < ts) n_hydro = ts;
                }
            }
            last_record = actual;
            /* save_output_maps.m:24-27, 148-258: a map record when the record state advances */
            if (do_maps) {
                const double recm = cfg->record_time_maps_min;
                int32_t actual_m = 0;
                for (int32_t j = 0; (double)j * recm <= routing_time + 1e-12; ++j)
                    if ((double)j * recm < t_save) actual_m = j + 1;
                const int32_t delta_m = actual_m - last_record_maps;
                last_record_maps = actual_m;
                if (delta_m > 0 && out->rec && out->n_rec < out->n_rec_max) {
                    double *R0 = out->rec + (int64_t)out->n_rec * HP2D_MAP_FIELDS * n;
                    for (int64_t c = 0; c < n; ++c) {
                        /* snisb_velocity_hazard_snapshot (save_output_maps.m:832-848) */
                        const double dm = fmax(d_t[c] / 1000.0, 0.0);
                        double v = vraster[c];
                        if (!isfinite(v)) v = 0.0;
                        double hz = dm * v;
                        if (dm < 0.01) hz = 0.0;
                        if (!isfinite(hz)) hz = 0.0;
                        R0[HP2D_MAP_DEPTH * n + c] = d_t[c];
                        R0[HP2D_MAP_VELOCITY * n + c] = is_nan[c] ? NAN : v;
                        R0[HP2D_MAP_HAZARD * n + c] = is_nan[c] ? NAN : hz;
                        R0[HP2D_MAP_RISK * n + c] = do_hi ? risk[c] : NAN;
                        R0[HP2D_MAP_INFILTRATION * n + c] = do_inf ? I_t[c] : NAN;
                        R0[HP2D_MAP_GW_DEPTH * n + c] = do_gw ? gw_h[c] - z[c] + gw_sd[c] : NAN;
                        if (do_wq) {
                            /* I_tot_end_cell of the step (Local_Inertial_Model_D4.m:513-514)
                             * over the refreshed dt (save_output_maps.m:254-256) */
                            const double itot = fabs(outflow[c] + outflow[n + c] + outflow[2 * n + c]) *
                                                (dt_used_min * 60.0) / 1000.0 * 1.0 / 3600.0 * (R * R);
                            R0[HP2D_MAP_POL_CONC * n + c] = Pc[c];
                            R0[HP2D_MAP_POL_MASS * n + c] = B[c];
                            R0[HP2D_MAP_POL_LOAD * n + c] = 1.0 / 1000.0 * Pc[c] * itot / (time_step * 60.0) / (R * R);
                        } else {
                            R0[HP2D_MAP_POL_CONC * n + c] = NAN;
                            R0[HP2D_MAP_POL_MASS * n + c] = NAN;
                            R0[HP2D_MAP_POL_LOAD * n + c] = NAN;
                        }
                    }
                    if (out->rec_t) out->rec_t[out->n_rec] = t;
                    out->n_rec++;
                }
            }
            /* :588-600 dmax */
            ORACLE_PAR
            for (int64_t c = 0; c < n; ++c) dmax[c] = (k == 1) ? d_t[c] : fmax(d_t[c], dmax[c]);
            /* :611-617 */
            runoff_mm = qsum * time_step / 60.0 / drainage_area + runoff_mm;
            /* FP64 running ledgers (every step) */
            cum_rain += rain_vol_step;
            cum_inflow += inflow_vol_step;
            cum_inf += inf_vol_step;
            cum_out += qsum / 1000.0 / 3600.0 * (dt_used_min * 60.0);
            gw_elapsed_min += dt_used_min;
            /* :766-774 */
            memcpy(d_p, d_t, (size_t)n * sizeof(double));
            memcpy(I_p, I_t, (size_t)n * sizeof(double));
            memcpy(outflow_prev, outflow, (size_t)3 * n * sizeof(double));
            if (n_dt < out->n_dt_max) out->dt_hist[n_dt] = tcr;
            n_dt++;
            /* :779-787 */
            t = tcr / 60.0 + t;
            k = k + 1;
            accepted++;
            progress = (t / routing_time) * 100.0;
            int do_report = 0;
            if (progress >= next_report || progress >= 100.0) {
                do_report = 1;
                next_report += cfg->report_every_percent;
            }
            if (do_report) {
                /* mass_balance_check.m:6-73 + system_mass_balance.m:27-36 */
                double *tmp = zalloc(n);
                for (int64_t c = 0; c < n; ++c) tmp[c] = qof[c] * area_nan[c];
                outflow_volume = nansum2(ny, nx, tmp) / 1000.0 / 3600.0 * time_step * 60.0 + outflow_volume;
                for (int64_t c = 0; c < n; ++c) {  /* nansum(BC_States.inflow/1000 .* C_a) */
                    const int32_t g = fs.n_gauges > 0 ? in->inflow_gauge[c] : 0;
                    const double q = (g >= 1 && g <= fs.n_gauges) ? fs.agg[g - 1] : 0.0;
                    tmp[c] = q / 1000.0 * area_nan[c];
                }
                double inflow_part = nansum2(ny, nx, tmp);
                for (int64_t c = 0; c < n; ++c)
                    tmp[c] = (fs.n_maps > 0 ? fs.dp_map[c]
                                            : (pat ? pat[c] * delta_p_agg : delta_p_agg)) /
                             1000.0 * area_nan[c];
                double Pv = nansum2(ny, nx, tmp);
                double inflow_vol = inflow_part + Pv + 0.0;
                inflow_volume = inflow_vol + inflow_volume;
                double Qin = inflow_vol - Pv;
                for (int64_t c = 0; c < n; ++c) tmp[c] = qof[c] * area_nan[c];
                double Qout = nansum2(ny, nx, tmp) / 1000.0 / 3600.0 * time_step * 60.0;
                Qout = Qout + 0.0; /* recharge_rate = 0 without groundwater */
                for (int64_t c = 0; c < n; ++c) tmp[c] = area_nan[c] * d_t[c] / 1000.0;
                double S_p = nansum2(ny, nx, tmp);
                double S_UZ = 0.0;
                if (do_inf) {
                    for (int64_t c = 0; c < n; ++c) tmp[c] = area_nan[c] * I_t[c] / 1000.0;
                    S_UZ = nansum2(ny, nx, tmp);
                }
                double S_GW = 0.0;
                if (do_gw) { /* mass_balance_check.m:66-70 */
                    for (int64_t c = 0; c < n; ++c)
                        tmp[c] = area_nan[c] * gw_Sy[c] * (gw_h[c] - (z[c] - gw_sd[c]));
                    S_GW = nansum2(ny, nx, tmp);
                    for (int64_t c = 0; c < n; ++c) tmp[c] = area_nan[c] * gw_pn[c];
                    S_GW = S_GW + nansum2(ny, nx, tmp);
                }
                double fluxes = Pv + Qin - 0.0 - 0.0 - 0.0 - Qout - 0.0;
                double dS = (0.0 + S_p + S_UZ + S_GW + 0.0) - S_prev;
                S_prev = S_prev + dS;
                volume_error = dS - fluxes;
                if (out->volume_error_hist && n_reports < out->n_report_max)
                    out->volume_error_hist[n_reports] = volume_error;
                n_reports++;
                free(tmp);
            }
            /* water-quality instability check at the end of the try block (:919-922):
             * the error lands in the catch block after the step was accepted */
            late = do_wq && (tmin_wq < 0 || isnan(tmin_wq) || isinf(tmin_wq));
        }
        if (failed || late) {
            /* catch block, Main_While.m:926-993 */
            rejected++;
            t = t - time_step;
            t_previous = t;
            double dmx = NAN;
            for (int64_t c = 0; c < n; ++c) {
                dmx = fmax(dmx, d_tot[c] / 1000.0);
            }
            double dmx2 = NAN;
            for (int64_t c = 0; c < n; ++c) dmx2 = fmax(dmx2, d_p[c] / 1000.0);
            double wave_celerity = sqrt(9.81 * fmax(dmx, dmx2));
            double max_vel = NAN;
            for (int64_t c = 0; c < n; ++c) max_vel = fmax(max_vel, vraster[c]);
            catch_index = catch_index + 1;
            double factor = 1.0 / catch_index / cfg->factor_reduction;
            double nt;
            if (cfg->flag_adaptive_timestepping == 1)
                nt = factor * (cfg->alfa_min * R / wave_celerity);
            else
                nt = factor * (cfg->alfa_min * R / (max_vel + wave_celerity));
            double dt_wb = nt;
            nt = fmin(nt, dt_wb);
            if (catch_index > 1) max_dt = fmin(nt * 1.5, max_dt);
            nt = fmin(nt, max_dt);
            if (isinf(nt)) nt = min_dt;
            time_step = nt / 60.0;
            t = t + time_step;
            memcpy(d_t, d_p, (size_t)n * sizeof(double));
            memcpy(I_t, I_p, (size_t)n * sizeof(double));
            out->rollback_dt_s = nt;
            out->rollback_max_dt_s = max_dt;
            if (out->stop_after_rollback) break;
            {
                double dt_step = fmax(t - t_previous, 0.0);
                delta_p_agg = 0.0;
                if (dt_step > 0) {
                    double avg = oracle_rain_avg(in->rain_time_min, in->rain_intensity_mmh,
                                                 in->n_rain, t_previous, t, &cursor);
                    if (isnan(avg)) avg = 0.0;
                    delta_p_agg = avg * dt_step / 60.0;
                }
                if (fs.n_maps > 0) forcing_rain_maps(in, n, is_nan, t_previous, t, &fs);
                if (fs.n_gauges > 0) forcing_inflow(in, t_previous, t, &fs);
            }
        }
    }

    /* outputs */
    if (out->d_t) memcpy(out->d_t, d_t, (size_t)n * sizeof(double));
    if (out->outflow) memcpy(out->outflow, outflow, (size_t)3 * n * sizeof(double));
    if (out->I_t && do_inf) memcpy(out->I_t, I_t, (size_t)n * sizeof(double));
    if (out->dmax) memcpy(out->dmax, dmax, (size_t)n * sizeof(double));
    if (do_wq) {
        if (out->B_t) memcpy(out->B_t, B, (size_t)n * sizeof(double));
        if (out->P_conc) memcpy(out->P_conc, Pc, (size_t)n * sizeof(double));
        if (out->Tot_Washed) memcpy(out->Tot_Washed, TW, (size_t)n * sizeof(double));
    }
    if (do_gw && out->gw_h) memcpy(out->gw_h, gw_h, (size_t)n * sizeof(double));
    if (do_hi && out->max_risk) memcpy(out->max_risk, max_risk, (size_t)n * sizeof(double));
    out->n_hydro = n_hydro;
    hp2d_run_report *rp = &out->rep;
    rp->n_map_records = out->n_rec;
    rp->steps = accepted;
    rp->rejected_steps = rejected;
    rp->t_min = t;
    rp->time_step_s = time_step * 60.0;
    rp->max_velocity = max_velocity;
    rp->volume_error_m3 = volume_error;
    rp->n_reports = n_reports;
    rp->outflow_volume_m3 = outflow_volume;
    rp->inflow_volume_m3 = inflow_volume;
    rp->outlet_runoff_volume_mm = runoff_mm;
    rp->cum_rain_m3 = cum_rain;
    rp->cum_inflow_m3 = cum_inflow;
    rp->cum_infiltration_m3 = cum_inf;
    rp->cum_outflow_m3 = cum_out;
    {
        double *tmp = zalloc(n);
        for (int64_t c = 0; c < n; ++c) tmp[c] = area_nan[c] * d_t[c] / 1000.0;
        rp->storage_m3 = nansum2(ny, nx, tmp);
        free(tmp);
    }
    rp->catch_index = catch_index;
    if (do_wq) {
        double sB = 0.0;
        for (int64_t c = 0; c < n; ++c)
            if (!isnan(B[c])) sB += B[c];
        rp->wq_mass_kg = sB;
        rp->wq_mass_lost_kg = mass_lost;
        rp->wq_out_conc_mgl = out_conc;
        rp->wq_tmin_s = tmin_wq;
        rp->wq_mass_outlet_mg = mass_outlet;
        rp->wq_vol_outlet = vol_outlet;
        rp->wq_substep_steps = wq_split;
    }
    if (do_gw) {
        double *tmp = zalloc(n);
        for (int64_t c = 0; c < n; ++c) tmp[c] = area_nan[c] * gw_Sy[c] * (gw_h[c] - (z[c] - gw_sd[c]));
        rp->gw_storage_m3 = nansum2(ny, nx, tmp);
        free(tmp);
        rp->gw_total_recharge_m3 = gs.total_recharge_m3;
        rp->gw_total_exfiltration_m3 = gs.total_exfiltration_m3;
        rp->gw_error_m3 = gw_err;
        rp->gw_updates = gs.n_updates;
    }
    free(z); free(nman); free(n2); free(d_t); free(d_p); free(d_tot);
    free(outflow); free(outflow_prev); free(Hf); free(qr); free(qu); free(qof); free(dmax);
    free(I_t); free(I_p); free(Tclk); free(vraster); free(area_nan); free(ga_f); free(is_nan);
    free(oidx);
    free(ga);
    free(B); free(C3m); free(C4m); free(Pc); free(TW); free(ql); free(qd);
    free(ca_q); free(ca_qin); free(ca_itot);
    free(gw_h); free(gw_sd); free(gw_ts); free(gw_tr); free(gw_nv); free(gw_al); free(gw_ks);
    free(gw_kg); free(gw_Sy); free(gw_q); free(gw_pn); free(gw_pr); free(gw_pc); free(gw_cum);
    free(gw_dom);
    free(fs.dp_map);
    free(vtot); free(max_risk); free(risk);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Rainfall_Interpolator.m:40-52 with idw_function.m:18-63 (p = 2, rad = Inf, */
/* L = 2): out(i,j) = sum_g W_g F_g / sum_g W_g, W_g = 1 / D_g^p, D_g the     */
/* distance from (x_grid(j), y_grid(i)) to gauge g, D = 0 -> eps, D > rad ->  */
/* Inf.  out is ny_grid x nx_grid column-major.                              */
/* ------------------------------------------------------------------------ */
void oracle_idw(int32_t n_obs, const double *x_obs, const double *y_obs, const double *f_obs,
                int32_t nx_grid, const double *x_grid, int32_t ny_grid, const double *y_grid,
                double p, double rad, double *out) {
    for (int32_t j = 0; j < nx_grid; ++j)
        for (int32_t i = 0; i < ny_grid; ++i) {
            double num = 0.0, den = 0.0;
            for (int32_t g = 0; g < n_obs; ++g) {
                const double dx = fabs(x_grid[j] - x_obs[g]);
                const double dy = fabs(y_grid[i] - y_obs[g]);
                double D = pow(pow(dx, 2.0) + pow(dy, 2.0), 1.0 / 2.0);
                if (D == 0) D = DBL_EPSILON;
                if (D > rad) D = INFINITY;
                const double W = 1.0 / pow(D, p);
                num += W * f_obs[g];
                den += W;
            }
            out[(int64_t)i + (int64_t)j * ny_grid] = num / den;
        }
}
