/*
 * hp2d_gwm.c - FP64 CPU restatement of the groundwater module step
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Follows HydroPol2D_Functions/Groundwater_Module.m:37-404 on the branch
 * flag_groundwater_modeling = 1 without layered soil (so no capillary rise),
 * with initialize_groundwater_scheduler.m:1-66 (defaults and state),
 * simulate_groundwater_recharge.m:55-215 (oracle_gw_recharge),
 * estimate_groundwater_stable_timestep.m:1-38 and Boussinesq_2D_explicit.m
 * (oracle_boussinesq_2d).  MATLAB's NaN-omitting min / max are fmin / fmax;
 * nansum(nansum(x)) is a sum of column sums.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hp2d_oracle.h"

static double colsum(int32_t ny, int32_t nx, const double *x) {
    double tot = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double s = 0.0;
        for (int32_t i = 0; i < ny; ++i) {
            const double v = x[(int64_t)i + (int64_t)j * ny];
            if (!isnan(v)) s += v;
        }
        tot += s;
    }
    return tot;
}

int32_t oracle_groundwater_module(const oracle_gwm_params *p, oracle_gwm_state *g, double *h_t,
                                  double *I_t, double *d_t, double *cumulative_recharge,
                                  double *max_gw_depth, double *q_exf, double *error,
                                  int32_t *ran, double *dt_gw_out) {
    const int32_t ny = p->ny, nx = p->nx;
    const int64_t n = (int64_t)ny * nx;
    if (n <= 0) return -1;
    const double dt_s = p->time_step_min * 60.0;
    const double A = p->resolution * p->resolution;
    double *tmp = malloc(sizeof(double) * n), *Rdown = malloc(sizeof(double) * n);
    double *S_new = malloc(sizeof(double) * n), *uzmax = malloc(sizeof(double) * n);
    double *zero = malloc(sizeof(double) * n), *S0 = malloc(sizeof(double) * n);
    double *rfg = malloc(sizeof(double) * n), *z_bed = malloc(sizeof(double) * n);
    double *Ksat_gw = malloc(sizeof(double) * n), *hsoil = malloc(sizeof(double) * n);
    /* scheduler state defaults (initialize_groundwater_scheduler.m:38-66) */
    const double target_dt_s = fmax(p->target_dt_min * 60.0, DBL_EPSILON);
    const double min_dt_s = fmax(p->min_dt_min * 60.0, DBL_EPSILON);
    const double max_head_change = fmax(p->max_head_change_m, DBL_EPSILON);
    const double courant = fmax(fmin(p->courant, 1.0), DBL_EPSILON);
    /* :42-62 zero_matrix, zwt, UZ capacity */
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        const int nan = isnan(p->elevation[c]);
        zero[c] = nan ? NAN : 0.0;
        const double gw_depth = h_t[c] - (p->elevation[c] - p->soil_depth[c]);
        double zwt = p->soil_depth[c] - gw_depth;
        zwt = fmax(zwt, 0.0);
        zwt = fmin(zwt, p->soil_depth[c]);
        uzmax[c] = zwt * (p->theta_sat[c] - p->theta_r[c]);
        S0[c] = I_t[c] / 1000.0;
        z_bed[c] = p->elevation[c] - p->soil_depth[c];
        Ksat_gw[c] = p->ksat_gw_mmh[c] / 1000.0 / 3600.0;
        hsoil[c] = p->soil_depth[c];
    }
    /* :85-104 vadose recharge with no infiltration input (inf_ms = zero_matrix) */
    double *minS = calloc((size_t)n, sizeof(double));
    oracle_gw_recharge(n, zero, S0, dt_s, minS, uzmax, p->theta_sat, p->theta_r, p->n_vg,
                       p->alpha_vg, p->l_vg, p->ksat_mmh, Rdown, S_new, cumulative_recharge);
    free(minS);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) I_t[c] = S_new[c] * 1000.0;
    /* :124-150 exchange (capillary rate = 0 without layered soil) and pending state */
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        const int nan = isnan(p->elevation[c]);
        const double cap = zero[c];
        double rr = Rdown[c] - cap;
        if (nan) rr = NAN;
        const double net = rr * dt_s;
        const double down = fmax(Rdown[c], 0.0) * dt_s;
        const double capd = fmax(cap, 0.0) * dt_s;
        g->pending_net[c] = g->pending_net[c] + net;
        g->pending_recharge[c] = g->pending_recharge[c] + down;
        g->pending_capillary[c] = g->pending_capillary[c] + capd;
        const double area = nan ? NAN : A;
        tmp[c] = down * area;
    }
    g->elapsed_s = g->elapsed_s + dt_s;
    g->total_recharge_m3 = g->total_recharge_m3 + colsum(ny, nx, tmp);
    for (int64_t c = 0; c < n; ++c) tmp[c] = 0.0 * (isnan(p->elevation[c]) ? NAN : A);
    g->total_capillary_m3 = g->total_capillary_m3 + colsum(ny, nx, tmp);
    /* estimate_groundwater_stable_timestep.m:1-38 */
    double dt_stable = target_dt_s;
    if (p->flag_baseflow == 1) {
        double maxD = NAN;
        for (int64_t c = 0; c < n; ++c) {
            const double Sy = fmax(p->Sy[c], 1e-6);
            double K = p->ksat_gw_mmh[c] / 1000.0 / 3600.0;
            if (!isfinite(K)) K = NAN;
            double sat = fmax(h_t[c] - z_bed[c], 1e-4);
            if (isnan(p->elevation[c])) sat = NAN;
            maxD = fmax(maxD, K * sat / Sy);
        }
        if (isfinite(maxD) && maxD > 0) {
            const double L = p->resolution;
            dt_stable = courant * (L * L) / maxD / 4.0;
            dt_stable = fmin(fmax(dt_stable, 1.0), target_dt_s);
        }
    }
    /* :166-194 scheduler */
    double target = fmin(target_dt_s, dt_stable);
    target = fmax(target, min_dt_s);
    double hc_max = NAN;
    for (int64_t c = 0; c < n; ++c) {
        const double Sy_head = fmax(p->Sy[c], 1e-6);
        hc_max = fmax(hc_max, fabs(g->pending_net[c]) / fmax(Sy_head, 1e-9));
    }
    const int head_due = hc_max >= max_head_change;
    const int final_step = p->t_min >= (p->routing_time_min - 1e-9);
    const int due = p->flag_async == 0 ? 1 : (g->elapsed_s >= target || head_due || final_step);
    const double dt_gw = fmax(g->elapsed_s, DBL_EPSILON);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        rfg[c] = due ? g->pending_net[c] / dt_gw : zero[c];
        if (due && isnan(p->elevation[c])) rfg[c] = NAN;
        q_exf[c] = zero[c];
    }
    double err = 0.0;
    int did = 0;
    if (due && p->flag_baseflow == 1) {
        double *ux = malloc(sizeof(double) * n), *uy = malloc(sizeof(double) * n);
        double *hn = malloc(sizeof(double) * n);
        int32_t nsub = 0;
        int32_t rc = oracle_boussinesq_2d(ny, nx, dt_gw, p->resolution, p->resolution, h_t, z_bed,
                                          p->Sy, rfg, Ksat_gw, p->courant, hsoil, p->domain, NULL,
                                          NULL, hn, ux, uy, q_exf, &err, &nsub);
        free(ux);
        free(uy);
        if (rc != 0) {
            free(hn);
            return rc;
        }
        memcpy(h_t, hn, sizeof(double) * n);
        free(hn);
        for (int64_t c = 0; c < n; ++c) {
            double dn = d_t[c] + dt_gw * (q_exf[c] * 1000.0);
            if (dn < 0) dn = 0.0;
            d_t[c] = dn;
        }
        did = 1;
    } else if (due) {
        for (int64_t c = 0; c < n; ++c) {
            const double Sy_l = fmax(p->Sy[c], 1e-6);
            const double rd = rfg[c] * dt_gw;
            const double dh = rd / Sy_l;
            double hc = h_t[c] + dh;
            hc = fmax(hc, z_bed[c]);
            const double ex_head = fmax(hc - p->elevation[c], 0.0);
            const double ex_water = ex_head * Sy_l;
            h_t[c] = fmin(hc, p->elevation[c]);
            double dn = d_t[c] + ex_water * 1000.0;
            if (dn < 0) dn = 0.0;
            d_t[c] = dn;
        }
        did = 1;
    }
    if (did) {
        for (int64_t c = 0; c < n; ++c) {
            const double area = isnan(p->elevation[c]) ? NAN : A;
            tmp[c] = fmax(q_exf[c], 0.0) * dt_gw * area;
        }
        g->total_exfiltration_m3 = g->total_exfiltration_m3 + colsum(ny, nx, tmp);
        g->n_updates += 1;
        for (int64_t c = 0; c < n; ++c) {
            g->pending_net[c] = 0.0;
            g->pending_recharge[c] = 0.0;
            g->pending_capillary[c] = 0.0;
        }
        g->elapsed_s = 0.0;
        g->last_dt_s = dt_gw;
        /* :380-395 unsaturated-zone excess after the water table moved */
        for (int64_t c = 0; c < n; ++c) {
            double zwt = p->elevation[c] - h_t[c];
            zwt = fmax(zwt, 0.0);
            zwt = fmin(zwt, p->soil_depth[c]);
            const double uz_new = zwt * (p->theta_sat[c] - p->theta_r[c]);
            const double Sm = I_t[c] / 1000.0;
            tmp[c] = fmax(Sm - uz_new, 0.0);
        }
        int any = 0;
        for (int64_t c = 0; c < n; ++c)
            if (tmp[c] > 0) any = 1;
        if (any) {
            for (int64_t c = 0; c < n; ++c) {
                const double Sm = I_t[c] / 1000.0;
                double It = (Sm - tmp[c]) * 1000.0;
                I_t[c] = fmax(It, 0.0);
                double dn = d_t[c] + tmp[c] * 1000.0;
                if (dn < 0) dn = 0.0;
                d_t[c] = dn;
            }
        }
    } else {
        g->last_dt_s = 0.0;
    }
    for (int64_t c = 0; c < n; ++c)
        if (max_gw_depth)
            max_gw_depth[c] = fmax(max_gw_depth[c], h_t[c] - (p->elevation[c] - p->soil_depth[c]));
    *error = err;
    *ran = did;
    *dt_gw_out = dt_gw;
    free(tmp); free(Rdown); free(S_new); free(uzmax); free(zero); free(S0); free(rfg);
    free(z_bed); free(Ksat_gw); free(hsoil);
    return 0;
}
