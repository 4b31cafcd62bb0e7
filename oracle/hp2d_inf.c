/*
 * hp2d_inf.c - FP64 CPU restatement of the mainline infiltration step
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Follows HydroPol2D_Functions/Infiltration_Module.m:101-583 on the branch
 * flag_infiltration = 1, no subgrid, no layered soil: the van Genuchten /
 * Mualem surface-interface capacity with a Darcy driving head, limited by the
 * supply and the remaining unsaturated-zone storage, then the I_t / d_t update
 * and the FP64 storage change errors(4).  MATLAB's NaN-omitting min / max are
 * fmin / fmax; nansum(nansum(x)) is a column-major sum of column sums.
 */
#include <math.h>

#include "hp2d_oracle.h"

static double nansum2(int32_t ny, int32_t nx, const double *x, double scale) {
    double tot = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double s = 0.0;
        for (int32_t i = 0; i < ny; ++i) {
            const double v = scale * x[(int64_t)i + (int64_t)j * ny] / 1000.0;
            if (!isnan(v)) s += v;
        }
        tot += s;
    }
    return tot;
}

/* One cell of the closure (:182-473): d [mm] the effective depth (updated,
 * :467-470), I [mm] the unsaturated-zone storage (updated, :456-463), zwt_in
 * the depth to the water table Soil_Depth - GW_Depth [m] before the max(., 0)
 * of :198; Ltop already floored at 1e-6 (:247) and wK clamped to [0, 1].
 * Returns f [mm/h]. */
double oracle_vg_cell(double dt_h, double A, double C_a, double zwt_in, double ts, double tr,
                      double nv, double a, double ksat_mmh, int imp, double l_vg, double Ltop,
                      double dh_max, double wK, double *d, double *I) {
    const double Lmix = Ltop;
    const double eff = *d;
    const double i_a = eff / fmax(dt_h, 1e-12);
    const double h0 = fmax(eff, 0.0) / 1000.0;
    const double zwt = fmax(zwt_in, 0.0);
    const double uz_max = zwt * (ts - tr);
    const double S = fmax(*I, 0.0) / 1000.0;
    const double S_rem = fmax(uz_max - S, 0.0);
    const int no_uz = (zwt <= 0.0) || (uz_max <= 0.0);
    const double m = 1.0 - 1.0 / nv;
    const double Ks = ksat_mmh / 1000.0 / 3600.0;
    const double zwt_safe = fmax(zwt, 1e-6);
    double th_b = tr + (S / zwt_safe);
    th_b = fmax(th_b, tr);
    th_b = fmin(th_b, ts);
    double Se_b = (th_b - tr) / fmax(ts - tr, 1e-12);
    Se_b = fmin(fmax(Se_b, 1e-6), 1.0);
    double h_b = -(1.0 / a) * pow(pow(Se_b, -1.0 / m) - 1.0, 1.0 / nv);
    if (Se_b >= 0.999999) h_b = 0.0;
    const double w_avail = fmax(eff, 0.0) / 1000.0;
    const double w_top_def = Lmix * fmax(ts - th_b, 0.0);
    double w_wet = fmin(w_avail, S_rem);
    w_wet = fmin(w_wet, w_top_def);
    w_wet = fmax(w_wet, 0.0);
    double th_t = th_b + w_wet / Lmix;
    th_t = fmax(th_t, tr);
    th_t = fmin(th_t, ts);
    double Se_t = (th_t - tr) / fmax(ts - tr, 1e-12);
    Se_t = fmin(fmax(Se_t, 1e-6), 1.0);
    double term = 1.0 - pow(Se_t, 1.0 / m);
    term = fmin(fmax(term, 0.0), 1.0);
    double Kr = pow(Se_t, l_vg) * pow(1.0 - pow(term, m), 2.0);
    Kr = fmin(fmax(Kr, 0.0), 1.0);
    const double Ktop = Ks * Kr;
    const double Ksurf = Ks + 0.0;
    const int wet = (i_a > 0.0) || (h0 > 0.0);
    double Ksoil = Ktop;
    if (wet) Ksoil = wK * Ksurf + (1.0 - wK) * Ktop;
    Ksoil = fmin(fmax(Ksoil, 0.0), Ksurf);
    double dh = h0 - h_b;
    dh = fmax(dh, 0.0);
    dh = fmin(dh, dh_max);
    const double grad = (dh / Ltop) + 1.0;
    double C = Ksoil * grad;
    C = C * 1000.0 * 3600.0;
    const double f_store = (S_rem * 1000.0) / fmax(dt_h, 1e-12);
    double f = fmin(i_a, C);
    f = fmin(f, f_store);
    if (imp) f = 0.0;
    if (no_uz) f = 0.0;
    if (isnan(f)) f = 0.0;
    f = fmax(f, 0.0);
    const double inf_cell = f * dt_h;
    double It = *I + inf_cell;
    It = fmin(It, uz_max * 1000.0);
    It = fmax(It, 0.0);
    *I = It;
    const double dn = *d - inf_cell * A / C_a;
    *d = dn < 0 ? 0.0 : dn;
    return f;
}

int32_t oracle_infiltration(int32_t ny, int32_t nx, double time_step_min, double resolution,
                            double C_a, const double *elevation, const double *h_t,
                            const double *soil_depth, const double *theta_sat,
                            const double *theta_r, const double *n_vg, const double *alpha_vg,
                            const double *ksat_mmh, const uint8_t *idx_imp, double l_vg,
                            double Ltop_in, double dh_max, double Ksurf_weight, double *d_t,
                            double *I_t, double *I_p, double *f_out, double *cum_inf,
                            double *inf_volume, double *dS_inf) {
    const int64_t n = (int64_t)ny * nx;
    const double A = resolution * resolution;  /* Coarse_Area (:102) */
    const double S_UZ_0 = nansum2(ny, nx, I_t, A);
    const double S_p_0 = nansum2(ny, nx, d_t, A);
    const double S_0 = S_UZ_0 + S_p_0;
    const double dt_h = time_step_min / 60.0;
    const double Ltop = fmax(Ltop_in, 1e-6);
    const double wK = fmin(fmax(Ksurf_weight, 0.0), 1.0);
    double vol = 0.0;
    /* inf_volume = 1/1000 * nansum(nansum(inf_volume_cell .* Coarse_Area)) (:473-474) */
    double col_acc = 0.0;
    for (int64_t c = 0; c < n; ++c) {
        /* GW_Depth = h_t - (elevation - Soil_Depth); zwt = Soil_Depth - GW_Depth (:194-197) */
        const double gw_depth = h_t[c] - (elevation[c] - soil_depth[c]);
        const double zwt = soil_depth[c] - gw_depth;
        I_p[c] = I_t[c];
        const double f = oracle_vg_cell(dt_h, A, C_a, zwt, theta_sat[c], theta_r[c], n_vg[c],
                                        alpha_vg[c], ksat_mmh[c], idx_imp && idx_imp[c], l_vg,
                                        Ltop, dh_max, wK, &d_t[c], &I_t[c]);
        f_out[c] = f;
        if (cum_inf) cum_inf[c] = cum_inf[c] + f * (time_step_min / 60.0) / 1000.0;
        const double v = (f * dt_h) * A;
        if (!isnan(v)) col_acc += v;
        if ((c + 1) % ny == 0) { /* end of a column */
            vol += col_acc;
            col_acc = 0.0;
        }
    }
    *inf_volume = 1.0 / 1000.0 * vol;
    const double S_UZ_t = nansum2(ny, nx, I_t, A);
    const double S_p_t = nansum2(ny, nx, d_t, A);
    *dS_inf = (S_UZ_t + S_p_t) - S_0;
    return 0;
}
