/*
 * hp2d_gw.c - FP64 CPU restatement of the explicit Boussinesq groundwater
 * solver and the vadose recharge bucket (TEST INFRASTRUCTURE ONLY).
 *
 * oracle_boussinesq_2d follows HydroPol2D_Functions/Boussinesq_2D_explicit.m
 * :59-195 with its local functions compute_fluxes_conservative (:315-349),
 * compute_flux_divergence (:351-371), compute_stable_dt (:452-462) and
 * compute_boussinesq_velocities (:567-584), statement by statement:
 *   - H = max(h - z0, 0) omits NaN, so H is never NaN and the reference's
 *     "Fx(isnan(H)) = 0" never fires; a NaN head gives a NaN face flux;
 *   - the flux cap uses the left/upper cell as donor whatever the sign;
 *   - the sub-step loop resets kk every iteration, so the last-step trim
 *     only fires when ceil(T/dt) == 3, and it changes the dt that advances t,
 *     not the dt just used; dt never grows back;
 *   - the returned error mixes the last sub-step's recharge and exfiltration
 *     with the storage change over the whole call.
 * The river exchange is disabled in the reference (q_river = 0) and the
 * perimeter argument is unused; neither is taken here.
 *
 * oracle_gw_recharge follows simulate_groundwater_recharge.m:55-215.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hp2d_oracle.h"

#define IX(i, j) ((int64_t)(i) + (int64_t)(j) * ny)

/* compute_fluxes_conservative: Fx[c] is the face (i,j)|(i,j+1), Fy[c] the face
 * (i,j)|(i+1,j); the last column / row hold 0 and are never read. */
static void gw_fluxes(int32_t ny, int32_t nx, const double *h, const double *z0, const double *K,
                      const double *Sy, double dx, double dy, double dt, double *Fx, double *Fy) {
    ORACLE_PAR
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t c = IX(i, j);
            const double H = fmax(h[c] - z0[c], 0.0);
            Fx[c] = 0.0;
            Fy[c] = 0.0;
            if (j < nx - 1) {
                const int64_t r = IX(i, j + 1);
                const double Hr = fmax(h[r] - z0[r], 0.0);
                const double Hx = 0.5 * (H + Hr);
                const double Kx = 0.5 * (K[c] + K[r]);
                const double dHdx = (h[r] - h[c]) / dx;
                double F = -Kx * Hx * dHdx;
                const double cap = Sy[c] * H / dt;
                const double sgn = isnan(F) ? NAN : (F > 0) - (F < 0);
                F = sgn * fmin(fabs(F), cap);
                if (isnan(Hr)) F = 0.0;
                Fx[c] = F;
            }
            if (i < ny - 1) {
                const int64_t d = IX(i + 1, j);
                const double Hd = fmax(h[d] - z0[d], 0.0);
                const double Hy = 0.5 * (H + Hd);
                const double Ky = 0.5 * (K[c] + K[d]);
                const double dHdy = (h[d] - h[c]) / dy;
                double F = -Ky * Hy * dHdy;
                const double cap = Sy[c] * H / dt;
                const double sgn = isnan(F) ? NAN : (F > 0) - (F < 0);
                F = sgn * fmin(fabs(F), cap);
                if (isnan(Hd)) F = 0.0;
                Fy[c] = F;
            }
        }
}

/* compute_flux_divergence */
static double gw_div(int32_t ny, int32_t nx, const double *Fx, const double *Fy, double dx,
                     double dy, int32_t i, int32_t j) {
    const int64_t c = IX(i, j);
    double div_x, div_y;
    if (j == 0)
        div_x = Fx[c] / dx;
    else if (j == nx - 1)
        div_x = -Fx[c - ny] / dx;
    else
        div_x = (Fx[c] - Fx[c - ny]) / dx;
    if (i == 0)
        div_y = Fy[c] / dy;
    else if (i == ny - 1)
        div_y = -Fy[c - 1] / dy;
    else
        div_y = (Fy[c] - Fy[c - 1]) / dy;
    return div_x + div_y;
}

/* compute_boussinesq_velocities (+ the catchment mask when `mask` is given);
 * returns compute_stable_dt of the result */
static double gw_velocities(int32_t ny, int32_t nx, const double *h, const double *K, double dx,
                            double dy, double courant, const uint8_t *mask, double *ux,
                            double *uy) {
    double dtmin = INFINITY;
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t c = IX(i, j);
            double vx = NAN, vy = NAN;
            if (j < nx - 1) vx = -(0.5 * (K[c] + K[c + ny])) * ((h[c + ny] - h[c]) / dx);
            if (i < ny - 1) vy = -(0.5 * (K[c] + K[c + 1])) * ((h[c + 1] - h[c]) / dy);
            if (mask && !mask[c]) {
                vx = NAN;
                vy = NAN;
            }
            if (ux) ux[c] = vx;
            if (uy) uy[c] = vy;
            const double cx = courant * dx / fabs(vx), cy = courant * dy / fabs(vy);
            if (isfinite(cx) && cx > 0) dtmin = fmin(dtmin, cx);
            if (isfinite(cy) && cy > 0) dtmin = fmin(dtmin, cy);
        }
    return dtmin;
}

/* nansum(nansum(x)): column sums skipping NaN, then their sum */
static double nansum2(int32_t ny, int32_t nx, const double *x) {
    double tot = 0.0;
    for (int32_t j = 0; j < nx; ++j) {
        double s = 0.0;
        for (int32_t i = 0; i < ny; ++i)
            if (!isnan(x[IX(i, j)])) s += x[IX(i, j)];
        tot += s;
    }
    return tot;
}

int32_t oracle_boussinesq_2d(int32_t ny, int32_t nx, double dt_in, double dx, double dy,
                             const double *h, const double *z0, const double *Sy, const double *R,
                             const double *K, double courant, const double *h_soil,
                             const uint8_t *catchment, const uint8_t *dirichlet,
                             const double *h_dirichlet, double *h_t, double *u_x, double *u_y,
                             double *q_exf, double *error, int32_t *n_sub) {
    if (ny < 2 || nx < 2 || !(dt_in > 0)) return -2;
    const int64_t n = (int64_t)ny * nx;
    double *h_surf = malloc(sizeof(double) * n), *hp = malloc(sizeof(double) * n);
    double *Fx1 = malloc(sizeof(double) * n), *Fy1 = malloc(sizeof(double) * n);
    double *Fx2 = malloc(sizeof(double) * n), *Fy2 = malloc(sizeof(double) * n);
    double *tmp = malloc(sizeof(double) * n);
    for (int64_t c = 0; c < n; ++c) {
        h_surf[c] = z0[c] + h_soil[c];
        q_exf[c] = 0.0;
        h_t[c] = h[c];
    }
    double t = 0.0, T = dt_in, dt = dt_in, err = 0.0;
    int32_t subs = 0;
    double dt_gw = gw_velocities(ny, nx, h_t, K, dx, dy, courant, NULL, u_x, u_y);
    const double A = dx * dy;
    while (t < T) {
        const double n_steps = ceil(T / dt);
        int kk = 1;
        dt = fmin(dt_gw, dt);
        /* predictor */
        gw_fluxes(ny, nx, h_t, z0, K, Sy, dx, dy, dt, Fx1, Fy1);
        ORACLE_PAR
        for (int32_t j = 0; j < nx; ++j)
            for (int32_t i = 0; i < ny; ++i) {
                const int64_t c = IX(i, j);
                double v = h_t[c] + dt / Sy[c] * (-gw_div(ny, nx, Fx1, Fy1, dx, dy, i, j));
                if (dirichlet && dirichlet[c]) v = h_dirichlet[c];
                hp[c] = fmax(v, z0[c]);
            }
        /* corrector and averaged fluxes */
        gw_fluxes(ny, nx, hp, z0, K, Sy, dx, dy, dt, Fx2, Fy2);
        ORACLE_PAR
        for (int64_t c = 0; c < n; ++c) {
            Fx2[c] = 0.5 * (Fx1[c] + Fx2[c]);
            Fy2[c] = 0.5 * (Fy1[c] + Fy2[c]);
        }
        ORACLE_PAR
        for (int32_t j = 0; j < nx; ++j)
            for (int32_t i = 0; i < ny; ++i) {
                const int64_t c = IX(i, j);
                double v = h_t[c] + dt / Sy[c] * (-gw_div(ny, nx, Fx2, Fy2, dx, dy, i, j) + R[c]);
                if (dirichlet && dirichlet[c]) v = h_dirichlet[c];
                v = fmax(v, z0[c]);
                q_exf[c] = fmax(0.0, v - h_surf[c]) * Sy[c] / dt;
                tmp[c] = fmin(v, h_surf[c]);
            }
        dt_gw = gw_velocities(ny, nx, tmp, K, dx, dy, courant, catchment, u_x, u_y);
        ORACLE_PAR
        for (int64_t c = 0; c < n; ++c) {
            h_t[c] = (catchment && !catchment[c]) ? NAN : tmp[c];
            if (catchment && !catchment[c]) q_exf[c] = NAN;
        }
        /* mass balance of the last sub-step against the storage at entry */
        for (int64_t c = 0; c < n; ++c) tmp[c] = R[c] * dt;
        const double recharge_mass = nansum2(ny, nx, tmp) * A;
        for (int64_t c = 0; c < n; ++c) tmp[c] = q_exf[c] * dt;
        const double exfil_mass = nansum2(ny, nx, tmp) * A;
        for (int64_t c = 0; c < n; ++c) tmp[c] = Sy[c] * fmax(h[c] - z0[c], 0.0);
        const double storage_prev = nansum2(ny, nx, tmp) * A;
        for (int64_t c = 0; c < n; ++c) tmp[c] = Sy[c] * fmax(h_t[c] - z0[c], 0.0);
        const double storage_curr = nansum2(ny, nx, tmp) * A;
        err = recharge_mass - exfil_mass - (storage_curr - storage_prev);
        ++subs;
        kk = kk + 1;
        if ((double)kk == n_steps - 1) dt = T - kk * dt;
        t = t + dt;
        if (subs > 10000000) break;
    }
    *error = err;
    *n_sub = subs;
    free(h_surf); free(hp); free(Fx1); free(Fy1); free(Fx2); free(Fy2); free(tmp);
    return subs > 10000000 ? -1 : 0;
}

/* simulate_groundwater_recharge.m:55-215 (van Genuchten / Mualem bucket) */
int32_t oracle_gw_recharge(int64_t n, const double *infiltration_rate, const double *S0,
                           double dt, const double *min_S, const double *max_S,
                           const double *theta_s, const double *theta_r, const double *n_vg,
                           const double *alpha_vg, double l_vg, const double *ksat_mmh,
                           double *recharge_rate, double *S_out, double *cumulative_recharge) {
    dt = fmax(dt, 1e-12);
    ORACLE_PAR
    for (int64_t c = 0; c < n; ++c) {
        const double nv = n_vg[c], m = 1.0 - 1.0 / nv, a = alpha_vg[c];
        const double ts = theta_s[c], tr = theta_r[c];
        const double Ks = ksat_mmh[c] / 1000.0 / 3600.0;
        double inf_r = isfinite(infiltration_rate[c]) ? infiltration_rate[c] : 0.0;
        const double s0 = isfinite(S0[c]) ? S0[c] : 0.0; /* sanitised in place (:99) */
        const double S_init = s0;
        double mn = isfinite(min_S[c]) ? min_S[c] : 0.0;
        double mx = isfinite(max_S[c]) ? max_S[c] : 0.0;
        const double min_eff = fmin(mn, mx);
        const int idx = isfinite(s0) && isfinite(mx) && isfinite(min_eff);
        double S = idx ? s0 : 0.0;
        S = fmax(S, min_eff);
        S = fmin(S, mx);
        const double por_eff = fmax(ts - tr, 1e-12);
        double zwt = mx / por_eff;
        zwt = fmax(zwt, 0.0);
        const double Lgw = fmax(0.5 * zwt, 1e-6);
        const double zwt_safe = fmax(zwt, 1e-6);
        double theta = tr + S / zwt_safe;
        theta = fmax(theta, tr);
        theta = fmin(theta, ts);
        double Se = (theta - tr) / fmax(ts - tr, 1e-12);
        Se = fmin(fmax(Se, 1e-6), 1.0);
        double h_soil = -(1.0 / a) * pow(pow(Se, -1.0 / m) - 1.0, 1.0 / nv);
        if (Se >= 0.999999) h_soil = 0.0;
        const double term = 1.0 - pow(Se, 1.0 / m);
        double Kr = pow(Se, l_vg) * pow(1.0 - pow(term, m), 2.0);
        Kr = fmin(fmax(Kr, 0.0), 1.0);
        const double Ksoil = Ks * Kr;
        double grad = 1.0 + h_soil / Lgw;
        grad = fmax(grad, -1.0);
        grad = fmin(grad, 2.0);
        double Rr = Ksoil * grad;
        const int no_uz = mx <= 0.0;
        if (no_uz) Rr = 0.0;
        double S_new = S + dt * (inf_r - Rr);
        S_new = fmax(S_new, min_eff);
        S_new = fmin(S_new, mx);
        if (no_uz) S_new = 0.0;
        double R_mb = inf_r - (S_new - S) / dt;
        if (no_uz) R_mb = 0.0;
        if (!idx) {
            R_mb = 0.0;
            S_new = S_init;
        }
        S_out[c] = S_new;
        recharge_rate[c] = R_mb;
        if (cumulative_recharge)
            cumulative_recharge[c] = cumulative_recharge[c] + fmax(R_mb, 0.0) * dt * 1000.0;
    }
    return 0;
}
