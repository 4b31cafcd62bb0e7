/*
 * hp2d_fm.c - FP64 CPU restatement of Full_Momentum_Model_D4 (TEST
 * INFRASTRUCTURE ONLY).
 *
 * Follows HydroPol2D_Functions/Full_Momentum_Model_D4.m:1-724 on its
 * ordinary-grid branch (use_lookup_subgrid false, flag_reservoir = 0) with the
 * settings fixed there: source-term friction predictor on, post-flux friction
 * corrector off, h_min = 0, max_velocity = 10 m/s.  Optional inputs mirror the
 * reference's Outlet_Properties: flag_HR_full_momentum (hydrostatic
 * reconstruction) and the four outlet face masks (face-based outlets and the
 * exclusion of those faces from the reflective walls).
 *
 * Arrays are column-major ny x nx; whole-array statements of the reference
 * become loops over every cell, kept in the reference's order.  The y faces
 * are stored as the reference stores them: row i holds the face between rows
 * i-1 (N) and i (S), positive from S to N.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hp2d_oracle.h"

#define IX(i, j) ((int64_t)(i) + (int64_t)(j) * ny)

/* clean_and_limit_momentum (:756-773) */
static void clean_limit(int64_t n, double *hu, double *hv, const double *h, double h_dry,
                        const uint8_t *active, double vmax) {
    for (int64_t c = 0; c < n; ++c) {
        if (!isfinite(hu[c])) hu[c] = 0.0;
        if (!isfinite(hv[c])) hv[c] = 0.0;
        const int wet = active[c] && h[c] > h_dry;
        if (!wet) {
            hu[c] = 0.0;
            hv[c] = 0.0;
            continue;
        }
        const double speed = sqrt(hu[c] * hu[c] + hv[c] * hv[c]) / fmax(h[c], h_dry);
        if (speed > vmax) {
            const double fac = vmax / fmax(speed, 1e-12);
            hu[c] = hu[c] * fac;
            hv[c] = hv[c] * fac;
        }
    }
}

/* hll_normal_flux (:896-955) on one face */
static void hll(double hL, double qnL, double qtL, double hR, double qnR, double qtR, double g,
                double h_dry, double *Fh, double *Fqn, double *Fqt) {
    const int wetL = hL > h_dry, wetR = hR > h_dry;
    const double uL = wetL ? qnL / hL : 0.0, vtL = wetL ? qtL / hL : 0.0;
    const double uR = wetR ? qnR / hR : 0.0, vtR = wetR ? qtR / hR : 0.0;
    (void)vtL;
    (void)vtR;
    const double cL = sqrt(g * fmax(hL, 0.0)), cR = sqrt(g * fmax(hR, 0.0));
    const double FL_h = qnL, FL_qn = qnL * uL + 0.5 * g * (hL * hL), FL_qt = qtL * uL;
    const double FR_h = qnR, FR_qn = qnR * uR + 0.5 * g * (hR * hR), FR_qt = qtR * uR;
    const double sL = fmin(uL - cL, uR - cR), sR = fmax(uL + cL, uR + cR);
    double fh = 0.0, fn = 0.0, ft = 0.0;
    if (sL >= 0) { /* idxL, then idxR overrides (the reference's assignment order) */
        fh = FL_h; fn = FL_qn; ft = FL_qt;
    }
    if (sR <= 0) {
        fh = FR_h; fn = FR_qn; ft = FR_qt;
    }
    if (!(sL >= 0 || sR <= 0) && sR > sL) {
        const double den = sR - sL;
        fh = (sR * FL_h - sL * FR_h + sL * sR * (hR - hL)) / den;
        fn = (sR * FL_qn - sL * FR_qn + sL * sR * (qnR - qnL)) / den;
        ft = (sR * FL_qt - sL * FR_qt + sL * sR * (qtR - qtL)) / den;
    }
    if ((!wetL && !wetR) || !isfinite(fh) || !isfinite(fn) || !isfinite(ft)) fh = fn = ft = 0.0;
    *Fh = fh;
    *Fqn = fn;
    *Fqt = ft;
}

/* apply_manning_friction_exact_implicit (:977-999) */
static void friction_exact(int64_t n, const double *hus, const double *hvs, const double *h,
                           const double *n2, double dt, double g, double h_dry,
                           const uint8_t *active, double *hu, double *hv) {
    for (int64_t c = 0; c < n; ++c) {
        hu[c] = 0.0;
        hv[c] = 0.0;
        const int wet = active[c] && h[c] > h_dry;
        double Q = sqrt(hus[c] * hus[c] + hvs[c] * hvs[c]);
        if (!isfinite(Q)) Q = 0.0;
        double Cf = 0.0;
        if (wet) Cf = g * n2[c] / pow(fmax(h[c], h_dry), 7.0 / 3.0);
        if (!isfinite(Cf) || Cf < 0) Cf = 0.0;
        const double a = dt * Cf;
        if (wet && Q > 0) {
            const double r = 2.0 * Q / (1.0 + sqrt(1.0 + 4.0 * a * Q));
            const double s = r / fmax(Q, 1e-30);
            hu[c] = hus[c] * s;
            hv[c] = hvs[c] * s;
        }
        if (!wet) {
            hu[c] = 0.0;
            hv[c] = 0.0;
        }
        if (!isfinite(hu[c])) hu[c] = 0.0;
        if (!isfinite(hv[c])) hv[c] = 0.0;
    }
}

/* the cell-based and face-based outlet sink of one cell set (:1035-1052 / :330-385) */
static void outlet_sink(int64_t m, const int64_t *cells, int accumulate, double *d_t,
                        double *outlet_flow, double *Hf3, double *hu, double *hv, const double *Ca,
                        const double *n2, const double *rough, int32_t outlet_type,
                        double slope_outlet, double time_step_min, double dx, double g) {
    const double h_min = 0.0;
    double *q = malloc(sizeof(double) * (m ? m : 1)), *hb = malloc(sizeof(double) * (m ? m : 1));
    for (int64_t k = 0; k < m; ++k) {
        const int64_t c = cells[k];
        double h_out = fmax(d_t[c], 0.0) / 1000.0;
        if (!isfinite(h_out)) h_out = 0.0;
        double area = Ca[c];
        if (!isfinite(area) || area <= 0) area = dx * dx;
        double sqrtS;
        if (outlet_type == 1) sqrtS = sqrt(fabs(slope_outlet));
        else {
            const double hs = fmax(h_out, 1e-12);
            sqrtS = sqrt(g * n2[c] * pow(hs, -1.0 / 3.0));
            if (!isfinite(sqrtS)) sqrtS = 0.0;
        }
        Hf3[c] = h_out;
        double nout = rough[c];
        if (!isfinite(nout) || nout <= 0) nout = sqrt(fmax(n2[c], 1e-12));
        double qo = (1.0 / nout) * dx * pow(h_out, 5.0 / 3.0) * sqrtS / area * 1000.0 * 3600.0;
        const double qmax = fmax(d_t[c] - 1000.0 * h_min, 0.0) / (time_step_min / 60.0);
        qo = fmin(qo, qmax);
        if (!isfinite(qo)) qo = 0.0;
        q[k] = qo;
        hb[k] = fmax(d_t[c] / 1000.0, 0.0);
    }
    for (int64_t k = 0; k < m; ++k) {
        const int64_t c = cells[k];
        outlet_flow[c] = accumulate ? outlet_flow[c] + q[k] : q[k];
        d_t[c] = d_t[c] - q[k] * (time_step_min / 60.0);
        d_t[c] = fmax(d_t[c], 1000.0 * h_min);
        const double ha = fmax(d_t[c] / 1000.0, 0.0);
        const double sc = hb[k] > 1e-12 ? ha / hb[k] : 1.0;
        hu[c] = hu[c] * sc;
        hv[c] = hv[c] * sc;
    }
    free(q);
    free(hb);
}

int32_t oracle_full_momentum(int32_t ny, int32_t nx, const double *z, const double *d_tot,
                             const double *rough, const double *n2_in, const double *cell_area,
                             double time_step_min, double R, const uint8_t *outlet_index,
                             int32_t outlet_type, double slope_outlet, const int64_t *outlet_idx,
                             int32_t n_outlet, double d_tolerance, const double *outflow_prev,
                             int32_t n_prev_pages, const uint8_t *idx_nan, int32_t flag_HR,
                             const uint8_t *const *face_masks, double *outflow, double *d_t,
                             double *I_tot, double *Hf, double *C_a_out) {
    const int64_t n = (int64_t)ny * nx;
    const double g = 9.81, dx = R, dt = time_step_min * 60.0;
    const double h_dry = fmax(d_tolerance / 1000.0, 1.0e-12), vmax = 10.0;
    double *Ca = malloc(sizeof(double) * n), *h = malloc(sizeof(double) * n);
    double *n2 = malloc(sizeof(double) * n);
    uint8_t *act = malloc((size_t)n);
    double *hu = calloc((size_t)n, sizeof(double)), *hv = calloc((size_t)n, sizeof(double));
    for (int64_t c = 0; c < n; ++c) {
        Ca[c] = cell_area[c];
        if (!isfinite(Ca[c]) || Ca[c] <= 0) Ca[c] = dx * dx; /* :69 */
        act[c] = isfinite(z[c]) && !(idx_nan && idx_nan[c]);  /* :76 */
        if (outlet_index && outlet_index[c] && isfinite(z[c])) act[c] = 1; /* :79 */
        h[c] = fmax(d_tot[c] / 1000.0, 0.0);
        if (!act[c]) h[c] = 0.0;
        n2[c] = n2_in ? n2_in[c] : rough[c] * rough[c];
        if (!isfinite(n2[c]) || n2[c] < 0) n2[c] = 0.0;
    }
    /* reflective wall masks (:1053-1097) */
    int any_face = 0;
    uint8_t *wall[4]; /* left, right, up, down */
    for (int k = 0; k < 4; ++k) wall[k] = calloc((size_t)n, 1);
    const uint8_t *fr = face_masks ? face_masks[0] : NULL, *fl = face_masks ? face_masks[1] : NULL;
    const uint8_t *fu = face_masks ? face_masks[2] : NULL, *fd = face_masks ? face_masks[3] : NULL;
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t c = IX(i, j);
            const int aR = j + 1 < nx && act[IX(i, j + 1)], aL = j > 0 && act[IX(i, j - 1)];
            const int aU = i > 0 && act[IX(i - 1, j)], aD = i + 1 < ny && act[IX(i + 1, j)];
            wall[1][c] = act[c] && !aR && !(fr && fr[c]);
            wall[0][c] = act[c] && !aL && !(fl && fl[c]);
            wall[2][c] = act[c] && !aU && !(fu && fu[c]);
            wall[3][c] = act[c] && !aD && !(fd && fd[c]);
        }
    /* recover_momentum_from_outflow (:725-747) */
    if (outflow_prev && n_prev_pages > 0) {
        for (int64_t c = 0; c < n; ++c) {
            if (n_prev_pages >= 5) {
                hu[c] = outflow_prev[3 * n + c];
                hv[c] = outflow_prev[4 * n + c];
            } else {
                hu[c] = outflow_prev[c] / 1000.0 / 3600.0 * Ca[c] / dx;
                if (n_prev_pages >= 2) hv[c] = outflow_prev[n + c] / 1000.0 / 3600.0 * Ca[c] / dx;
            }
        }
    }
    for (int64_t c = 0; c < n; ++c) {
        if (!isfinite(hu[c])) hu[c] = 0.0;
        if (!isfinite(hv[c])) hv[c] = 0.0;
        if (!act[c] || h[c] <= h_dry) hu[c] = hv[c] = 0.0;
    }
    clean_limit(n, hu, hv, h, h_dry, act, vmax);
    /* source-term friction predictor (:98-110, compute_centered_bed_gradients :957-976) */
    double *hus = malloc(sizeof(double) * n), *hvs = malloc(sizeof(double) * n);
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t c = IX(i, j);
            const int aR = j + 1 < nx && act[IX(i, j + 1)], aL = j > 0 && act[IX(i, j - 1)];
            const int aU = i > 0 && act[IX(i - 1, j)], aD = i + 1 < ny && act[IX(i + 1, j)];
            double gx = 0.0, gy = 0.0;
            if (act[c] && aR && aL) gx = (z[IX(i, j + 1)] - z[IX(i, j - 1)]) / (2 * dx);
            else if (act[c] && aR) gx = (z[IX(i, j + 1)] - z[c]) / dx;
            else if (act[c] && aL) gx = (z[c] - z[IX(i, j - 1)]) / dx;
            if (act[c] && aU && aD) gy = (z[IX(i - 1, j)] - z[IX(i + 1, j)]) / (2 * dx);
            else if (act[c] && aU) gy = (z[IX(i - 1, j)] - z[c]) / dx;
            else if (act[c] && aD) gy = (z[c] - z[IX(i + 1, j)]) / dx;
            if (!act[c] || !isfinite(gx)) gx = 0.0;
            if (!act[c] || !isfinite(gy)) gy = 0.0;
            hus[c] = hu[c] - dt * g * h[c] * gx;
            hvs[c] = hv[c] - dt * g * h[c] * gy;
            if (!act[c]) hus[c] = hvs[c] = 0.0;
        }
    double *huf = malloc(sizeof(double) * n), *hvf = malloc(sizeof(double) * n);
    friction_exact(n, hus, hvs, h, n2, dt, g, h_dry, act, huf, hvf);
    clean_limit(n, huf, hvf, h, h_dry, act, vmax);
    /* fluxes (:777-844 hydrostatic reconstruction, :1136-1203 raw) */
    double *Fxh = calloc((size_t)n, sizeof(double)), *FxuL = calloc((size_t)n, sizeof(double));
    double *FxuR = calloc((size_t)n, sizeof(double)), *Fxv = calloc((size_t)n, sizeof(double));
    double *Fyh = calloc((size_t)n, sizeof(double)), *FyvS = calloc((size_t)n, sizeof(double));
    double *FyvN = calloc((size_t)n, sizeof(double)), *Fyu = calloc((size_t)n, sizeof(double));
    double *Hfx = calloc((size_t)n, sizeof(double)), *Hfy = calloc((size_t)n, sizeof(double));
    for (int32_t j = 0; j + 1 < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t a = IX(i, j), b = IX(i, j + 1);
            const int fa = act[a] && act[b];
            double fh, fn, ft, cL = 0.0, cR = 0.0;
            if (flag_HR) {
                const double zf = fmax(z[a], z[b]);
                const double hLs = fmax(z[a] + h[a] - zf, 0.0), hRs = fmax(z[b] + h[b] - zf, 0.0);
                const double uL = h[a] > h_dry ? huf[a] / h[a] : 0.0, vL = h[a] > h_dry ? hvf[a] / h[a] : 0.0;
                const double uR = h[b] > h_dry ? huf[b] / h[b] : 0.0, vR = h[b] > h_dry ? hvf[b] / h[b] : 0.0;
                hll(hLs, hLs * uL, hLs * vL, hRs, hRs * uR, hRs * vR, g, h_dry, &fh, &fn, &ft);
                cL = 0.5 * g * (h[a] * h[a] - hLs * hLs);
                cR = 0.5 * g * (h[b] * h[b] - hRs * hRs);
                Hfx[a] = fmax(hLs, hRs);
            } else {
                hll(h[a], huf[a], hvf[a], h[b], huf[b], hvf[b], g, h_dry, &fh, &fn, &ft);
                Hfx[a] = fmax(h[a], h[b]);
            }
            if (!fa) fh = fn = ft = cL = cR = 0.0;
            Fxh[a] = fh;
            FxuL[a] = fn + cL;
            FxuR[a] = fn + cR;
            Fxv[a] = ft;
        }
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 1; i < ny; ++i) {
            const int64_t s = IX(i, j), nn = IX(i - 1, j);
            const int fa = act[s] && act[nn];
            double fh, fn, ft, cS = 0.0, cN = 0.0;
            if (flag_HR) {
                const double zf = fmax(z[s], z[nn]);
                const double hSs = fmax(z[s] + h[s] - zf, 0.0), hNs = fmax(z[nn] + h[nn] - zf, 0.0);
                const double uS = h[s] > h_dry ? huf[s] / h[s] : 0.0, vS = h[s] > h_dry ? hvf[s] / h[s] : 0.0;
                const double uN = h[nn] > h_dry ? huf[nn] / h[nn] : 0.0, vN = h[nn] > h_dry ? hvf[nn] / h[nn] : 0.0;
                hll(hSs, hSs * vS, hSs * uS, hNs, hNs * vN, hNs * uN, g, h_dry, &fh, &fn, &ft);
                cS = 0.5 * g * (h[s] * h[s] - hSs * hSs);
                cN = 0.5 * g * (h[nn] * h[nn] - hNs * hNs);
                Hfy[s] = fmax(hSs, hNs);
            } else {
                hll(h[s], hvf[s], huf[s], h[nn], hvf[nn], huf[nn], g, h_dry, &fh, &fn, &ft);
                Hfy[s] = fmax(h[s], h[nn]);
            }
            if (!fa) fh = fn = ft = cS = cN = 0.0;
            Fyh[s] = fh;
            FyvS[s] = fn + cS;
            FyvN[s] = fn + cN;
            Fyu[s] = ft;
        }
    /* apply_draining_limiter (:957-1013) */
    double *outV = calloc((size_t)n, sizeof(double)), *theta = malloc(sizeof(double) * n);
    for (int32_t j = 0; j + 1 < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) outV[IX(i, j)] += dt * dx * fmax(Fxh[IX(i, j)], 0.0);
    for (int32_t j = 1; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) outV[IX(i, j)] += dt * dx * fmax(-Fxh[IX(i, j - 1)], 0.0);
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 1; i < ny; ++i) outV[IX(i, j)] += dt * dx * fmax(Fyh[IX(i, j)], 0.0);
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i + 1 < ny; ++i) outV[IX(i, j)] += dt * dx * fmax(-Fyh[IX(i + 1, j)], 0.0);
    for (int64_t c = 0; c < n; ++c) {
        double avail = fmax(h[c] - h_dry, 0.0) * Ca[c];
        if (!act[c]) avail = 0.0;
        double th = 1.0;
        if (act[c] && outV[c] > avail && outV[c] > 0) th = avail / fmax(outV[c], 1e-30);
        if (!isfinite(th)) th = 0.0;
        theta[c] = fmax(fmin(th, 1.0), 0.0);
    }
    for (int32_t j = 0; j + 1 < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t a = IX(i, j);
            const double sc = Fxh[a] < 0 ? theta[IX(i, j + 1)] : theta[a];
            Fxh[a] *= sc; FxuL[a] *= sc; FxuR[a] *= sc; Fxv[a] *= sc;
        }
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 1; i < ny; ++i) {
            const int64_t s = IX(i, j);
            const double sc = Fyh[s] < 0 ? theta[IX(i - 1, j)] : theta[s];
            Fyh[s] *= sc; FyvS[s] *= sc; FyvN[s] *= sc; Fyu[s] *= sc;
        }
    /* :120-158 volume and momentum updates */
    double *Vol = malloc(sizeof(double) * n), *hun = malloc(sizeof(double) * n), *hvn = malloc(sizeof(double) * n);
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) {
            const int64_t c = IX(i, j);
            const double west = j > 0 ? Fxh[IX(i, j - 1)] : 0.0;
            const double south = i + 1 < ny ? Fyh[IX(i + 1, j)] : 0.0;
            Vol[c] = dt * dx * (west - Fxh[c] - Fyh[c] + south);
            hun[c] = huf[c];
            hvn[c] = hvf[c];
        }
    const double dtdx = dt * dx;
    for (int32_t j = 0; j + 1 < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) hun[IX(i, j)] = hun[IX(i, j)] - dtdx * FxuL[IX(i, j)] / Ca[IX(i, j)];
    for (int32_t j = 1; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) hun[IX(i, j)] = hun[IX(i, j)] + dtdx * FxuR[IX(i, j - 1)] / Ca[IX(i, j)];
    for (int32_t j = 0; j + 1 < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) hvn[IX(i, j)] = hvn[IX(i, j)] - dtdx * Fxv[IX(i, j)] / Ca[IX(i, j)];
    for (int32_t j = 1; j < nx; ++j)
        for (int32_t i = 0; i < ny; ++i) hvn[IX(i, j)] = hvn[IX(i, j)] + dtdx * Fxv[IX(i, j - 1)] / Ca[IX(i, j)];
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 1; i < ny; ++i) hun[IX(i, j)] = hun[IX(i, j)] - dtdx * Fyu[IX(i, j)] / Ca[IX(i, j)];
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i + 1 < ny; ++i) hun[IX(i, j)] = hun[IX(i, j)] + dtdx * Fyu[IX(i + 1, j)] / Ca[IX(i, j)];
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 1; i < ny; ++i) hvn[IX(i, j)] = hvn[IX(i, j)] - dtdx * FyvS[IX(i, j)] / Ca[IX(i, j)];
    for (int32_t j = 0; j < nx; ++j)
        for (int32_t i = 0; i + 1 < ny; ++i) hvn[IX(i, j)] = hvn[IX(i, j)] + dtdx * FyvN[IX(i + 1, j)] / Ca[IX(i, j)];
    /* apply_reflective_wall_pressure_fluxes (:1098-1134) */
    for (int64_t c = 0; c < n; ++c) {
        double p = (act[c] && h[c] > h_dry) ? 0.5 * g * (h[c] * h[c]) : 0.0;
        if (!isfinite(p)) p = 0.0;
        double coef = dt * dx / Ca[c];
        if (!isfinite(coef) || coef < 0) coef = 0.0;
        if (wall[0][c]) hun[c] = hun[c] + coef * p;
        if (wall[1][c]) hun[c] = hun[c] - coef * p;
        if (wall[2][c]) hvn[c] = hvn[c] - coef * p;
        if (wall[3][c]) hvn[c] = hvn[c] + coef * p;
        if (!isfinite(hun[c])) hun[c] = 0.0;
        if (!isfinite(hvn[c])) hvn[c] = 0.0;
    }
    /* :170-198 */
    double *hn = malloc(sizeof(double) * n);
    for (int64_t c = 0; c < n; ++c) {
        hn[c] = h[c] + Vol[c] / Ca[c];
        if (!act[c]) hn[c] = 0.0;
        hn[c] = fmax(hn[c], 0.0);
        if (h[c] > h_dry) {
            const double sc = fmin(1.0, hn[c] / fmax(h[c], h_dry));
            hun[c] = hun[c] * sc;
            hvn[c] = hvn[c] * sc;
        }
    }
    clean_limit(n, hun, hvn, hn, h_dry, act, vmax);
    double *ofl = calloc((size_t)n, sizeof(double)), *Hf3 = calloc((size_t)n, sizeof(double));
    for (int64_t c = 0; c < n; ++c) {
        d_t[c] = d_tot[c] + 1000.0 * Vol[c] / Ca[c];
        if (!act[c]) d_t[c] = NAN;
        else d_t[c] = fmax(d_t[c], 0.0);
        double ox = Fxh[c] * dx / Ca[c] * 1000.0 * 3600.0, oy = Fyh[c] * dx / Ca[c] * 1000.0 * 3600.0;
        if (!isfinite(ox)) ox = 0.0;
        if (!isfinite(oy)) oy = 0.0;
        if (!act[c]) ox = oy = 0.0;
        outflow[c] = ox;
        outflow[n + c] = oy;
    }
    /* outlets (:247-330) */
    int64_t *cells = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    if (face_masks) {
        for (int k = 0; k < 4; ++k)
            for (int64_t c = 0; c < n; ++c)
                if (face_masks[k] && face_masks[k][c] && act[c]) any_face = 1;
    }
    if (any_face) {
        for (int k = 0; k < 4; ++k) { /* right, left, up, down */
            int64_t m = 0;
            if (face_masks[k])
                for (int64_t c = 0; c < n; ++c)
                    if (face_masks[k][c] && act[c]) cells[m++] = c;
            if (m) outlet_sink(m, cells, 1, d_t, ofl, Hf3, hun, hvn, Ca, n2, rough, outlet_type,
                               slope_outlet, time_step_min, dx, g);
        }
    } else if (n_outlet > 0) {
        outlet_sink(n_outlet, outlet_idx, 0, d_t, ofl, Hf3, hun, hvn, Ca, n2, rough, outlet_type,
                    slope_outlet, time_step_min, dx, g);
    }
    /* :386-411 */
    double *hf = malloc(sizeof(double) * n);
    for (int64_t c = 0; c < n; ++c) {
        hf[c] = fmax(d_t[c] / 1000.0, 0.0);
        if (!act[c]) hf[c] = 0.0;
    }
    clean_limit(n, hun, hvn, hf, h_dry, act, vmax);
    for (int64_t c = 0; c < n; ++c) {
        outflow[2 * n + c] = ofl[c];
        outflow[3 * n + c] = hun[c];
        outflow[4 * n + c] = hvn[c];
        Hf[c] = Hfx[c];
        Hf[n + c] = Hfy[c];
        Hf[2 * n + c] = Hf3[c];
        const double s = outflow[c] + outflow[n + c] + outflow[2 * n + c];
        I_tot[c] = act[c] ? fabs(s) * dt / 1000.0 * 1.0 / 3600.0 * Ca[c] : NAN;
        if (C_a_out) C_a_out[c] = Ca[c];
    }
    free(Ca); free(h); free(n2); free(act); free(hu); free(hv); free(hus); free(hvs); free(huf);
    free(hvf); free(Fxh); free(FxuL); free(FxuR); free(Fxv); free(Fyh); free(FyvS); free(FyvN);
    free(Fyu); free(Hfx); free(Hfy); free(outV); free(theta); free(Vol); free(hun); free(hvn);
    free(hn); free(ofl); free(Hf3); free(cells); free(hf);
    for (int k = 0; k < 4; ++k) free(wall[k]);
    return 0;
}
