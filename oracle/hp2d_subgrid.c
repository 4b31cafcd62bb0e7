/*
 * hp2d_subgrid.c - FP64 CPU restatement of the lookup-table subgrid local
 * inertial model (TEST INFRASTRUCTURE ONLY).
 *
 *   oracle_subgrid_build        Subgrid_Properties_Lookup.m:1-510 (SFINCS-style
 *                               cell, velocity-point and boundary tables)
 *   oracle_subgrid_volume       hp2d_subgrid_apply_volume_change.m (sfincs_exact)
 *   oracle_li_subgrid           Local_Inertial_Model_D4_Subgrid.m:1-317
 *                               (flag_reservoir = 0)
 * with the table queries of hp2d_sfincs_cell_volume_from_zs.m,
 * hp2d_sfincs_zs_from_cell_volume.m, hp2d_sfincs_cell_wet_area_from_zs.m,
 * hp2d_sfincs_velocity_state.m and hp2d_sfincs_boundary_velocity_state.m.
 *
 * Tables follow hp2d_subgrid_tables (include/hp2d.h): column-major, level k of
 * a face or cell table at offset k * n_entries.  Whole-grid statements are
 * restated as loops over every entry; MATLAB's max/min omit NaN as fmax/fmin.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hp2d_oracle.h"

#define EPS DBL_EPSILON

/* ---------------------------------------------------------------- builder -- */

static double mat_linspace(double d1, double d2, int n, int k) {
    const int n1 = n - 1;
    if (d1 == d2) return d1;
    if (k == 0) return d1;
    if (k == n1) return d2;
    return d1 + ((double)k * (d2 - d1)) / n1;
}

static double fine_volume(const double *z, int m, double zs, double fa) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += fmax(zs - z[i], 0.0);
    return s * fa;
}

/* invert_volume_exact (:208-231) */
static double invert_volume(const double *z, int m, double V, double zmin, double zmax,
                            double volmax, double fa) {
    if (!isfinite(V) || V <= 0) return zmin;
    if (volmax <= EPS || zmax <= zmin) return zmax + V / (m * fa);
    if (V >= volmax) return zmax + (V - volmax) / (m * fa);
    double lo = zmin, hi = zmax;
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (fine_volume(z, m, mid, fa) < V) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* a velocity-point table from pooled fine elevations and roughness (:296-345) */
static void table_from_vectors(const double *z, const double *n, int m, double zmin, double zmax,
                               int nl, double *havg, double *nrep, double *pwet, double *navg,
                               double *ffit, double *tzmin, double *tzmax) {
    for (int k = 0; k < nl; ++k) havg[k] = nrep[k] = pwet[k] = NAN;
    *navg = *ffit = *tzmin = *tzmax = NAN;
    if (m == 0 || !isfinite(zmin) || !isfinite(zmax)) return;
    for (int k = 0; k < nl; ++k) {
        const double zu = zmax > zmin ? mat_linspace(zmin, zmax, nl, k) : zmin;
        double sh = 0.0, sw = 0.0, sd = 0.0;
        for (int i = 0; i < m; ++i) {
            const double h = fmax(zu - z[i], 0.0);
            sw += (double)(h > 0);
            sh += h;
            sd += pow(fmax(zu - fmax(z[i], zmin), 0.0), 5.0 / 3.0) / n[i];
        }
        pwet[k] = sw / m;
        havg[k] = sh / m;
        const double denom = sd / m;
        nrep[k] = (havg[k] > 0 && denom > 0 && isfinite(denom)) ? pow(havg[k], 5.0 / 3.0) / denom
                                                                : INFINITY;
    }
    double sn = 0.0;
    for (int i = 0; i < m; ++i) sn += n[i];
    const double na = sn / m, nM = nrep[nl - 1], HG_M = havg[nl - 1];
    const double dzfit = fmax(zmax - zmin, EPS), zfit = zmax + dzfit;
    double sf = 0.0;
    for (int i = 0; i < m; ++i) sf += pow(fmax(zfit - fmax(z[i], zmin), 0.0), 5.0 / 3.0) / n[i];
    const double denom_fit = sf / m;
    const double nfit = (denom_fit > 0 && isfinite(denom_fit))
                            ? pow(HG_M + zfit - zmax, 5.0 / 3.0) / denom_fit : nM;
    double beta = 0.0;
    if (isfinite(na) && isfinite(nM) && isfinite(nfit) && fabs(na - nfit) > EPS)
        beta = ((na - nM) / (na - nfit) - 1.0) / dzfit;
    if (!isfinite(beta)) beta = 0.0;
    *navg = na;
    *ffit = beta;
    *tzmin = zmin;
    *tzmax = zmax;
}

typedef struct {
    const double *dem, *rough;
    int fny, fnx, rough_fine, ny;
    int r0, r1, c0, c1; /* rows [r0, r1), cols [c0, c1) of the fine grid */
    int coarse_i, coarse_j;
} patch_t;

static double patch_z(const patch_t *p, int r, int c) { return p->dem[r + (int64_t)c * p->fny]; }
static double patch_n(const patch_t *p, int r, int c) {
    if (p->rough_fine) return p->rough[r + (int64_t)c * p->fny];
    return p->rough[p->coarse_i + (int64_t)p->coarse_j * p->ny];
}
static double clean_n(double n) { return (!isfinite(n) || n <= 0) ? NAN : n; }

/* the valid (finite z, finite positive n) entries of a sub-rectangle, column-major */
static int gather_valid(const patch_t *p, int r0, int r1, int c0, int c1, double *z, double *n) {
    int m = 0;
    for (int c = c0; c < c1; ++c)
        for (int r = r0; r < r1; ++r) {
            const double zz = patch_z(p, r, c), nn = clean_n(patch_n(p, r, c));
            if (isfinite(zz) && isfinite(nn) && nn > 0) {
                z[m] = zz;
                n[m] = nn;
                ++m;
            }
        }
    return m;
}

static int half_of(int len) {
    int h = (int)round(len / 2.0);
    return h < 1 ? 1 : h;
}

static void put_face(hp2d_subgrid_faces *F, int64_t nf, int64_t f, int nl, const double *z,
                     const double *n, int m, double zmin, double zmax) {
    double havg[256], nrep[256], pwet[256], navg, ffit, tzmin, tzmax;
    table_from_vectors(z, n, m, zmin, zmax, nl, havg, nrep, pwet, &navg, &ffit, &tzmin, &tzmax);
    F->zmin[f] = tzmin;
    F->zmax[f] = tzmax;
    F->navg[f] = navg;
    F->ffit[f] = ffit;
    for (int k = 0; k < nl; ++k) {
        F->havg[f + k * nf] = havg[k];
        F->nrep[f + k * nf] = nrep[k];
        F->pwet[f + k * nf] = pwet[k];
    }
}

static void fill_face_nan(hp2d_subgrid_faces *F, int64_t nf, int nl) {
    for (int64_t f = 0; f < nf; ++f) {
        F->zmin[f] = F->zmax[f] = F->navg[f] = F->ffit[f] = NAN;
        for (int k = 0; k < nl; ++k) F->havg[f + k * nf] = F->nrep[f + k * nf] = F->pwet[f + k * nf] = NAN;
    }
}

static double vmin_of(const double *z, int m) {
    double v = INFINITY;
    for (int i = 0; i < m; ++i) v = fmin(v, z[i]);
    return v;
}
static double vmax_of(const double *z, int m) {
    double v = -INFINITY;
    for (int i = 0; i < m; ++i) v = fmax(v, z[i]);
    return v;
}

int32_t oracle_subgrid_build(const double *dem, int32_t fny, int32_t fnx, double cellsize,
                             const double *rough, int32_t rough_is_fine, double coarse_res,
                             int32_t ny, int32_t nx, int32_t nl, hp2d_subgrid_tables *T) {
    if (nl < 2 || nl > 256) return -1;
    const double r = coarse_res / cellsize;
    const long r_round = lround(r);
    if (r_round % 2 != 0) return -2; /* :42-46 */
    const int64_t n = (int64_t)ny * nx;
    const int64_t nux = (int64_t)ny * (nx > 1 ? nx - 1 : 1), nvy = (int64_t)(ny > 1 ? ny - 1 : 1) * nx;
    const double fa = cellsize * cellsize;
    T->nr_levels = nl;
    T->cell_area = coarse_res * coarse_res;
    for (int64_t c = 0; c < n; ++c) {
        T->z_zmin[c] = T->z_zmax[c] = T->z_volmax[c] = NAN;
        for (int k = 0; k < nl; ++k) T->z_level[c + k * n] = T->z_volume[c + k * n] = T->z_area[c + k * n] = NAN;
    }
    fill_face_nan(&T->u, nux, nl);
    fill_face_nan(&T->v, nvy, nl);
    for (int s = 0; s < 4; ++s) fill_face_nan(&T->side[s], n, nl);
    /* patch bounds (:107-110) */
    int *rs = malloc(sizeof(int) * ny), *re = malloc(sizeof(int) * ny);
    int *cs = malloc(sizeof(int) * nx), *ce = malloc(sizeof(int) * nx);
    for (int i = 0; i < ny; ++i) {
        rs[i] = (int)floor(i * coarse_res / cellsize);
        re[i] = (int)floor((i + 1) * coarse_res / cellsize);
        if (re[i] > fny) re[i] = fny;
    }
    for (int j = 0; j < nx; ++j) {
        cs[j] = (int)floor(j * coarse_res / cellsize);
        ce[j] = (int)floor((j + 1) * coarse_res / cellsize);
        if (ce[j] > fnx) ce[j] = fnx;
    }
    const int cap = 2 * (int)(r_round + 2) * (int)(r_round + 2) + 16;
    double *z = malloc(sizeof(double) * cap), *nn = malloc(sizeof(double) * cap);
    double *z2 = malloc(sizeof(double) * cap), *n2 = malloc(sizeof(double) * cap);
#define PATCH(i, j) ((patch_t){dem, rough, fny, fnx, rough_is_fine, ny, rs[i], re[i], cs[j], ce[j], i, j})
    /* cell tables (:117-139, sfincs_cell_table :186-207) */
    for (int j = 0; j < nx; ++j)
        for (int i = 0; i < ny; ++i) {
            const patch_t p = PATCH(i, j);
            int m = 0;
            for (int c = p.c0; c < p.c1; ++c)
                for (int rr = p.r0; rr < p.r1; ++rr) {
                    const double zz = patch_z(&p, rr, c);
                    if (isfinite(zz)) z[m++] = zz;
                }
            const int64_t cc = i + (int64_t)j * ny;
            if (m == 0) continue;
            const double zmin = vmin_of(z, m), zmax = vmax_of(z, m);
            double volmax = 0.0;
            for (int q = 0; q < m; ++q) volmax += fmax(zmax - z[q], 0.0);
            volmax *= fa;
            T->z_zmin[cc] = zmin;
            T->z_zmax[cc] = zmax;
            T->z_volmax[cc] = volmax;
            for (int k = 0; k < nl; ++k) {
                const double Vk = mat_linspace(0.0, volmax, nl, k);
                const double zl = invert_volume(z, m, Vk, zmin, zmax, volmax, fa);
                double wet = 0.0;
                for (int q = 0; q < m; ++q) wet += (double)(zl > z[q]);
                T->z_level[cc + k * n] = zl;
                T->z_volume[cc + k * n] = Vk;
                T->z_area[cc + k * n] = wet * fa;
            }
        }
    /* u faces between columns j and j+1 (:141-158; velocity_point_support 'u') */
    for (int j = 0; j + 1 < nx; ++j)
        for (int i = 0; i < ny; ++i) {
            const patch_t A = PATCH(i, j), B = PATCH(i, j + 1);
            const int64_t f = i + (int64_t)j * ny;
            if (A.c1 <= A.c0 || B.c1 <= B.c0 || A.r1 <= A.r0 || B.r1 <= B.r0) continue;
            const int hA = half_of(A.c1 - A.c0), hB = half_of(B.c1 - B.c0);
            const int mA = gather_valid(&A, A.r0, A.r1, A.c1 - hA, A.c1, z, nn);
            const int mB = gather_valid(&B, B.r0, B.r1, B.c0, B.c0 + hB, z2, n2);
            if (mA == 0 || mB == 0) continue;
            const double zmin = fmax(vmin_of(z, mA), vmin_of(z2, mB));
            const double zmax = fmax(vmax_of(z, mA), vmax_of(z2, mB));
            memcpy(z + mA, z2, sizeof(double) * mB);
            memcpy(nn + mA, n2, sizeof(double) * mB);
            put_face(&T->u, nux, f, nl, z, nn, mA + mB, zmin, zmax);
        }
    /* v faces between rows i and i+1 (:159-176; 'v') */
    for (int j = 0; j < nx; ++j)
        for (int i = 0; i + 1 < ny; ++i) {
            const patch_t A = PATCH(i, j), B = PATCH(i + 1, j);
            const int64_t f = i + (int64_t)j * (ny - 1);
            if (A.c1 <= A.c0 || B.c1 <= B.c0 || A.r1 <= A.r0 || B.r1 <= B.r0) continue;
            const int hA = half_of(A.r1 - A.r0), hB = half_of(B.r1 - B.r0);
            const int mA = gather_valid(&A, A.r1 - hA, A.r1, A.c0, A.c1, z, nn);
            const int mB = gather_valid(&B, B.r0, B.r0 + hB, B.c0, B.c1, z2, n2);
            if (mA == 0 || mB == 0) continue;
            const double zmin = fmax(vmin_of(z, mA), vmin_of(z2, mB));
            const double zmax = fmax(vmax_of(z, mA), vmax_of(z2, mB));
            memcpy(z + mA, z2, sizeof(double) * mB);
            memcpy(nn + mA, n2, sizeof(double) * mB);
            put_face(&T->v, nvy, f, nl, z, nn, mA + mB, zmin, zmax);
        }
    /* boundary tables north, south, west, east (:177-193; boundary_velocity_support) */
    for (int j = 0; j < nx; ++j)
        for (int i = 0; i < ny; ++i) {
            const patch_t p = PATCH(i, j);
            const int64_t cc = i + (int64_t)j * ny;
            if (p.c1 <= p.c0 || p.r1 <= p.r0) continue;
            int anyf = 0;
            for (int c = p.c0; c < p.c1 && !anyf; ++c)
                for (int rr = p.r0; rr < p.r1; ++rr)
                    if (isfinite(patch_z(&p, rr, c))) { anyf = 1; break; }
            if (!anyf) continue;
            const int hr = half_of(p.r1 - p.r0), hc = half_of(p.c1 - p.c0);
            for (int s = 0; s < 4; ++s) {
                int m;
                if (s == 0) m = gather_valid(&p, p.r0, p.r0 + hr, p.c0, p.c1, z, nn);      /* north */
                else if (s == 1) m = gather_valid(&p, p.r1 - hr, p.r1, p.c0, p.c1, z, nn); /* south */
                else if (s == 2) m = gather_valid(&p, p.r0, p.r1, p.c0, p.c0 + hc, z, nn); /* west */
                else m = gather_valid(&p, p.r0, p.r1, p.c1 - hc, p.c1, z, nn);            /* east */
                if (m == 0) continue;
                put_face(&T->side[s], n, cc, nl, z, nn, m, vmin_of(z, m), vmax_of(z, m));
            }
        }
#undef PATCH
    free(rs); free(re); free(cs); free(ce); free(z); free(nn); free(z2); free(n2);
    return 0;
}

/* ---------------------------------------------------------------- queries -- */

/* lookup_cell_table_by_z / lookup_cell_table_by_volume: the first bracketing
 * interval of the levels, else the first / last level */
static double cell_lookup(const double *xlev, const double *ylev, int64_t n, int64_t c, int nl,
                          double x) {
    for (int k = 0; k + 1 < nl; ++k) {
        const double x1 = xlev[c + k * n], x2 = xlev[c + (k + 1) * n];
        const double y1 = ylev[c + k * n], y2 = ylev[c + (k + 1) * n];
        if (isfinite(x1) && isfinite(x2) && isfinite(y1) && isfinite(y2) && x2 > x1 && x >= x1 &&
            x <= x2) {
            const double w = (x - x1) / (x2 - x1);
            return y1 + w * (y2 - y1);
        }
    }
    const double xf = xlev[c], yf = ylev[c];
    const double xl = xlev[c + (int64_t)(nl - 1) * n], yl = ylev[c + (int64_t)(nl - 1) * n];
    if (isfinite(xf) && isfinite(yf) && x <= xf) return yf;
    if (isfinite(xl) && isfinite(yl) && x >= xl) return yl;
    if (isfinite(yl)) return yl;
    return NAN; /* unassigned: zeros for by_z, NaN for by_volume (the callers fix it) */
}

static double cell_volume_from_zs(const hp2d_subgrid_tables *T, int64_t n, int64_t c, double zs) {
    const double zmin = T->z_zmin[c], zmax = T->z_zmax[c];
    double V = 0.0;
    const int valid = isfinite(zs) && isfinite(zmin) && isfinite(zmax);
    if (valid) {
        if (zs <= zmin) V = 0.0;
        else if (zs >= zmax) V = T->z_volmax[c] + (zs - zmax) * T->cell_area;
        else {
            V = cell_lookup(T->z_level, T->z_volume, n, c, T->nr_levels, zs);
            if (isnan(V)) V = 0.0;
        }
    }
    if (!isfinite(V)) V = 0.0;
    return fmax(V, 0.0);
}

static double zs_from_cell_volume(const hp2d_subgrid_tables *T, int64_t n, int64_t c, double V) {
    const double zmin = T->z_zmin[c], zmax = T->z_zmax[c];
    V = fmax(V, 0.0);
    double zs = zmin;
    const int valid = isfinite(V) && isfinite(zmin) && isfinite(zmax);
    if (!valid) return NAN;
    if (V >= T->z_volmax[c]) zs = zmax + (V - T->z_volmax[c]) / T->cell_area;
    else if (V > 0) zs = cell_lookup(T->z_volume, T->z_level, n, c, T->nr_levels, V);
    return zs;
}

static double wet_area_from_zs(const hp2d_subgrid_tables *T, int64_t n, int64_t c, double zs) {
    const double zmin = T->z_zmin[c], zmax = T->z_zmax[c];
    double A = 0.0;
    const int valid = isfinite(zs) && isfinite(zmin) && isfinite(zmax);
    if (valid) {
        if (zs >= zmax) A = T->cell_area;
        else if (zs > zmin) {
            A = cell_lookup(T->z_level, T->z_area, n, c, T->nr_levels, zs);
            if (isnan(A)) A = 0.0;
        }
    }
    if (!isfinite(A)) A = 0.0;
    return fmax(fmin(A, T->cell_area), 0.0);
}

/* lookup_levels of hp2d_sfincs_velocity_state.m */
static double level_lookup(const double *tab, int64_t nf, int64_t f, int nl, double zmin,
                           double zmax, double zu) {
    if (zmax <= zmin) return tab[f + (int64_t)(nl - 1) * nf];
    double pos = (zu - zmin) / (zmax - zmin);
    pos = fmax(fmin(pos, 1.0), 0.0);
    const double fidx = pos * (nl - 1) + 1.0;
    double i1 = floor(fidx);
    i1 = fmax(fmin(i1, (double)(nl - 1)), 1.0);
    const double w = fidx - i1;
    const double v1 = tab[f + (int64_t)(i1 - 1) * nf], v2 = tab[f + (int64_t)i1 * nf];
    return v1 + w * (v2 - v1);
}

/* query_table_set: face depth HG, representative n, wet fraction phi */
static void face_state(const hp2d_subgrid_faces *F, int64_t nf, int64_t f, int nl, double zu,
                       double *HG, double *nr, double *phi) {
    const double zmin = F->zmin[f], zmax = F->zmax[f];
    double havg = 0.0, nrep = INFINITY, pwet = 0.0;
    const int valid = isfinite(zu) && isfinite(zmin) && isfinite(zmax);
    if (valid && zu >= zmin && zu < zmax) {
        havg = level_lookup(F->havg, nf, f, nl, zmin, zmax, zu);
        nrep = level_lookup(F->nrep, nf, f, nl, zmin, zmax, zu);
        pwet = level_lookup(F->pwet, nf, f, nl, zmin, zmax, zu);
    }
    if (valid && zu >= zmax) {
        const double hM = F->havg[f + (int64_t)(nl - 1) * nf];
        double nM = F->nrep[f + (int64_t)(nl - 1) * nf];
        const double navg = F->navg[f];
        const double dz = zu - zmax;
        havg = hM + dz;
        if (!isfinite(nM) || nM <= 0) nM = navg;
        const double denom = F->ffit[f] * dz + 1.0;
        double na = navg - (navg - nM) / denom;
        if (!isfinite(na) || na <= 0) na = navg;
        nrep = na;
        pwet = 1.0;
    }
    if (valid && zu < zmin) {
        havg = 0.0;
        pwet = 0.0;
        nrep = INFINITY;
    }
    if (!isfinite(havg)) havg = 0.0;
    if (!isfinite(pwet)) pwet = 0.0;
    if (!isfinite(nrep) || nrep <= 0) nrep = INFINITY;
    *HG = fmax(havg, 0.0);
    *nr = nrep;
    *phi = fmax(fmin(pwet, 1.0), 0.0);
}

void oracle_subgrid_queries(const hp2d_subgrid_tables *T, int32_t ny, int32_t nx, const double *zs,
                            double *V, double *zs_back, double *A) {
    const int64_t n = (int64_t)ny * nx;
    for (int64_t c = 0; c < n; ++c) {
        if (V) V[c] = cell_volume_from_zs(T, n, c, zs[c]);
        if (zs_back) zs_back[c] = zs_from_cell_volume(T, n, c, V ? V[c] : cell_volume_from_zs(T, n, c, zs[c]));
        if (A) A[c] = wet_area_from_zs(T, n, c, zs[c]);
    }
}

void oracle_subgrid_face_state(const hp2d_subgrid_faces *F, int64_t nf, int32_t nl,
                               const double *zu, double *HG, double *nr, double *phi) {
    for (int64_t f = 0; f < nf; ++f) face_state(F, nf, f, nl, zu[f], &HG[f], &nr[f], &phi[f]);
}

/* ------------------------------------------------ volume change (rainfall) -- */

void oracle_subgrid_volume(const hp2d_subgrid_tables *T, int32_t ny, int32_t nx,
                           const double *drep_old, const double *delta_mm, double cell_area,
                           double *drep_new, double *V_old, double *V_new, double *A_new,
                           double *residual) {
    const int64_t n = (int64_t)ny * nx;
    for (int64_t c = 0; c < n; ++c) {
        const double dv = delta_mm[c] / 1000.0 * cell_area;
        const double zs_old = T->z_zmin[c] + fmax(drep_old[c], 0.0);
        const double vo = cell_volume_from_zs(T, n, c, zs_old);
        const double vn = fmax(vo + dv, 0.0);
        const double zs_new = zs_from_cell_volume(T, n, c, vn);
        drep_new[c] = fmax(zs_new - T->z_zmin[c], 0.0);
        if (V_old) V_old[c] = vo;
        if (V_new) V_new[c] = vn;
        if (A_new) A_new[c] = wet_area_from_zs(T, n, c, zs_new);
        if (residual) residual[c] = (vo + dv < 0) ? vn : vn - vo - dv;
    }
}

/* -------------------------------------------------------------------- step -- */

/* sfincs_lie_step (:210-219) */
static double lie_step(double qold, double HG, double nrep, double slope, double dt, double g,
                       int active) {
    if (!active) return 0.0;
    const double den = 1.0 + g * dt * (nrep * nrep) * fabs(qold) / fmax(pow(HG, 7.0 / 3.0), EPS);
    const double rhs = qold - g * dt * HG * slope;
    const double q = rhs / den;
    return isfinite(q) ? q : 0.0;
}

int32_t oracle_li_subgrid(int32_t ny, int32_t nx, const hp2d_subgrid_tables *T,
                          const double *d_tot, double time_step_min, double R,
                          int32_t outlet_type, double slope_outlet, const int64_t *outlet_idx,
                          const int32_t *outlet_side, int32_t n_outlet,
                          const double *outflow_prev, int32_t n_prev_pages,
                          const uint8_t *idx_nan, double *outflow, double *d_t,
                          double *I_tot_end_cell, double *Hf_out, double *C_a_out,
                          double *eta_out, double *V_out, int32_t *n_sub) {
    const int64_t n = (int64_t)ny * nx;
    const double dt_total = time_step_min * 60.0, g = 9.81;
    const int nl = T->nr_levels;
    const int64_t nux = (int64_t)ny * (nx > 1 ? nx - 1 : 1), nvy = (int64_t)(ny > 1 ? ny - 1 : 1) * nx;
    double *eta = malloc(sizeof(double) * n), *V = malloc(sizeof(double) * n);
    double *Ca = malloc(sizeof(double) * n), *qG = calloc((size_t)(2 * n), sizeof(double));
    double *qc = calloc((size_t)(2 * n), sizeof(double)), *Q = calloc((size_t)(2 * n), sizeof(double));
    double *Hx = calloc((size_t)n, sizeof(double)), *Hy = calloc((size_t)n, sizeof(double));
    double *Hxc = calloc((size_t)n, sizeof(double)), *Hyc = calloc((size_t)n, sizeof(double));
    double *scale = malloc(sizeof(double) * n), *Vc = malloc(sizeof(double) * n);
    double *outvol = calloc((size_t)n, sizeof(double)), *Hf = calloc((size_t)(3 * n), sizeof(double));
    double *Hsub3 = calloc((size_t)n, sizeof(double)), *oflow = calloc((size_t)n, sizeof(double));
    for (int64_t c = 0; c < n; ++c) { /* :28-33 */
        eta[c] = T->z_zmin[c] + fmax(d_tot[c] / 1000.0, 0.0);
        V[c] = cell_volume_from_zs(T, n, c, eta[c]);
        Ca[c] = wet_area_from_zs(T, n, c, eta[c]);
    }
    if (outflow_prev && n_prev_pages > 0) { /* :37-43 */
        const int np = n_prev_pages < 2 ? n_prev_pages : 2;
        for (int p = 0; p < np; ++p)
            for (int64_t c = 0; c < n; ++c) qG[p * n + c] = outflow_prev[p * n + c] / 1000.0 / 3600.0 * R;
    }
    double dt_remaining = dt_total, dt_try = dt_total;
    const double min_sub_dt = fmax(dt_total / 64.0, 1e-4);
    int subs = 0, guard = 0;
    while (dt_remaining > 1e-12) {
        if (++guard > 1000000) break;
        const double dt_sub = fmin(dt_try, dt_remaining);
        /* sfincs_flux_update (:168-208) */
        memset(qc, 0, sizeof(double) * 2 * n);
        memset(Hxc, 0, sizeof(double) * n);
        memset(Hyc, 0, sizeof(double) * n);
        for (int32_t j = 0; j + 1 < nx; ++j)
            for (int32_t i = 0; i < ny; ++i) {
                const int64_t a = i + (int64_t)j * ny, b = a + ny, f = a;
                const double zu = fmax(eta[a], eta[b]);
                double HG, nr, phi;
                face_state(&T->u, nux, f, nl, zu, &HG, &nr, &phi);
                const double slope = (eta[b] - eta[a]) / R;
                const int inact = idx_nan && (idx_nan[a] || idx_nan[b]);
                const int active = !inact && HG > 0 && phi > 0 && isfinite(nr) && nr > 0 && isfinite(slope);
                qc[a] = lie_step(qG[a], HG, nr, slope, dt_sub, g, active);
                Hxc[a] = HG;
            }
        for (int32_t j = 0; j < nx; ++j)
            for (int32_t i = 0; i + 1 < ny; ++i) {
                const int64_t a = i + (int64_t)j * ny, b = a + 1, f = i + (int64_t)j * (ny - 1);
                const double zu = fmax(eta[a], eta[b]);
                double HG, nr, phi;
                face_state(&T->v, nvy, f, nl, zu, &HG, &nr, &phi);
                const double slope = (eta[b] - eta[a]) / R;
                const int inact = idx_nan && (idx_nan[a] || idx_nan[b]);
                const int active = !inact && HG > 0 && phi > 0 && isfinite(nr) && nr > 0 && isfinite(slope);
                qc[n + a] = lie_step(qG[n + a], HG, nr, slope, dt_sub, g, active);
                Hyc[a] = HG;
            }
        /* :62-64 and limit_face_discharge_by_available_volume (:132-166) */
        for (int64_t k = 0; k < 2 * n; ++k) Q[k] = qc[k] * R;
        for (int32_t j = 0; j < nx; ++j)
            for (int32_t i = 0; i < ny; ++i) {
                const int64_t c = i + (int64_t)j * ny;
                const double Qw = j > 0 ? Q[c - ny] : 0.0, Qe = Q[c];
                const double Qn = i > 0 ? Q[n + c - 1] : 0.0, Qs = Q[n + c];
                const double out = fmax(Qe, 0.0) + fmax(-Qw, 0.0) + fmax(Qs, 0.0) + fmax(-Qn, 0.0);
                double s = 1.0;
                if (out * dt_sub > fmax(V[c], 0.0) && out > 0) s = fmax(V[c], 0.0) / (out * dt_sub);
                scale[c] = fmax(fmin(s, 1.0), 0.0);
            }
        for (int32_t j = 0; j + 1 < nx; ++j)
            for (int32_t i = 0; i < ny; ++i) {
                const int64_t c = i + (int64_t)j * ny;
                Q[c] = Q[c] * (Q[c] >= 0 ? scale[c] : scale[c + ny]);
            }
        for (int32_t j = 0; j < nx; ++j)
            for (int32_t i = 0; i + 1 < ny; ++i) {
                const int64_t c = i + (int64_t)j * ny;
                Q[n + c] = Q[n + c] * (Q[n + c] >= 0 ? scale[c] : scale[c + 1]);
            }
        for (int64_t k = 0; k < 2 * n; ++k) qc[k] = Q[k] / R;
        /* :65-82 */
        int negative = 0;
        for (int32_t j = 0; j < nx; ++j)
            for (int32_t i = 0; i < ny; ++i) {
                const int64_t c = i + (int64_t)j * ny;
                const double Qw = j > 0 ? Q[c - ny] : 0.0, Qe = Q[c];
                const double Qn = i > 0 ? Q[n + c - 1] : 0.0, Qs = Q[n + c];
                Vc[c] = V[c] + dt_sub * (Qw - Qe + Qn - Qs);
                if (Vc[c] < -1e-9) negative = 1;
            }
        if (negative && dt_sub > min_sub_dt) {
            dt_try = fmax(0.5 * dt_sub, min_sub_dt);
            continue;
        }
        ++subs;
        for (int64_t c = 0; c < n; ++c) { /* :84-86 */
            V[c] = fmax(Vc[c], 0.0);
            eta[c] = zs_from_cell_volume(T, n, c, V[c]);
            Ca[c] = wet_area_from_zs(T, n, c, eta[c]);
            oflow[c] = 0.0;
            Hsub3[c] = 0.0;
        }
        if (n_outlet > 0) { /* apply_sfincs_outlet (:242-284) */
            double *Qo = calloc((size_t)n_outlet, sizeof(double)), *Ho = calloc((size_t)n_outlet, sizeof(double));
            for (int32_t k = 0; k < n_outlet; ++k) {
                const int64_t c = outlet_idx[k];
                double HG, nr, phi;
                face_state(&T->side[outlet_side[k]], n, c, nl, eta[c], &HG, &nr, &phi);
                if (HG > 0 && isfinite(nr) && nr > 0) {
                    const double qGo = outlet_type == 1
                                           ? (1.0 / nr) * pow(HG, 5.0 / 3.0) * sqrt(fmax(fabs(slope_outlet), 0.0))
                                           : HG * sqrt(g * fmax(HG, 0.0));
                    Qo[k] = qGo * R;
                    Ho[k] = HG;
                }
            }
            for (int32_t k = 0; k < n_outlet; ++k) {
                const int64_t c = outlet_idx[k];
                const double avail = fmax(V[c], 0.0) / dt_sub;
                Qo[k] = fmin(fmax(Qo[k], 0.0), avail);
            }
            double *Vnew = malloc(sizeof(double) * n_outlet);
            for (int32_t k = 0; k < n_outlet; ++k) {
                const int64_t c = outlet_idx[k];
                oflow[c] = Qo[k] / (R * R) * 1000.0 * 3600.0;
                Vnew[k] = fmax(V[c] - Qo[k] * dt_sub, 0.0);
            }
            for (int32_t k = 0; k < n_outlet; ++k) V[outlet_idx[k]] = Vnew[k];
            for (int64_t c = 0; c < n; ++c) {
                eta[c] = zs_from_cell_volume(T, n, c, V[c]);
                Ca[c] = wet_area_from_zs(T, n, c, eta[c]);
            }
            for (int32_t k = 0; k < n_outlet; ++k) Hsub3[outlet_idx[k]] = Ho[k];
            free(Qo); free(Ho); free(Vnew);
        }
        for (int64_t c = 0; c < n; ++c) { /* :99-107 */
            outvol[c] = outvol[c] + oflow[c] / 1000.0 / 3600.0 * (R * R) * dt_sub;
            Hf[c] = Hxc[c];
            Hf[n + c] = Hyc[c];
            Hf[2 * n + c] = Hsub3[c];
        }
        memcpy(qG, qc, sizeof(double) * 2 * n);
        memcpy(Hx, Hxc, sizeof(double) * n);
        memcpy(Hy, Hyc, sizeof(double) * n);
        dt_remaining = dt_remaining - dt_sub;
        dt_try = fmin(dt_try * 1.25, fmax(dt_remaining, min_sub_dt));
    }
    /* :113-124 */
    for (int64_t c = 0; c < n; ++c) {
        const double o1 = qG[c] / R * 1000.0 * 3600.0, o2 = qG[n + c] / R * 1000.0 * 3600.0;
        const double o3 = outvol[c] / (R * R) * 1000.0 * 3600.0 / dt_total;
        outflow[c] = o1;
        outflow[n + c] = o2;
        outflow[2 * n + c] = o3;
        d_t[c] = fmax(eta[c] - T->z_zmin[c], 0.0) * 1000.0;
        I_tot_end_cell[c] = fabs(o1 + o2 + o3) * dt_total / 1000.0 / 3600.0 * (R * R);
        if (C_a_out) C_a_out[c] = Ca[c];
        if (eta_out) eta_out[c] = eta[c];
        if (V_out) V_out[c] = V[c];
    }
    if (Hf_out) memcpy(Hf_out, Hf, sizeof(double) * 3 * n);
    if (n_sub) *n_sub = subs;
    free(eta); free(V); free(Ca); free(qG); free(qc); free(Q); free(Hx); free(Hy); free(Hxc);
    free(Hyc); free(scale); free(Vc); free(outvol); free(Hf); free(Hsub3); free(oflow);
    return guard > 1000000 ? -1 : 0;
}
