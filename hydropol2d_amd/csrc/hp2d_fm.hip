// hp2d_fm.hip - Full_Momentum_Model_D4 (Full_Momentum_Model_D4.m:1-724, ordinary
// grid, flag_reservoir = 0) for gfx950: HLL finite volumes for depth and both
// momentum components, with the reference's fixed settings (source-term
// friction predictor with the exact implicit Manning solve, no post-flux
// corrector, 10 m/s velocity cap), optional hydrostatic reconstruction,
// the draining limiter, reflective-wall pressure and cell- or face-based
// outlets.
//
// Four passes of one lane per cell over the column-major grid:
//   fm_prep_kernel    activity, depth, momentum memory, bed-slope predictor,
//                     implicit friction, velocity cap (the flux momentum);
//   fm_flux_kernel    the HLL fluxes of the two faces a cell stores (its east
//                     x face; its north y face, stored at the southern cell as
//                     the reference stores it, positive from S to N);
//   fm_theta_kernel   the draining-limiter factor of each cell;
//   fm_update_kernel  limited fluxes of the cell's four faces, volume and
//                     momentum update, walls, outlets and every output.
// No reduction crosses cells, so a step is four launches and no host wait.
#include <cmath>
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_handle.hpp"
#include "hp2d_internal.hpp"

namespace hp2d {
namespace {

constexpr int kFmBlock = 256;

struct FmArgs {
    int ny, nx, n_prev, flag_HR, outlet_type, face_outlets;
    long long n;
    double dx, dt, ts_min, g, h_dry, vmax, slope_outlet, cell_area;
    const double *z, *d_tot, *rough, *n2in, *Ca_in, *prev;
    const uint8_t *idx_nan, *outlet_index, *outlet_cell;
    const uint8_t *face[4];  // right, left, up, down (NULL: none)
    double *h, *Ca, *n2, *huf, *hvf;
    uint8_t *act;
    double *Fxh, *FxuL, *FxuR, *Fxv, *Fyh, *FyvS, *FyvN, *Fyu, *Hfx, *Hfy, *theta;
    double *outflow, *d_t, *Itot, *Hf, *Ca_out;
};

__device__ __forceinline__ bool active_of(const FmArgs &a, long long c) {
    const bool fin = isfinite(a.z[c]);
    bool act = fin && !(a.idx_nan && a.idx_nan[c]);                // :76
    if (a.outlet_index && a.outlet_index[c] && fin) act = true;    // :79
    return act;
}

// clean_and_limit_momentum (:756-773)
__device__ __forceinline__ void clean_limit(double &hu, double &hv, double h, bool act, double h_dry,
                                            double vmax) {
    if (!isfinite(hu)) hu = 0.0;
    if (!isfinite(hv)) hv = 0.0;
    if (!(act && h > h_dry)) {
        hu = 0.0;
        hv = 0.0;
        return;
    }
    const double speed = sqrt(hu * hu + hv * hv) / fmax(h, h_dry);
    if (speed > vmax) {
        const double fac = vmax / fmax(speed, 1e-12);
        hu = hu * fac;
        hv = hv * fac;
    }
}

// hll_normal_flux (:896-955)
__device__ void hll(double hL, double qnL, double qtL, double hR, double qnR, double qtR, double g,
                    double h_dry, double &Fh, double &Fqn, double &Fqt) {
    const bool wetL = hL > h_dry, wetR = hR > h_dry;
    const double uL = wetL ? qnL / hL : 0.0, uR = wetR ? qnR / hR : 0.0;
    const double cL = sqrt(g * fmax(hL, 0.0)), cR = sqrt(g * fmax(hR, 0.0));
    const double FL_h = qnL, FL_qn = qnL * uL + 0.5 * g * (hL * hL), FL_qt = qtL * uL;
    const double FR_h = qnR, FR_qn = qnR * uR + 0.5 * g * (hR * hR), FR_qt = qtR * uR;
    const double sL = fmin(uL - cL, uR - cR), sR = fmax(uL + cL, uR + cR);
    double fh = 0.0, fn = 0.0, ft = 0.0;
    if (sL >= 0) {  // idxL; idxR assigned after it wins where both hold
        fh = FL_h;
        fn = FL_qn;
        ft = FL_qt;
    }
    if (sR <= 0) {
        fh = FR_h;
        fn = FR_qn;
        ft = FR_qt;
    }
    if (!(sL >= 0 || sR <= 0) && sR > sL) {
        const double den = sR - sL;
        fh = (sR * FL_h - sL * FR_h + sL * sR * (hR - hL)) / den;
        fn = (sR * FL_qn - sL * FR_qn + sL * sR * (qnR - qnL)) / den;
        ft = (sR * FL_qt - sL * FR_qt + sL * sR * (qtR - qtL)) / den;
    }
    if ((!wetL && !wetR) || !isfinite(fh) || !isfinite(fn) || !isfinite(ft)) fh = fn = ft = 0.0;
    Fh = fh;
    Fqn = fn;
    Fqt = ft;
}

__global__ __launch_bounds__(kFmBlock) void fm_prep_kernel(FmArgs a) {
    const long long c = (long long)xcd_logical_tile() * kFmBlock + threadIdx.x;
    if (c >= a.n) return;
    const int ny = a.ny, nx = a.nx, i = (int)(c % ny), j = (int)(c / ny);
    const bool act = active_of(a, c);
    double Ca = a.Ca_in ? a.Ca_in[c] : a.cell_area;
    if (!isfinite(Ca) || Ca <= 0) Ca = a.dx * a.dx;  // :69
    double h = fmax(a.d_tot[c] / 1000.0, 0.0);
    if (!act) h = 0.0;
    double n2 = a.n2in ? a.n2in[c] : a.rough[c] * a.rough[c];
    if (!isfinite(n2) || n2 < 0) n2 = 0.0;
    // recover_momentum_from_outflow (:725-747)
    double hu = 0.0, hv = 0.0;
    if (a.prev && a.n_prev > 0) {
        if (a.n_prev >= 5) {
            hu = a.prev[3 * a.n + c];
            hv = a.prev[4 * a.n + c];
        } else {
            hu = a.prev[c] / 1000.0 / 3600.0 * Ca / a.dx;
            if (a.n_prev >= 2) hv = a.prev[a.n + c] / 1000.0 / 3600.0 * Ca / a.dx;
        }
    }
    if (!isfinite(hu)) hu = 0.0;
    if (!isfinite(hv)) hv = 0.0;
    if (!act || h <= a.h_dry) hu = hv = 0.0;
    clean_limit(hu, hv, h, act, a.h_dry, a.vmax);
    // compute_centered_bed_gradients (:957-976)
    const bool aR = j + 1 < nx && active_of(a, c + ny), aL = j > 0 && active_of(a, c - ny);
    const bool aU = i > 0 && active_of(a, c - 1), aD = i + 1 < ny && active_of(a, c + 1);
    const double dx = a.dx;
    double gx = 0.0, gy = 0.0;
    if (act && aR && aL) gx = (a.z[c + ny] - a.z[c - ny]) / (2 * dx);
    else if (act && aR) gx = (a.z[c + ny] - a.z[c]) / dx;
    else if (act && aL) gx = (a.z[c] - a.z[c - ny]) / dx;
    if (act && aU && aD) gy = (a.z[c - 1] - a.z[c + 1]) / (2 * dx);
    else if (act && aU) gy = (a.z[c - 1] - a.z[c]) / dx;
    else if (act && aD) gy = (a.z[c] - a.z[c + 1]) / dx;
    if (!act || !isfinite(gx)) gx = 0.0;
    if (!act || !isfinite(gy)) gy = 0.0;
    double hus = hu - a.dt * a.g * h * gx, hvs = hv - a.dt * a.g * h * gy;
    if (!act) hus = hvs = 0.0;
    // apply_manning_friction_exact_implicit (:977-999)
    const bool wet = act && h > a.h_dry;
    double Q = sqrt(hus * hus + hvs * hvs);
    if (!isfinite(Q)) Q = 0.0;
    double Cf = wet ? a.g * n2 / pow(fmax(h, a.h_dry), 7.0 / 3.0) : 0.0;
    if (!isfinite(Cf) || Cf < 0) Cf = 0.0;
    const double aa = a.dt * Cf;
    double huf = 0.0, hvf = 0.0;
    if (wet && Q > 0) {
        const double r = 2.0 * Q / (1.0 + sqrt(1.0 + 4.0 * aa * Q));
        const double s = r / fmax(Q, 1e-30);
        huf = hus * s;
        hvf = hvs * s;
    }
    if (!isfinite(huf)) huf = 0.0;
    if (!isfinite(hvf)) hvf = 0.0;
    clean_limit(huf, hvf, h, act, a.h_dry, a.vmax);
    a.h[c] = h;
    a.Ca[c] = Ca;
    a.n2[c] = n2;
    a.act[c] = act ? 1 : 0;
    a.huf[c] = huf;
    a.hvf[c] = hvf;
}

__global__ __launch_bounds__(kFmBlock) void fm_flux_kernel(FmArgs a) {
    const long long c = (long long)xcd_logical_tile() * kFmBlock + threadIdx.x;
    if (c >= a.n) return;
    const int ny = a.ny, nx = a.nx, i = (int)(c % ny), j = (int)(c / ny);
    const double g = a.g, hd = a.h_dry;
    // x face between c and its east neighbour (:777-844 / :1136-1203)
    double fh = 0.0, fn = 0.0, ft = 0.0, cL = 0.0, cR = 0.0, hfx = 0.0;
    if (j + 1 < nx) {
        const long long b = c + ny;
        const double ha = a.h[c], hb = a.h[b];
        if (a.flag_HR) {
            const double za = a.z[c], zb = a.z[b], zf = fmax(za, zb);
            const double hLs = fmax(za + ha - zf, 0.0), hRs = fmax(zb + hb - zf, 0.0);
            const double uL = ha > hd ? a.huf[c] / ha : 0.0, vL = ha > hd ? a.hvf[c] / ha : 0.0;
            const double uR = hb > hd ? a.huf[b] / hb : 0.0, vR = hb > hd ? a.hvf[b] / hb : 0.0;
            hll(hLs, hLs * uL, hLs * vL, hRs, hRs * uR, hRs * vR, g, hd, fh, fn, ft);
            cL = 0.5 * g * (ha * ha - hLs * hLs);
            cR = 0.5 * g * (hb * hb - hRs * hRs);
            hfx = fmax(hLs, hRs);
        } else {
            hll(ha, a.huf[c], a.hvf[c], hb, a.huf[b], a.hvf[b], g, hd, fh, fn, ft);
            hfx = fmax(ha, hb);
        }
        if (!(a.act[c] && a.act[b])) fh = fn = ft = cL = cR = 0.0;
    }
    a.Fxh[c] = fh;
    a.FxuL[c] = fn + cL;
    a.FxuR[c] = fn + cR;
    a.Fxv[c] = ft;
    a.Hfx[c] = hfx;
    // y face between the northern cell c-1 and c (S), stored at c
    double gh = 0.0, gn = 0.0, gt = 0.0, cS = 0.0, cN = 0.0, hfy = 0.0;
    if (i > 0) {
        const long long nn = c - 1;
        const double hs = a.h[c], hN = a.h[nn];
        if (a.flag_HR) {
            const double zs = a.z[c], zn = a.z[nn], zf = fmax(zs, zn);
            const double hSs = fmax(zs + hs - zf, 0.0), hNs = fmax(zn + hN - zf, 0.0);
            const double uS = hs > hd ? a.huf[c] / hs : 0.0, vS = hs > hd ? a.hvf[c] / hs : 0.0;
            const double uN = hN > hd ? a.huf[nn] / hN : 0.0, vN = hN > hd ? a.hvf[nn] / hN : 0.0;
            hll(hSs, hSs * vS, hSs * uS, hNs, hNs * vN, hNs * uN, g, hd, gh, gn, gt);
            cS = 0.5 * g * (hs * hs - hSs * hSs);
            cN = 0.5 * g * (hN * hN - hNs * hNs);
            hfy = fmax(hSs, hNs);
        } else {
            hll(hs, a.hvf[c], a.huf[c], hN, a.hvf[nn], a.huf[nn], g, hd, gh, gn, gt);
            hfy = fmax(hs, hN);
        }
        if (!(a.act[c] && a.act[nn])) gh = gn = gt = cS = cN = 0.0;
    }
    a.Fyh[c] = gh;
    a.FyvS[c] = gn + cS;
    a.FyvN[c] = gn + cN;
    a.Fyu[c] = gt;
    a.Hfy[c] = hfy;
}

// apply_draining_limiter (:957-1013): theta of each cell
__global__ __launch_bounds__(kFmBlock) void fm_theta_kernel(FmArgs a) {
    const long long c = (long long)xcd_logical_tile() * kFmBlock + threadIdx.x;
    if (c >= a.n) return;
    const int ny = a.ny, nx = a.nx, i = (int)(c % ny), j = (int)(c / ny);
    const double k = a.dt * a.dx;
    double outV = 0.0;
    if (j + 1 < nx) outV += k * fmax(a.Fxh[c], 0.0);
    if (j > 0) outV += k * fmax(-a.Fxh[c - ny], 0.0);
    if (i > 0) outV += k * fmax(a.Fyh[c], 0.0);
    if (i + 1 < ny) outV += k * fmax(-a.Fyh[c + 1], 0.0);
    const bool act = a.act[c];
    double avail = fmax(a.h[c] - a.h_dry, 0.0) * a.Ca[c];
    if (!act) avail = 0.0;
    double th = 1.0;
    if (act && outV > avail && outV > 0) th = avail / fmax(outV, 1e-30);
    if (!isfinite(th)) th = 0.0;
    a.theta[c] = fmax(fmin(th, 1.0), 0.0);
}

// the outlet sink of one cell for one pass (:330-385 / apply_outlet_face_sink)
__device__ void outlet_pass(const FmArgs &a, long long c, bool accumulate, double Ca, double n2,
                            double &d_t, double &oflow, double &Hf3, double &hu, double &hv) {
    double h_out = fmax(d_t, 0.0) / 1000.0;
    if (!isfinite(h_out)) h_out = 0.0;
    double area = Ca;
    if (!isfinite(area) || area <= 0) area = a.dx * a.dx;
    double sqrtS;
    if (a.outlet_type == 1) {
        sqrtS = sqrt(fabs(a.slope_outlet));
    } else {
        const double hs = fmax(h_out, 1e-12);
        sqrtS = sqrt(a.g * n2 * pow(hs, -1.0 / 3.0));
        if (!isfinite(sqrtS)) sqrtS = 0.0;
    }
    Hf3 = h_out;
    double nout = a.rough[c];
    if (!isfinite(nout) || nout <= 0) nout = sqrt(fmax(n2, 1e-12));
    double q = (1.0 / nout) * a.dx * pow(h_out, 5.0 / 3.0) * sqrtS / area * 1000.0 * 3600.0;
    const double qmax = fmax(d_t - 1000.0 * 0.0, 0.0) / (a.ts_min / 60.0);
    q = fmin(q, qmax);
    if (!isfinite(q)) q = 0.0;
    const double hb = fmax(d_t / 1000.0, 0.0);
    oflow = accumulate ? oflow + q : q;
    d_t = d_t - q * (a.ts_min / 60.0);
    d_t = fmax(d_t, 1000.0 * 0.0);
    const double ha = fmax(d_t / 1000.0, 0.0);
    const double sc = hb > 1e-12 ? ha / hb : 1.0;
    hu = hu * sc;
    hv = hv * sc;
}

__global__ __launch_bounds__(kFmBlock) void fm_update_kernel(FmArgs a) {
    const long long c = (long long)xcd_logical_tile() * kFmBlock + threadIdx.x;
    if (c >= a.n) return;
    const int ny = a.ny, nx = a.nx, i = (int)(c % ny), j = (int)(c / ny);
    const long long n = a.n;
    const bool act = a.act[c];
    const double Ca = a.Ca[c], h = a.h[c], dtdx = a.dt * a.dx;
    // limited flux factors of the four faces (donor theta)
    auto sx = [&](long long f) { return a.Fxh[f] < 0 ? a.theta[f + ny] : a.theta[f]; };
    auto sy = [&](long long f) { return a.Fyh[f] < 0 ? a.theta[f - 1] : a.theta[f]; };
    const bool hasE = j + 1 < nx, hasW = j > 0, hasN = i > 0, hasS = i + 1 < ny;
    const double kE = hasE ? sx(c) : 0.0, kW = hasW ? sx(c - ny) : 0.0;
    const double kN = hasN ? sy(c) : 0.0, kS = hasS ? sy(c + 1) : 0.0;
    const double FE = hasE ? a.Fxh[c] * kE : 0.0, FW = hasW ? a.Fxh[c - ny] * kW : 0.0;
    const double FN = hasN ? a.Fyh[c] * kN : 0.0, FS = hasS ? a.Fyh[c + 1] * kS : 0.0;
    const double Vol = a.dt * a.dx * (FW - FE - FN + FS);  // :111-114
    double hu = a.huf[c], hv = a.hvf[c];                   // :115-119
    if (hasE) hu = hu - dtdx * (a.FxuL[c] * kE) / Ca;
    if (hasW) hu = hu + dtdx * (a.FxuR[c - ny] * kW) / Ca;
    if (hasE) hv = hv - dtdx * (a.Fxv[c] * kE) / Ca;
    if (hasW) hv = hv + dtdx * (a.Fxv[c - ny] * kW) / Ca;
    if (hasN) hu = hu - dtdx * (a.Fyu[c] * kN) / Ca;
    if (hasS) hu = hu + dtdx * (a.Fyu[c + 1] * kS) / Ca;
    if (hasN) hv = hv - dtdx * (a.FyvS[c] * kN) / Ca;
    if (hasS) hv = hv + dtdx * (a.FyvN[c + 1] * kS) / Ca;
    // apply_reflective_wall_pressure_fluxes (:1098-1134)
    const bool aR = hasE && a.act[c + ny], aL = hasW && a.act[c - ny];
    const bool aU = hasN && a.act[c - 1], aD = hasS && a.act[c + 1];
    double p = (act && h > a.h_dry) ? 0.5 * a.g * (h * h) : 0.0;
    if (!isfinite(p)) p = 0.0;
    double coef = a.dt * a.dx / Ca;
    if (!isfinite(coef) || coef < 0) coef = 0.0;
    const bool fR = a.face[0] && a.face[0][c], fL = a.face[1] && a.face[1][c];
    const bool fU = a.face[2] && a.face[2][c], fD = a.face[3] && a.face[3][c];
    if (act && !aL && !fL) hu = hu + coef * p;
    if (act && !aR && !fR) hu = hu - coef * p;
    if (act && !aU && !fU) hv = hv - coef * p;
    if (act && !aD && !fD) hv = hv + coef * p;
    if (!isfinite(hu)) hu = 0.0;
    if (!isfinite(hv)) hv = 0.0;
    // :170-198
    double hn = h + Vol / Ca;
    if (!act) hn = 0.0;
    hn = fmax(hn, 0.0);
    if (h > a.h_dry) {
        const double sc = fmin(1.0, hn / fmax(h, a.h_dry));
        hu = hu * sc;
        hv = hv * sc;
    }
    clean_limit(hu, hv, hn, act, a.h_dry, a.vmax);
    double d = a.d_tot[c] + 1000.0 * Vol / Ca;
    d = act ? fmax(d, 0.0) : NAN;
    double ox = FE * a.dx / Ca * 1000.0 * 3600.0, oy = FN * a.dx / Ca * 1000.0 * 3600.0;
    if (!isfinite(ox)) ox = 0.0;
    if (!isfinite(oy)) oy = 0.0;
    if (!act) ox = oy = 0.0;
    // outlets (:247-330)
    double oflow = 0.0, Hf3 = 0.0;
    if (a.face_outlets) {
        for (int k = 0; k < 4; ++k)
            if (a.face[k] && a.face[k][c] && act)
                outlet_pass(a, c, true, Ca, a.n2[c], d, oflow, Hf3, hu, hv);
    } else if (a.outlet_cell && a.outlet_cell[c]) {
        outlet_pass(a, c, false, Ca, a.n2[c], d, oflow, Hf3, hu, hv);
    }
    // :386-411
    double hf = fmax(d / 1000.0, 0.0);
    if (!act) hf = 0.0;
    clean_limit(hu, hv, hf, act, a.h_dry, a.vmax);
    a.d_t[c] = d;
    a.outflow[c] = ox;
    a.outflow[n + c] = oy;
    a.outflow[2 * n + c] = oflow;
    a.outflow[3 * n + c] = hu;
    a.outflow[4 * n + c] = hv;
    a.Hf[c] = a.Hfx[c];
    a.Hf[n + c] = a.Hfy[c];
    a.Hf[2 * n + c] = Hf3;
    const double s = ox + oy + oflow;
    a.Itot[c] = act ? fabs(s) * a.dt / 1000.0 * 1.0 / 3600.0 * Ca : NAN;
    a.Ca_out[c] = Ca;
}

int fm_grid(long long n) { return (int)((n + kFmBlock - 1) / kFmBlock); }

}  // namespace
}  // namespace hp2d

using namespace hp2d;

extern "C" hp2d_status hp2d_full_momentum_d4(
    hp2d_handle *h, const double *z, const double *d_tot, const double *roughness_cell,
    const double *roughness_squared, double cell_area, const double *cell_area_grid,
    double time_step_min, double resolution, const uint8_t *outlet_index, int32_t outlet_type,
    double slope_outlet, const int32_t *row_outlet, const int32_t *col_outlet, int32_t n_outlet,
    double d_tolerance, const double *outflow_prev, int32_t n_prev_pages, const uint8_t *idx_nan,
    int32_t flag_HR_full_momentum, const uint8_t *face_right, const uint8_t *face_left,
    const uint8_t *face_up, const uint8_t *face_down, double *outflow, double *d_t,
    double *I_tot_end_cell, double *Hf, double *C_a) {
    if (!h || !z || !d_tot || !roughness_cell || n_outlet < 0 ||
        (n_outlet > 0 && (!row_outlet || !col_outlet)) || n_prev_pages < 0 ||
        (n_prev_pages > 0 && !outflow_prev) || !(resolution > 0))
        return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const long long n = (long long)ny * nx;
    const size_t nb = (size_t)n * 8;
    // cell-based outlet mask (repeats collapse, as the reference's vector assignment does)
    std::vector<uint8_t> ocell((size_t)n, 0);
    for (int k = 0; k < n_outlet; ++k) {
        if (row_outlet[k] < 1 || row_outlet[k] > ny || col_outlet[k] < 1 || col_outlet[k] > nx)
            return hp2d_handle_fail(h, "hp2d_full_momentum_d4: outlet row/col off the grid");
        ocell[(size_t)(row_outlet[k] - 1) + (size_t)(col_outlet[k] - 1) * ny] = 1;
    }
    // face-based outlets when a face mask marks an active cell (:247-266)
    const uint8_t *faces[4] = {face_right, face_left, face_up, face_down};
    int face_outlets = 0;
    for (int k = 0; k < 4 && !face_outlets; ++k)
        if (faces[k])
            for (long long c = 0; c < n; ++c) {
                const bool fin = std::isfinite(z[c]);
                const bool act = (fin && !(idx_nan && idx_nan[c])) ||
                                 (outlet_index && outlet_index[c] && fin);
                if (faces[k][c] && act) {
                    face_outlets = 1;
                    break;
                }
            }
    // workspace (doubles): z d_tot rough n2 Ca prev[np] | h Ca n2 huf hvf | 10 flux arrays +
    // theta | outflow[5] d_t Itot Hf[3] Ca_out ; then bytes: idx_nan outlet_index ocell act faces[4]
    const int np = n_prev_pages;
    const size_t nd = 5 + (size_t)np + 5 + 11 + 5 + 2 + 3 + 1;
    char *ws = (char *)hp2d_handle_workspace(h, nd * nb + 8 * (size_t)n + 512);
    if (!ws) return HP2D_E_NOMEM;
    double *w = (double *)ws;
    auto take = [&](size_t k) {
        double *q = w;
        w += k * n;
        return q;
    };
    FmArgs a{};
    a.ny = ny;
    a.nx = nx;
    a.n = n;
    a.n_prev = np;
    a.flag_HR = flag_HR_full_momentum == 1 ? 1 : 0;
    a.outlet_type = outlet_type;
    a.face_outlets = face_outlets;
    a.dx = resolution;
    a.dt = time_step_min * 60.0;
    a.ts_min = time_step_min;
    a.g = 9.81;
    a.h_dry = std::fmax(d_tolerance / 1000.0, 1.0e-12);
    a.vmax = 10.0;
    a.slope_outlet = slope_outlet;
    a.cell_area = cell_area;
    double *dz = take(1), *dd = take(1), *dr = take(1), *dn2 = take(1), *dCa = take(1);
    double *dprev = take(np);
    a.z = dz;
    a.d_tot = dd;
    a.rough = dr;
    a.n2in = roughness_squared ? dn2 : nullptr;
    a.Ca_in = cell_area_grid ? dCa : nullptr;
    a.prev = np > 0 ? dprev : nullptr;
    a.h = take(1);
    a.Ca = take(1);
    a.n2 = take(1);
    a.huf = take(1);
    a.hvf = take(1);
    a.Fxh = take(1);
    a.FxuL = take(1);
    a.FxuR = take(1);
    a.Fxv = take(1);
    a.Fyh = take(1);
    a.FyvS = take(1);
    a.FyvN = take(1);
    a.Fyu = take(1);
    a.Hfx = take(1);
    a.Hfy = take(1);
    a.theta = take(1);
    a.outflow = take(5);
    a.d_t = take(1);
    a.Itot = take(1);
    a.Hf = take(3);
    a.Ca_out = take(1);
    uint8_t *bytes = (uint8_t *)w;
    uint8_t *dnan = bytes, *doi = bytes + n, *doc = bytes + 2 * n, *dact = bytes + 3 * n;
    uint8_t *dface = bytes + 4 * n;
    a.idx_nan = idx_nan ? dnan : nullptr;
    a.outlet_index = outlet_index ? doi : nullptr;
    a.outlet_cell = n_outlet > 0 ? doc : nullptr;
    a.act = dact;
    for (int k = 0; k < 4; ++k) a.face[k] = faces[k] ? dface + k * n : nullptr;
    hipStream_t s = hp2d_handle_stream(h);
    hipError_t e = hipSuccess;
    auto up = [&](void *dst, const void *src, size_t bytes_) {
        if (e == hipSuccess && src) e = hipMemcpyAsync(dst, src, bytes_, hipMemcpyHostToDevice, s);
    };
    up(dz, z, nb);
    up(dd, d_tot, nb);
    up(dr, roughness_cell, nb);
    up(dn2, roughness_squared, nb);
    up(dCa, cell_area_grid, nb);
    up(dprev, outflow_prev, (size_t)np * nb);
    up(dnan, idx_nan, (size_t)n);
    up(doi, outlet_index, (size_t)n);
    if (n_outlet > 0) up(doc, ocell.data(), (size_t)n);
    for (int k = 0; k < 4; ++k) up(dface + k * n, faces[k], (size_t)n);
    const int G = fm_grid(n);
    if (e == hipSuccess) e = hp2d_handle_mark(h, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fm_prep_kernel, dim3(G), dim3(kFmBlock), 0, s, a);
        hipLaunchKernelGGL(fm_flux_kernel, dim3(G), dim3(kFmBlock), 0, s, a);
        hipLaunchKernelGGL(fm_theta_kernel, dim3(G), dim3(kFmBlock), 0, s, a);
        hipLaunchKernelGGL(fm_update_kernel, dim3(G), dim3(kFmBlock), 0, s, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hp2d_handle_mark(h, 1);
    auto down = [&](double *dst, const double *src, size_t bytes_) {
        if (e == hipSuccess && dst) e = hipMemcpyAsync(dst, src, bytes_, hipMemcpyDeviceToHost, s);
    };
    down(outflow, a.outflow, 5 * nb);
    down(d_t, a.d_t, nb);
    down(I_tot_end_cell, a.Itot, nb);
    down(Hf, a.Hf, 3 * nb);
    down(C_a, a.Ca_out, nb);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_full_momentum_d4: ") + hipGetErrorString(e));
    return HP2D_OK;
}
