// hp2d_inf.hip - mainline infiltration step (Infiltration_Module.m:101-583,
// flag_infiltration = 1, no subgrid, no layered soil) for gfx950.
//
// One elementwise kernel evaluates every cell: the van Genuchten / Mualem
// state of the unsaturated-zone bucket, the wetted top layer's conductivity,
// the surface-interface capacity with its Darcy driving head, the supply and
// storage limits, and the I_t / d_t / cumulative updates.  The FP64 storage
// sums of errors(4) and inf_volume are per-column sums in row order (one
// thread per column), finished on the host in column order: the association
// of the reference's nansum(nansum(.)).
#include <cmath>
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_colred.hpp"
#include "hp2d_handle.hpp"

namespace hp2d {

struct InfArgs {
    int ny, nx;
    long long n;
    double dt_h, A, C_a, Ltop, dh_max, wK, l_vg, time_step_min;
    const double *elev, *h_t, *soil_depth, *theta_s, *theta_r, *n_vg, *alpha, *ksat;
    const uint8_t *imp;
    double *d_t, *I_t, *I_p, *f, *cum, *inf_cell;
};

__global__ __launch_bounds__(256) void infiltration_kernel(InfArgs a) {
    const long long c = blockIdx.x * 256LL + threadIdx.x;
    if (c >= a.n) return;
    const double eff = a.d_t[c];
    const double i_a = eff / fmax(a.dt_h, 1e-12);
    const double h0 = fmax(eff, 0.0) / 1000.0;
    const double gw_depth = a.h_t[c] - (a.elev[c] - a.soil_depth[c]);
    double zwt = a.soil_depth[c] - gw_depth;
    zwt = fmax(zwt, 0.0);
    const double ts = a.theta_s[c], tr = a.theta_r[c];
    const double uz_max = zwt * (ts - tr);
    const double I_in = a.I_t[c];
    const double S = fmax(I_in, 0.0) / 1000.0;
    const double S_rem = fmax(uz_max - S, 0.0);
    const bool no_uz = (zwt <= 0.0) || (uz_max <= 0.0);
    const double nv = a.n_vg[c], al = a.alpha[c];
    const double m = 1.0 - 1.0 / nv;
    const double Ks = a.ksat[c] / 1000.0 / 3600.0;
    const double zwt_safe = fmax(zwt, 1e-6);
    double th_b = tr + (S / zwt_safe);
    th_b = fmax(th_b, tr);
    th_b = fmin(th_b, ts);
    double Se_b = (th_b - tr) / fmax(ts - tr, 1e-12);
    Se_b = fmin(fmax(Se_b, 1e-6), 1.0);
    double h_b = -(1.0 / al) * pow(pow(Se_b, -1.0 / m) - 1.0, 1.0 / nv);
    if (Se_b >= 0.999999) h_b = 0.0;
    const double w_avail = fmax(eff, 0.0) / 1000.0;
    const double w_top_def = a.Ltop * fmax(ts - th_b, 0.0);
    double w_wet = fmin(w_avail, S_rem);
    w_wet = fmin(w_wet, w_top_def);
    w_wet = fmax(w_wet, 0.0);
    double th_t = th_b + w_wet / a.Ltop;
    th_t = fmax(th_t, tr);
    th_t = fmin(th_t, ts);
    double Se_t = (th_t - tr) / fmax(ts - tr, 1e-12);
    Se_t = fmin(fmax(Se_t, 1e-6), 1.0);
    double term = 1.0 - pow(Se_t, 1.0 / m);
    term = fmin(fmax(term, 0.0), 1.0);
    double Kr = pow(Se_t, a.l_vg) * pow(1.0 - pow(term, m), 2.0);
    Kr = fmin(fmax(Kr, 0.0), 1.0);
    const double Ktop = Ks * Kr;
    const double Ksurf = Ks + 0.0;
    const bool wet = (i_a > 0.0) || (h0 > 0.0);
    double Ksoil = Ktop;
    if (wet) Ksoil = a.wK * Ksurf + (1.0 - a.wK) * Ktop;
    Ksoil = fmin(fmax(Ksoil, 0.0), Ksurf);
    double dh = h0 - h_b;
    dh = fmax(dh, 0.0);
    dh = fmin(dh, a.dh_max);
    const double grad = (dh / a.Ltop) + 1.0;
    double C = Ksoil * grad;
    C = C * 1000.0 * 3600.0;
    const double f_store = (S_rem * 1000.0) / fmax(a.dt_h, 1e-12);
    double f = fmin(i_a, C);
    f = fmin(f, f_store);
    if (a.imp && a.imp[c]) f = 0.0;
    if (no_uz) f = 0.0;
    if (isnan(f)) f = 0.0;
    f = fmax(f, 0.0);
    a.f[c] = f;
    a.I_p[c] = I_in;
    const double inf_cell = f * a.dt_h;
    double It = I_in + inf_cell;
    It = fmin(It, uz_max * 1000.0);
    It = fmax(It, 0.0);
    a.I_t[c] = It;
    const double dn = eff - inf_cell * a.A / a.C_a;
    a.d_t[c] = dn < 0.0 ? 0.0 : dn;
    if (a.cum) a.cum[c] = a.cum[c] + f * (a.time_step_min / 60.0) / 1000.0;
    a.inf_cell[c] = inf_cell * a.A;
}

// per column j: [0] nansum(A*I/1000), [1] nansum(A*d/1000) of the given
// arrays; with v != NULL also [2] nansum(v)
__global__ __launch_bounds__(kColThreads) void inf_colsum_kernel(int ny, int nx, double A,
                                                                 const double *I, const double *d,
                                                                 const double *v, double *out) {
    const bool is_max[3] = {false, false, false};
    column_reduce<3>(ny, blockIdx.x, is_max, [&](long long c, double (&x)[3]) {
        x[0] = A * I[c] / 1000.0;
        x[1] = A * d[c] / 1000.0;
        x[2] = v ? v[c] : 0.0;
    }, out);
}

}  // namespace hp2d

extern "C" hp2d_status hp2d_infiltration_module(
    hp2d_handle *h, double time_step_min, double resolution, double C_a, const double *elevation,
    const double *h_t, const double *soil_depth, const double *theta_sat, const double *theta_r,
    const double *n_vg, const double *alpha_vg, const double *ksat_mmh, const uint8_t *idx_imp,
    double l_vg, double Ltop, double dh_max, double Ksurf_weight, const double *d_t,
    const double *I_t, double *d_t_out, double *I_t_out, double *I_p, double *f, double *cum_inf,
    double *inf_volume, double *dS_inf) {
    using namespace hp2d;
    if (!h || !elevation || !h_t || !soil_depth || !theta_sat || !theta_r || !n_vg || !alpha_vg ||
        !ksat_mmh || !d_t || !I_t || !d_t_out || !I_t_out || !f || !(resolution > 0.0) ||
        !(C_a > 0.0))
        return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const size_t n = (size_t)ny * nx, nb = n * sizeof(double);
    hipStream_t s = hp2d_handle_stream(h);
    // workspace: 8 inputs + d, I, I_p, f, cum, inf_cell | colsums before/after | mask
    const size_t n_arrays = 14;
    const size_t col_off = n_arrays * nb;
    const size_t mask_off = col_off + 6 * (size_t)nx * sizeof(double);
    char *ws = (char *)hp2d_handle_workspace(h, mask_off + n + 256);
    if (!ws) return HP2D_E_NOMEM;
    double *p = (double *)ws;
    auto take = [&]() {
        double *r = p;
        p += n;
        return r;
    };
    InfArgs a;
    a.ny = ny;
    a.nx = nx;
    a.n = (long long)n;
    a.dt_h = time_step_min / 60.0;
    a.A = resolution * resolution;
    a.C_a = C_a;
    a.Ltop = std::fmax(Ltop, 1e-6);
    a.dh_max = dh_max;
    a.wK = std::fmin(std::fmax(Ksurf_weight, 0.0), 1.0);
    a.l_vg = l_vg;
    a.time_step_min = time_step_min;
    double *de = take(), *dh = take(), *dsd = take(), *dts = take(), *dtr = take(), *dn = take(),
           *dal = take(), *dks = take(), *dd = take(), *dI = take(), *dIp = take(), *df = take(),
           *dcum = take(), *dinf = take();
    double *col0 = (double *)(ws + col_off), *col1 = col0 + 3 * (size_t)nx;
    uint8_t *dimp = idx_imp ? (uint8_t *)(ws + mask_off) : nullptr;
    a.elev = de; a.h_t = dh; a.soil_depth = dsd; a.theta_s = dts; a.theta_r = dtr; a.n_vg = dn;
    a.alpha = dal; a.ksat = dks; a.imp = dimp;
    a.d_t = dd; a.I_t = dI; a.I_p = dIp; a.f = df; a.cum = cum_inf ? dcum : nullptr; a.inf_cell = dinf;
    hipError_t e = hipSuccess;
    auto H2D = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    auto D2H = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    H2D(de, elevation, nb); H2D(dh, h_t, nb); H2D(dsd, soil_depth, nb); H2D(dts, theta_sat, nb);
    H2D(dtr, theta_r, nb); H2D(dn, n_vg, nb); H2D(dal, alpha_vg, nb); H2D(dks, ksat_mmh, nb);
    H2D(dd, d_t, nb); H2D(dI, I_t, nb);
    if (cum_inf) H2D(dcum, cum_inf, nb);
    if (dimp) H2D(dimp, idx_imp, n);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(inf_colsum_kernel, dim3(nx), dim3(kColThreads), 0, s, ny, nx, a.A,
                           (const double *)dI, (const double *)dd, (const double *)nullptr, col0);
        hipLaunchKernelGGL(infiltration_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(inf_colsum_kernel, dim3(nx), dim3(kColThreads), 0, s, ny, nx, a.A,
                           (const double *)dI, (const double *)dd, (const double *)dinf, col1);
        e = hipGetLastError();
    }
    std::vector<double> c0(3 * (size_t)nx), c1(3 * (size_t)nx);
    D2H(c0.data(), col0, c0.size() * 8);
    D2H(c1.data(), col1, c1.size() * 8);
    D2H(d_t_out, dd, nb);
    D2H(I_t_out, dI, nb);
    D2H(f, df, nb);
    if (I_p) D2H(I_p, dIp, nb);
    if (cum_inf) D2H(cum_inf, dcum, nb);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_infiltration_module: ") + hipGetErrorString(e));
    double s_uz0 = 0, s_p0 = 0, s_uz1 = 0, s_p1 = 0, vol = 0;
    for (int j = 0; j < nx; ++j) {
        s_uz0 += c0[3 * j];
        s_p0 += c0[3 * j + 1];
        s_uz1 += c1[3 * j];
        s_p1 += c1[3 * j + 1];
        vol += c1[3 * j + 2];
    }
    if (inf_volume) *inf_volume = 1.0 / 1000.0 * vol;
    if (dS_inf) *dS_inf = (s_uz1 + s_p1) - (s_uz0 + s_p0);
    return HP2D_OK;
}
