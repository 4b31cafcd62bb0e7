// hp2d_ca.hip - cellular-automata D4 routing (CA_Routing.m:19-171) for gfx950.
//
// One lane per cell, radius-1 stencil: the 4 neighbour elevations and depths
// come from coalesced loads of the adjacent rows (same column, +-1 row) and the
// adjacent columns (+-ny).  The weights, Manning velocity, transferable
// volume and the 5 outflow rates of a cell are computed in registers.  As in
// the reference, the inflow gather and depth update are the caller's
// (HydroPol2D_Main_While.m:631-660).
#include <vector>

#include "hp2d.h"
#include "hp2d_handle.hpp"
#include "hp2d_internal.hpp"

namespace hp2d {

struct CaArgs {
    int ny, nx;
    double R, cell_area, time_step_min;
    int outlet_type, flag_critical, reset_prev;  // reset_prev: sum(I_tot_end_previous) == 0
    double slope_outlet;
    const double *z, *d_tot, *rough, *h0, *Iprev;
    const uint8_t *outlet;
    double *ql, *qr, *qu, *qd, *qo, *d_t, *Itot;
};

__global__ __launch_bounds__(256) void ca_kernel(CaArgs a) {
    const long long n = (long long)a.ny * a.nx;
    const long long c = blockIdx.x * 256LL + threadIdx.x;
    if (c >= n) return;
    const int ny = a.ny, nx = a.nx;
    const int i = (int)(c % ny), j = (int)(c / ny);
    const double R = a.R, h_min = 1e-6;
    const double depth = a.d_tot[c] / 1000.0;
    const double wse = a.z[c] + depth;
    double dh[5];
    // :51-54 (neighbours outside the grid read as elevation 0, depth 0)
    dh[0] = (j > 0) ? wse - a.z[c - ny] - a.d_tot[c - ny] / 1000.0 : wse - 0.0 - 0.0;
    dh[1] = (j < nx - 1) ? wse - a.z[c + ny] - a.d_tot[c + ny] / 1000.0 : wse - 0.0 - 0.0;
    dh[2] = (i > 0) ? wse - a.z[c - 1] - a.d_tot[c - 1] / 1000.0 : wse - 0.0 - 0.0;
    dh[3] = (i < ny - 1) ? wse - a.z[c + 1] - a.d_tot[c + 1] / 1000.0 : wse - 0.0 - 0.0;
    // :58-72 outlet pseudo-neighbour
    const bool is_out = a.outlet && a.outlet[c];
    if (a.outlet_type == 1) {
        dh[4] = a.slope_outlet * R * (is_out ? 1.0 : 0.0);
    } else {
        const double S0 = is_out ? (pow(depth, -1.0 / 6.0)) * sqrt(9.81) * a.rough[c] : 0.0;
        dh[4] = S0 * R;
    }
    // :80-87
    if (j == 0) dh[0] = 0.0;
    if (j == nx - 1) dh[1] = 0.0;
    if (i == 0) dh[2] = 0.0;
    if (i == ny - 1) dh[3] = 0.0;
    for (int p = 0; p < 5; ++p)
        if (dh[p] < h_min || isnan(dh[p])) dh[p] = 0.0;
    // :90-106
    const double cbig = R * R * 1000.0;
    double vt = 0.0, vmin = NAN, maxh = NAN;
    for (int p = 0; p < 5; ++p) {
        const double v = a.cell_area * dh[p];
        vt += v;
        vmin = fmin(vmin, (dh[p] < h_min || isnan(dh[p])) ? cbig : v);
        maxh = fmax(maxh, dh[p]);
    }
    double w[5], wmax = NAN;
    for (int p = 0; p < 5; ++p) {
        double ww = (a.cell_area * dh[p]) / (vt + vmin);
        if (isnan(ww)) ww = 0.0;
        w[p] = ww;
        wmax = fmax(wmax, ww);
    }
    // :110-129
    const double h0 = a.h0 ? a.h0[c] : 0.0;
    const double man = 1.0 / a.rough[c] * pow(fmax(depth - h0 / 1000.0, 0.0), 2.0 / 3.0) *
                       sqrt(maxh / R);
    const double vm = (a.flag_critical == 1) ? fmin(sqrt(9.81 * depth), man) : man;
    const double Iprev = a.reset_prev ? 0.9 * a.cell_area * depth : a.Iprev[c];
    double I = fmin(depth * a.cell_area, vm * depth / wmax * (R * a.time_step_min * 60.0));
    I = fmin(I, Iprev + vmin);
    if (a.d_tot[c] < 1e-6 * 1000.0 || isnan(a.z[c])) I = 0.0;
    // :132-170
    const double dts = a.time_step_min * 60.0;
    double tot = 0.0, q[5];
    for (int p = 0; p < 5; ++p) {
        const double Ip = w[p] * I;
        tot += Ip;
        q[p] = Ip / dts / a.cell_area * 1000.0 * 3600.0;
    }
    if (a.ql) a.ql[c] = q[0];
    if (a.qr) a.qr[c] = q[1];
    if (a.qu) a.qu[c] = q[2];
    if (a.qd) a.qd[c] = q[3];
    if (a.qo) a.qo[c] = q[4];
    a.Itot[c] = tot;
    a.d_t[c] = a.d_tot[c] - tot / a.cell_area * 1000.0;
}

}  // namespace hp2d

extern "C" hp2d_status hp2d_ca_routing(hp2d_handle *h, const double *z, const double *d_tot,
                                       const double *roughness, double cell_area,
                                       double time_step_min, const double *h_0, double resolution,
                                       const double *I_prev, int32_t outlet_type,
                                       double slope_outlet, const int32_t *row_outlet,
                                       const int32_t *col_outlet, int32_t n_outlet,
                                       int32_t flag_critical, double *qout_left,
                                       double *qout_right, double *qout_up, double *qout_down,
                                       double *outlet_flow, double *d_t, double *I_tot_end_cell) {
    using namespace hp2d;
    if (!h || !z || !d_tot || !roughness || !I_prev || !d_t || !I_tot_end_cell ||
        (n_outlet > 0 && (!row_outlet || !col_outlet)))
        return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const size_t n = (size_t)ny * nx;
    std::vector<uint8_t> mask(n, 0);
    for (int k = 0; k < n_outlet; ++k) {
        if (row_outlet[k] < 1 || row_outlet[k] > ny || col_outlet[k] < 1 || col_outlet[k] > nx)
            return HP2D_E_ARG;
        mask[(size_t)(row_outlet[k] - 1) + (size_t)(col_outlet[k] - 1) * ny] = 1;
    }
    // :121 global condition, evaluated on the host copy (MATLAB sum order)
    double sprev = 0.0;
    for (int j = 0; j < nx; ++j) {
        double col = 0.0;
        for (int i = 0; i < ny; ++i) col += I_prev[(size_t)i + (size_t)j * ny];
        sprev += col;
    }
    hipStream_t s = hp2d_handle_stream(h);
    double *buf = nullptr;
    uint8_t *dmask = nullptr;
    const int nbuf = 12;
    hipError_t e = hipMalloc((void **)&buf, nbuf * n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&dmask, n);
    double *dz = buf, *dd = buf + n, *dr = buf + 2 * n, *dh0 = buf + 3 * n, *dIp = buf + 4 * n;
    double *o[7];
    for (int k = 0; k < 7; ++k) o[k] = buf + (5 + k) * n;
    if (e == hipSuccess) e = hipMemcpyAsync(dz, z, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dd, d_tot, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, roughness, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && h_0) e = hipMemcpyAsync(dh0, h_0, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dIp, I_prev, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dmask, mask.data(), n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        CaArgs a;
        a.ny = ny;
        a.nx = nx;
        a.R = resolution;
        a.cell_area = cell_area;
        a.time_step_min = time_step_min;
        a.outlet_type = outlet_type;
        a.flag_critical = flag_critical;
        a.reset_prev = (sprev == 0.0) ? 1 : 0;
        a.slope_outlet = slope_outlet;
        a.z = dz;
        a.d_tot = dd;
        a.rough = dr;
        a.h0 = h_0 ? dh0 : nullptr;
        a.Iprev = dIp;
        a.outlet = dmask;
        a.ql = o[0];
        a.qr = o[1];
        a.qu = o[2];
        a.qd = o[3];
        a.qo = o[4];
        a.d_t = o[5];
        a.Itot = o[6];
        const int blocks = (int)((n + 255) / 256);
        hipLaunchKernelGGL(ca_kernel, dim3(blocks), dim3(256), 0, s, a);
        e = hipGetLastError();
    }
    double *outs[7] = {qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t, I_tot_end_cell};
    for (int k = 0; k < 7 && e == hipSuccess; ++k)
        if (outs[k]) e = hipMemcpyAsync(outs[k], o[k], n * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (buf) hipFree(buf);
    if (dmask) hipFree(dmask);
    if (e != hipSuccess) return hp2d_handle_fail(h, std::string("hp2d_ca_routing: ") + hipGetErrorString(e));
    return HP2D_OK;
}
