// hp2d_d8.hip - the eight-direction (D8) routing drop-ins for gfx950:
//   Bates_Inertial_8D   (Bates_Inertial_8D.m:2-274, flag_reservoir = 0)
//   CA_Routing_8D       (CA_Routing_8D.m:2-188, flag_reservoir = 0)
//   the caller's inflow gather and depth update (HydroPol2D_Main_While.m:631-660)
//
// One lane per cell over the column-major grid, 256 cells per workgroup.  A
// lane reads its 3x3 neighbourhood of elevation and depth (the centre column
// and the rows above and below come from the same and the adjacent cache
// lines; the columns either side from the adjacent workgroups' lines, which
// the XCD-aware workgroup order keeps in the same L2), evaluates all nine
// flow pages in registers and writes each page as one coalesced store.  The
// page arrays follow the reference order: left, right, up, down, outlet, NE,
// SE, SW, NW.
//
// Edge pages.  The reference builds every page by concatenation and then
// zeroes the entries whose neighbour lies off the grid; a page therefore
// carries a water-surface difference exactly when its neighbour is on the
// grid.  Its face depth Hf on an off-grid face is the cell's own depth, except
// for a diagonal face on the first or last column whose row neighbour exists:
// there the concatenation places the depth of the cell one row up (NE, NW) or
// down (SE, SW) (Bates_Inertial_8D.m:76-79), which is kept.
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_handle.hpp"
#include "hp2d_internal.hpp"

namespace hp2d {

constexpr int kD8Block = 256;
// row / column offset of the neighbour a page flows to (0, 0 for the outlet)
__device__ __forceinline__ constexpr int page_di(int p) {
    return (p == 2 || p == 5 || p == 8) ? -1 : ((p == 3 || p == 6 || p == 7) ? 1 : 0);
}
__device__ __forceinline__ constexpr int page_dj(int p) {
    return (p == 0 || p == 7 || p == 8) ? -1 : ((p == 1 || p == 5 || p == 6) ? 1 : 0);
}

struct D8Args {
    int ny, nx;
    long long n;
    double R, A, time_step_min;
    int outlet_type, flag_critical, reset_prev;
    double slope_outlet, d_tolerance;
    const double *z, *d_tot, *d_p, *rough, *h0, *Iprev, *qprev;
    const uint8_t *outlet;
    double *q, *Hf, *d_t, *Itot;
};

// the 3x3 neighbourhood of elevation z and water surface y = z + depth;
// entries off the grid are NaN and never read
template <bool BATES>
__device__ __forceinline__ void load_nbhd(const D8Args &a, int i, int j, double (&y)[3][3],
                                          double (&z)[3][3]) {
    const double *dsrc = BATES ? a.d_p : a.d_tot;
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
        for (int di = -1; di <= 1; ++di) {
            const int ii = i + di, jj = j + dj;
            double zz = NAN, yy = NAN;
            if (ii >= 0 && ii < a.ny && jj >= 0 && jj < a.nx) {
                const long long k = (long long)ii + (long long)jj * a.ny;
                zz = a.z[k];
                double d = fmax(dsrc[k] / 1000.0, 0.0);  // max(d/1000, 0) omits NaN
                if (BATES && d <= 0.0) d = 0.0;          // Bates_Inertial_8D.m:43-44
                yy = zz + d;
            }
            z[di + 1][dj + 1] = zz;
            y[di + 1][dj + 1] = yy;
        }
}

// the outlet page's water-surface difference (Bates_Inertial_8D.m:84-98,
// CA_Routing_8D.m:59-73)
__device__ __forceinline__ double outlet_difference(const D8Args &a, long long c, double depth) {
    const bool is_out = a.outlet && a.outlet[c];
    if (a.outlet_type == 1) return a.slope_outlet * a.R * (is_out ? 1.0 : 0.0);
    const double S0 = is_out ? (pow(depth, -1.0 / 6.0)) * sqrt(9.81) * a.rough[c] : 0.0;
    return S0 * a.R;
}

__global__ __launch_bounds__(kD8Block) void bates8d_kernel(D8Args a) {
    const long long c = (long long)xcd_logical_tile() * kD8Block + threadIdx.x;
    if (c >= a.n) return;
    const int ny = a.ny;
    const int i = (int)(c % ny), j = (int)(c / ny);
    double y[3][3], z[3][3];
    load_nbhd<true>(a, i, j, y, z);
    const double yc = y[1][1], zc = z[1][1];
    const double depth_raw = fmax(a.d_p[c] / 1000.0, 0.0);  // :40
    const bool dry = !(depth_raw > 0.0);                     // mask (:43)
    const double depth = dry ? 0.0 : depth_raw;              // :44
    const bool nan_cell = isnan(zc);
    const double R = a.R, R2 = R * R, Ld = sqrt(2.0) * R, g = 9.81;
    const double dt = a.time_step_min * 60.0;
    const double n2 = a.rough[c] * a.rough[c];
    const double kq = 1.0 / 1000.0 * 1.0 / 3600.0;
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int di = page_di(p), dj = page_dj(p);
        const bool rin = i + di >= 0 && i + di < ny, cin = j + dj >= 0 && j + dj < a.nx;
        double ms, hf;
        if (p == 4) {
            ms = outlet_difference(a, c, depth);
            hf = yc - zc;
        } else if (rin && cin) {
            const double yn = y[di + 1][dj + 1], zn = z[di + 1][dj + 1];
            ms = yc - yn;
            hf = fmax(yc, yn) - fmax(zc, zn);
        } else {
            ms = 0.0;
            hf = (di != 0 && dj != 0 && rin) ? y[di + 1][1] - z[di + 1][1] : yc - zc;
        }
        if (ms < 0.0 || isnan(ms)) ms = 0.0;  // :116 (h_min = 0)
        ms = ms / (p < 5 ? R : Ld);           // :119-121
        if (ms < 0.0) ms = 0.0;               // :124
        const double S = -ms;                 // :129
        const double qp = a.qprev[p * a.n + c] / 1000.0 / 3600.0 * R2 / R;  // :131
        double q = (qp - g * hf * dt * S) / (1.0 + g * hf * dt * n2 * qp / pow(hf, 10.0 / 3.0));
        if (a.flag_critical == 1) q = fmin(q, hf * sqrt(g * hf));  // :209-212
        q = R * q / R2 * 1000.0 * 3600.0;                          // :215
        if (q < 0.0 || isnan(q) || isinf(q)) q = 0.0;               // :216-217
        if (dry) q = 0.0;                                           // :218
        if (nan_cell) q = NAN;                                      // :220
        a.q[p * a.n + c] = q;
        if (a.Hf) a.Hf[p * a.n + c] = hf;
        s += kq * q;
    }
    const double I = dt * s * R2;  // :223
    a.Itot[c] = I;
    a.d_t[c] = a.d_tot[c] - I / a.A * 1000.0;  // :269
}

__global__ __launch_bounds__(kD8Block) void ca8d_kernel(D8Args a) {
    const long long c = (long long)xcd_logical_tile() * kD8Block + threadIdx.x;
    if (c >= a.n) return;
    const int ny = a.ny;
    const int i = (int)(c % ny), j = (int)(c / ny);
    double y[3][3], z[3][3];
    load_nbhd<false>(a, i, j, y, z);
    const double yc = y[1][1], zc = z[1][1];
    const double depth = fmax(a.d_tot[c] / 1000.0, 0.0);  // :40
    const double R = a.R, Ld = sqrt(2.0) * R, A = a.A;
    const double d_t_min = a.d_tolerance / 1000.0, h_min = d_t_min;  // :29, :34
    double v[9];
    double best = 0.0;
    int idx = 0;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int di = page_di(p), dj = page_dj(p);
        const bool on = i + di >= 0 && i + di < ny && j + dj >= 0 && j + dj < a.nx;
        double ms = p == 4 ? outlet_difference(a, c, depth) : (on ? yc - y[di + 1][dj + 1] : 0.0);
        if (ms < h_min || isnan(ms)) ms = 0.0;  // :85
        const double L = p < 5 ? R : Ld;
        const double sl = ms / L;               // :88-90
        if (p == 0 || sl > best) {              // :91 (first of equal maxima)
            best = sl;
            idx = p;
        }
        v[p] = A * (sl * L);                    // :98-103
    }
    const double Dist = idx <= 4 ? R : sqrt(2.0) * R;  // :93-96
    const double cbig = R * R * 1000.0;                 // :108
    double vmin = NAN, vt = 0.0;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        vmin = fmin(vmin, v[p] + (v[p] == 0.0 ? 1.0 : 0.0) * cbig);  // :109
        vt += v[p];                                                 // :112
    }
    double wmax = NAN;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        double w = v[p] / (vt + vmin);  // :115
        if (isnan(w)) w = 0.0;          // :116
        v[p] = w;
        wmax = fmax(wmax, w);           // :117
    }
    const double h0 = a.h0 ? a.h0[c] : 0.0;
    const double man = 1.0 / a.rough[c] * pow(fmax(depth - h0 / 1000.0, 0.0), 2.0 / 3.0) * sqrt(best);
    const double vm = a.flag_critical == 1 ? fmin(sqrt(9.81 * depth), man) : man;  // :121-125
    const double Ip = a.reset_prev ? 0.9 * A * depth : a.Iprev[c];                // :132-134
    double I = fmin(depth * A, vm * depth / wmax * (Dist * a.time_step_min * 60.0));  // :136
    I = fmin(I, Ip + vmin);                                                           // :137
    if (a.d_tot[c] < d_t_min * 1000.0 || isnan(zc)) I = 0.0;                          // :138
    const double dts = a.time_step_min * 60.0;
    double tot = 0.0;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const double vol = v[p] * I;  // :141
        tot += vol;                   // :142
        a.q[p * a.n + c] = vol / dts / A * 1000.0 * 3600.0;  // :143
    }
    a.Itot[c] = tot;
    a.d_t[c] = a.d_tot[c] - tot / A * 1000.0;  // :185
}

// qin_t from the outflow pages in the order of HydroPol2D_Main_While.m:634-645;
// d_t += qin_t * time_step / 60 (:657-658) when d_t is given
__global__ __launch_bounds__(kD8Block) void gather_kernel(const double *q, int ny, int nx,
                                                          long long n, int d8, double ts_min,
                                                          double *qin, double *d_t) {
    const long long c = (long long)xcd_logical_tile() * kD8Block + threadIdx.x;
    if (c >= n) return;
    const int i = (int)(c % ny), j = (int)(c / ny);
    double s = 0.0;
    if (j > 0) s += q[1 * n + c - ny];           // right outflow of the left cell
    if (j < nx - 1) s += q[0 * n + c + ny];      // left outflow of the right cell
    if (i > 0) s += q[3 * n + c - 1];            // down outflow of the upper cell
    if (i < ny - 1) s += q[2 * n + c + 1];       // up outflow of the lower cell
    if (d8) {
        if (i > 0 && j < nx - 1) s += q[7 * n + c - 1 + ny];       // SW of the NE cell
        if (i < ny - 1 && j < nx - 1) s += q[8 * n + c + 1 + ny];  // NW of the SE cell
        if (i < ny - 1 && j > 0) s += q[5 * n + c + 1 - ny];       // NE of the SW cell
        if (i > 0 && j > 0) s += q[6 * n + c - 1 - ny];            // SE of the NW cell
    }
    if (qin) qin[c] = s;
    if (d_t) d_t[c] = d_t[c] + s * ts_min / 60.0;
}

namespace {

int d8_blocks(long long n) { return (int)((n + kD8Block - 1) / kD8Block); }

// the outlet mask from 1-based rows/cols (HP2D_E_ARG when one is off the grid)
bool outlet_mask(int ny, int nx, const int32_t *rows, const int32_t *cols, int32_t n_outlet,
                 std::vector<uint8_t> &mask) {
    mask.assign((size_t)ny * nx, 0);
    for (int k = 0; k < n_outlet; ++k) {
        if (rows[k] < 1 || rows[k] > ny || cols[k] < 1 || cols[k] > nx) return false;
        mask[(size_t)(rows[k] - 1) + (size_t)(cols[k] - 1) * ny] = 1;
    }
    return true;
}

struct Stager {
    hipStream_t s;
    hipError_t e = hipSuccess;
    void up(void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && src) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    }
    void down(void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && dst) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    }
};

}  // namespace

hipError_t launch_bates8d(const D8Args &a, hipStream_t s) {
    hipLaunchKernelGGL(bates8d_kernel, dim3(d8_blocks(a.n)), dim3(kD8Block), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ca8d(const D8Args &a, hipStream_t s) {
    hipLaunchKernelGGL(ca8d_kernel, dim3(d8_blocks(a.n)), dim3(kD8Block), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_gather(const double *q, int ny, int nx, int n_pages, double ts_min, double *qin,
                         double *d_t, hipStream_t s) {
    const long long n = (long long)ny * nx;
    hipLaunchKernelGGL(gather_kernel, dim3(d8_blocks(n)), dim3(kD8Block), 0, s, q, ny, nx, n,
                       n_pages == 9 ? 1 : 0, ts_min, qin, d_t);
    return hipGetLastError();
}

}  // namespace hp2d

extern "C" hp2d_status hp2d_bates_inertial_8d(
    hp2d_handle *h, const double *z, const double *d_tot, const double *d_p,
    const double *roughness_cell, double cell_area, double time_step_min, double resolution,
    int32_t outlet_type, double slope_outlet, const int32_t *row_outlet, const int32_t *col_outlet,
    int32_t n_outlet, const double *outflow_prev, int32_t flag_critical, double *outflow,
    double *d_t, double *I_tot_end_cell, double *Hf) {
    using namespace hp2d;
    if (!h || !z || !d_tot || !d_p || !roughness_cell || !outflow_prev || n_outlet < 0 ||
        (n_outlet > 0 && (!row_outlet || !col_outlet)) || (outlet_type != 1 && outlet_type != 2))
        return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const size_t n = (size_t)ny * nx, nb = n * sizeof(double);
    std::vector<uint8_t> mask;
    if (!outlet_mask(ny, nx, row_outlet, col_outlet, n_outlet, mask))
        return hp2d_handle_fail(h, "hp2d_bates_inertial_8d: outlet row/col off the grid");
    // workspace: z d_tot d_p rough | qprev[9] | q[9] | Hf[9] | d_t Itot | mask
    char *ws = (char *)hp2d_handle_workspace(h, 33 * nb + n + 256);
    if (!ws) return HP2D_E_NOMEM;
    double *w = (double *)ws;
    D8Args a{};
    a.ny = ny;
    a.nx = nx;
    a.n = (long long)n;
    a.R = resolution;
    a.A = cell_area;
    a.time_step_min = time_step_min;
    a.outlet_type = outlet_type;
    a.flag_critical = flag_critical;
    a.slope_outlet = slope_outlet;
    double *dz = w, *dd = w + n, *dp = w + 2 * n, *dr = w + 3 * n, *dq0 = w + 4 * n;
    double *dq = w + 13 * n, *dH = w + 22 * n, *ddt = w + 31 * n, *dI = w + 32 * n;
    uint8_t *dm = (uint8_t *)(w + 33 * n);
    a.z = dz;
    a.d_tot = dd;
    a.d_p = dp;
    a.rough = dr;
    a.qprev = dq0;
    a.outlet = dm;
    a.q = dq;
    a.Hf = dH;
    a.d_t = ddt;
    a.Itot = dI;
    Stager io{hp2d_handle_stream(h)};
    io.up(dz, z, nb);
    io.up(dd, d_tot, nb);
    io.up(dp, d_p, nb);
    io.up(dr, roughness_cell, nb);
    io.up(dq0, outflow_prev, 9 * nb);
    io.up(dm, mask.data(), n);
    if (io.e == hipSuccess) io.e = hp2d_handle_mark(h, 0);
    if (io.e == hipSuccess) io.e = launch_bates8d(a, io.s);
    if (io.e == hipSuccess) io.e = hp2d_handle_mark(h, 1);
    io.down(outflow, dq, 9 * nb);
    io.down(Hf, dH, 9 * nb);
    io.down(d_t, ddt, nb);
    io.down(I_tot_end_cell, dI, nb);
    if (io.e == hipSuccess) io.e = hipStreamSynchronize(io.s);
    if (io.e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_bates_inertial_8d: ") + hipGetErrorString(io.e));
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_ca_routing_8d(
    hp2d_handle *h, const double *z, const double *d_tot, const double *roughness,
    double cell_area, double time_step_min, const double *h_0, double resolution,
    const double *I_tot_end_cell_prev, int32_t outlet_type, double slope_outlet,
    const int32_t *row_outlet, const int32_t *col_outlet, int32_t n_outlet, int32_t flag_critical,
    double d_tolerance, double *outflow, double *d_t, double *I_tot_end_cell) {
    using namespace hp2d;
    if (!h || !z || !d_tot || !roughness || !I_tot_end_cell_prev || n_outlet < 0 ||
        (n_outlet > 0 && (!row_outlet || !col_outlet)) || (outlet_type != 1 && outlet_type != 2))
        return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const size_t n = (size_t)ny * nx, nb = n * sizeof(double);
    std::vector<uint8_t> mask;
    if (!outlet_mask(ny, nx, row_outlet, col_outlet, n_outlet, mask))
        return hp2d_handle_fail(h, "hp2d_ca_routing_8d: outlet row/col off the grid");
    // :132 global condition on the host copy, in sum(sum()) order
    double sprev = 0.0;
    for (int j = 0; j < nx; ++j) {
        double col = 0.0;
        for (int i = 0; i < ny; ++i) col += I_tot_end_cell_prev[(size_t)i + (size_t)j * ny];
        sprev += col;
    }
    // workspace: z d_tot rough h0 Iprev | q[9] | d_t Itot | mask
    char *ws = (char *)hp2d_handle_workspace(h, 16 * nb + n + 256);
    if (!ws) return HP2D_E_NOMEM;
    double *w = (double *)ws;
    D8Args a{};
    a.ny = ny;
    a.nx = nx;
    a.n = (long long)n;
    a.R = resolution;
    a.A = cell_area;
    a.time_step_min = time_step_min;
    a.outlet_type = outlet_type;
    a.flag_critical = flag_critical;
    a.reset_prev = sprev == 0.0 ? 1 : 0;
    a.slope_outlet = slope_outlet;
    a.d_tolerance = d_tolerance;
    double *dz = w, *dd = w + n, *dr = w + 2 * n, *dh0 = w + 3 * n, *dIp = w + 4 * n;
    double *dq = w + 5 * n, *ddt = w + 14 * n, *dI = w + 15 * n;
    uint8_t *dm = (uint8_t *)(w + 16 * n);
    a.z = dz;
    a.d_tot = dd;
    a.rough = dr;
    a.h0 = h_0 ? dh0 : nullptr;
    a.Iprev = dIp;
    a.outlet = dm;
    a.q = dq;
    a.d_t = ddt;
    a.Itot = dI;
    Stager io{hp2d_handle_stream(h)};
    io.up(dz, z, nb);
    io.up(dd, d_tot, nb);
    io.up(dr, roughness, nb);
    io.up(dh0, h_0, nb);
    io.up(dIp, I_tot_end_cell_prev, nb);
    io.up(dm, mask.data(), n);
    if (io.e == hipSuccess) io.e = hp2d_handle_mark(h, 0);
    if (io.e == hipSuccess) io.e = launch_ca8d(a, io.s);
    if (io.e == hipSuccess) io.e = hp2d_handle_mark(h, 1);
    io.down(outflow, dq, 9 * nb);
    io.down(d_t, ddt, nb);
    io.down(I_tot_end_cell, dI, nb);
    if (io.e == hipSuccess) io.e = hipStreamSynchronize(io.s);
    if (io.e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_ca_routing_8d: ") + hipGetErrorString(io.e));
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_inflow_gather(hp2d_handle *h, const double *outflow, int32_t n_pages,
                                          double time_step_min, double *qin_t, double *d_t) {
    using namespace hp2d;
    if (!h || !outflow || (n_pages != 5 && n_pages != 9)) return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const size_t n = (size_t)ny * nx, nb = n * sizeof(double);
    char *ws = (char *)hp2d_handle_workspace(h, (n_pages + 2) * nb + 256);
    if (!ws) return HP2D_E_NOMEM;
    double *dq = (double *)ws, *dqin = dq + n_pages * n, *ddt = dqin + n;
    Stager io{hp2d_handle_stream(h)};
    io.up(dq, outflow, n_pages * nb);
    io.up(ddt, d_t, nb);
    if (io.e == hipSuccess) io.e = hp2d_handle_mark(h, 0);
    if (io.e == hipSuccess)
        io.e = launch_gather(dq, ny, nx, n_pages, time_step_min, dqin, d_t ? ddt : nullptr, io.s);
    if (io.e == hipSuccess) io.e = hp2d_handle_mark(h, 1);
    io.down(qin_t, dqin, nb);
    io.down(d_t, ddt, nb);
    if (io.e == hipSuccess) io.e = hipStreamSynchronize(io.s);
    if (io.e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_inflow_gather: ") + hipGetErrorString(io.e));
    return HP2D_OK;
}
