// hp2d_simca.hip - the device loop's step with cellular-automata D4 routing
// (flags.flag_CA = 1) for gfx950.
//
// Per step, three passes over the grid, each one lane per cell in a
// grid-stride loop whose chunks are dealt to the XCDs in contiguous runs (the
// neighbours one column away, ny cells off, then mostly sit in the same L2):
//   1. hydrology: rainfall, infiltration (Green-Ampt or van Genuchten), inflow
//      hydrographs and the depth clean-up of Hydrological_Model.m:15-55 -> d_tot;
//   2. CA_Routing.m:19-171 on d_tot: the five outflow rates, I_tot_end_cell
//      and the depth after the outflows;
//   3. the inflow gather and mass balance of HydroPol2D_Main_While.m:631-660,
//      then the per-cell terms of refreshing_timestep.m:31-53, 93-108, 146-150,
//      180-242 (the CA branch: face velocities from the outflows over the cell
//      depth), the depth maximum and the outlet records.
// Each pass writes its fields of the step's per-workgroup partials, which the
// finalize kernel reduces as it does the local-inertial step's.
#include "hp2d_forcing.hpp"
#include "hp2d_hydro.hpp"
#include "hp2d_internal.hpp"
#include "hp2d_simca.hpp"

namespace hp2d {

namespace {

constexpr int kCaThreads = 256;
constexpr int kCaWaves = kCaThreads / kWave;

struct StepView {
    long long k;
    double t, dt_min;
    int src, dst, fail_step;
    bool may_report;
};

template <typename T> __device__ __forceinline__ bool step_view(const LiArgs<T> &a, StepView &v) {
    const SimState *st = a.st;
    if (st->done || st->paused || st->hold) return false;
    v.k = st->k;
    v.t = st->t;
    v.dt_min = st->time_step_min;
    v.src = (int)((v.k + 1) & 1);
    v.dst = (int)(v.k & 1);
    v.fail_step = (st->inject_k == v.k);
    // a mass-balance report can only follow this step if t + max_dt reaches the
    // next threshold (HydroPol2D_Main_While.m:787-803): only then are storages summed
    const double prog_max = ((st->t + st->max_dt / 60.0) / a.routing_time) * 100.0;
    v.may_report = (prog_max >= st->next_report) || (prog_max >= 100.0);
    return true;
}

// Fields of the partials, reduced over the block in a fixed order: f_sum
// summed, the others by max (or min where is_min).
struct BlockRed {
    double v[kPartCount][kCaWaves];
};

__device__ __forceinline__ void block_write(BlockRed &red, const int *fields, const bool *is_sum,
                                            const bool *is_min, int n_fields, double *partials,
                                            long long cap, int slot) {
    const int t = threadIdx.x;
    if (t < n_fields) {
        const int f = fields[t];
        double acc = red.v[f][0];
        for (int w = 1; w < kCaWaves; ++w) {
            const double x = red.v[f][w];
            acc = is_sum[t] ? acc + x : (is_min[t] ? fmin(acc, x) : fmax(acc, x));
        }
        partials[(long long)f * cap + slot] = acc;
    }
}

// rate x ^ y for x >= 0 (FP64: pow; FP32: exp2 / log2)
__device__ __forceinline__ double ca_pow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float ca_pow(float x, float y) { return vg_pow(x, y); }

}  // namespace

// ---------------------------------------------------------------------------
// Pass 1: the hydrological prologue of the step (Hydrological_Model.m:15-55)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kCaThreads) void ca_hydro_kernel(CaSimArgs<T> A) {
    const LiArgs<T> &a = A.li;
    StepView sv;
    if (!step_view(a, sv)) return;
    constexpr bool f64 = sizeof(T) == 8;
    const SimState *st = a.st;
    const T dt_h = T(sv.dt_min / 60.0);
    const T inv_dt_h = frcp(dmax(dt_h, T(1e-12)));
    const T rain = T(st->delta_p_agg);
    const T *d_src = a.d_buf[sv.src];
    const T *I_src = a.I_buf[sv.src], *Tc_src = a.T_buf[sv.src];
    T *I_dst = a.I_buf[sv.dst], *Tc_dst = a.T_buf[sv.dst];
    const long long n = (long long)a.ny * a.nx;
    // rainfall maps of the step (update_spatial_BC.m:1361-1408)
    int mn = 0;
    long long moff[kMaxMapTerms];
    T mw[kMaxMapTerms], mdtot = T(0), mk = T(0);
    if (a.rain_maps) {
        mn = st->map_n;
        for (int j = 0; j < kMaxMapTerms; ++j) {
            moff[j] = (long long)st->map_idx[j] * n;
            mw[j] = T(st->map_w[j]);
        }
        mdtot = T(st->map_dtot);
        mk = T(st->map_k);
    }
    // FP32 class tables in LDS
    __shared__ float4 tab[2 * kMaxSoilClasses];
    const bool inf = a.ga != nullptr || a.vg != nullptr;
    if (!f64 && inf) {
        const float4 *src = reinterpret_cast<const float4 *>(a.ga ? a.ga_f32 : a.vg_f32);
        for (int i = threadIdx.x; i < 2 * a.n_soil_classes; i += kCaThreads) tab[i] = src[i];
        __syncthreads();
    }
    double s_uz = 0.0, s_inf = 0.0;
    T dtmax = T(NAN), dpmax = T(NAN);
    int neg = 0;
    const int chunk = xcd_logical_tile();
    for (long long c = (long long)chunk * kCaThreads + threadIdx.x; c < n;
         c += (long long)gridDim.x * kCaThreads) {
        const T d0 = d_src[c];
        const bool nan_cell = f64 ? isnan(a.z[c]) : isnan(a.n2s[c]);
        T eff;
        if (a.rain_maps) {
            T acc = T(0);
            for (int j = 0; j < mn; ++j) acc = acc + mw[j] * a.rain_maps[moff[j] + c];
            eff = mn > 0 ? acc / mdtot * mk : T(0);
        } else {
            eff = a.rain_pattern ? a.rain_pattern[c] * rain : rain;
        }
        if (nan_cell) eff = T(NAN);
        T dd = d0 + eff;
        if (inf) {
            const int cls = a.soil_class ? a.soil_class[c] : 0;
            T In, f;
            if (a.ga) {
                T ksat, psi, dth, Lu_mm, k_out, Tr;
                if (f64) {
                    const GaClass g = a.ga[cls];
                    ksat = T(g.ksat); psi = T(g.psi); dth = T(g.dtheta);
                    Lu_mm = T(g.Lu_mm); k_out = T(g.k_out); Tr = T(g.Tr);
                } else {
                    const float4 g0 = tab[2 * cls], g1 = tab[2 * cls + 1];
                    ksat = g0.x; psi = g0.y; dth = g0.z; Lu_mm = g0.w; k_out = g1.x; Tr = g1.y;
                }
                T Tn = Tc_src[c];
                f = ga_infiltrate<T>(ksat, psi, dth, Lu_mm, k_out, Tr, a.I_min, sv.k == 1, dt_h,
                                     inv_dt_h, dd, I_src[c], Tn, In);
                if (nan_cell) {
                    In = T(NAN);
                    f = T(NAN);
                }
                dd = f64 ? dd - f * dt_h : fmadd(-f, dt_h, dd);
                Tc_dst[c] = Tn;
            } else {
                In = I_src[c];
                if (!nan_cell) {
                    f = vg_infiltrate<T>(a, tab, cls, a.vg_zwt[c], dt_h, inv_dt_h, dd, In);
                } else {
                    In = T(NAN);
                    f = T(NAN);
                }
            }
            I_dst[c] = In;
            if (!sv.fail_step) {
                if (!isnan(f)) s_inf += (double)(f * dt_h);
                if (sv.may_report && !isnan(In)) s_uz += (double)In;
            }
        }
        // Hydrological_Model.m:30: the inflow hydrographs on their gauge cells
        if (a.inflow_gauge) {
            const int g = a.inflow_gauge[c];
            if (g > 0) dd = dd + T(st->inflow_agg[g - 1]);
        }
        if (dd < T(-1e-8)) neg = 1;
        if (dd < T(1e-6)) dd = T(0);  // :49
        if (dd < T(0)) dd = T(0);     // :52
        A.dtot[c] = dd;
        if (sv.fail_step || a.track_dmax) {  // rollback celerity inputs (Main_While.m:932)
            dtmax = dmax(dtmax, dd);
            dpmax = dmax(dpmax, d0);
        }
    }
    __shared__ BlockRed red;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const double uz = wave_sum(s_uz), in = wave_sum(s_inf);
    const T wdt = wave_fmax(dtmax), wdp = wave_fmax(dpmax);
    const int any_neg = __any(neg);
    if (lane == 0) {
        red.v[kPartUz][w] = uz;
        red.v[kPartInfil][w] = in;
        red.v[kPartDtotMax][w] = (double)wdt;
        red.v[kPartDprevMax][w] = (double)wdp;
        red.v[kPartNeg][w] = any_neg ? 1.0 : 0.0;
    }
    __syncthreads();
    const int fields[5] = {kPartUz, kPartInfil, kPartDtotMax, kPartDprevMax, kPartNeg};
    const bool is_sum[5] = {true, true, false, false, false};
    const bool is_min[5] = {false, false, false, false, false};
    block_write(red, fields, is_sum, is_min, 5, a.partials, a.part_cap, chunk);
}

// ---------------------------------------------------------------------------
// Pass 2: CA_Routing.m:19-171 on the depth after the hydrology
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kCaThreads) void ca_route_kernel(CaSimArgs<T> A) {
    const LiArgs<T> &a = A.li;
    StepView sv;
    if (!step_view(a, sv)) return;
    constexpr bool f64 = sizeof(T) == 8;
    const int ny = a.ny, nx = a.nx;
    const long long n = (long long)ny * nx;
    T *d_dst = a.d_buf[sv.dst], *qr = a.qx_buf[sv.dst], *qu = a.qy_buf[sv.dst];
    const T *dtot = A.dtot;
    const T R = a.R, Ac = a.cell_area, h_min = T(1e-6);
    const T tsm = T(sv.dt_min);
    // :121: the limiter restarts from 0.9 A d when the last I_tot_end_cell sum to 0
    const bool reset = !(a.st->ca_itot_any > 0.0);
    int any = 0;
    const int chunk = xcd_logical_tile();
    for (long long c = (long long)chunk * kCaThreads + threadIdx.x; c < n;
         c += (long long)gridDim.x * kCaThreads) {
        const int i = (int)(c % ny), j = (int)(c / ny);
        const T dc = dtot[c];
        const T depth = f64 ? dc / T(1000) : dc * T(1e-3);
        const T n2 = a.n2s[c];
        const bool nan_cell = f64 ? isnan(a.z[c]) : isnan(n2);
        T dh[5];
        // :51-54, with the faces on the grid boundary closed (:80-87)
        if (f64) {
            const T wse = a.z[c] + depth;
            dh[0] = (j > 0) ? wse - a.z[c - ny] - dtot[c - ny] / T(1000) : T(0);
            dh[1] = (j < nx - 1) ? wse - a.z[c + ny] - dtot[c + ny] / T(1000) : T(0);
            dh[2] = (i > 0) ? wse - a.z[c - 1] - dtot[c - 1] / T(1000) : T(0);
            dh[3] = (i < ny - 1) ? wse - a.z[c + 1] - dtot[c + 1] / T(1000) : T(0);
        } else {
            // dzx(i,j) = z(i,j+1) - z(i,j), dzy(i,j) = z(i-1,j) - z(i,j)
            dh[0] = (j > 0) ? (a.dzx[c - ny] + depth) - dtot[c - ny] * T(1e-3) : T(0);
            dh[1] = (j < nx - 1) ? (depth - a.dzx[c]) - dtot[c + ny] * T(1e-3) : T(0);
            dh[2] = (i > 0) ? (depth - a.dzy[c]) - dtot[c - 1] * T(1e-3) : T(0);
            dh[3] = (i < ny - 1) ? (a.dzy[c + 1] + depth) - dtot[c + 1] * T(1e-3) : T(0);
        }
        // :58-72 the outlet pseudo-neighbour (outlet cells carry the sign bit of n^2)
        const bool is_out = signbit(n2);
        const T rough = sqrt(fabs(n2));
        if (a.outlet_type == 1) {
            dh[4] = A.slope_outlet * R * (is_out ? T(1) : T(0));
        } else {
            const T S0 = is_out ? ca_pow(depth, T(-1.0 / 6.0)) * sqrt(T(9.81)) * rough : T(0);
            dh[4] = S0 * R;
        }
#pragma unroll
        for (int p = 0; p < 5; ++p)
            if (dh[p] < h_min || isnan(dh[p])) dh[p] = T(0);
        // :90-106 weights
        const T cbig = R * R * T(1000);
        T vt = T(0), vmin = T(NAN), maxh = T(NAN);
#pragma unroll
        for (int p = 0; p < 5; ++p) {
            const T v = Ac * dh[p];
            vt += v;
            vmin = fmin(vmin, (dh[p] < h_min || isnan(dh[p])) ? cbig : v);
            maxh = fmax(maxh, dh[p]);
        }
        T w[5], wmax = T(NAN);
#pragma unroll
        for (int p = 0; p < 5; ++p) {
            T ww = (Ac * dh[p]) / (vt + vmin);
            if (isnan(ww)) ww = T(0);
            w[p] = ww;
            wmax = fmax(wmax, ww);
        }
        // :110-129 velocity and the cell's outflow volume
        const T h0 = A.h0 ? A.h0[c] : T(0);
        const T man = T(1) / rough * ca_pow(fmax(depth - h0 / T(1000), T(0)), T(2.0 / 3.0)) *
                      sqrt(maxh / R);
        const T vm = a.flag_critical == 1 ? fmin(sqrt(T(9.81) * depth), man) : man;
        const T Iprev = reset ? T(0.9) * Ac * depth : A.itot[c];
        T I = fmin(depth * Ac, vm * depth / wmax * (R * tsm * T(60)));
        I = fmin(I, Iprev + vmin);
        if (dc < T(1e-6) * T(1000) || nan_cell) I = T(0);
        // :132-170 outflow rates [mm/h]
        const T dts = tsm * T(60);
        T tot = T(0), q[5];
#pragma unroll
        for (int p = 0; p < 5; ++p) {
            const T Ip = w[p] * I;
            tot += Ip;
            q[p] = Ip / dts / Ac * T(1000) * T(3600);
        }
        A.ql[c] = q[0];
        qr[c] = q[1];
        qu[c] = q[2];
        A.qd[c] = q[3];
        A.qo[c] = q[4];
        A.itot[c] = tot;
        d_dst[c] = dc - tot / Ac * T(1000);
        if (!(tot == T(0))) any = 1;
    }
    __shared__ BlockRed red;
    const int any_w = __any(any);
    if ((threadIdx.x & (kWave - 1)) == 0) red.v[kPartCaItot][threadIdx.x / kWave] = any_w ? 1.0 : 0.0;
    __syncthreads();
    const int fields[1] = {kPartCaItot};
    const bool is_sum[1] = {false}, is_min[1] = {false};
    block_write(red, fields, is_sum, is_min, 1, a.partials, a.part_cap, chunk);
}

// ---------------------------------------------------------------------------
// Pass 3: the inflow gather and mass balance (Main_While.m:631-660), the
// per-cell terms of the time step, the depth maximum and the outlet records
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kCaThreads) void ca_gather_kernel(CaSimArgs<T> A) {
    const LiArgs<T> &a = A.li;
    StepView sv;
    if (!step_view(a, sv)) return;
    constexpr bool f64 = sizeof(T) == 8;
    const int ny = a.ny, nx = a.nx;
    const long long n = (long long)ny * nx;
    T *d = a.d_buf[sv.dst];
    const T *qr = a.qx_buf[sv.dst], *qu = a.qy_buf[sv.dst];
    const T tsm = T(sv.dt_min);
    const T R = a.R, tol = A.depth_tol_mm;
    const T t_end_min = T(sv.t + sv.dt_min);
    const bool track = a.dmax_map != nullptr && !sv.fail_step;
    const bool risk_on = a.max_risk != nullptr && !sv.fail_step;
    // a map record can only follow this step if t + max_dt passes the next record
    // time (the new dt never exceeds max_dt): only then are the velocities stored
    const bool may_snap = a.snap_vras != nullptr &&
                          (a.st->t + a.st->max_dt / 60.0 > (double)a.st->map_last * a.rec_maps);
    T vkey = T(NAN), wkey = T(NAN), rmax = T(NAN);
    double s_store = 0.0;
    const int chunk = xcd_logical_tile();
    for (long long c = (long long)chunk * kCaThreads + threadIdx.x; c < n;
         c += (long long)gridDim.x * kCaThreads) {
        const int i = (int)(c % ny), j = (int)(c / ny);
        // :634-637 in that order, then :657-658
        T qin = T(0);
        if (j > 0) qin = qin + qr[c - ny];
        if (j < nx - 1) qin = qin + A.ql[c + ny];
        if (i > 0) qin = qin + A.qd[c - 1];
        if (i < ny - 1) qin = qin + qu[c + 1];
        const T dn = d[c] + qin * tsm / T(60);
        d[c] = dn;
        // refreshing_timestep.m:31-53: face velocities over the cell depth
        const T safe = dn < tol ? T(1e12) : dn;
        T vl, vr, vu, vd;
        if (f64) {
            vl = (A.ql[c] / T(1000) / T(3600)) * R / (safe / T(1000));
            vr = (qr[c] / T(1000) / T(3600)) * R / (safe / T(1000));
            vu = (qu[c] / T(1000) / T(3600)) * R / (safe / T(1000));
            vd = (A.qd[c] / T(1000) / T(3600)) * R / (safe / T(1000));
        } else {
            const T k = a.k_mmh_q * frcp(safe * T(1e-3));  // R / 3.6e6 / (d / 1000)
            vl = A.ql[c] * k;
            vr = qr[c] * k;
            vu = qu[c] * k;
            vd = A.qd[c] * k;
        }
        // :93-108, 146-150, 180-190 (CA: the components are the plain differences)
        const T raster = fmax(fmax(fmax(fabs(vl), fabs(vr)), fabs(vu)), fabs(vd));
        const T rc = vr - vl, uc = vu - vd;
        T vtot;
        bool stagnant;
        if (f64) {
            vtot = hypot(rc, uc);
            stagnant = vtot < T(0.01);
            vkey = fmax(vkey, vtot);
        } else {  // the FP32 key is the squared speed
            vtot = fmadd(rc, rc, uc * uc);
            stagnant = vtot < T(1e-4);
            vkey = fmax(vkey, vtot);
        }
        if (!a.adaptive) rmax = fmax(rmax, raster);
        if (risk_on || may_snap) {
            const T vt = f64 ? vtot : sqrt(vtot);  // velocities.total_velocity
            if (risk_on) {
                // Human_Instability_Module.m:8-25 and its running maximum
                const T h = f64 ? dn / T(1000) : dn * T(1e-3);
                const T F_person = a.hi_weight * a.hi_g;
                const T F_buoy = (a.hi_w1 * a.hi_w2) * h * a.hi_ro * a.hi_g;
                const T avail = a.hi_mu * (fmax(F_person - F_buoy, T(0)));
                const T hydro = fmax(T(1) / T(2) * (a.hi_Cd * a.hi_ro * a.hi_w1 * h * (vt * vt)), T(0));
                T risk = T(0);
                if (avail > T(0)) risk = hydro / avail;
                else if (h > T(0)) risk = T(1);
                risk = fmin(fmax(risk, T(0)), T(1));
                if (isnan(dn)) risk = T(NAN);
                const T prev = (sv.k == 1) ? (isnan(dn) ? T(NAN) : T(0)) : a.max_risk[c];
                const T nm = fmax(risk, prev);
                if (sv.k == 1 || nm != prev) a.max_risk[c] = nm;
            }
            if (may_snap) {
                a.snap_vras[c] = raster;
                a.snap_vtot[c] = vt;
            }
        }
        // :227-238: the wave candidate over the non-stagnant cells (depth d / 1000;
        // a NaN depth counts as 0, as max(NaN, 0) does)
        const T dts = f64 ? dn / T(1000) : dn * T(1e-3);
        if (!stagnant) {
            if (a.adaptive) {
                wkey = fmax(wkey, fmax(dts, T(0)));
            } else {
                T cw = sqrt(T(9.81) * fmax(dts, T(0)));
                cw = fmax(cw, T(1e-10));
                if (!isinf(cw)) wkey = fmin(wkey, a.alfa_min * R / (raster + cw));
            }
        }
        if (sv.may_report && !isnan(dn)) s_store += (double)dn;
        // save_output_maps.m:588-600 (and the time of the peak)
        if (track) {
            const T prev = a.dmax_map[c];
            const T nm = (sv.k == 1) ? dn : fmax(dn, prev);
            if (sv.k == 1 || nm != prev) a.dmax_map[c] = nm;
            if (a.tpeak_map && (sv.k == 1 || dn > prev)) a.tpeak_map[c] = t_end_min;
        }
        // outlet records (outlet_flow and the depth at the outlets)
        if (a.n_outlet > 0 && signbit(a.n2s[c])) {
            const int slot = find_outlet(a.out_idx, a.n_outlet, c);
            if (slot >= 0) {
                a.out_q[slot] = A.qo[c];
                a.out_d[slot] = dn;
            }
        }
    }
    __shared__ BlockRed red;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const double sto = wave_sum(s_store);
    const T wv = wave_fmax(vkey), wr = wave_fmax(rmax);
    const T ww = a.adaptive ? wave_fmax(wkey) : wave_fmin(wkey);
    if (lane == 0) {
        red.v[kPartStorage][w] = sto;
        red.v[kPartVmax][w] = (double)wv;
        red.v[kPartDtw][w] = (double)ww;
        red.v[kPartRaster][w] = (double)wr;
        red.v[kPartDnewMax][w] = NAN;
    }
    __syncthreads();
    const int fields[5] = {kPartStorage, kPartVmax, kPartDtw, kPartRaster, kPartDnewMax};
    const bool is_sum[5] = {true, false, false, false, false};
    const bool is_min[5] = {false, false, !a.adaptive, false, false};
    block_write(red, fields, is_sum, is_min, 5, a.partials, a.part_cap, chunk);
}

template <typename T> hipError_t launch_ca_sim(const CaSimArgs<T> &a, hipStream_t s) {
    const dim3 grid((unsigned)a.blocks), block(kCaThreads);
    hipLaunchKernelGGL((ca_hydro_kernel<T>), grid, block, 0, s, a);
    hipLaunchKernelGGL((ca_route_kernel<T>), grid, block, 0, s, a);
    hipLaunchKernelGGL((ca_gather_kernel<T>), grid, block, 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_ca_sim<float>(const CaSimArgs<float> &, hipStream_t);
template hipError_t launch_ca_sim<double>(const CaSimArgs<double> &, hipStream_t);

}  // namespace hp2d
