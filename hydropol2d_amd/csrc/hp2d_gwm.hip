// hp2d_gwm.hip - the groundwater module step (Groundwater_Module.m:37-404, no
// layered soil) for gfx950.
//
// Per-cell work runs in elementwise kernels: the van Genuchten / Mualem
// recharge bucket (simulate_groundwater_recharge.m:55-215) with the pending
// exchange, the scheduler's per-cell terms (head change, diffusivity), the
// local head update, the return of exfiltrated and excess water to the
// surface.  Reductions are per-column FP64 sums / NaN-omitting maxima, folded
// on the host in column order, where the scheduler's scalar decision is taken.
// The Boussinesq update runs the device solver of hp2d_gw.hip in place.
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_colred.hpp"
#include "hp2d_gw.hpp"
#include "hp2d_handle.hpp"

namespace hp2d {

// the aquifer's specific yield and conductivity [mm/h] of a cell: the map,
// or the uniform value
__device__ __forceinline__ double gwm_Sy(const GwmArgs &a, long long c) {
    return a.Sy ? a.Sy[c] : a.Sy_u;
}
__device__ __forceinline__ double gwm_ksat_gw(const GwmArgs &a, long long c) {
    return a.ksat_gw ? a.ksat_gw[c] : a.ksat_gw_u;
}

// soil depth of a cell [m]: the map, or the uniform value
__device__ __forceinline__ double gwm_soil_depth(const GwmArgs &a, long long c) {
    return a.soil_depth ? a.soil_depth[c] : a.soil_depth_u;
}

// the soil's van Genuchten parameters and saturated conductivity [mm/h] of a
// cell: the per-cell maps, or its soil class's entries of the class tables
struct GwmSoil {
    double ts, tr, nv, alpha, ksat;
};
__device__ __forceinline__ GwmSoil gwm_soil(const GwmArgs &a, long long c) {
    GwmSoil s;
    if (a.soil_class) {
        const int k = a.soil_class[c];
        s.ts = a.cls_ts[k];
        s.tr = a.cls_tr[k];
        s.nv = a.cls_nv[k];
        s.alpha = a.cls_alpha[k];
        s.ksat = a.cls_ksat[k];
    } else {
        s.ts = a.ts[c];
        s.tr = a.tr[c];
        s.nv = a.nv[c];
        s.alpha = a.alpha[c];
        s.ksat = a.ksat[c];
    }
    return s;
}

// The surface depth d [mm] and the vadose storage I [mm] of a cell: the
// module's FP64 arrays, or an FP32 run's own buffers (read as FP64, rounded
// once per kernel on the way back)
__device__ __forceinline__ double gwm_I(const GwmArgs &a, long long c) {
    return a.I32 ? (double)a.I32[c] : a.I[c];
}
__device__ __forceinline__ void gwm_set_I(const GwmArgs &a, long long c, double v) {
    if (a.I32) a.I32[c] = (float)v;
    else a.I[c] = v;
}
__device__ __forceinline__ double gwm_d(const GwmArgs &a, long long c) {
    return a.d32 ? (double)a.d32[c] : a.d[c];
}
__device__ __forceinline__ void gwm_set_d(const GwmArgs &a, long long c, double v) {
    if (a.d32) a.d32[c] = (float)v;
    else a.d[c] = v;
}

// recharge bucket, pending exchange and the scheduler's per-cell terms of
// cell c; v receives [recharge volume, pending head change, diffusivity]
template <bool F32>
__device__ __forceinline__ void gwm_recharge_cell(const GwmArgs &a, long long c, double dt_rfg,
                                                  double (&v)[3]) {
    const bool nan = isnan(a.elev[c]);
    const double zero = nan ? NAN : 0.0;
    const double sd = gwm_soil_depth(a, c);
    const double gw_depth = a.h[c] - (a.elev[c] - sd);
    double zwt = sd - gw_depth;
    zwt = fmax(zwt, 0.0);
    zwt = fmin(zwt, sd);
    const GwmSoil soil = gwm_soil(a, c);
    const double ts = soil.ts, tr = soil.tr;
    const double uzmax = zwt * (ts - tr);
    // simulate_groundwater_recharge with infiltration = zero_matrix, min = 0
    const double dt = fmax(a.dt_s, 1e-12);
    const double nvg = soil.nv, m = 1.0 - 1.0 / nvg, al = soil.alpha;
    const double Ks = soil.ksat / 1000.0 / 3600.0;
    const double inf_r = isfinite(zero) ? zero : 0.0;
    const double S0 = gwm_I(a, c) / 1000.0;
    const double s0 = isfinite(S0) ? S0 : 0.0;
    const double mx = isfinite(uzmax) ? uzmax : 0.0;
    const double min_eff = fmin(0.0, mx);
    double S = s0;
    S = fmax(S, min_eff);
    S = fmin(S, mx);
    const double por_eff = fmax(ts - tr, 1e-12);
    double zb = mx / por_eff;
    zb = fmax(zb, 0.0);
    const double Lgw = fmax(0.5 * zb, 1e-6);
    const double zb_safe = fmax(zb, 1e-6);
    double theta = tr + S / zb_safe;
    theta = fmax(theta, tr);
    theta = fmin(theta, ts);
    double Se = (theta - tr) / fmax(ts - tr, 1e-12);
    Se = fmin(fmax(Se, 1e-6), 1.0);
    double hs, Kr;
    if constexpr (F32) {
        // FP32 run: the van Genuchten head and the Mualem conductivity in FP32
        // with the hardware log2 / exp2 (their results feed a storage that is
        // itself FP32; the bucket balance below stays FP64)
        const float sf = (float)Se, mf = (float)m;
        const float l2s = __builtin_amdgcn_logf(sf);
        const float se_inv_m = __builtin_amdgcn_exp2f((-1.0f / mf) * l2s);
        const float hsf = -(1.0f / (float)al) *
                          __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(se_inv_m - 1.0f) / (float)nvg);
        const float term = 1.0f - 1.0f / se_inv_m;
        const float y = 1.0f - __builtin_amdgcn_exp2f(mf * __builtin_amdgcn_logf(term));
        const float se_l = a.l_vg == 0.5 ? sqrtf(sf) : __builtin_amdgcn_exp2f((float)a.l_vg * l2s);
        hs = (double)hsf;
        Kr = (double)(se_l * (y * y));
    } else {
        // Se^(-1/m) is taken once and Se^(1/m) is its reciprocal; the Mualem
        // factor Se^l is a square root for the usual l = 1/2 (each within an ulp
        // of pow: the recharge pass is bound by its FP64 powers)
        const double se_inv_m = pow(Se, -1.0 / m);
        hs = -(1.0 / al) * pow(se_inv_m - 1.0, 1.0 / nvg);
        const double term = 1.0 - 1.0 / se_inv_m;
        const double y = 1.0 - pow(term, m);
        const double se_l = a.l_vg == 0.5 ? sqrt(Se) : pow(Se, a.l_vg);
        Kr = se_l * (y * y);  // (.)^2: y*y is the correctly rounded square
    }
    if (Se >= 0.999999) hs = 0.0;
    Kr = fmin(fmax(Kr, 0.0), 1.0);
    double grad = 1.0 + hs / Lgw;
    grad = fmax(grad, -1.0);
    grad = fmin(grad, 2.0);
    double Rr = (Ks * Kr) * grad;
    const bool no_uz = mx <= 0.0;
    if (no_uz) Rr = 0.0;
    double S_new = S + dt * (inf_r - Rr);
    S_new = fmax(S_new, min_eff);
    S_new = fmin(S_new, mx);
    if (no_uz) S_new = 0.0;
    double R_mb = inf_r - (S_new - S) / dt;
    if (no_uz) R_mb = 0.0;
    gwm_set_I(a, c, S_new * 1000.0);
    if (a.cum) a.cum[c] = a.cum[c] + fmax(R_mb, 0.0) * dt * 1000.0;
    // exchange (the capillary rate is the zero matrix without layered soil)
    double rr = R_mb - zero;
    if (nan) rr = NAN;
    const double down = fmax(R_mb, 0.0) * a.dt_s;
    const double pn = a.pn[c] + rr * a.dt_s;
    a.pn[c] = pn;
    if (a.pr) a.pr[c] = a.pr[c] + down;
    if (a.pc) a.pc[c] = a.pc[c] + fmax(zero, 0.0) * a.dt_s;
    if (dt_rfg > 0.0) {  // the update is known to follow: recharge_for_gw and q_exf (:199-203)
        a.rfg[c] = nan ? NAN : pn / dt_rfg;
        a.q[c] = nan ? NAN : 0.0;
    }
    v[0] = down * (nan ? NAN : a.A);
    // scheduler terms: pending head change and diffusivity K H / Sy
    const double Sy_head = fmax(gwm_Sy(a, c), 1e-6);
    v[1] = fabs(pn) / fmax(Sy_head, 1e-9);
    const double zbed = a.elev[c] - sd;
    double K = gwm_ksat_gw(a, c) / 1000.0 / 3600.0;
    if (!isfinite(K)) K = NAN;
    double sat = fmax(a.h[c] - zbed, 1e-4);
    if (nan) sat = NAN;
    v[2] = K * sat / Sy_head;
    if (!a.statics_ready) {
        a.z_bed[c] = zbed;
        if (a.Kgw) a.Kgw[c] = gwm_ksat_gw(a, c) / 1000.0 / 3600.0;
    }
}

// One block per column: the recharge of its cells and the column's
// [0] nansum of the recharge volume, [1] NaN-omitting max of the pending head
// change, [2] that of the diffusivity (the scheduler's inputs).  dt_rfg > 0:
// the update's interval when the update is known to follow (synchronous
// module), whose recharge rate and zeroed exfiltration are then written here.
// F32: an FP32 run (I32 set), whose variant holds no FP64 powers and so
// needs fewer registers.
template <bool F32>
__global__ __launch_bounds__(kColThreads) void gwm_recharge_kernel(GwmArgs a, double dt_rfg,
                                                                  double *col) {
    const bool is_max[3] = {false, true, true};
    column_reduce<3>(a.ny, blockIdx.x, is_max,
                     [&](long long c, double (&v)[3]) { gwm_recharge_cell<F32>(a, c, dt_rfg, v); }, col);
}

// recharge_for_gw (:199-203) and q_exf = zero_matrix
__global__ __launch_bounds__(256) void gwm_rfg_kernel(GwmArgs a, double dt_gw) {
    const long long c = blockIdx.x * 256LL + threadIdx.x;
    if (c >= a.n) return;
    const bool nan = isnan(a.elev[c]);
    a.rfg[c] = nan ? NAN : a.pn[c] / dt_gw;
    a.q[c] = nan ? NAN : 0.0;
}

// local water-table update without baseflow (:252-263)
__global__ __launch_bounds__(256) void gwm_local_kernel(GwmArgs a, double dt_gw) {
    const long long c = blockIdx.x * 256LL + threadIdx.x;
    if (c >= a.n) return;
    const double Sy_l = fmax(gwm_Sy(a, c), 1e-6);
    const double dh = a.rfg[c] * dt_gw / Sy_l;
    double hc = a.h[c] + dh;
    hc = fmax(hc, a.z_bed[c]);
    const double ex_water = fmax(hc - a.elev[c], 0.0) * Sy_l;
    a.h[c] = fmin(hc, a.elev[c]);
    double dn = gwm_d(a, c) + ex_water * 1000.0;
    if (dn < 0.0) dn = 0.0;
    gwm_set_d(a, c, dn);
}

// after an update: exfiltrated volume, pending reset, unsaturated-zone excess
// (:266-395), one block per column with the column's [0] nansum of the
// exfiltrated volume, [1] NaN-omitting max of the excess and [2] nansum of the
// aquifer storage A*Sy*(h - z_bed) (the head is final here; the pending term
// of the storage is zero after the reset)
// With h_new / q_new (after a Boussinesq update) it first takes over the new
// head and q_exf and moves the exfiltration onto the surface (:248-250), so
// the head and q_exf are not read back.
__global__ __launch_bounds__(kColThreads) void gwm_post_kernel(GwmArgs a, double dt_gw, double *col,
                                                              const double *h_new,
                                                              const double *q_new) {
    const bool is_max[3] = {false, true, false};
    column_reduce<3>(a.ny, blockIdx.x, is_max, [&](long long c, double (&v)[3]) {
        const bool nan = isnan(a.elev[c]);
        double q, h;
        if (h_new) {
            h = h_new[c];
            q = q_new[c];
            a.h[c] = h;
            a.q[c] = q;
            double dn = gwm_d(a, c) + dt_gw * (q * 1000.0);
            if (dn < 0.0) dn = 0.0;
            gwm_set_d(a, c, dn);
        } else {
            h = a.h[c];
            q = a.q[c];
        }
        v[0] = fmax(q, 0.0) * dt_gw * (nan ? NAN : a.A);
        a.pn[c] = 0.0;
        if (a.pr) a.pr[c] = 0.0;
        if (a.pc) a.pc[c] = 0.0;
        double zwt = a.elev[c] - h;
        zwt = fmax(zwt, 0.0);
        zwt = fmin(zwt, gwm_soil_depth(a, c));
        const GwmSoil soil = gwm_soil(a, c);
        const double uz_new = zwt * (soil.ts - soil.tr);
        const double ex = fmax(gwm_I(a, c) / 1000.0 - uz_new, 0.0);  // excess [m]
        a.t2[c] = ex;
        v[1] = ex;
        const double area = nan ? NAN : a.A;
        v[2] = area * gwm_Sy(a, c) * (h - (a.elev[c] - gwm_soil_depth(a, c)));
    }, col);
}

__global__ __launch_bounds__(256) void gwm_excess_kernel(GwmArgs a) {
    const long long c = blockIdx.x * 256LL + threadIdx.x;
    if (c >= a.n) return;
    const double Sm = gwm_I(a, c) / 1000.0, ex = a.t2[c];
    gwm_set_I(a, c, fmax((Sm - ex) * 1000.0, 0.0));
    double dn = gwm_d(a, c) + ex * 1000.0;
    if (dn < 0.0) dn = 0.0;
    gwm_set_d(a, c, dn);
}

__global__ __launch_bounds__(256) void gwm_maxdepth_kernel(GwmArgs a) {
    const long long c = blockIdx.x * 256LL + threadIdx.x;
    if (c >= a.n || !a.maxd) return;
    a.maxd[c] = fmax(a.maxd[c], a.h[c] - (a.elev[c] - gwm_soil_depth(a, c)));
}

// per column: sum of A*Sy*(h - z_bed) and of A*pending_net (mass_balance_check.m:66-70)
__global__ __launch_bounds__(kColThreads) void gwm_storage_kernel(GwmArgs a, double *out) {
    const bool is_max[2] = {false, false};
    column_reduce<2>(a.ny, blockIdx.x, is_max, [&](long long c, double (&v)[2]) {
        const double area = isnan(a.elev[c]) ? NAN : a.A;
        v[0] = area * gwm_Sy(a, c) * (a.h[c] - (a.elev[c] - gwm_soil_depth(a, c)));
        v[1] = area * a.pn[c];
    }, out);
}

hp2d_status gwm_run(hp2d_handle *h, const hp2d_gw_params *p, hp2d_gw_state *g, GwmArgs a,
                    double *col, const uint8_t *dom, double *error, int32_t *ran, double *s_gw) {
    const int ny = a.ny, nx = a.nx;
    const size_t n = (size_t)a.n;
    hipStream_t s = hp2d_handle_stream(h);
    a.dt_s = p->time_step_min * 60.0;
    a.A = p->resolution * p->resolution;
    a.l_vg = p->l_vg;
    hipError_t e = hipSuccess;
    auto D2H = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    const dim3 G((unsigned)((n + 255) / 256)), B(256);
    std::vector<double> hc(3 * (size_t)nx);
    auto fold = [&](double &sum, double &m2, double &m3) {
        sum = 0.0;
        m2 = NAN;
        m3 = NAN;
        for (int j = 0; j < nx; ++j) {
            sum += hc[3 * (size_t)j];
            m2 = std::fmax(m2, hc[3 * (size_t)j + 1]);
            m3 = std::fmax(m3, hc[3 * (size_t)j + 2]);
        }
    };
    // synchronous module: the update follows every call, over the interval known now
    const bool sync = p->flag_async == 0;
    const double dt_sync = std::fmax(g->elapsed_s + a.dt_s, DBL_EPSILON);
    if (a.I32)
        hipLaunchKernelGGL(gwm_recharge_kernel<true>, dim3(nx), dim3(kColThreads), 0, s, a,
                           sync ? dt_sync : 0.0, col);
    else
        hipLaunchKernelGGL(gwm_recharge_kernel<false>, dim3(nx), dim3(kColThreads), 0, s, a,
                           sync ? dt_sync : 0.0, col);
    e = hipGetLastError();
    D2H(hc.data(), col, hc.size() * 8);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("groundwater module: ") + hipGetErrorString(e));
    double rech_sum, hc_max, maxD;
    fold(rech_sum, hc_max, maxD);
    g->elapsed_s = g->elapsed_s + a.dt_s;
    g->total_recharge_m3 = g->total_recharge_m3 + rech_sum;
    g->total_capillary_m3 = g->total_capillary_m3 + 0.0;
    // scheduler (initialize_groundwater_scheduler.m:62-65, estimate_...m, :166-194)
    const double target_dt_s = std::fmax(p->target_dt_min * 60.0, DBL_EPSILON);
    const double min_dt_s = std::fmax(p->min_dt_min * 60.0, DBL_EPSILON);
    const double max_head_change = std::fmax(p->max_head_change_m, DBL_EPSILON);
    const double courant_c = std::fmax(std::fmin(p->courant, 1.0), DBL_EPSILON);
    double dt_stable = target_dt_s;
    if (p->flag_baseflow == 1 && std::isfinite(maxD) && maxD > 0) {
        const double L = p->resolution;
        dt_stable = courant_c * (L * L) / maxD / 4.0;
        dt_stable = std::fmin(std::fmax(dt_stable, 1.0), target_dt_s);
    }
    double target = std::fmin(target_dt_s, dt_stable);
    target = std::fmax(target, min_dt_s);
    const bool head_due = hc_max >= max_head_change;
    const bool final_step = p->t_min >= (p->routing_time_min - 1e-9);
    const bool due = p->flag_async == 0 ? true : (g->elapsed_s >= target || head_due || final_step);
    const double dt_gw = std::fmax(g->elapsed_s, DBL_EPSILON);
    double err = 0.0;
    double s_after = 0.0;       // aquifer storage after the update (post pass)
    const double *h_new = nullptr, *q_new = nullptr;  // Boussinesq head and q_exf
    bool have_storage = false;
    if (due) {
        if (!sync) {
            hipLaunchKernelGGL(gwm_rfg_kernel, G, B, 0, s, a, dt_gw);
            e = hipGetLastError();
        }
        if (p->flag_baseflow == 1) {
            char *wsb = (char *)hp2d_handle_workspace(h, gw_scratch_bytes(ny, nx) + 256);
            if (!wsb) return HP2D_E_NOMEM;
            GwScratch w = gw_scratch(wsb, ny, nx);
            w.ux = w.uy = nullptr;  // the module does not return the Darcy velocities
            GwArgs ga;
            ga.ny = ny;
            ga.nx = nx;
            ga.n = (long long)n;
            ga.dx = p->resolution;
            ga.dy = p->resolution;
            ga.courant = p->courant;
            ga.z0 = a.z_bed;
            ga.Sy = a.Sy;
            ga.Sy_u = a.Sy_u;
            ga.R = a.rfg;
            ga.K = a.Kgw;
            ga.K_u = a.Kgw_u;
            ga.h_soil = a.soil_depth;
            ga.h_soil_u = a.soil_depth_u;
            ga.h_dir = nullptr;
            ga.catchment = dom;
            ga.dirichlet = nullptr;
            ga.ctl = w.ctl;
            ga.fast = a.fast_boussinesq;
            if (e == hipSuccess) {
                hp2d_status st = gw_run(h, ga, dt_gw, a.h, w, &err, nullptr, &h_new);
                if (st != HP2D_OK) return st;
                q_new = w.q;  // the head and q_exf are taken over by the post pass
            }
        } else if (e == hipSuccess) {
            hipLaunchKernelGGL(gwm_local_kernel, G, B, 0, s, a, dt_gw);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(gwm_post_kernel, dim3(nx), dim3(kColThreads), 0, s, a, dt_gw, col,
                               (const double *)h_new, (const double *)q_new);
            e = hipGetLastError();
        }
        D2H(hc.data(), col, hc.size() * 8);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess)
            return hp2d_handle_fail(h, std::string("groundwater module: ") + hipGetErrorString(e));
        double exf = 0.0, ex_max = NAN;
        for (int j = 0; j < nx; ++j) {
            exf += hc[3 * (size_t)j];
            ex_max = std::fmax(ex_max, hc[3 * (size_t)j + 1]);
            s_after += hc[3 * (size_t)j + 2];
        }
        have_storage = true;
        g->total_exfiltration_m3 = g->total_exfiltration_m3 + exf;
        g->n_updates += 1;
        g->elapsed_s = 0.0;
        g->last_dt_s = dt_gw;
        if (ex_max > 0) {  // any(excess_m(:) > 0)
            hipLaunchKernelGGL(gwm_excess_kernel, G, B, 0, s, a);
            e = hipGetLastError();
        }
    } else {
        g->last_dt_s = 0.0;
        hipLaunchKernelGGL(gwm_rfg_kernel, G, B, 0, s, a, 1.0);  // q_exf = zero_matrix
        e = hipGetLastError();
    }
    if (e == hipSuccess && a.maxd) {
        hipLaunchKernelGGL(gwm_maxdepth_kernel, G, B, 0, s, a);
        e = hipGetLastError();
    }
    if (s_gw && !have_storage && e == hipSuccess) {
        hipLaunchKernelGGL(gwm_storage_kernel, dim3(nx), dim3(kColThreads), 0, s, a, col);
        e = hipGetLastError();
        D2H(hc.data(), col, 2 * (size_t)nx * 8);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("groundwater module: ") + hipGetErrorString(e));
    if (s_gw && have_storage) {  // the pending term is zero after the update
        *s_gw = s_after + 0.0;
    } else if (s_gw) {  // nansum(nansum(A.*Sy.*(h - z_bed))) + nansum(nansum(A.*pending))
        double sg = 0.0, sp = 0.0;
        for (int j = 0; j < nx; ++j) {
            sg += hc[2 * (size_t)j];
            sp += hc[2 * (size_t)j + 1];
        }
        *s_gw = sg + sp;
    }
    *error = err;
    *ran = due ? 1 : 0;
    return HP2D_OK;
}

}  // namespace hp2d

extern "C" hp2d_status hp2d_groundwater_module(hp2d_handle *h, const hp2d_gw_params *p,
                                               hp2d_gw_state *g, double *h_t, double *I_t,
                                               double *d_t, double *cumulative_recharge,
                                               double *max_gw_depth, double *q_exf, double *error,
                                               int32_t *ran) {
    using namespace hp2d;
    if (!h || !p || !g || !h_t || !I_t || !d_t || !q_exf || !error || !ran || !p->elevation ||
        !p->soil_depth || !p->theta_sat || !p->theta_r || !p->n_vg || !p->alpha_vg ||
        !p->ksat_mmh || !p->ksat_gw_mmh || !p->Sy || !g->pending_net || !g->pending_recharge ||
        !g->pending_capillary || !(p->resolution > 0.0))
        return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    if (ny != p->ny || nx != p->nx || (p->flag_baseflow == 1 && (ny < 2 || nx < 2))) return HP2D_E_ARG;
    const size_t n = (size_t)ny * nx, nb = n * sizeof(double);
    hipStream_t s = hp2d_handle_stream(h);
    // module arrays in the second workspace: 9 inputs, 8 state, 5 temporaries | colsums | mask
    const size_t n_arr = 22;
    const size_t col_off = n_arr * nb;
    const size_t mask_off = col_off + 3 * (size_t)nx * sizeof(double);
    char *ws = (char *)hp2d_handle_workspace2(h, mask_off + n + 256);
    if (!ws) return HP2D_E_NOMEM;
    double *q0 = (double *)ws;
    auto arr = [&](int k) { return q0 + (size_t)k * n; };
    GwmArgs a;
    a.ny = ny;
    a.nx = nx;
    a.n = (long long)n;
    double *in[9] = {arr(0), arr(1), arr(2), arr(3), arr(4), arr(5), arr(6), arr(7), arr(8)};
    a.elev = in[0]; a.soil_depth = in[1]; a.ts = in[2]; a.tr = in[3]; a.nv = in[4]; a.alpha = in[5];
    a.ksat = in[6]; a.ksat_gw = in[7]; a.Sy = in[8];
    a.h = arr(9); a.I = arr(10); a.d = arr(11); a.cum = cumulative_recharge ? arr(12) : nullptr;
    a.maxd = max_gw_depth ? arr(13) : nullptr; a.pn = arr(14); a.pr = arr(15); a.pc = arr(16);
    a.z_bed = arr(17); a.Kgw = arr(18); a.rfg = arr(19); a.q = arr(20); a.t2 = arr(21);
    double *col = (double *)(ws + col_off);
    uint8_t *dom = p->domain ? (uint8_t *)(ws + mask_off) : nullptr;
    hipError_t e = hipSuccess;
    auto H2D = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    auto D2H = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    const double *hin[9] = {p->elevation, p->soil_depth, p->theta_sat, p->theta_r, p->n_vg,
                            p->alpha_vg, p->ksat_mmh, p->ksat_gw_mmh, p->Sy};
    for (int k = 0; k < 9; ++k) H2D(in[k], hin[k], nb);
    H2D(a.h, h_t, nb);
    H2D(a.I, I_t, nb);
    H2D(a.d, d_t, nb);
    if (a.cum) H2D(a.cum, cumulative_recharge, nb);
    if (a.maxd) H2D(a.maxd, max_gw_depth, nb);
    H2D(a.pn, g->pending_net, nb);
    H2D(a.pr, g->pending_recharge, nb);
    H2D(a.pc, g->pending_capillary, nb);
    if (dom) H2D(dom, p->domain, n);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_groundwater_module: ") + hipGetErrorString(e));
    a.fast_boussinesq = prec == HP2D_FP32;  // an FP32 handle: the FP32 runs' sub-steps
    double err = 0.0;
    int32_t did = 0;
    st = gwm_run(h, p, g, a, col, dom, &err, &did, nullptr);
    if (st != HP2D_OK) return st;
    D2H(h_t, a.h, nb);
    D2H(I_t, a.I, nb);
    D2H(d_t, a.d, nb);
    D2H(q_exf, a.q, nb);
    if (a.cum) D2H(cumulative_recharge, a.cum, nb);
    if (a.maxd) D2H(max_gw_depth, a.maxd, nb);
    D2H(g->pending_net, a.pn, nb);
    D2H(g->pending_recharge, a.pr, nb);
    D2H(g->pending_capillary, a.pc, nb);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_groundwater_module: ") + hipGetErrorString(e));
    *error = err;
    *ran = did;
    return HP2D_OK;
}
