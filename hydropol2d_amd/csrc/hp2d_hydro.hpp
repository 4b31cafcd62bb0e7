// hp2d_hydro.hpp - per-cell arithmetic shared by the step kernels: precision
// helpers and the infiltration closures of the hydrological prologue
// (Hydrological_Model.m:24), with the reference's operation order in FP64.
#pragma once

#include "hp2d_internal.hpp"
#include "hp2d_li.hpp"

namespace hp2d {

// ---------------------------------------------------------------------------
// precision-specific arithmetic
// ---------------------------------------------------------------------------
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// a*b + c: in the reference's order (two roundings) in FP64, fused in FP32
__device__ __forceinline__ double fmadd(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// mm -> m
__device__ __forceinline__ double mm2m(double x) { return x / 1000.0; }
__device__ __forceinline__ float mm2m(float x) { return x * 1e-3f; }

// The slot of linear index key in the sorted outlet list, or -1
__device__ __forceinline__ int find_outlet(const long long *idx, int n, long long key) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const long long v = idx[mid];
        if (v == key) return mid;
        if (v < key) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// The legacy explicit Green-Ampt closure (HydroPol2D_Main_While_Ensemble.m:243-298)
// on one cell of the class (ksat [mm/h], psi [mm], dtheta, Lu [mm], k_out
// [mm/h], Tr [h]): dd the depth after the rain [mm], Ip the infiltrated depth
// of the last step [mm], Tn the recovery clock [h] (updated); In receives the
// new infiltrated depth.  Returns f [mm/h].
template <typename T>
__device__ __forceinline__ T ga_infiltrate(T ksat, T psi, T dth, T Lu_mm, T k_out, T Tr, T Imin,
                                           bool first_step, T dt_h, T inv_dt_h, T dd, T Ip, T &Tn,
                                           T &In) {
    constexpr bool f64 = sizeof(T) == 8;
    const T i_a = f64 ? dd / dmax(dt_h, T(1e-12)) : dd * inv_dt_h;
    T C = f64 ? ksat * (T(1) + ((dd + psi) * dth) / Ip)
              : ksat * fmadd((dd + psi) * dth, frcp(Ip), T(1));
    T f;
    if (first_step) {
        f = dmin(C, i_a);
        In = dmax(f64 ? Ip + f * dt_h : fmadd(f, dt_h, Ip), Imin);
        Tn = Tr;
    } else {
        if (Ip > Lu_mm) C = ksat;
        const T idxC = (i_a <= C) ? T(1) : T(0);
        Tn = Tn - dt_h;
        if (i_a > C) Tn = Tr;
        f = dmin(C, i_a);
        In = dmax(f64 ? Ip + f * dt_h - k_out * idxC * dt_h
                      : fmadd(f, dt_h, Ip) - k_out * idxC * dt_h,
                  Imin);
        if (Tn < T(0)) In = Imin;
    }
    return f;
}

// x^y for x >= 0 in FP32 as exp2(y log2 x) on the hardware's log / exp
// (0^y = 0 for y > 0, +Inf for y < 0)
__device__ __forceinline__ float vg_pow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// The van Genuchten / Mualem closure of Infiltration_Module.m:182-473
// (flag_infiltration = 1, no subgrid, no layered soil) on one cell: dd is the
// depth after the rain [mm] (the step's eff_depth, updated: :467-470), I the
// unsaturated-zone storage [mm] (updated: :456-463), zw the depth to the water
// table zwt = Soil_Depth - (h_t - (elevation - Soil_Depth)) [m] (:194-197) with
// the sign bit set on impervious cells (idx_imp, :440).  Returns f [mm/h].
// FP64 follows the reference's operations in order; FP32 takes the powers as
// exp2 / log2 and the divisions as reciprocal multiplies.
template <typename T>
__device__ __forceinline__ T vg_infiltrate(const LiArgs<T> &a, const float4 *lds, int cls, T zw,
                                           T dt_h, T inv_dt_h, T &dd, T &I) {
    constexpr bool f64 = sizeof(T) == 8;
    const bool imp = signbit(zw);
    const T eff = dd;
    T zwt = fabs(zw);
    T ts, tr, f;
    const T Ltop = a.vg_Ltop, Lmix = a.vg_Ltop;
    if constexpr (f64) {
        const VgClass g = a.vg[cls];
        ts = g.ts;
        tr = g.tr;
        const T i_a = eff / fmax(dt_h, T(1e-12));                         // :182
        const T h0 = fmax(eff, T(0)) / T(1000);                           // :185
        zwt = fmax(zwt, T(0));                                            // :197-198
        const T uz_max = zwt * (ts - tr);                                 // :206
        const T S = fmax(I, T(0)) / T(1000);                              // :210
        const T S_rem = fmax(uz_max - S, T(0));                           // :216
        const bool no_uz = (zwt <= T(0)) || (uz_max <= T(0));            // :220
        const T nv = g.nv, al = g.alpha;
        const T m = T(1) - T(1) / nv;
        const T Ks = g.ksat / T(1000) / T(3600);
        const T zwt_safe = fmax(zwt, T(1e-6));
        T th_b = tr + (S / zwt_safe);
        th_b = fmin(fmax(th_b, tr), ts);
        T Se_b = (th_b - tr) / fmax(ts - tr, T(1e-12));
        Se_b = fmin(fmax(Se_b, T(1e-6)), T(1));
        T h_b = -(T(1) / al) * pow(pow(Se_b, -T(1) / m) - T(1), T(1) / nv);
        if (Se_b >= T(0.999999)) h_b = T(0);
        const T w_avail = fmax(eff, T(0)) / T(1000);
        const T w_top_def = Lmix * fmax(ts - th_b, T(0));
        T w_wet = fmin(w_avail, S_rem);
        w_wet = fmax(fmin(w_wet, w_top_def), T(0));
        T th_t = th_b + w_wet / Lmix;
        th_t = fmin(fmax(th_t, tr), ts);
        T Se_t = (th_t - tr) / fmax(ts - tr, T(1e-12));
        Se_t = fmin(fmax(Se_t, T(1e-6)), T(1));
        T term = T(1) - pow(Se_t, T(1) / m);
        term = fmin(fmax(term, T(0)), T(1));
        T Kr = pow(Se_t, a.vg_l) * pow(T(1) - pow(term, m), T(2));
        Kr = fmin(fmax(Kr, T(0)), T(1));
        const T Ktop = Ks * Kr, Ksurf = Ks;
        const bool wet = (i_a > T(0)) || (h0 > T(0));
        T Ksoil = wet ? a.vg_wK * Ksurf + (T(1) - a.vg_wK) * Ktop : Ktop;
        Ksoil = fmin(fmax(Ksoil, T(0)), Ksurf);
        T dh = fmax(h0 - h_b, T(0));
        dh = fmin(dh, a.vg_dhmax);
        const T grad = (dh / Ltop) + T(1);
        T C = Ksoil * grad;
        C = C * T(1000) * T(3600);
        const T f_store = (S_rem * T(1000)) / fmax(dt_h, T(1e-12));
        f = fmin(fmin(i_a, C), f_store);
        if (imp || no_uz || isnan(f)) f = T(0);
        f = fmax(f, T(0));
        const T inf_cell = f * dt_h;
        T It = fmin(I + inf_cell, uz_max * T(1000));
        I = fmax(It, T(0));
        const T dn = dd - inf_cell * a.cell_area / a.C_a;
        dd = dn < T(0) ? T(0) : dn;
    } else {
        // {tr, ts, 1/n, m, 1/m, -1/m, 1/alpha, Ks [m/s]}
        const float4 g0 = lds[2 * cls], g1 = lds[2 * cls + 1];
        tr = g0.x;
        ts = g0.y;
        const float inv_n = g0.z, m = g0.w, inv_m = g1.x, ninv_m = g1.y, inv_al = g1.z, Ks = g1.w;
        const float i_a = eff * inv_dt_h;
        const float h0 = fmaxf(eff, 0.0f) * 1e-3f;
        zwt = fmaxf(zwt, 0.0f);
        const float dts = ts - tr;
        const float uz_max = zwt * dts;
        const float S = fmaxf(I, 0.0f) * 1e-3f;
        const float S_rem = fmaxf(uz_max - S, 0.0f);
        const bool no_uz = (zwt <= 0.0f) || (uz_max <= 0.0f);
        const float inv_dts = frcp(fmaxf(dts, 1e-12f));
        float th_b = tr + S * frcp(fmaxf(zwt, 1e-6f));
        th_b = fminf(fmaxf(th_b, tr), ts);
        const float Se_b = fminf(fmaxf((th_b - tr) * inv_dts, 1e-6f), 1.0f);
        float h_b = -inv_al * vg_pow(vg_pow(Se_b, ninv_m) - 1.0f, inv_n);
        if (Se_b >= 0.999999f) h_b = 0.0f;
        const float w_avail = h0;
        const float w_top_def = Lmix * fmaxf(ts - th_b, 0.0f);
        const float w_wet = fmaxf(fminf(fminf(w_avail, S_rem), w_top_def), 0.0f);
        float th_t = th_b + w_wet * frcp(Lmix);
        th_t = fminf(fmaxf(th_t, tr), ts);
        const float Se_t = fminf(fmaxf((th_t - tr) * inv_dts, 1e-6f), 1.0f);
        // Se_top^(1/m) and Se_top^l share one log2
        const float lse = __builtin_amdgcn_logf(Se_t);
        const float term = fminf(fmaxf(1.0f - __builtin_amdgcn_exp2f(inv_m * lse), 0.0f), 1.0f);
        const float u = 1.0f - vg_pow(term, m);
        const float Kr = fminf(fmaxf(__builtin_amdgcn_exp2f(a.vg_l * lse) * (u * u), 0.0f), 1.0f);
        const float Ktop = Ks * Kr;
        const bool wet = (i_a > 0.0f) || (h0 > 0.0f);
        float Ksoil = wet ? fmadd(a.vg_wK, Ks - Ktop, Ktop) : Ktop;
        Ksoil = fminf(fmaxf(Ksoil, 0.0f), Ks);
        const float dh = fminf(fmaxf(h0 - h_b, 0.0f), a.vg_dhmax);
        const float C = Ksoil * fmadd(dh, frcp(Ltop), 1.0f) * 3.6e6f;
        const float f_store = S_rem * 1000.0f * inv_dt_h;
        f = fminf(fminf(i_a, C), f_store);
        if (imp || no_uz || isnan(f)) f = 0.0f;
        f = fmaxf(f, 0.0f);
        const float inf_cell = f * dt_h;
        I = fmaxf(fminf(I + inf_cell, uz_max * 1000.0f), 0.0f);
        const float dn = dd - inf_cell * a.cell_area / a.C_a;
        dd = dn < 0.0f ? 0.0f : dn;
    }
    return f;
}

}  // namespace hp2d
