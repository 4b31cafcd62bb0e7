// hp2d_li.hip - fused local-inertial D4 step for gfx950.
//
// One wavefront owns a slab of 64 grid rows (MATLAB index i, the contiguous
// one) and marches along the columns j of its chunk.  Lanes 2..61 finalise
// rows; lanes 0,1,62,63 are the radius-2 halo that the donor draining limiter
// needs (a face flux depends on its donor cell's scale, which depends on all
// four faces of that donor).  Every per-cell array is read once per column
// with a 64-lane coalesced load; the 2-row halo is re-read from L2, and
// neighbours across lanes come from cross-lane shuffles (no LDS).  Per column
// the wave computes, with a lag of two columns:
//   raw face fluxes  (Inertial_Solver.m:23-26 + caps, Local_Inertial_Model_D4.m:123-142,198,263-304)
//   donor scales     (Local_Inertial_Model_D4.m:799-829)
//   limited fluxes   (:831-857), continuity (:313-315,:389-415), outlet (:418-498)
// and, in simulation mode, the Green-Ampt / rainfall prologue of
// Hydrological_Model.m (folded into the column loads) and the per-cell terms of
// refreshing_timestep.m:63-242 reduced to per-wave partials.
#include "hp2d_internal.hpp"
#include "hp2d_li.hpp"

namespace hp2d {

// Face slope and flow depth.  FP64 keeps the reference's operation order
// (y = z + h, Sx/Sy :123-127, Hf :133-134).  FP32 evaluates the same
// expressions with the elevation difference taken first, so the flow depth of
// a shallow face keeps the precision of the cell depths instead of that of
// the elevations (both are the same real-number expression).
__device__ __forceinline__ void face_geom(double zA, double zB, double hA, double hB, double R,
                                          double &S, double &Hf) {
    const double yA = zA + hA, yB = zB + hB;
    S = (yB - yA) / R;
    Hf = fmax(fmax(yA, yB) - fmax(zA, zB), 0.0);
}
__device__ __forceinline__ void face_geom(float zA, float zB, float hA, float hB, float R,
                                          float &S, float &Hf) {
    const float zm = fmaxf(zA, zB);
    S = ((zB - zA) + (hB - hA)) / R;
    Hf = fmaxf(fmaxf(hA + (zA - zm), hB + (zB - zm)), 0.0f);
}

template <typename T>
__device__ __forceinline__ T face_flux(T qp_mmh, T zA, T zB, T hA, T hB, T n2, T R, T R2, T dt,
                                       int crit, T &Hf_out) {
    // A is the cell that stores the face, B the other side; S = (yB - yA)/R
    // reproduces Sx (:123) for x faces and Sy (:127) for y faces.
    const T g = T(9.81);
    T S, Hf;
    face_geom(zA, zB, hA, hB, R, S, Hf);
    T q_p = qp_mmh / T(1000) / T(3600) * R2 / R;
    T num = q_p - g * Hf * dt * S;
    T den = T(1) + g * dt * n2 * fabs(q_p) / pow73(Hf);
    T q = num / den;
    if (!isfinite(q)) q = T(0);
    if (crit) {
        T cv = Hf * sqrt(T(9.81) * Hf);
        q = dmin(dmax(q, -cv), cv);
    }
    T thr = Hf * T(10);
    q = dmin(dmax(q, -thr), thr);
    Hf_out = Hf;
    return q * R;  // outflow_rate [m3/s]
}

// Per-column loads of one lane, fetched one column ahead of the compute.
template <typename T> struct ColLoad {
    T z, d, n2, qx, qy, pat, I, Tc;
    int soil;
};

template <typename T, int MODE>
__device__ __forceinline__ ColLoad<T> load_column(const LiArgs<T> &a, const T *d_src,
                                                  const T *qx_src, const T *qy_src,
                                                  const T *I_src, const T *Tc_src, bool in,
                                                  long long id) {
    ColLoad<T> v;
    v.z = v.d = v.n2 = v.qx = v.qy = v.I = v.Tc = T(0);
    v.pat = T(1);
    v.soil = 0;
    if (in) {
        v.z = a.z[id];
        v.d = d_src[id];
        v.n2 = a.n2s[id];
        v.qx = qx_src[id];
        v.qy = qy_src[id];
        if (MODE == kModeSim) {
            if (a.rain_pattern) v.pat = a.rain_pattern[id];
            if (a.ga) {
                v.I = I_src[id];
                v.Tc = Tc_src[id];
                if (a.soil_class) v.soil = a.soil_class[id];
            }
        }
    }
    return v;
}

template <typename T>
__device__ __forceinline__ int find_outlet(const long long *idx, int n, long long key) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        long long v = idx[mid];
        if (v == key) return mid;
        if (v < key) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void li_step_kernel(LiArgs<T> a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int slab = blockIdx.x * a.waves_per_block + wv;
    if (slab >= a.n_slabs) return;
    const int ny = a.ny, nx = a.nx;

    long long kstep = 1;
    T dt_s, dt_h, rain = T(0);
    int fail_step = 0;
    const T *d_src = a.d_src, *qx_src = a.qx_src, *qy_src = a.qy_src;
    T *d_dst = a.d_dst, *qx_dst = a.qx_dst, *qy_dst = a.qy_dst;
    const T *I_src = a.I_src, *Tc_src = a.T_src;
    T *I_dst = a.I_dst, *Tc_dst = a.T_dst;
    if (MODE == kModeSim) {
        const SimState *st = a.st;
        if (st->done || st->paused) return;
        kstep = st->k;
        dt_s = T(st->time_step_min * 60.0);
        dt_h = T(st->time_step_min / 60.0);
        rain = T(st->delta_p_agg);
        fail_step = (st->inject_k == kstep);
        const int src = (int)((kstep + 1) & 1), dst = (int)(kstep & 1);
        d_src = a.d_buf[src];
        d_dst = a.d_buf[dst];
        qx_src = a.qx_buf[src];
        qx_dst = a.qx_buf[dst];
        qy_src = a.qy_buf[src];
        qy_dst = a.qy_buf[dst];
        I_src = a.I_buf[src];
        I_dst = a.I_buf[dst];
        Tc_src = a.T_buf[src];
        Tc_dst = a.T_buf[dst];
    } else {
        dt_s = a.dt_s;
        dt_h = a.dt_h;
    }

    const T R = a.R, R2 = a.R2, A = a.cell_area, Ca = a.C_a;
    const int r = slab * kRowsOut - kHalo + lane;
    const bool row_in = (r >= 0) && (r < ny);
    const bool row_own = row_in && lane >= kHalo && lane < kWave - kHalo;
    const int c0 = blockIdx.y * a.tj;
    const int c1 = min(c0 + a.tj, nx);
    const T avail_den = dmax(dt_s, eps_of<T>());

    // pipeline registers (suffix = column lag behind the load column c)
    T h1 = 0, z1 = 0, n21 = 0, qxp1 = 0, d1 = 0, n2s1 = 0;   // column c-1
    T rawQx2 = 0, rawQy1 = 0, rawQy2 = 0, Hfx2 = 0, Hfy1 = 0, Hfy2 = 0;
    T d2 = 0, scale2 = 0, finQx3 = 0, ox3 = 0, n2s2 = 0;

    // per-lane reductions
    T vmax = T(NAN), dtw = T(NAN), rmax = T(NAN), dtmax = T(NAN), dpmax = T(NAN);
    double s_store = 0.0, s_uz = 0.0, s_rain = 0.0, s_inf = 0.0;
    int neg = 0;
    const T alfaR = a.alfa_min * R;

    ColLoad<T> nxt = load_column<T, MODE>(a, d_src, qx_src, qy_src, I_src, Tc_src,
                                          row_in && c0 - 2 >= 0 && c0 - 2 < nx,
                                          (long long)r + (long long)(c0 - 2) * ny);
    for (int c = c0 - 2; c <= c1 + 1; ++c) {
        const bool col_in = (c >= 0) && (c < nx);
        const bool cell_in = row_in && col_in;
        const long long id = (long long)r + (long long)c * ny;
        const ColLoad<T> cur = nxt;
        // issue the next column's loads before computing this one
        if (c + 1 <= c1 + 1)
            nxt = load_column<T, MODE>(a, d_src, qx_src, qy_src, I_src, Tc_src,
                                       row_in && c + 1 >= 0 && c + 1 < nx, id + ny);
        T zc = cur.z, dc = cur.d, n2c = cur.n2, qxc = cur.qx, qyc = cur.qy, dprev = 0;
        // ---- hydrology prologue: Hydrological_Model.m:15-55 (sim mode) ----
        if (MODE == kModeSim && cell_in) {
            dprev = dc;
            const bool nan_cell = isnan(zc);
            T eff = a.rain_pattern ? cur.pat * rain : rain;
            if (nan_cell) eff = T(NAN);
            T dd = dc + eff;
            const bool own = row_own && c >= c0 && c < c1;
            if (own && !nan_cell && !fail_step) s_rain += (double)eff / 1000.0 * (double)A;
            if (a.ga) {
                const GaClass g = a.ga[cur.soil];
                const T ksat = T(g.ksat), psi = T(g.psi), dth = T(g.dtheta), Imin = a.I_min;
                const T Ip = cur.I;
                T Tn = cur.Tc;
                const T i_a = dd / dmax(dt_h, T(1e-12));
                T C = ksat * (T(1) + ((dd + psi) * dth) / Ip);
                T f, In;
                if (kstep == 1) {
                    f = dmin(C, i_a);
                    In = dmax(Ip + f * dt_h, Imin);
                    Tn = T(g.Tr);
                } else {
                    if (Ip > T(g.Lu_mm)) C = ksat;
                    const T idxC = (i_a <= C) ? T(1) : T(0);
                    Tn = Tn - dt_h;
                    if (i_a > C) Tn = T(g.Tr);
                    f = dmin(C, i_a);
                    In = dmax(Ip + f * dt_h - T(g.k_out) * idxC * dt_h, Imin);
                    if (Tn < T(0)) In = Imin;
                }
                if (nan_cell) {
                    In = T(NAN);
                    f = T(NAN);
                }
                dd = dd - f * dt_h;
                if (own) {
                    I_dst[id] = In;
                    Tc_dst[id] = Tn;
                    if (!fail_step) {
                        if (!isnan(f)) s_inf += (double)(f * dt_h) / 1000.0 * (double)A;
                        if (!isnan(In)) s_uz += (double)A * (double)In / 1000.0;
                    }
                }
            }
            if (dd < T(-1e-8)) neg = 1;
            if (dd < T(1e-6)) dd = T(0);  // Hydrological_Model.m:49
            if (dd < T(0)) dd = T(0);     // :52
            dc = dd;
            if (own) {  // rollback celerity inputs (HydroPol2D_Main_While.m:932)
                dtmax = dmax(dtmax, dc / T(1000));
                dpmax = dmax(dpmax, dprev / T(1000));
            }
        }
        const T hc = dmax(dc / T(1000), T(0));  // depth_cell (:31)

        // ---- raw y-face flux at (r,c): face between rows r-1 and r ----
        const T h_up = __shfl_up(hc, 1), z_up = __shfl_up(zc, 1);
        T rawQyc = 0, Hfyc = 0;
        if (cell_in && r >= 1)
            rawQyc = face_flux<T>(qyc, zc, z_up, hc, h_up, fabs(n2c), R, R2, dt_s, a.flag_critical, Hfyc);

        // ---- raw x-face flux at (r,c-1): face between columns c-1 and c ----
        T rawQx1 = 0, Hfx1 = 0;
        if (row_in && c - 1 >= 0 && c - 1 < nx - 1)
            rawQx1 = face_flux<T>(qxp1, z1, zc, h1, hc, fabs(n21), R, R2, dt_s, a.flag_critical, Hfx1);

        // ---- donor scale of cell (r,c-1): Local_Inertial_Model_D4.m:814-829 ----
        const T qy1_dn = __shfl_down(rawQy1, 1);
        T outr = T(0) + dmax(rawQx1, T(0));
        if (c - 1 >= 1) outr += dmax(-rawQx2, T(0));
        outr += dmax(rawQy1, T(0));
        if (r < ny - 1) outr += dmax(-qy1_dn, T(0));
        const T avail = dmax(d1 / T(1000), T(0)) * A / avail_den;
        T s1 = T(1);
        if (outr > avail && outr > T(0)) s1 = avail / outr;
        if (!isfinite(s1)) s1 = T(0);
        s1 = dmax(dmin(s1, T(1)), T(0));

        // ---- limited fluxes of output column o = c-2 (:831-857) ----
        const int o = c - 2;
        T fqx = rawQx2;
        if (fqx > T(0)) fqx = fqx * scale2;
        else if (fqx < T(0)) fqx = fqx * s1;
        if (o == nx - 1) fqx = T(0);
        if (!isfinite(fqx)) fqx = T(0);
        const T s2_up = __shfl_up(scale2, 1);
        T fqy = rawQy2;
        if (fqy > T(0)) fqy = fqy * scale2;
        else if (fqy < T(0)) fqy = fqy * s2_up;
        if (r == 0) fqy = T(0);
        if (!isfinite(fqy)) fqy = T(0);
        const T fqy_dn = __shfl_down(fqy, 1);

        // ---- continuity (:313-315, :390, :415) ----
        const T west = (o >= 1) ? finQx3 : T(0);
        const T south = (r < ny - 1) ? fqy_dn : T(0);
        const T vol = dt_s * (west - fqx - fqy + south);
        T dn = d2 + T(1000) * vol / Ca;
        if (dn < T(0)) dn = T(0);

        const bool out_cell = row_own && o >= c0 && o < c1;
        const long long oid = (long long)r + (long long)o * ny;
        // ---- outlet (:418-498) ----
        T hf3 = T(0), qo = T(0);
        if (out_cell && signbit(n2s2) && a.n_outlet > 0) {
            const int slot = find_outlet<T>(a.out_idx, a.n_outlet, oid);
            if (slot >= 0) {
                const T h_out = dmax(dn, T(0)) / T(1000);
                T sqrtS;
                if (a.outlet_type == 1) {
                    sqrtS = a.sqrtS_type1;
                } else {
                    const T h_safe = dmax(h_out, T(1e-12));
                    sqrtS = sqrt(T(9.81) * fabs(n2s2) * powm13(h_safe));
                    if (!isfinite(sqrtS)) sqrtS = T(0);
                }
                hf3 = h_out;
                qo = (T(1) / a.out_n[slot]) * R * pow53(h_out) * sqrtS / A * T(1000) * T(3600);
                const T qmax = dmax(dn - T(1000) * T(0), T(0)) / dt_h;
                qo = dmin(qo, qmax);
                if (!isfinite(qo)) qo = T(0);
                dn = dn - qo * dt_h;
                if (a.out_q) a.out_q[slot] = qo;
                if (a.out_d) a.out_d[slot] = dn;
            }
        }
        const T ox = fqx / A * T(1000) * T(3600);
        const T oy = fqy / A * T(1000) * T(3600);

        if (MODE == kModeSim) {
            // ---- refreshing_timestep.m:63-238 per-cell terms ----
            const T f1 = Hfx2 * T(1000), f2 = Hfy2 * T(1000), f3 = hf3 * T(1000);
            const T tol = a.depth_tol_mm;
            const T s1d = (f2 < tol) ? T(1e12) : f2;
            const T vu = (oy / T(1000) / T(3600)) * R / (s1d / T(1000));
            const T vu_dn = __shfl_down(vu, 1);
            if (out_cell && !fail_step) {
                const T dts = dmax(dmax(f1 / T(1000), f2 / T(1000)), f3 / T(1000));
                const T sxd = (f1 < tol) ? T(1e12) : f1;
                const T vr = (ox / T(1000) / T(3600)) * R / (sxd / T(1000));
                const T vl = (o >= 1) ? -vr : -T(0);  // :73 unshifted quirk
                const T vd = (r < ny - 1) ? -vu_dn : -T(0);
                const T raster = dmax(dmax(dmax(fabs(vl), fabs(vr)), fabs(vu)), fabs(vd));
                const T rc = T(0.5) * (vr - vl), uc = T(0.5) * (vu - vd);
                const T tot = hypot(rc, uc);
                vmax = dmax(vmax, tot);
                rmax = dmax(rmax, raster);
                T cw = sqrt(T(9.81) * dmax(dts, T(0)));
                cw = dmax(cw, T(1e-10));
                if (isinf(cw)) cw = T(NAN);
                if (tot < T(0.01)) cw = T(NAN);
                const T cand = a.adaptive ? alfaR / cw : alfaR / (raster + cw);
                dtw = dmin(dtw, cand);
                if (!isnan(dn)) s_store += (double)A * (double)dn / 1000.0;
            }
            if (out_cell) {
                d_dst[oid] = dn;
                qx_dst[oid] = ox;
                qy_dst[oid] = oy;
                if (a.dmax_map && !fail_step) {
                    const T prevmax = a.dmax_map[oid];
                    a.dmax_map[oid] = (kstep == 1) ? dn : dmax(dn, prevmax);
                }
            }
        } else {
            const T oy_dn = __shfl_down(oy, 1);
            if (out_cell) {
                const long long n = (long long)ny * nx;
                d_dst[oid] = dn;
                if (qx_dst) qx_dst[oid] = ox;
                if (qy_dst) qy_dst[oid] = oy;
                if (a.q3) a.q3[oid] = qo;
                if (a.hfx) a.hfx[oid] = Hfx2;
                if (a.hfy) a.hfy[oid] = Hfy2;
                if (a.hf3) a.hf3[oid] = hf3;
                if (a.outlet_flow) a.outlet_flow[oid] = qo;
                if (a.qright) a.qright[oid] = ox;
                if (a.qup) a.qup[oid] = oy;
                if (a.qleft) a.qleft[oid] = -((o >= 1) ? ox3 : T(0));
                if (a.qdown) a.qdown[oid] = -((r < ny - 1) ? oy_dn : T(0));
                if (a.itot)
                    a.itot[oid] = fabs(ox + oy + qo) * dt_s / T(1000) * T(1) / T(3600) * R2;
                (void)n;
            }
        }

        // ---- shift the pipeline ----
        finQx3 = fqx;
        ox3 = ox;
        rawQx2 = rawQx1;
        rawQy2 = rawQy1;
        rawQy1 = rawQyc;
        Hfx2 = Hfx1;
        Hfy2 = Hfy1;
        Hfy1 = Hfyc;
        scale2 = s1;
        d2 = d1;
        d1 = dc;
        h1 = hc;
        z1 = zc;
        n21 = n2c;
        n2s2 = n2s1;
        n2s1 = n2c;
        qxp1 = qxc;
    }

    if (MODE == kModeSim) {
        // per-wave partials (deterministic butterfly reductions)
        const T wv_vmax = wave_fmax(vmax), wv_dtw = wave_fmin(dtw), wv_r = wave_fmax(rmax);
        const T wv_dt = wave_fmax(dtmax), wv_dp = wave_fmax(dpmax);
        const double st_sum = wave_sum(s_store), uz_sum = wave_sum(s_uz);
        const double rn_sum = wave_sum(s_rain), in_sum = wave_sum(s_inf);
        const int any_neg = __any(neg);
        if (lane == 0) {
            const long long idx = (long long)blockIdx.y * a.n_slabs + slab;
            double *P = a.partials;
            const long long cap = a.part_cap;
            P[kPartStorage * cap + idx] = st_sum;
            P[kPartUz * cap + idx] = uz_sum;
            P[kPartRain * cap + idx] = rn_sum;
            P[kPartInfil * cap + idx] = in_sum;
            P[kPartVmax * cap + idx] = (double)wv_vmax;
            P[kPartDtw * cap + idx] = (double)wv_dtw;
            P[kPartRaster * cap + idx] = (double)wv_r;
            P[kPartDtotMax * cap + idx] = (double)wv_dt;
            P[kPartDprevMax * cap + idx] = (double)wv_dp;
            P[kPartNeg * cap + idx] = any_neg ? 1.0 : 0.0;
        }
    }
}

template <typename T>
hipError_t launch_li_step(const LiArgs<T> &a, int mode, hipStream_t s) {
    dim3 block(kWave * a.waves_per_block);
    dim3 grid((a.n_slabs + a.waves_per_block - 1) / a.waves_per_block, (a.nx + a.tj - 1) / a.tj);
    if (mode == kModeSim)
        hipLaunchKernelGGL((li_step_kernel<T, kModeSim>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((li_step_kernel<T, kModeDropIn>), grid, block, 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_li_step<float>(const LiArgs<float> &, int, hipStream_t);
template hipError_t launch_li_step<double>(const LiArgs<double> &, int, hipStream_t);

}  // namespace hp2d
