// hp2d_li.hip - fused local-inertial D4 step for gfx950.
//
// One wavefront owns a slab of 64 grid rows (MATLAB index i, the contiguous
// one) and marches along the columns j of its chunk.  Lanes 2..61 finalise
// rows; lanes 0,1,62,63 are the radius-2 halo that the donor draining limiter
// needs (a face flux depends on its donor cell's scale, which depends on all
// four faces of that donor).  Every per-cell array is read once per column
// with a 64-lane coalesced load issued one column ahead of the compute; the
// 2-row halo is re-read from L2, and neighbours across lanes come from
// wavefront-shift DPP moves (no LDS).  Per column the wave computes, with a
// lag of two columns:
//   raw face fluxes  (Inertial_Solver.m:23-26 + caps, Local_Inertial_Model_D4.m:123-142,198,263-304)
//   donor scales     (Local_Inertial_Model_D4.m:799-829)
//   limited fluxes   (:831-857), continuity (:313-315,:389-415), outlet (:418-498)
// and, in simulation mode, the rainfall / Green-Ampt prologue of
// Hydrological_Model.m (folded into the column loads) and the per-cell terms of
// refreshing_timestep.m:63-242 reduced to per-wave partials.
//
// Precision policy: the double instantiation keeps the reference's operation
// order (it is the parity mode); the float instantiation evaluates the same
// expressions with reciprocal multiplies and the hardware log2/exp2/rcp.
#include "hp2d_forcing.hpp"
#include "hp2d_internal.hpp"
#include "hp2d_li.hpp"

namespace hp2d {

// ---------------------------------------------------------------------------
// cross-lane moves: lane i receives lane i-1 (up) / lane i+1 (down)
// (DPP wave_shr:1 / wave_shl:1; the end lanes keep 0 and are halo lanes)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_up(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_down(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_up(double v) { return __shfl_up(v, 1); }
__device__ __forceinline__ double lane_down(double v) { return __shfl_down(v, 1); }

// ---------------------------------------------------------------------------
// precision-specific arithmetic
// ---------------------------------------------------------------------------
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// mm -> m
__device__ __forceinline__ double mm2m(double x) { return x / 1000.0; }
__device__ __forceinline__ float mm2m(float x) { return x * 1e-3f; }

// Face slope and flow depth.  FP64: y = z + h, Sx/Sy (:123-127), Hf (:133-134)
// in the reference's order.  FP32: the same expressions with the elevation
// difference taken first, so a shallow face keeps the precision of the cell
// depths instead of that of the elevations.
__device__ __forceinline__ void face_geom(double zA, double zB, double hA, double hB,
                                          const LiArgs<double> &a, double &S, double &Hf) {
    const double yA = zA + hA, yB = zB + hB;
    S = (yB - yA) / a.R;
    Hf = fmax(fmax(yA, yB) - fmax(zA, zB), 0.0);
}
__device__ __forceinline__ void face_geom(float zA, float zB, float hA, float hB,
                                          const LiArgs<float> &a, float &S, float &Hf) {
    const float zm = fmaxf(zA, zB);
    S = ((zB - zA) + (hB - hA)) * a.inv_R;
    Hf = fmaxf(fmaxf(hA + (zA - zm), hB + (zB - zm)), 0.0f);
}

// FP32 from the face's elevation difference dz = zB - zA, taken once in FP64
// when the run is uploaded (the elevations themselves never enter FP32, so a
// 200 m relief does not cost the face its millimetres): S = (dz + hB - hA)/R,
// Hf = max(hA + (zA - zm), hB + (zB - zm), 0) with zm = max(zA, zB).  A NaN
// dz (a cell outside the domain) gives a NaN slope, whose flux the caller
// zeroes, and the depth of the valid side, as the elevation form does.
__device__ __forceinline__ void face_geom_dz(float dz, float hA, float hB,
                                             const LiArgs<float> &a, float &S, float &Hf) {
    S = (dz + (hB - hA)) * a.inv_R;
    Hf = fmaxf(fmaxf(hA - fmaxf(dz, 0.0f), hB + fminf(dz, 0.0f)), 0.0f);
}

// stored mm/h -> m2/s (:198)
__device__ __forceinline__ double mmh_to_q(double v, const LiArgs<double> &a) {
    return v / 1000.0 / 3600.0 * a.R2 / a.R;
}
__device__ __forceinline__ float mmh_to_q(float v, const LiArgs<float> &a) { return v * a.k_mmh_q; }

// Bates update (Inertial_Solver.m:24-25)
__device__ __forceinline__ double bates(double q_p, double Hf, double S, double n2, double dt) {
    const double g = 9.81;
    return (q_p - g * Hf * dt * S) / (1.0 + g * dt * n2 * fabs(q_p) / pow(Hf, 7.0 / 3.0));
}
__device__ __forceinline__ float bates(float q_p, float Hf, float S, float n2, float dt) {
    const float g = 9.81f;
    const float p73 = __builtin_amdgcn_exp2f((7.0f / 3.0f) * __builtin_amdgcn_logf(Hf));
    const float den = 1.0f + g * dt * n2 * fabsf(q_p) * __builtin_amdgcn_rcpf(p73);
    return (q_p - g * Hf * dt * S) * __builtin_amdgcn_rcpf(den);
}

// face discharge [m3/s] -> stored mm/h of cell area (:309)
__device__ __forceinline__ double q_to_mmh(double Q, const LiArgs<double> &a) {
    return Q / a.cell_area * 1000.0 * 3600.0;
}
__device__ __forceinline__ float q_to_mmh(float Q, const LiArgs<float> &a) { return Q * a.k_q_mmh; }

// FP64: zA, zB are the elevations; FP32: zA is unused and zB is dz = zB - zA
template <typename T, bool CRIT>
__device__ __forceinline__ T face_flux(T qp_mmh, T zA, T zB, T hA, T hB, T n2, T dt,
                                       const LiArgs<T> &a, T &Hf_out) {
    // A is the cell that stores the face, B the other side; S = (yB - yA)/R
    // reproduces Sx (:123) for x faces and Sy (:127) for y faces.
    T S, Hf;
    if constexpr (sizeof(T) == 8) face_geom(zA, zB, hA, hB, a, S, Hf);
    else face_geom_dz(zB, hA, hB, a, S, Hf);
    T q = bates(mmh_to_q(qp_mmh, a), Hf, S, n2, dt);
    if (!isfinite(q)) q = T(0);
    if (CRIT) {
        const T cv = Hf * sqrt(T(9.81) * Hf);
        q = dmin(dmax(q, -cv), cv);
    }
    const T thr = Hf * T(10);
    q = dmin(dmax(q, -thr), thr);
    Hf_out = Hf;
    return q * a.R;  // outflow_rate [m3/s]
}

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

// The step's optional inputs and outputs, as bits of a kernel's feature mask.
enum : int {
    kFeatGa = 1,        // Green-Ampt infiltration (I, T, soil class)
    kFeatPattern = 2,   // a static rainfall pattern
    kFeatMaps = 4,      // rainfall maps
    kFeatGauge = 8,     // inflow hydrographs
    kFeatRisk = 16,     // the human-instability risk map
    kFeatSnap = 32,     // the map stacks' velocity snapshots
    kFeatDmax = 64,     // the maximum-depth map
    kFeatTpeak = 128,   // the time-to-peak map
    kFeatRuntime = -1,  // all read from the arguments
};
template <int FEAT> __device__ __forceinline__ bool feat_on(int bit, bool runtime) {
    if constexpr (FEAT < 0) return runtime;
    else return (FEAT & bit) != 0;
}

// The run's optional inputs and outputs: fixed at compile time in the
// specialised kernels (FEAT >= 0), read from the arguments otherwise.
struct Feats {
    bool ga, pat, maps, gauge, risk, snap, dmax, tpeak;
};
template <int FEAT, typename T> __device__ __forceinline__ Feats feats_of(const LiArgs<T> &a) {
    Feats f;
    f.ga = feat_on<FEAT>(kFeatGa, a.ga != nullptr);
    f.pat = feat_on<FEAT>(kFeatPattern, a.rain_pattern != nullptr);
    f.maps = feat_on<FEAT>(kFeatMaps, a.rain_maps != nullptr);
    f.gauge = feat_on<FEAT>(kFeatGauge, a.inflow_gauge != nullptr);
    f.risk = feat_on<FEAT>(kFeatRisk, a.max_risk != nullptr);
    f.snap = feat_on<FEAT>(kFeatSnap, a.snap_vras != nullptr);
    f.dmax = feat_on<FEAT>(kFeatDmax, a.dmax_map != nullptr);
    f.tpeak = feat_on<FEAT>(kFeatTpeak, a.tpeak_map != nullptr);
    return f;
}

// Per-column loads of one lane, fetched one column ahead of the compute.
// z: the elevation (FP64), or dz of the cell's x face (FP32); zy: dz of its
// y face (FP32 only)
template <typename T> struct ColLoad {
    T z, zy, d, n2, qx, qy, pat, I, Tc;
    int soil, gauge;
};

// The rainfall maps' terms of this step (wave-uniform): the offsets of the
// overlapping maps, their overlaps [min], t - t_previous [min] and dt/60 [h].
template <typename T> struct MapStep {
    int n;
    unsigned off[kMaxMapTerms];
    T w[kMaxMapTerms];
    T dtot, k;
};

template <typename T, int MODE>
__device__ __forceinline__ ColLoad<T> load_column(const LiArgs<T> &a, const Feats &F,
                                                  const T *d_src, const T *qx_src,
                                                  const T *qy_src, const T *I_src,
                                                  const T *Tc_src, bool in, unsigned id,
                                                  const MapStep<T> &ms) {
    ColLoad<T> v;
    v.z = v.zy = v.d = v.n2 = v.qx = v.qy = v.I = v.Tc = T(0);
    v.pat = T(1);
    v.soil = v.gauge = 0;
    if (in) {
        if constexpr (sizeof(T) == 8) {
            v.z = a.z[id];
        } else {
            v.z = a.dzx[id];
            v.zy = a.dzy[id];
        }
        v.d = d_src[id];
        v.n2 = a.n2s[id];
        v.qx = qx_src[id];
        v.qy = qy_src[id];
        if (MODE == kModeSim) {
            if (F.maps) {
                // the step's rain depth from the maps (update_spatial_BC.m:1361-1408):
                // (sum_i overlap_i * map_i) / (t - t_prev) * dt / 60
                T acc = T(0);
                for (int j = 0; j < ms.n; ++j) acc = acc + ms.w[j] * a.rain_maps[ms.off[j] + id];
                v.pat = ms.n > 0 ? acc / ms.dtot * ms.k : T(0);
            } else if (F.pat) {
                v.pat = a.rain_pattern[id];
            }
            if (F.gauge) v.gauge = a.inflow_gauge[id];
            if (F.ga) {
                v.I = I_src[id];
                v.Tc = Tc_src[id];
                if (a.soil_class) v.soil = a.soil_class[id];
            }
        }
    }
    return v;
}

// One workgroup's tile of one grid: `logical` counts the grid's tiles, row
// blocks fastest.
template <typename T, int MODE, bool CRIT, int FEAT>
__device__ __forceinline__ void li_step_tile(const LiArgs<T> &a, const int logical) {
    constexpr bool f64 = sizeof(T) == 8;
    const Feats F = feats_of<FEAT>(a);
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int n_bx = (a.n_slabs + a.waves_per_block - 1) / a.waves_per_block;
    const int bx = logical % n_bx, by = logical / n_bx;
    const int slab = bx * a.waves_per_block + wv;
    const bool active = slab < a.n_slabs;
    const int ny = a.ny, nx = a.nx;

    long long kstep = 1;
    T dt_s, dt_h, rain = T(0), t_end_min = T(0);
    MapStep<T> ms;
    ms.n = 0;
    ms.dtot = ms.k = T(0);
    int fail_step = 0;
    bool may_report = false, may_snap = false;
    const T *d_src = a.d_src, *qx_src = a.qx_src, *qy_src = a.qy_src;
    T *d_dst = a.d_dst, *qx_dst = a.qx_dst, *qy_dst = a.qy_dst;
    const T *I_src = a.I_src, *Tc_src = a.T_src;
    T *I_dst = a.I_dst, *Tc_dst = a.T_dst;
    if (MODE == kModeSim) {
        const SimState *st = a.st;
        if (st->done || st->paused || st->hold) return;
        kstep = st->k;
        dt_s = T(st->time_step_min * 60.0);
        dt_h = T(st->time_step_min / 60.0);
        rain = T(st->delta_p_agg);
        t_end_min = T(st->t + st->time_step_min);
        fail_step = (st->inject_k == kstep);
        if (F.maps) {
            const unsigned cells = (unsigned)ny * (unsigned)nx;
            ms.n = st->map_n;
            for (int j = 0; j < kMaxMapTerms; ++j) {
                ms.off[j] = (unsigned)st->map_idx[j] * cells;
                ms.w[j] = T(st->map_w[j]);
            }
            ms.dtot = T(st->map_dtot);
            ms.k = T(st->map_k);
        }
        // a mass-balance report can only follow this step if t + max_dt reaches the
        // next report threshold (the new dt never exceeds max_dt); only then are the
        // storage sums needed (HydroPol2D_Main_While.m:787-803)
        const double prog_max = ((st->t + st->max_dt / 60.0) / a.routing_time) * 100.0;
        may_report = (prog_max >= st->next_report) || (prog_max >= 100.0);
        // a map record can only follow this step if t + max_dt passes the next
        // record time (the new dt never exceeds max_dt); only then are the
        // velocities stored
        may_snap = F.snap && (st->t + st->max_dt / 60.0 > (double)st->map_last * a.rec_maps);
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

    // Green-Ampt class table (FP32 path) staged in LDS: the per-cell class
    // gather then costs an LDS read instead of a dependent global load
    __shared__ float4 ga_lds[2 * kMaxSoilClasses];
    if (MODE == kModeSim && !f64 && F.ga) {
        for (int i = threadIdx.x; i < 2 * a.n_soil_classes; i += blockDim.x)
            ga_lds[i] = reinterpret_cast<const float4 *>(a.ga_f32)[i];
        __syncthreads();
    }

    const T A = a.cell_area;
    const int r = slab * kRowsOut - kHalo + lane;
    const bool row_in = (r >= 0) && (r < ny);
    const bool row_own = row_in && lane >= kHalo && lane < kWave - kHalo;
    const int c0 = by * a.tj;
    const int c1 = min(c0 + a.tj, nx);
    const T avail_den = dmax(dt_s, eps_of<T>());
    const T inv_dt_h = frcp(dmax(dt_h, T(1e-12)));  // FP32 path only
    const T k_avail = A * frcp(avail_den);            // FP32 path only
    const T k_cont = T(1000) * frcp(a.C_a);           // FP32 path only

    // pipeline registers (suffix = column lag behind the load column c)
    T h1 = 0, z1 = 0, n21 = 0, qxp1 = 0, d1 = 0, n2s1 = 0;   // column c-1
    T rawQx2 = 0, rawQy1 = 0, rawQy2 = 0, Hfx2 = 0, Hfy1 = 0, Hfy2 = 0;
    T d2 = 0, scale2 = 0, finQx3 = 0, ox3 = 0, n2s2 = 0;

    // per-lane reductions
    T vkey = T(NAN), wkey = T(NAN), rmax = T(NAN), dtmax = T(NAN), dpmax = T(NAN), dnmax = T(NAN);
    double s_store = 0.0, s_uz = 0.0, s_inf = 0.0;
    int neg = 0;

    // 32-bit element offsets (grids are limited to 2^32 cells at creation)
    ColLoad<T> nxt = load_column<T, MODE>(a, F, d_src, qx_src, qy_src, I_src, Tc_src,
                                          active && row_in && c0 - 2 >= 0 && c0 - 2 < nx,
                                          (unsigned)r + (unsigned)(c0 - 2) * (unsigned)ny, ms);
    const int c_end = active ? c1 + 1 : c0 - 3;  // inactive waves skip the march
#pragma clang loop unroll_count(2)
    for (int c = c0 - 2; c <= c_end; ++c) {
        const bool col_in = (c >= 0) && (c < nx);
        const bool cell_in = row_in && col_in;
        const unsigned id = (unsigned)r + (unsigned)c * (unsigned)ny;
        const ColLoad<T> cur = nxt;
        // issue the next column's loads before computing this one
        if (c + 1 <= c1 + 1)
            nxt = load_column<T, MODE>(a, F, d_src, qx_src, qy_src, I_src, Tc_src,
                                       row_in && c + 1 >= 0 && c + 1 < nx, id + (unsigned)ny, ms);
        const T zc = cur.z, n2c = cur.n2, qxc = cur.qx, qyc = cur.qy;
        T dc = cur.d;
        // cells outside the domain: NaN elevation (FP64) or NaN n^2 (FP32)
        const bool nan_cell = f64 ? isnan(zc) : isnan(n2c);

        // ---- rainfall + Green-Ampt prologue: Hydrological_Model.m:15-55 ----
        if (MODE == kModeSim && cell_in) {
            const T dprev = dc;
            T eff = F.maps ? cur.pat : (F.pat ? cur.pat * rain : rain);
            if (nan_cell) eff = T(NAN);
            T dd = dc + eff;
            const bool own = row_own && c >= c0 && c < c1;
            if (F.ga) {
                // legacy explicit Green-Ampt closure in the Infiltration_Module slot
                // (HydroPol2D_Main_While_Ensemble.m:243-298)
                T ksat, psi, dth, Lu_mm, k_out, Tr;
                if (f64) {
                    const GaClass g = a.ga[cur.soil];
                    ksat = T(g.ksat); psi = T(g.psi); dth = T(g.dtheta);
                    Lu_mm = T(g.Lu_mm); k_out = T(g.k_out); Tr = T(g.Tr);
                } else {
                    const float4 g0 = ga_lds[2 * cur.soil];
                    const float4 g1 = ga_lds[2 * cur.soil + 1];
                    ksat = g0.x; psi = g0.y; dth = g0.z; Lu_mm = g0.w; k_out = g1.x; Tr = g1.y;
                }
                const T Imin = a.I_min;
                const T Ip = cur.I;
                T Tn = cur.Tc;
                const T i_a = f64 ? dd / dmax(dt_h, T(1e-12)) : dd * inv_dt_h;
                T C = f64 ? ksat * (T(1) + ((dd + psi) * dth) / Ip)
                          : ksat * (T(1) + ((dd + psi) * dth) * frcp(Ip));
                T f, In;
                if (kstep == 1) {
                    f = dmin(C, i_a);
                    In = dmax(Ip + f * dt_h, Imin);
                    Tn = Tr;
                } else {
                    if (Ip > Lu_mm) C = ksat;
                    const T idxC = (i_a <= C) ? T(1) : T(0);
                    Tn = Tn - dt_h;
                    if (i_a > C) Tn = Tr;
                    f = dmin(C, i_a);
                    In = dmax(Ip + f * dt_h - k_out * idxC * dt_h, Imin);
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
                        if (!isnan(f)) s_inf += (double)(f * dt_h);
                        if (may_report && !isnan(In)) s_uz += (double)In;
                    }
                }
            }
            // Hydrological_Model.m:30: the inflow hydrographs on their gauge cells
            if (F.gauge && cur.gauge > 0) dd = dd + T(a.st->inflow_agg[cur.gauge - 1]);
            if (dd < T(-1e-8)) neg = 1;
            if (dd < T(1e-6)) dd = T(0);  // Hydrological_Model.m:49
            if (dd < T(0)) dd = T(0);     // :52
            dc = dd;
            if (own && (fail_step || a.track_dmax)) {  // rollback celerity inputs (HydroPol2D_Main_While.m:932)
                dtmax = dmax(dtmax, dc);
                dpmax = dmax(dpmax, dprev);
            }
        }
        const T hc = dmax(mm2m(dc), T(0));  // depth_cell (:31)

        // ---- raw y-face flux at (r,c): face between rows r-1 and r ----
        const T h_up = lane_up(hc);
        const T z_up = f64 ? lane_up(zc) : cur.zy;  // FP32: dz = z(r-1) - z(r)
        T rawQyc = 0, Hfyc = 0;
        if (cell_in && r >= 1)
            rawQyc = face_flux<T, CRIT>(qyc, zc, z_up, hc, h_up, fabs(n2c), dt_s, a, Hfyc);

        // ---- raw x-face flux at (r,c-1): face between columns c-1 and c ----
        T rawQx1 = 0, Hfx1 = 0;
        // (FP32: z1 is dz = z(c) - z(c-1), stored with column c-1)
        if (row_in && c - 1 >= 0 && c - 1 < nx - 1)
            rawQx1 = face_flux<T, CRIT>(qxp1, z1, f64 ? zc : z1, h1, hc, fabs(n21), dt_s, a, Hfx1);

        // ---- donor scale of cell (r,c-1): Local_Inertial_Model_D4.m:814-829 ----
        const T qy1_dn = lane_down(rawQy1);
        T outr = T(0) + dmax(rawQx1, T(0));
        if (c - 1 >= 1) outr += dmax(-rawQx2, T(0));
        outr += dmax(rawQy1, T(0));
        if (r < ny - 1) outr += dmax(-qy1_dn, T(0));
        const T avail = f64 ? dmax(d1 / T(1000), T(0)) * A / avail_den
                            : dmax(d1 * T(1e-3), T(0)) * k_avail;
        T s1 = T(1);
        if (outr > avail && outr > T(0)) s1 = f64 ? avail / outr : avail * frcp(outr);
        if (!isfinite(s1)) s1 = T(0);
        s1 = dmax(dmin(s1, T(1)), T(0));

        // ---- limited fluxes of output column o = c-2 (:831-857) ----
        const int o = c - 2;
        // donor-scaled with a select of the multiplier (a zero flux is never
        // scaled, as the masked assignments of :834-851 leave it untouched)
        T fqx = rawQx2 * ((rawQx2 > T(0)) ? scale2 : ((rawQx2 < T(0)) ? s1 : T(1)));
        if (o == nx - 1) fqx = T(0);
        if (!isfinite(fqx)) fqx = T(0);
        const T s2_up = lane_up(scale2);
        T fqy = rawQy2 * ((rawQy2 > T(0)) ? scale2 : ((rawQy2 < T(0)) ? s2_up : T(1)));
        if (r == 0) fqy = T(0);
        if (!isfinite(fqy)) fqy = T(0);
        const T fqy_dn = lane_down(fqy);

        // ---- continuity (:313-315, :390, :415) ----
        const T west = (o >= 1) ? finQx3 : T(0);
        const T south = (r < ny - 1) ? fqy_dn : T(0);
        const T vol = dt_s * (west - fqx - fqy + south);
        T dn = f64 ? d2 + T(1000) * vol / a.C_a : d2 + vol * k_cont;
        if (dn < T(0)) dn = T(0);

        const bool out_cell = row_own && o >= c0 && o < c1;
        const unsigned oid = (unsigned)r + (unsigned)o * (unsigned)ny;
        // ---- outlet (:418-498) ----
        T hf3 = T(0), qo = T(0);
        if (out_cell && signbit(n2s2) && a.n_outlet > 0) {
            const int slot = find_outlet(a.out_idx, a.n_outlet, (long long)oid);
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
                qo = (T(1) / a.out_n[slot]) * a.R * pow53(h_out) * sqrtS / A * T(1000) * T(3600);
                const T qmax = dmax(dn - T(1000) * T(0), T(0)) / dt_h;
                qo = dmin(qo, qmax);
                if (!isfinite(qo)) qo = T(0);
                dn = dn - qo * dt_h;
                if (a.out_q) a.out_q[slot] = qo;
                if (a.out_d) a.out_d[slot] = dn;
            }
        }
        const T ox = q_to_mmh(fqx, a);
        const T oy = q_to_mmh(fqy, a);

        if (MODE == kModeSim) {
            // ---- refreshing_timestep.m:63-238 per-cell terms ----
            const T tol = a.depth_tol_mm;
            T vu;
            if (f64) {
                const T f2 = Hfy2 * T(1000);
                const T s2d = (f2 < tol) ? T(1e12) : f2;
                vu = (oy / T(1000) / T(3600)) * a.R / (s2d / T(1000));
            } else {
                const T hy = (Hfy2 * T(1000) < tol) ? T(1e9) : Hfy2;
                vu = fqy * a.inv_R * frcp(hy);
            }
            const T vu_dn = lane_down(vu);
            if (out_cell && !fail_step) {
                T vr, dts;
                if (f64) {
                    const T f1 = Hfx2 * T(1000), f2 = Hfy2 * T(1000), f3 = hf3 * T(1000);
                    dts = dmax(dmax(f1 / T(1000), f2 / T(1000)), f3 / T(1000));
                    const T s1d = (f1 < tol) ? T(1e12) : f1;
                    vr = (ox / T(1000) / T(3600)) * a.R / (s1d / T(1000));
                } else {
                    dts = dmax(dmax(Hfx2, Hfy2), hf3);
                    const T hx = (Hfx2 * T(1000) < tol) ? T(1e9) : Hfx2;
                    vr = fqx * a.inv_R * frcp(hx);
                }
                const T vl = (o >= 1) ? -vr : -T(0);  // :73 unshifted quirk
                const T vd = (r < ny - 1) ? -vu_dn : -T(0);
                // velocity_raster only enters dt (non-adaptive) and the rollback celerity
                const T raster = a.adaptive ? T(0)
                                            : dmax(dmax(dmax(fabs(vl), fabs(vr)), fabs(vu)), fabs(vd));
                const T rc = T(0.5) * (vr - vl), uc = T(0.5) * (vu - vd);
                bool stagnant;
                T vtot = T(0);  // velocities.total_velocity (Human_Instability_Module.m:9)
                if (f64) {
                    const T tot = hypot(rc, uc);
                    vkey = dmax(vkey, tot);
                    stagnant = tot < T(0.01);
                    vtot = tot;
                } else {
                    const T tot2 = rc * rc + uc * uc;  // compared squared, rooted once per step
                    vkey = dmax(vkey, tot2);
                    stagnant = tot2 < T(1e-4);
                    if (F.risk || may_snap) vtot = sqrt(tot2);
                }
                if (F.risk) {
                    // Human_Instability_Module.m:8-25 and its running maximum
                    const T h = mm2m(dn);
                    const T F_person = a.hi_weight * a.hi_g;
                    const T F_buoy = (a.hi_w1 * a.hi_w2) * h * a.hi_ro * a.hi_g;
                    const T avail = a.hi_mu * (dmax(F_person - F_buoy, T(0)));
                    const T hydro = dmax(T(1) / T(2) * (a.hi_Cd * a.hi_ro * a.hi_w1 * h * (vtot * vtot)), T(0));
                    T risk = T(0);
                    if (avail > T(0)) risk = hydro / avail;
                    else if (h > T(0)) risk = T(1);
                    risk = dmin(dmax(risk, T(0)), T(1));
                    if (isnan(dn)) risk = T(NAN);
                    const T prev = (kstep == 1) ? (isnan(dn) ? T(NAN) : T(0)) : a.max_risk[oid];
                    const T nm = dmax(risk, prev);
                    if (kstep == 1 || nm != prev) a.max_risk[oid] = nm;
                }
                if (may_snap) {
                    a.snap_vras[oid] = dmax(dmax(dmax(fabs(vl), fabs(vr)), fabs(vu)), fabs(vd));
                    a.snap_vtot[oid] = vtot;
                }
                if (!a.adaptive) rmax = dmax(rmax, raster);
                if (!stagnant) {
                    if (a.adaptive) {
                        wkey = dmax(wkey, dts);  // min(alfa*dx/c) = alfa*dx/c(max depth)
                    } else {
                        T cw = sqrt(T(9.81) * dmax(dts, T(0)));
                        cw = dmax(cw, T(1e-10));
                        if (!isinf(cw)) wkey = dmin(wkey, a.alfa_min * a.R / (raster + cw));
                    }
                }
                if (may_report && !isnan(dn)) s_store += (double)dn;
                if (a.track_dmax) dnmax = dmax(dnmax, dn);
            }
            if (out_cell) {
                d_dst[oid] = dn;
                qx_dst[oid] = ox;
                qy_dst[oid] = oy;
                if (F.dmax && !fail_step) {
                    const T prevmax = a.dmax_map[oid];
                    const T newmax = (kstep == 1) ? dn : dmax(dn, prevmax);
                    // stored only when it changes (a NaN old value always differs):
                    // a cell that is not rising costs no write
                    if (kstep == 1 || newmax != prevmax) a.dmax_map[oid] = newmax;
                    // time to peak: end time of the step that last raised the maximum
                    if (F.tpeak && (kstep == 1 || dn > prevmax)) a.tpeak_map[oid] = t_end_min;
                }
            }
        } else {
            const T oy_dn = lane_down(oy);
            if (out_cell) {
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
                    a.itot[oid] = fabs(ox + oy + qo) * dt_s / T(1000) * T(1) / T(3600) * a.R2;
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
        // per-wave butterfly reductions, then the waves of the block combined in
        // a fixed order (deterministic); volumes are reduced as sums of depths
        // and scaled by A/1000 by the finalize kernel
        __shared__ double red[kPartCount][kMaxWavesPerBlock];
        const T wv_v = wave_fmax(vkey), wv_r = wave_fmax(rmax);
        const T wv_w = a.adaptive ? wave_fmax(wkey) : wave_fmin(wkey);
        const T wv_dt = wave_fmax(dtmax), wv_dp = wave_fmax(dpmax), wv_dn = wave_fmax(dnmax);
        const double st_sum = wave_sum(s_store), uz_sum = wave_sum(s_uz);
        const double in_sum = wave_sum(s_inf);
        const int any_neg = __any(neg);
        if (lane == 0) {
            red[kPartStorage][wv] = st_sum;
            red[kPartUz][wv] = uz_sum;
            red[kPartInfil][wv] = in_sum;
            red[kPartVmax][wv] = (double)wv_v;
            red[kPartDtw][wv] = (double)wv_w;
            red[kPartRaster][wv] = (double)wv_r;
            red[kPartDtotMax][wv] = (double)wv_dt;
            red[kPartDprevMax][wv] = (double)wv_dp;
            red[kPartNeg][wv] = any_neg ? 1.0 : 0.0;
            red[kPartDnewMax][wv] = (double)wv_dn;
        }
        __syncthreads();
        if (threadIdx.x < kPartCount) {
            const int f = threadIdx.x;
            double acc = red[f][0];
            for (int w = 1; w < a.waves_per_block; ++w) {
                const double v = red[f][w];
                if (f == kPartStorage || f == kPartUz || f == kPartInfil) acc += v;
                else if (f == kPartDtw && !a.adaptive) acc = fmin(acc, v);
                else acc = fmax(acc, v);
            }
            a.partials[(long long)f * a.part_cap + logical] = acc;
        }
    }
}

template <typename T, int MODE, bool CRIT, int FEAT>
__global__ __launch_bounds__(256) void li_step_kernel(LiArgs<T> a) {
    li_step_tile<T, MODE, CRIT, FEAT>(a, xcd_logical_tile());
}

// Many grids in one launch (an ensemble): tile t of the launch belongs to grid
// tile_grid[t] and is that grid's tile t - grid_first_tile[grid].  Every grid
// keeps its own arguments, state, partials and dt; the block size is shared.
template <typename T, bool CRIT>
__global__ __launch_bounds__(256) void li_step_batch_kernel(const LiArgs<T> *__restrict__ args,
                                                            const int *__restrict__ tile_grid,
                                                            const int *__restrict__ grid_first_tile) {
    const int t = xcd_logical_tile();
    const int g = tile_grid[t];
    // read in place: the fields are wave-uniform scalar loads from a read-only
    // table, kept in the scalar cache
    const LiArgs<T> &a = args[g];
    li_step_tile<T, kModeSim, CRIT, kFeatRuntime>(a, t - grid_first_tile[g]);
}

// The feature mask of a run's arguments.
template <typename T> int li_feature_mask(const LiArgs<T> &a) {
    return (a.ga ? kFeatGa : 0) | (a.rain_pattern ? kFeatPattern : 0) |
           (a.rain_maps ? kFeatMaps : 0) | (a.inflow_gauge ? kFeatGauge : 0) |
           (a.max_risk ? kFeatRisk : 0) | (a.snap_vras ? kFeatSnap : 0) |
           (a.dmax_map ? kFeatDmax : 0) | (a.tpeak_map ? kFeatTpeak : 0);
}

// The simulation step: FP32 runs whose optional arrays are one of the common
// sets (rain as a pattern or lumped, with or without Green-Ampt, with the
// maximum-depth map) get a kernel specialised to that set, which keeps the
// arrays it does not have out of its code and its scalar registers; any other
// set runs the kernel that reads the set from the arguments.
template <typename T, bool CRIT>
void launch_li_sim(const LiArgs<T> &a, dim3 grid, dim3 block, hipStream_t s) {
    if constexpr (sizeof(T) == 4) {
        constexpr int kGP = kFeatGa | kFeatPattern | kFeatDmax, kG = kFeatGa | kFeatDmax;
        constexpr int kP = kFeatPattern | kFeatDmax, kD = kFeatDmax;
        switch (li_feature_mask(a)) {
        case kGP: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kGP>), grid, block, 0, s, a); return;
        case kG: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kG>), grid, block, 0, s, a); return;
        case kP: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kP>), grid, block, 0, s, a); return;
        case kD: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kD>), grid, block, 0, s, a); return;
        default: break;
        }
    }
    hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kFeatRuntime>), grid, block, 0, s, a);
}

template <typename T>
hipError_t launch_li_step(const LiArgs<T> &a, int mode, hipStream_t s) {
    dim3 block(kWave * a.waves_per_block);
    dim3 grid(((a.n_slabs + a.waves_per_block - 1) / a.waves_per_block) * ((a.nx + a.tj - 1) / a.tj));
    if (mode == kModeSim) {
        if (a.flag_critical) launch_li_sim<T, true>(a, grid, block, s);
        else launch_li_sim<T, false>(a, grid, block, s);
    } else {
        // the drop-in mode has none of the simulation's optional arrays
        if (a.flag_critical)
            hipLaunchKernelGGL((li_step_kernel<T, kModeDropIn, true, 0>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((li_step_kernel<T, kModeDropIn, false, 0>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_li_step_batch(const LiArgs<T> *args, const int *tile_grid,
                                const int *grid_first_tile, int n_tiles, int waves_per_block,
                                int critical, hipStream_t s) {
    dim3 block(kWave * waves_per_block), grid(n_tiles);
    if (critical)
        hipLaunchKernelGGL((li_step_batch_kernel<T, true>), grid, block, 0, s, args, tile_grid,
                           grid_first_tile);
    else
        hipLaunchKernelGGL((li_step_batch_kernel<T, false>), grid, block, 0, s, args, tile_grid,
                           grid_first_tile);
    return hipGetLastError();
}

template hipError_t launch_li_step_batch<float>(const LiArgs<float> *, const int *, const int *,
                                                int, int, int, hipStream_t);
template hipError_t launch_li_step_batch<double>(const LiArgs<double> *, const int *, const int *,
                                                 int, int, int, hipStream_t);
template hipError_t launch_li_step<float>(const LiArgs<float> &, int, hipStream_t);
template hipError_t launch_li_step<double>(const LiArgs<double> &, int, hipStream_t);

}  // namespace hp2d
