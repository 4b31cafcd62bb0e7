// hp2d_li.hip - fused local-inertial D4 step for gfx950.
//
// One wavefront owns a slab of 64*G grid rows (MATLAB index i, the contiguous
// one; each lane holds G = 1 or 2 consecutive rows) and marches along the
// columns j of its chunk.  The outer two rows on each side are the radius-2
// halo that the donor draining limiter needs (a face flux depends on its donor
// cell's scale, which depends on all four faces of that donor); the others are
// finalised.  Every per-cell array is read once per column with a coalesced
// load (8 bytes per lane in FP32 at G = 2) issued one column ahead of the
// compute; the 2-row halo is re-read from L2, and neighbours across lanes come
// from wavefront-shift DPP moves (no LDS).  Per column the wave computes, with a
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
#include "hp2d_hydro.hpp"
#include "hp2d_internal.hpp"
#include "hp2d_li.hpp"

#include <type_traits>

namespace hp2d {

// ---------------------------------------------------------------------------
// cross-lane moves: lane i receives lane i-1 (up) / lane i+1 (down)
// (DPP wave_shr:1 / wave_shl:1 with bound_ctrl: the end lanes receive 0 and
// are halo lanes)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_up(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v),
                                                              0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_down(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v),
                                                              0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ double lane_up(double v) { return __shfl_up(v, 1); }
__device__ __forceinline__ double lane_down(double v) { return __shfl_down(v, 1); }

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
    const float den = fmadd(g * dt * n2 * fabsf(q_p), __builtin_amdgcn_rcpf(p73), 1.0f);
    return fmadd(-(g * dt) * Hf, S, q_p) * __builtin_amdgcn_rcpf(den);
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
    kFeatVg = 256,      // van Genuchten infiltration (I, soil class, water-table depth)
    kFeatRuntime = -1,  // all read from the arguments
};
template <int FEAT> __device__ __forceinline__ bool feat_on(int bit, bool runtime) {
    if constexpr (FEAT < 0) return runtime;
    else return (FEAT & bit) != 0;
}

// The run's optional inputs and outputs: fixed at compile time in the
// specialised kernels (FEAT >= 0), read from the arguments otherwise.
struct Feats {
    bool ga, pat, maps, gauge, risk, snap, dmax, tpeak, vg;
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
    f.vg = feat_on<FEAT>(kFeatVg, a.vg != nullptr);
    return f;
}

// A lane holds G consecutive grid rows (G = 1 or 2; G = 2 needs an even row
// count, so that the rows of a lane are all inside the grid or all outside and
// their loads are aligned vector loads).  The neighbour above row k of a lane
// is row k-1 of the same lane, or the last row of the lane above (a DPP move);
// the neighbour below likewise.
template <int G, typename T> __device__ __forceinline__ T up_of(const T (&v)[G], int k) {
    return k > 0 ? v[k - 1] : lane_up(v[G - 1]);
}
template <int G, typename T> __device__ __forceinline__ T down_of(const T (&v)[G], int k) {
    return k < G - 1 ? v[k + 1] : lane_down(v[0]);
}

// Address of p[id].  With OFF32 (every array of the launch under 4 GiB) it is
// the array's base plus a 32-bit byte offset: the load then takes the base from
// SGPRs and one shared VGPR offset, with no 64-bit address arithmetic.
template <bool OFF32, typename E> __device__ __forceinline__ E *elem_ptr(E *p, unsigned id) {
    if constexpr (OFF32) {
        using B = typename std::conditional<std::is_const<E>::value, const char, char>::type;
        return reinterpret_cast<E *>(reinterpret_cast<B *>(p) + id * (unsigned)sizeof(E));
    } else {
        return p + id;
    }
}

// The G elements p[id .. id+G-1] (id a multiple of G): one vector access.
template <int G, bool OFF32, typename T>
__device__ __forceinline__ void ld_rows(const T *__restrict__ p, unsigned id, T (&o)[G]) {
    const T *q = elem_ptr<OFF32>(p, id);
    if constexpr (G == 2 && sizeof(T) == 4) {
        const float2 v = *reinterpret_cast<const float2 *>(q);
        o[0] = v.x;
        o[1] = v.y;
    } else if constexpr (G == 2 && sizeof(T) == 8) {
        const double2 v = *reinterpret_cast<const double2 *>(q);
        o[0] = v.x;
        o[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k) o[k] = q[k];
    }
}
template <int G, bool OFF32>
__device__ __forceinline__ void ld_rows_u8(const uint8_t *__restrict__ p, unsigned id, int (&o)[G]) {
    const uint8_t *q = elem_ptr<OFF32>(p, id);
    if constexpr (G == 2) {
        const unsigned v = *reinterpret_cast<const unsigned short *>(q);
        o[0] = (int)(v & 0xffu);
        o[1] = (int)(v >> 8);
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k) o[k] = q[k];
    }
}
template <int G, bool OFF32, typename T>
__device__ __forceinline__ void st_rows(T *__restrict__ p, unsigned id, const T (&v)[G]) {
    T *q = elem_ptr<OFF32>(p, id);
    if constexpr (G == 2 && sizeof(T) == 4) {
        *reinterpret_cast<float2 *>(q) = make_float2(v[0], v[1]);
    } else if constexpr (G == 2 && sizeof(T) == 8) {
        *reinterpret_cast<double2 *>(q) = make_double2(v[0], v[1]);
    } else {
#pragma unroll
        for (int k = 0; k < G; ++k) q[k] = v[k];
    }
}

// Per-column loads of one lane, fetched one column ahead of the compute.
// z: the elevation (FP64), or dz of the cell's x face (FP32); zy: dz of its
// y face (FP32 only)
template <typename T, int G> struct ColLoad {
    T z[G], zy[G], d[G], n2[G], qx[G], qy[G], pat[G], I[G], Tc[G];
    int soil[G], gauge[G];
    T dm[G];  // the depth maximum of the output column two behind (simulation mode)
};

// The rainfall maps' terms of this step (wave-uniform): the offsets of the
// overlapping maps, their overlaps [min], t - t_previous [min] and dt/60 [h].
template <typename T> struct MapStep {
    int n;
    unsigned off[kMaxMapTerms];
    T w[kMaxMapTerms];
    T dtot, k;
};

template <typename T, int MODE, int G, bool OFF32>
__device__ __forceinline__ ColLoad<T, G> load_column(const LiArgs<T> &a, const Feats &F,
                                                     const T *d_src, const T *qx_src,
                                                     const T *qy_src, const T *I_src,
                                                     const T *Tc_src, bool in, unsigned id,
                                                     const MapStep<T> &ms, bool dm_in = false) {
    ColLoad<T, G> v;
#pragma unroll
    for (int k = 0; k < G; ++k) {
        v.z[k] = v.zy[k] = v.d[k] = v.n2[k] = v.qx[k] = v.qy[k] = v.I[k] = v.Tc[k] = T(0);
        v.pat[k] = T(1);
        v.soil[k] = v.gauge[k] = 0;
        v.dm[k] = T(0);
    }
    // the depth maximum of column c - 1 when this loads column c + 1: the
    // output column of the step that computes c + 1
    if (MODE == kModeSim && dm_in) ld_rows<G, OFF32>(a.dmax_map, id - 2u * (unsigned)a.ny, v.dm);
    if (in) {
        if constexpr (sizeof(T) == 8) {
            ld_rows<G, OFF32>(a.z, id, v.z);
        } else {
            ld_rows<G, OFF32>(a.dzx, id, v.z);
            ld_rows<G, OFF32>(a.dzy, id, v.zy);
        }
        ld_rows<G, OFF32>(d_src, id, v.d);
        ld_rows<G, OFF32>(a.n2s, id, v.n2);
        ld_rows<G, OFF32>(qx_src, id, v.qx);
        ld_rows<G, OFF32>(qy_src, id, v.qy);
        if (MODE == kModeSim) {
            if (F.maps) {
                // the step's rain depth from the maps (update_spatial_BC.m:1361-1408):
                // (sum_i overlap_i * map_i) / (t - t_prev) * dt / 60
                T acc[G];
#pragma unroll
                for (int k = 0; k < G; ++k) acc[k] = T(0);
                for (int j = 0; j < ms.n; ++j) {
                    T m[G];
                    ld_rows<G, false>(a.rain_maps + ms.off[j], id, m);
#pragma unroll
                    for (int k = 0; k < G; ++k) acc[k] = acc[k] + ms.w[j] * m[k];
                }
#pragma unroll
                for (int k = 0; k < G; ++k) v.pat[k] = ms.n > 0 ? acc[k] / ms.dtot * ms.k : T(0);
            } else if (F.pat) {
                ld_rows<G, OFF32>(a.rain_pattern, id, v.pat);
            }
            if (F.gauge) ld_rows_u8<G, OFF32>(a.inflow_gauge, id, v.gauge);
            if (F.ga || F.vg) {
                ld_rows<G, OFF32>(I_src, id, v.I);
                // Green-Ampt: the recovery clock; van Genuchten: the water-table depth
                ld_rows<G, OFF32>(F.ga ? Tc_src : a.vg_zwt, id, v.Tc);
                if (a.soil_class) ld_rows_u8<G, OFF32>(a.soil_class, id, v.soil);
            }
        }
    }
    return v;
}

// One workgroup's tile of one grid: `logical` counts the grid's tiles, row
// blocks fastest.  A wave's slab is kWave*G rows, of which the outer kHalo on
// each side are halo.  OFF32: every per-cell array is under 4 GiB, so element
// addresses are 32-bit byte offsets from the array bases.
template <typename T, int MODE, bool CRIT, int FEAT, int G, bool OFF32>
__device__ __forceinline__ void li_step_tile(const LiArgs<T> &a, const int logical) {
    static_assert(kHalo % G == 0, "the rows of a lane are all halo or all owned");
    constexpr bool f64 = sizeof(T) == 8;
    constexpr int kOut = kWave * G - 2 * kHalo;  // rows finalised per slab
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
    if (MODE == kModeSim && !f64 && (F.ga || F.vg)) {
        const float4 *tab = reinterpret_cast<const float4 *>(F.ga ? a.ga_f32 : a.vg_f32);
        for (int i = threadIdx.x; i < 2 * a.n_soil_classes; i += blockDim.x) ga_lds[i] = tab[i];
        __syncthreads();
    }

    const T A = a.cell_area;
    // rows r0 .. r0+G-1 of the lane; the row count is a multiple of G, so they
    // are all inside the grid or all outside
    const int r0 = slab * kOut - kHalo + lane * G;
    const bool row_in = (r0 >= 0) && (r0 < ny);
    const bool row_own = row_in && lane * G >= kHalo && lane * G < kWave * G - kHalo;
    const int c0 = by * a.tj;
    const int c1 = min(c0 + a.tj, nx);
    const T avail_den = dmax(dt_s, eps_of<T>());
    const T inv_dt_h = frcp(dmax(dt_h, T(1e-12)));  // FP32 path only
    const T k_avail = A * frcp(avail_den);            // FP32 path only
    const T k_cont = T(1000) * frcp(a.C_a);           // FP32 path only

    // pipeline registers (suffix = column lag behind the load column c)
    T h1[G], z1[G], n21[G], qxp1[G], d1[G], n2s1[G];   // column c-1
    T rawQx2[G], rawQy1[G], rawQy2[G], Hfx2[G], Hfy1[G], Hfy2[G];
    T d2[G], scale2[G], finQx3[G], ox3[G], n2s2[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        h1[k] = z1[k] = n21[k] = qxp1[k] = d1[k] = n2s1[k] = T(0);
        rawQx2[k] = rawQy1[k] = rawQy2[k] = Hfx2[k] = Hfy1[k] = Hfy2[k] = T(0);
        d2[k] = scale2[k] = finQx3[k] = ox3[k] = n2s2[k] = T(0);
    }

    // per-lane reductions
    T vkey = T(NAN), wkey = T(NAN), rmax = T(NAN), dtmax = T(NAN), dpmax = T(NAN), dnmax = T(NAN);
    double s_store = 0.0, s_uz = 0.0, s_inf = 0.0;
    int neg = 0;

    // 32-bit element offsets (grids are limited to 2^32 cells at creation).
    // The loads run PF columns ahead of the compute (two at one row per lane:
    // a wave then keeps two columns of loads in flight), rotating through
    // PF + 1 buffers.
    constexpr int PF = G == 1 ? 2 : 1;
    ColLoad<T, G> buf0 = load_column<T, MODE, G, OFF32>(
        a, F, d_src, qx_src, qy_src, I_src, Tc_src, active && row_in && c0 - 2 >= 0 && c0 - 2 < nx,
        (unsigned)r0 + (unsigned)(c0 - 2) * (unsigned)ny, ms);
    ColLoad<T, G> buf1, buf2;
    if constexpr (PF == 2)
        buf1 = load_column<T, MODE, G, OFF32>(
            a, F, d_src, qx_src, qy_src, I_src, Tc_src, active && row_in && c0 - 1 >= 0 && c0 - 1 < nx,
            (unsigned)r0 + (unsigned)(c0 - 1) * (unsigned)ny, ms);

    // One column of the march: compute column c from `cur` while the loads of
    // column c + PF go into `nxt`; output column o = c - 2.  STAGE 0 is the
    // first column of the march (c0 - 2): only its depth after the rain and
    // infiltration enters the pipeline; STAGE 1 (c0 - 1) adds the raw faces
    // that the donor scale of c0 - 1 needs; STAGE 2 is the full column.
    auto column = [&](const int c, auto stage, const ColLoad<T, G> &cur, ColLoad<T, G> &nxt) {
        constexpr int STAGE = decltype(stage)::value;
        const bool col_in = (c >= 0) && (c < nx);
        const bool cell_in = row_in && col_in;
        const unsigned id = (unsigned)r0 + (unsigned)c * (unsigned)ny;
        // issue the next column's loads before computing this one
        if (c + PF <= c1 + 1)
            nxt = load_column<T, MODE, G, OFF32>(a, F, d_src, qx_src, qy_src, I_src, Tc_src,
                                                 row_in && c + PF >= 0 && c + PF < nx,
                                                 id + (unsigned)(PF * ny), ms,
                                                 F.dmax && !fail_step && row_own && c + PF - 2 >= c0 &&
                                                     c + PF - 2 < c1);
        T dc[G], hc[G];
        // ---- rainfall + Green-Ampt prologue: Hydrological_Model.m:15-55 ----
        {
            const bool own = row_own && c >= c0 && c < c1;
            T In_k[G], Tn_k[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                dc[k] = cur.d[k];
                // cells outside the domain: NaN elevation (FP64) or NaN n^2 (FP32)
                const bool nan_cell = f64 ? isnan(cur.z[k]) : isnan(cur.n2[k]);
                In_k[k] = Tn_k[k] = T(0);
                if (MODE == kModeSim && cell_in) {
                    const T dprev = dc[k];
                    T eff = F.maps ? cur.pat[k] : (F.pat ? cur.pat[k] * rain : rain);
                    if (nan_cell) eff = T(NAN);
                    T dd = dc[k] + eff;
                    if (F.ga) {
                        // legacy explicit Green-Ampt closure in the Infiltration_Module slot
                        // (HydroPol2D_Main_While_Ensemble.m:243-298)
                        T ksat, psi, dth, Lu_mm, k_out, Tr;
                        if (f64) {
                            const GaClass g = a.ga[cur.soil[k]];
                            ksat = T(g.ksat); psi = T(g.psi); dth = T(g.dtheta);
                            Lu_mm = T(g.Lu_mm); k_out = T(g.k_out); Tr = T(g.Tr);
                        } else {
                            const float4 g0 = ga_lds[2 * cur.soil[k]];
                            const float4 g1 = ga_lds[2 * cur.soil[k] + 1];
                            ksat = g0.x; psi = g0.y; dth = g0.z; Lu_mm = g0.w; k_out = g1.x; Tr = g1.y;
                        }
                        T Tn = cur.Tc[k], In;
                        T f = ga_infiltrate<T>(ksat, psi, dth, Lu_mm, k_out, Tr, a.I_min, kstep == 1,
                                               dt_h, inv_dt_h, dd, cur.I[k], Tn, In);
                        if (nan_cell) {
                            In = T(NAN);
                            f = T(NAN);
                        }
                        dd = f64 ? dd - f * dt_h : fmadd(-f, dt_h, dd);
                        In_k[k] = In;
                        Tn_k[k] = Tn;
                        if (own && !fail_step) {
                            if (!isnan(f)) s_inf += (double)(f * dt_h);
                            if (may_report && !isnan(In)) s_uz += (double)In;
                        }
                    }
                    if (F.vg) {
                        T f = T(0), In = cur.I[k];
                        if (!nan_cell) {
                            f = vg_infiltrate<T>(a, ga_lds, cur.soil[k], cur.Tc[k], dt_h, inv_dt_h,
                                                 dd, In);
                        } else {
                            In = T(NAN);
                            f = T(NAN);
                        }
                        In_k[k] = In;
                        if (own && !fail_step) {
                            if (!isnan(f)) s_inf += (double)(f * dt_h);
                            if (may_report && !isnan(In)) s_uz += (double)In;
                        }
                    }
                    // Hydrological_Model.m:30: the inflow hydrographs on their gauge cells
                    if (F.gauge && cur.gauge[k] > 0) dd = dd + T(a.st->inflow_agg[cur.gauge[k] - 1]);
                    if (dd < T(-1e-8)) neg = 1;
                    if (dd < T(1e-6)) dd = T(0);  // Hydrological_Model.m:49
                    if (dd < T(0)) dd = T(0);     // :52
                    dc[k] = dd;
                    if (own && (fail_step || a.track_dmax)) {  // rollback celerity inputs (HydroPol2D_Main_While.m:932)
                        dtmax = dmax(dtmax, dc[k]);
                        dpmax = dmax(dpmax, dprev);
                    }
                }
                hc[k] = dmax(mm2m(dc[k]), T(0));  // depth_cell (:31)
            }
            if (MODE == kModeSim && (F.ga || F.vg) && cell_in && own) {
                st_rows<G, OFF32>(I_dst, id, In_k);
                if (F.ga) st_rows<G, OFF32>(Tc_dst, id, Tn_k);
            }
        }

        // ---- raw y-face flux at (r,c): face between rows r-1 and r ----
        T rawQyc[G], Hfyc[G];
#pragma unroll
        for (int k = 0; k < G; ++k) rawQyc[k] = Hfyc[k] = T(0);
        if constexpr (STAGE >= 1) {
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const T h_up = up_of<G>(hc, k);
                const T z_up = f64 ? up_of<G>(cur.z, k) : cur.zy[k];  // FP32: dz = z(r-1) - z(r)
                if (cell_in && r0 + k >= 1)
                    rawQyc[k] = face_flux<T, CRIT>(cur.qy[k], cur.z[k], z_up, hc[k], h_up,
                                                   fabs(cur.n2[k]), dt_s, a, Hfyc[k]);
            }
        }

        // ---- raw x-face flux at (r,c-1): face between columns c-1 and c ----
        T rawQx1[G], Hfx1[G];
#pragma unroll
        for (int k = 0; k < G; ++k) rawQx1[k] = Hfx1[k] = T(0);
        if constexpr (STAGE >= 1) {
#pragma unroll
            for (int k = 0; k < G; ++k) {
                // (FP32: z1 is dz = z(c) - z(c-1), stored with column c-1)
                if (row_in && c - 1 >= 0 && c - 1 < nx - 1)
                    rawQx1[k] = face_flux<T, CRIT>(qxp1[k], z1[k], f64 ? cur.z[k] : z1[k], h1[k],
                                                   hc[k], fabs(n21[k]), dt_s, a, Hfx1[k]);
            }
        }

        // ---- donor scale of cell (r,c-1): Local_Inertial_Model_D4.m:814-829 ----
        T s1[G];
#pragma unroll
        for (int k = 0; k < G; ++k) s1[k] = T(0);
        if constexpr (STAGE >= 2) {
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const T qy1_dn = down_of<G>(rawQy1, k);
                T outr = T(0) + dmax(rawQx1[k], T(0));
                if (c - 1 >= 1) outr += dmax(-rawQx2[k], T(0));
                outr += dmax(rawQy1[k], T(0));
                if (r0 + k < ny - 1) outr += dmax(-qy1_dn, T(0));
                const T avail = f64 ? dmax(d1[k] / T(1000), T(0)) * A / avail_den
                                    : dmax(d1[k] * T(1e-3), T(0)) * k_avail;
                T sc = T(1);
                if (outr > avail && outr > T(0)) sc = f64 ? avail / outr : avail * frcp(outr);
                if (!isfinite(sc)) sc = T(0);
                s1[k] = dmax(dmin(sc, T(1)), T(0));
            }
        }

        T fqx[G], fqy[G], ox[G], oy[G];
#pragma unroll
        for (int k = 0; k < G; ++k) fqx[k] = fqy[k] = ox[k] = oy[k] = T(0);
        if constexpr (STAGE >= 2) {
            // ---- limited fluxes of output column o = c-2 (:831-857) ----
            const int o = c - 2;
#pragma unroll
            for (int k = 0; k < G; ++k) {
                // donor-scaled with a select of the multiplier (a zero flux is never
                // scaled, as the masked assignments of :834-851 leave it untouched)
                fqx[k] = rawQx2[k] * ((rawQx2[k] > T(0)) ? scale2[k] : ((rawQx2[k] < T(0)) ? s1[k] : T(1)));
                if (o == nx - 1) fqx[k] = T(0);
                if (!isfinite(fqx[k])) fqx[k] = T(0);
                const T s2_up = up_of<G>(scale2, k);
                fqy[k] = rawQy2[k] * ((rawQy2[k] > T(0)) ? scale2[k] : ((rawQy2[k] < T(0)) ? s2_up : T(1)));
                if (r0 + k == 0) fqy[k] = T(0);
                if (!isfinite(fqy[k])) fqy[k] = T(0);
            }

            // ---- continuity (:313-315, :390, :415) ----
            T dn[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const T fqy_dn = down_of<G>(fqy, k);
                const T west = (o >= 1) ? finQx3[k] : T(0);
                const T south = (r0 + k < ny - 1) ? fqy_dn : T(0);
                const T vol = dt_s * (west - fqx[k] - fqy[k] + south);
                dn[k] = f64 ? d2[k] + T(1000) * vol / a.C_a : fmadd(vol, k_cont, d2[k]);
                if (dn[k] < T(0)) dn[k] = T(0);
            }

            const bool out_cell = row_own && o >= c0 && o < c1;
            const unsigned oid = (unsigned)r0 + (unsigned)o * (unsigned)ny;
            // ---- outlet (:418-498) ----
            T hf3[G], qo[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                hf3[k] = qo[k] = T(0);
                if (out_cell && signbit(n2s2[k]) && a.n_outlet > 0) {
                    const int slot = find_outlet(a.out_idx, a.n_outlet, (long long)(oid + k));
                    if (slot >= 0) {
                        const T h_out = dmax(dn[k], T(0)) / T(1000);
                        T sqrtS;
                        if (a.outlet_type == 1) {
                            sqrtS = a.sqrtS_type1;
                        } else {
                            const T h_safe = dmax(h_out, T(1e-12));
                            sqrtS = sqrt(T(9.81) * fabs(n2s2[k]) * powm13(h_safe));
                            if (!isfinite(sqrtS)) sqrtS = T(0);
                        }
                        hf3[k] = h_out;
                        T q = (T(1) / a.out_n[slot]) * a.R * pow53(h_out) * sqrtS / A * T(1000) * T(3600);
                        const T qmax = dmax(dn[k] - T(1000) * T(0), T(0)) / dt_h;
                        q = dmin(q, qmax);
                        if (!isfinite(q)) q = T(0);
                        dn[k] = dn[k] - q * dt_h;
                        qo[k] = q;
                        if (a.out_q) a.out_q[slot] = q;
                        if (a.out_d) a.out_d[slot] = dn[k];
                    }
                }
                ox[k] = q_to_mmh(fqx[k], a);
                oy[k] = q_to_mmh(fqy[k], a);
            }

            if (MODE == kModeSim) {
                // ---- refreshing_timestep.m:63-238 per-cell terms ----
                const T tol = a.depth_tol_mm;
                T vu[G];
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    if (f64) {
                        const T f2 = Hfy2[k] * T(1000);
                        const T s2d = (f2 < tol) ? T(1e12) : f2;
                        vu[k] = (oy[k] / T(1000) / T(3600)) * a.R / (s2d / T(1000));
                    } else {
                        const T hy = (Hfy2[k] * T(1000) < tol) ? T(1e9) : Hfy2[k];
                        vu[k] = fqy[k] * a.inv_R * frcp(hy);
                    }
                }
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const T vu_dn = down_of<G>(vu, k);
                    if (out_cell && !fail_step) {
                        T vr, dts;
                        if (f64) {
                            const T f1 = Hfx2[k] * T(1000), f2 = Hfy2[k] * T(1000), f3 = hf3[k] * T(1000);
                            dts = dmax(dmax(f1 / T(1000), f2 / T(1000)), f3 / T(1000));
                            const T s1d = (f1 < tol) ? T(1e12) : f1;
                            vr = (ox[k] / T(1000) / T(3600)) * a.R / (s1d / T(1000));
                        } else {
                            dts = dmax(dmax(Hfx2[k], Hfy2[k]), hf3[k]);
                            const T hx = (Hfx2[k] * T(1000) < tol) ? T(1e9) : Hfx2[k];
                            vr = fqx[k] * a.inv_R * frcp(hx);
                        }
                        const T vl = (o >= 1) ? -vr : -T(0);  // :73 unshifted quirk
                        const T vd = (r0 + k < ny - 1) ? -vu_dn : -T(0);
                        // velocity_raster only enters dt (non-adaptive) and the rollback celerity
                        const T raster = a.adaptive ? T(0)
                                                    : dmax(dmax(dmax(fabs(vl), fabs(vr)), fabs(vu[k])), fabs(vd));
                        const T rc = T(0.5) * (vr - vl), uc = T(0.5) * (vu[k] - vd);
                        bool stagnant;
                        T vtot = T(0);  // velocities.total_velocity (Human_Instability_Module.m:9)
                        if (f64) {
                            const T tot = hypot(rc, uc);
                            vkey = dmax(vkey, tot);
                            stagnant = tot < T(0.01);
                            vtot = tot;
                        } else {
                            const T tot2 = fmadd(rc, rc, uc * uc);  // compared squared, rooted once per step
                            vkey = dmax(vkey, tot2);
                            stagnant = tot2 < T(1e-4);
                            if (F.risk || may_snap) vtot = sqrt(tot2);
                        }
                        if (F.risk) {
                            // Human_Instability_Module.m:8-25 and its running maximum
                            const T h = mm2m(dn[k]);
                            const T F_person = a.hi_weight * a.hi_g;
                            const T F_buoy = (a.hi_w1 * a.hi_w2) * h * a.hi_ro * a.hi_g;
                            const T avail = a.hi_mu * (dmax(F_person - F_buoy, T(0)));
                            const T hydro = dmax(T(1) / T(2) * (a.hi_Cd * a.hi_ro * a.hi_w1 * h * (vtot * vtot)), T(0));
                            T risk = T(0);
                            if (avail > T(0)) risk = hydro / avail;
                            else if (h > T(0)) risk = T(1);
                            risk = dmin(dmax(risk, T(0)), T(1));
                            if (isnan(dn[k])) risk = T(NAN);
                            const T prev = (kstep == 1) ? (isnan(dn[k]) ? T(NAN) : T(0)) : a.max_risk[oid + k];
                            const T nm = dmax(risk, prev);
                            if (kstep == 1 || nm != prev) a.max_risk[oid + k] = nm;
                        }
                        if (may_snap) {
                            a.snap_vras[oid + k] = dmax(dmax(dmax(fabs(vl), fabs(vr)), fabs(vu[k])), fabs(vd));
                            a.snap_vtot[oid + k] = vtot;
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
                        if (may_report && !isnan(dn[k])) s_store += (double)dn[k];
                        if (a.track_dmax) dnmax = dmax(dnmax, dn[k]);
                    }
                }
                if (out_cell) {
                    st_rows<G, OFF32>(d_dst, oid, dn);
                    st_rows<G, OFF32>(qx_dst, oid, ox);
                    st_rows<G, OFF32>(qy_dst, oid, oy);
                    if (F.dmax && !fail_step) {
                        const T(&prevmax)[G] = cur.dm;  // fetched with the column two ahead
#pragma unroll
                        for (int k = 0; k < G; ++k) {
                            const T newmax = (kstep == 1) ? dn[k] : dmax(dn[k], prevmax[k]);
                            // stored only when it changes (a NaN old value always differs):
                            // a cell that is not rising costs no write
                            if (kstep == 1 || newmax != prevmax[k]) *elem_ptr<OFF32>(a.dmax_map, oid + k) = newmax;
                            // time to peak: end time of the step that last raised the maximum
                            if (F.tpeak && (kstep == 1 || dn[k] > prevmax[k])) a.tpeak_map[oid + k] = t_end_min;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const T oy_dn = down_of<G>(oy, k);
                    if (out_cell) {
                        const unsigned q = oid + (unsigned)k;
                        d_dst[q] = dn[k];
                        if (qx_dst) qx_dst[q] = ox[k];
                        if (qy_dst) qy_dst[q] = oy[k];
                        if (a.q3) a.q3[q] = qo[k];
                        if (a.hfx) a.hfx[q] = Hfx2[k];
                        if (a.hfy) a.hfy[q] = Hfy2[k];
                        if (a.hf3) a.hf3[q] = hf3[k];
                        if (a.outlet_flow) a.outlet_flow[q] = qo[k];
                        if (a.qright) a.qright[q] = ox[k];
                        if (a.qup) a.qup[q] = oy[k];
                        if (a.qleft) a.qleft[q] = -((o >= 1) ? ox3[k] : T(0));
                        if (a.qdown) a.qdown[q] = -((r0 + k < ny - 1) ? oy_dn : T(0));
                        if (a.itot)
                            a.itot[q] = fabs(ox[k] + oy[k] + qo[k]) * dt_s / T(1000) * T(1) / T(3600) * a.R2;
                    }
                }
            }
        }

        // ---- shift the pipeline ----
#pragma unroll
        for (int k = 0; k < G; ++k) {
            finQx3[k] = fqx[k];
            ox3[k] = ox[k];
            rawQx2[k] = rawQx1[k];
            rawQy2[k] = rawQy1[k];
            rawQy1[k] = rawQyc[k];
            Hfx2[k] = Hfx1[k];
            Hfy2[k] = Hfy1[k];
            Hfy1[k] = Hfyc[k];
            scale2[k] = s1[k];
            d2[k] = d1[k];
            d1[k] = dc[k];
            h1[k] = hc[k];
            z1[k] = cur.z[k];
            n21[k] = cur.n2[k];
            n2s2[k] = n2s1[k];
            n2s1[k] = cur.n2[k];
            qxp1[k] = cur.qx[k];
        }
    };

    if (active) {
        using S0 = std::integral_constant<int, 0>;
        using S1 = std::integral_constant<int, 1>;
        using S2 = std::integral_constant<int, 2>;
        int c = c0;
        if constexpr (PF == 2) {
            // three columns per trip, so that the buffers and the pipeline
            // registers rotate roles instead of being copied
            column(c0 - 2, S0{}, buf0, buf2);
            column(c0 - 1, S1{}, buf1, buf0);
            for (; c + 2 <= c1 + 1; c += 3) {
                column(c, S2{}, buf2, buf1);
                column(c + 1, S2{}, buf0, buf2);
                column(c + 2, S2{}, buf1, buf0);
            }
            if (c <= c1 + 1) column(c, S2{}, buf2, buf1);
            if (c + 1 <= c1 + 1) column(c + 1, S2{}, buf0, buf2);
        } else {
            // at two rows per lane the registers of a rotation would lower
            // the occupancy
            column(c0 - 2, S0{}, buf0, buf1);
            column(c0 - 1, S1{}, buf1, buf0);
            for (; c <= c1 + 1; ++c) {
                column(c, S2{}, buf0, buf1);
                buf0 = buf1;
            }
        }
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
            red[kPartCaItot][wv] = 0.0;
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

// The simulation kernels specialised to a feature set (FEAT >= 0) address
// the per-cell arrays with 32-bit byte offsets: their launcher takes them only
// for arrays under 4 GiB.
template <typename T, int MODE, bool CRIT, int FEAT, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((G == 1 && sizeof(T) == 4 && FEAT >= 0) ? 5 : 1))) void li_step_kernel(LiArgs<T> a) {
    constexpr bool kOff32 = MODE == kModeSim && FEAT >= 0;
    li_step_tile<T, MODE, CRIT, FEAT, G, kOff32>(a, xcd_logical_tile());
}

// Many grids in one launch (an ensemble): tile t of the launch belongs to grid
// tile_grid[t] and is that grid's tile t - grid_first_tile[grid].  Every grid
// keeps its own arguments, state, partials and dt; the block size is shared,
// and a lane holds one row.
template <typename T, bool CRIT>
__global__ __launch_bounds__(256) void li_step_batch_kernel(const LiArgs<T> *__restrict__ args,
                                                            const int *__restrict__ tile_grid,
                                                            const int *__restrict__ grid_first_tile) {
    const int t = xcd_logical_tile();
    const int g = tile_grid[t];
    // read in place: the fields are wave-uniform scalar loads from a read-only
    // table, kept in the scalar cache
    const LiArgs<T> &a = args[g];
    li_step_tile<T, kModeSim, CRIT, kFeatRuntime, 1, false>(a, t - grid_first_tile[g]);
}

// The feature mask of a run's arguments.
template <typename T> int li_feature_mask(const LiArgs<T> &a) {
    return (a.ga ? kFeatGa : 0) | (a.rain_pattern ? kFeatPattern : 0) |
           (a.rain_maps ? kFeatMaps : 0) | (a.inflow_gauge ? kFeatGauge : 0) |
           (a.max_risk ? kFeatRisk : 0) | (a.snap_vras ? kFeatSnap : 0) |
           (a.dmax_map ? kFeatDmax : 0) | (a.tpeak_map ? kFeatTpeak : 0) |
           (a.vg ? kFeatVg : 0);
}

// The simulation step: FP32 runs whose optional arrays are one of the common
// sets (rain as a pattern or lumped, with or without Green-Ampt, with the
// maximum-depth map) get a kernel specialised to that set, which keeps the
// arrays it does not have out of its code and its scalar registers; any other
// set runs the kernel that reads the set from the arguments.
template <typename T, bool CRIT, int G>
void launch_li_sim(const LiArgs<T> &a, dim3 grid, dim3 block, hipStream_t s) {
    // (the specialised kernels address with 32-bit byte offsets)
    const bool off32 = (unsigned long long)a.ny * (unsigned long long)a.nx * sizeof(T) < (1ull << 32);
    if constexpr (sizeof(T) == 4) {
        constexpr int kGP = kFeatGa | kFeatPattern | kFeatDmax, kG = kFeatGa | kFeatDmax;
        constexpr int kP = kFeatPattern | kFeatDmax, kD = kFeatDmax;
        constexpr int kVP = kFeatVg | kFeatPattern | kFeatDmax, kV = kFeatVg | kFeatDmax;
        switch (off32 ? li_feature_mask(a) : -1) {
        case kGP: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kGP, G>), grid, block, 0, s, a); return;
        case kVP: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kVP, G>), grid, block, 0, s, a); return;
        case kV: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kV, G>), grid, block, 0, s, a); return;
        case kG: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kG, G>), grid, block, 0, s, a); return;
        case kP: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kP, G>), grid, block, 0, s, a); return;
        case kD: hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kD, G>), grid, block, 0, s, a); return;
        default: break;
        }
    }
    hipLaunchKernelGGL((li_step_kernel<T, kModeSim, CRIT, kFeatRuntime, G>), grid, block, 0, s, a);
}

template <typename T, int G>
void launch_li_rows(const LiArgs<T> &a, int mode, dim3 grid, dim3 block, hipStream_t s) {
    if (mode == kModeSim) {
        if (a.flag_critical) launch_li_sim<T, true, G>(a, grid, block, s);
        else launch_li_sim<T, false, G>(a, grid, block, s);
    } else {
        // the drop-in mode has none of the simulation's optional arrays
        if (a.flag_critical)
            hipLaunchKernelGGL((li_step_kernel<T, kModeDropIn, true, 0, G>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((li_step_kernel<T, kModeDropIn, false, 0, G>), grid, block, 0, s, a);
    }
}

template <typename T>
hipError_t launch_li_step(const LiArgs<T> &a, int mode, hipStream_t s) {
    if (!li_rows_ok(a.ny, a.rows) || a.n_slabs != li_n_slabs(a.ny, a.rows)) return hipErrorInvalidValue;
    dim3 block(kWave * a.waves_per_block);
    dim3 grid(((a.n_slabs + a.waves_per_block - 1) / a.waves_per_block) * ((a.nx + a.tj - 1) / a.tj));
    if (a.rows == 2) launch_li_rows<T, 2>(a, mode, grid, block, s);
    else launch_li_rows<T, 1>(a, mode, grid, block, s);
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
