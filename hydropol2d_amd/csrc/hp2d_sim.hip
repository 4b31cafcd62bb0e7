// hp2d_sim.hip - device-side time control of the HydroPol2D step loop.
//
// One 256-thread block per step.  It reduces the per-wave partials of the
// fused step in a fixed order (deterministic), then runs, in FP64 on one
// lane, the scalar tail of one iteration of HydroPol2D_Main_While.m:403-994:
//   refreshing_timestep bucket + Courant dt   (:693-705, refreshing_timestep.m:17-355)
//   first report check                         (:711-715)
//   update_spatial_BC rainfall integral        (:721, update_spatial_BC.m:422-448)
//   t_previous / t_save                        (:734-740)
//   save_output_maps hydrograph, runoff        (save_output_maps.m:266-269,409-411,481-490,611-617)
//   t += dt, k += 1, report check + mass_balance_check (:779-803, mass_balance_check.m:6-73)
// or, for a rejected step, the catch block (:926-993).
#include "hp2d_forcing.hpp"
#include "hp2d_internal.hpp"
#include "hp2d_sim.hpp"

namespace hp2d {

constexpr int kFinThreads = 256;
constexpr int kFinWaves = kFinThreads / kWave;
constexpr int kRainWindow = 256;  // rainfall-series entries staged in LDS per step

// The rainfall series with entries [w0, w0 + wn) (and the first two, for the
// spacing) staged in LDS; the serial lookup reaches HBM only past the window.
struct SeriesWindow {
    const double *t, *v;        // the whole series in device memory
    const double *lt, *lv;      // the window in LDS
    const double *head;         // t[0], t[1] in LDS
    int w0, wn;
    __device__ double time(int k) const {
        if (k - w0 >= 0 && k - w0 < wn) return lt[k - w0];
        if (k < 2) return head[k];
        return t[k];
    }
    __device__ double value(int k) const { return (k - w0 >= 0 && k - w0 < wn) ? lv[k - w0] : v[k]; }
};

// sum over the valid cells of the rainfall maps' depth of the step whose
// terms the state holds [mm x cells]
template <typename T>
__device__ inline double map_depth_sum(const FinArgs<T> &a, const SimState *st) {
    if (st->map_n == 0) return 0.0;
    double acc = 0.0;
    for (int j = 0; j < st->map_n; ++j) acc = acc + st->map_w[j] * a.map_sum[st->map_idx[j]];
    return acc / st->map_dtot * st->map_k;
}

// sum over the valid cells of the inflow depths the state holds [mm x cells]
template <typename T>
__device__ inline double inflow_depth_sum(const FinArgs<T> &a, const SimState *st) {
    double acc = 0.0;
    for (int g = 0; g < a.n_gauges; ++g) acc += st->inflow_agg[g] * a.gauge_cells[g];
    return acc;
}

// update_spatial_BC's rainfall maps and inflow hydrographs for (t0, t1]
template <typename T>
__device__ inline void update_maps_inflow(const FinArgs<T> &a, SimState *st, double t0, double t1) {
    if (a.n_maps > 0) {
        MapTerms mt;
        int cur = st->map_cursor;
        if (!map_terms(a.map_t, a.n_maps, t0, t1, &cur, &mt)) st->map_err = 1;
        st->map_cursor = cur;
        st->map_n = mt.n;
        for (int j = 0; j < kMaxMapTerms; ++j) {
            st->map_idx[j] = j < mt.n ? mt.idx[j] : 0;
            st->map_w[j] = j < mt.n ? mt.w[j] : 0.0;
        }
        st->map_dtot = mt.dt_total;
        st->map_k = mt.k;
    }
    if (a.n_gauges > 0) {
        int cur = st->inflow_cursor;
        inflow_depths(a.inflow_t, a.inflow_v, a.n_inflow_times, a.n_gauges, t0, t1, &cur,
                      st->inflow_agg);
        st->inflow_cursor = cur;
    }
}

template <typename T>
__device__ __forceinline__ void finalize_body(const FinArgs<T> &a) {
    // The scalar tail below is one thread's chain of dependent reads and
    // writes of the run state.  It works on a copy in LDS, loaded and stored
    // back in one pass, so the chain pays LDS latency instead of L2 latency.
    __shared__ __attribute__((aligned(16))) SimState sst;
    {
        const unsigned *src = reinterpret_cast<const unsigned *>(a.st);
        unsigned *dst = reinterpret_cast<unsigned *>(&sst);
        for (int i = threadIdx.x; i < (int)(sizeof(SimState) / 4); i += kFinThreads) dst[i] = src[i];
    }
    __syncthreads();
    SimState *st = &sst;
    if (st->done || st->paused || st->hold) return;
    // the rainfall series around the cursor, for the serial lookup below (the
    // block barrier after the reductions orders these writes before it)
    __shared__ double rain_lt[kRainWindow], rain_lv[kRainWindow], rain_head[2];
    SeriesWindow ser{a.rain_t, a.rain_v, rain_lt, rain_lv, rain_head, 0, 0};
    if (a.n_rain > 0) {
        ser.w0 = st->cursor - 2 > 0 ? st->cursor - 2 : 0;
        ser.wn = a.n_rain - ser.w0 < kRainWindow ? a.n_rain - ser.w0 : kRainWindow;
        for (int i = threadIdx.x; i < ser.wn; i += kFinThreads) {
            rain_lt[i] = a.rain_t[ser.w0 + i];
            rain_lv[i] = a.rain_v[ser.w0 + i];
        }
        if (threadIdx.x < 2 && threadIdx.x < a.n_rain) rain_head[threadIdx.x] = a.rain_t[threadIdx.x];
    }
    __shared__ double sh[kFinWaves][15];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    const double *P = a.partials;
    const long long cap = a.part_cap;
    double storage = 0, uz = 0, rain = 0, infil = 0, vmax = NAN, dtw = NAN, raster = NAN;
    double dtm = NAN, dpm = NAN, dnm = NAN, neg = 0, qsum = 0, dsum = 0, qraw = 0, cai = 0;
#pragma unroll 4
    for (int i = t; i < a.n_partials; i += kFinThreads) {
        storage += P[kPartStorage * cap + i];
        uz += P[kPartUz * cap + i];
        infil += P[kPartInfil * cap + i];
        vmax = fmax(vmax, P[kPartVmax * cap + i]);
        dtw = a.adaptive ? fmax(dtw, P[kPartDtw * cap + i]) : fmin(dtw, P[kPartDtw * cap + i]);
        raster = fmax(raster, P[kPartRaster * cap + i]);
        dtm = fmax(dtm, P[kPartDtotMax * cap + i]);
        dpm = fmax(dpm, P[kPartDprevMax * cap + i]);
        neg = fmax(neg, P[kPartNeg * cap + i]);
        if (a.flag_wq) dnm = fmax(dnm, P[kPartDnewMax * cap + i]);
        if (a.ca) cai = fmax(cai, P[kPartCaItot * cap + i]);
    }
    for (int i = t; i < a.n_outlet; i += kFinThreads) {
        qsum += (double)a.out_q[i] * a.A;
        dsum += (double)a.out_d[i];
        qraw += (double)a.out_q[i];
    }
    storage = wave_sum(storage); uz = wave_sum(uz); rain = wave_sum(rain);
    infil = wave_sum(infil); qsum = wave_sum(qsum); dsum = wave_sum(dsum);
    vmax = wave_fmax(vmax); dtw = a.adaptive ? wave_fmax(dtw) : wave_fmin(dtw);
    raster = wave_fmax(raster);
    dtm = wave_fmax(dtm); dpm = wave_fmax(dpm); neg = wave_fmax(neg);
    dnm = wave_fmax(dnm); qraw = wave_sum(qraw); cai = wave_fmax(cai);
    if (lane == 0) {
        sh[w][0] = storage; sh[w][1] = uz; sh[w][2] = rain; sh[w][3] = infil;
        sh[w][4] = vmax; sh[w][5] = dtw; sh[w][6] = raster; sh[w][7] = dtm;
        sh[w][8] = dpm; sh[w][9] = neg; sh[w][10] = qsum; sh[w][11] = dsum;
        sh[w][12] = dnm; sh[w][13] = qraw; sh[w][14] = cai;
    }
    __syncthreads();
    if (t != 0) return;
    // the serial tail works on a private copy of the run state (registers:
    // each dependent read-modify-write of a field costs no LDS round trip).
    // The rainfall-map and inflow fields, which are indexed by run-time
    // values, stay in the LDS copy: a dynamically indexed array in the private
    // copy would move the whole copy to scratch memory.
    SimState loc = sst;
    st = &loc;
    storage = uz = rain = infil = qsum = dsum = qraw = cai = 0.0;
    vmax = dtw = raster = dtm = dpm = dnm = NAN;
    neg = 0.0;
    for (int v = 0; v < kFinWaves; ++v) {  // fixed order: deterministic
        storage += sh[v][0]; uz += sh[v][1]; rain += sh[v][2]; infil += sh[v][3];
        vmax = fmax(vmax, sh[v][4]);
        dtw = a.adaptive ? fmax(dtw, sh[v][5]) : fmin(dtw, sh[v][5]);
        raster = fmax(raster, sh[v][6]);
        dtm = fmax(dtm, sh[v][7]); dpm = fmax(dpm, sh[v][8]); neg = fmax(neg, sh[v][9]);
        qsum += sh[v][10]; dsum += sh[v][11];
        dnm = fmax(dnm, sh[v][12]); qraw += sh[v][13];
        cai = fmax(cai, sh[v][14]);
    }
    // the routing's I_tot_end_cell are kept even when the step is rejected
    if (a.ca) st->ca_itot_any = cai;
    // volumes were reduced as sums of depths [mm]
    const double a1000 = a.A / 1000.0;
    storage *= a1000;
    uz *= a1000;
    infil *= a1000;
    if (a.vmax_squared) vmax = sqrt(vmax);
    dtm /= 1000.0;  // max depths were reduced in mm
    dpm /= 1000.0;
    dnm /= 1000.0;
    if (a.adaptive && !isnan(dtw)) {
        // refreshing_timestep.m:227-233 on the deepest non-stagnant cell:
        // alfa*dx/c is monotone in the depth, so its minimum is taken there
        double cw = sqrt(9.81 * fmax(dtw, 0.0));
        cw = fmax(cw, 1e-10);
        dtw = isinf(cw) ? NAN : a.alfa_min * a.R / cw;
    }

    if (a.dry_run) {  // launch-shape timing: reduce only, leave the state alone
        if (t == 0) a.partials_out[0] = storage + uz + infil + vmax + dtw + raster + qsum;
        return;
    }
    const long long k = st->k;
    bool failed = (st->inject_k == k);
    if (a.flag_wq && !failed) {
        // the wash-off of this step ran (HydroPol2D_Main_While.m:683-689): its
        // buffers become current whatever follows, B_t is not rolled back
        st->wq_n += 1;
        st->wq_out_conc = fmax(1000.0 * st->wq_sum_wout / (qraw / 1000.0 * a.A), 0.0);
        // :691-695: more than 1.2 x the initial mass raises an error (rollback)
        if (st->wq_sum_b > 1.2 * st->wq_initial_mass) failed = true;
    }
    // the same decision the step kernel took: storage sums exist only then
    const double prog_max = ((st->t + st->max_dt / 60.0) / a.routing_time) * 100.0;
    const bool may_report = (prog_max >= st->next_report) || (prog_max >= 100.0);
    // rainfall volume of this step: the applied depth over the valid cells
    rain = failed ? 0.0
                  : (a.n_maps > 0 ? map_depth_sum(a, &sst)
                                  : st->delta_p_agg * (a.has_pattern ? a.sum_pattern : a.n_valid)) *
                        (a.A / 1000.0);
    const double inflow_step = failed ? 0.0 : inflow_depth_sum(a, &sst) * (a.A / 1000.0);
    const double routing_time = a.routing_time;
    if (neg > 0) {  // Hydrological_Model.m:45-47
        st->catch_index += 1;
        st->neg_events += 1;
    }
    st->snap_event = 0;
    if (!failed) {
        const double dt_used_min = st->time_step_min;
        st->dt_used_min = dt_used_min;
        double t_now = st->t;
        // ---------------- refreshing_timestep ----------------
        const long long pos_save = (long long)ceil((t_now * 60.0) / a.tsc);
        const double time_save = (double)(pos_save - 1) * a.tsc / 60.0;
        const double delta_save = time_save - st->time_save_previous;
        st->time_save_previous = time_save;
        double tcr;
        if (delta_save > 0 || k == 1) {
            const double R = a.R;
            double dt_adv;
            if (vmax > 0) {
                dt_adv = 1.0 * R / fmax(vmax, 1e-10);
                const double tsf = fmax(a.alfa_max - 2.0 * fmax(vmax - 5.0, 0.0), a.alfa_min);
                dt_adv = dt_adv * tsf;
                dt_adv = fmin(dt_adv, st->max_dt);
            } else {
                dt_adv = st->max_dt;
            }
            double dw = dtw;
            if (isnan(dw)) dw = st->max_dt;
            double nt = a.flag_timestep == 1 ? fmin(dt_adv, dw) : dw;
            if (a.flag_wq) nt = fmin(1.0 * st->tmin_wq, nt);
            if (isnan(nt) || isinf(nt)) nt = st->max_dt;
            tcr = fmax(a.inc * floor(nt / a.inc), st->min_dt);
            tcr = fmin(tcr, st->max_dt);
            st->time_step_min = tcr / 60.0;
            st->max_velocity = vmax;
            st->max_vel_raster = raster;
        } else {
            tcr = st->time_step_min * 60.0;
        }
        st->tcr = tcr;
        // ---------------- first report check ----------------
        if (st->progress >= st->next_report || st->progress >= 100.0)
            st->next_report += a.report_pct;
        // ---------------- update_spatial_BC ----------------
        int cur = st->cursor;
        st->delta_p_agg = rain_depth(ser, a.n_rain, st->t_previous, t_now, &cur);
        st->cursor = cur;
        update_maps_inflow(a, &sst, st->t_previous, t_now);
        // ---------------- EMC (:724-731), with the refreshed time_step ----------------
        if (a.flag_wq) {
            const double qo = qraw / 1000.0 / 3600.0 * 1000.0;
            st->wq_mass_outlet = fmax(st->wq_mass_outlet + st->wq_out_conc * qo * (st->time_step_min * 60.0), 0.0);
            st->wq_vol_outlet = fmax(qo * (st->time_step_min * 60.0) + st->wq_vol_outlet, 0.0);
        }
        // ---------------- t_previous / t_save ----------------
        st->t_previous = (k == 1) ? tcr / 60.0 : t_now;
        const double t_save = t_now + tcr / 60.0;
        // ---------------- save_output_maps ----------------
        int actual = (int)floor(t_save / a.rec) + 1;
        if (actual > a.n_rec) actual = a.n_rec;
        if (actual < 1) actual = 1;
        while (actual < a.n_rec && (double)actual * a.rec <= t_save) ++actual;
        while (actual > 1 && (double)(actual - 1) * a.rec > t_save) --actual;
        const double dmean = a.n_outlet > 0 ? dsum / a.n_outlet : NAN;
        if (k == 1) {
            if (a.n_hydro_max > 0) {
                a.hq[0] = qsum / 1000.0 / 3600.0;
                a.ht[0] = tcr / 60.0;
                a.hd[0] = dmean;
                if (st->n_hydro < 1) st->n_hydro = 1;
            }
        } else if (actual - st->last_record > 0) {
            if (actual - 1 < a.n_hydro_max) {
                a.hq[actual - 1] = qsum / 1000.0 / 3600.0;
                a.ht[actual - 1] = t_now;
                a.hd[actual - 1] = dmean;
                if (st->n_hydro < actual) st->n_hydro = actual;
            }
        }
        st->last_record = actual;
        // ---------------- map stacks (save_output_maps.m:24-27) ----------------
        if (a.rec_maps > 0) {
            // records j >= 0 at j * rec_maps <= routing_time; the state is the
            // number of them before t_save
            int jmax = (int)floor(routing_time / a.rec_maps + 1e-12);
            while (jmax > 0 && (double)jmax * a.rec_maps > routing_time + 1e-12) --jmax;
            int am = 0;
            if (t_save > 0) {
                am = (int)ceil(t_save / a.rec_maps);
                while (am > 0 && (double)(am - 1) * a.rec_maps >= t_save) --am;
                while (am <= jmax && (double)am * a.rec_maps < t_save) ++am;
                if (am > jmax + 1) am = jmax + 1;
            }
            const int dm = am - st->map_last;
            st->map_last = am;
            if (dm > 0 && st->snap_n < a.snap_slots) {
                st->snap_event = 1;
                a.snap_t[st->snap_n] = t_now;
                st->snap_n += 1;
                st->n_map_records += 1;
            }
        }
        st->runoff_mm = qsum * st->time_step_min / 60.0 / a.drainage_area + st->runoff_mm;
        st->cum_rain += rain;
        st->cum_inflow += inflow_step;
        st->cum_inf += infil;
        st->cum_out += qsum / 1000.0 / 3600.0 * (dt_used_min * 60.0);
        if (a.flag_gw) st->gw_elapsed_min += dt_used_min;
        if (may_report) {
            st->storage = storage;
            st->uz = uz;
        }
        if (a.n_dt_max > 0) a.dt_hist[st->accepted % a.n_dt_max] = tcr;
        // ---------------- advance ----------------
        t_now = tcr / 60.0 + t_now;
        st->t = t_now;
        st->k = k + 1;
        st->accepted += 1;
        st->progress = (t_now / routing_time) * 100.0;
        bool do_report = false;
        if (st->progress >= st->next_report || st->progress >= 100.0) {
            do_report = true;
            st->next_report += a.report_pct;
        }
        if (do_report) {
            const double ts = st->time_step_min;
            st->outflow_volume = qsum / 1000.0 / 3600.0 * ts * 60.0 + st->outflow_volume;
            const double Pv = a.n_maps > 0 ? map_depth_sum(a, &sst) / 1000.0 * a.A
                              : a.has_pattern ? st->delta_p_agg * a.sum_pattern / 1000.0 * a.A
                                              : a.n_valid * (st->delta_p_agg / 1000.0 * a.A);
            const double inflow_part = inflow_depth_sum(a, &sst) / 1000.0 * a.A;
            const double inflow_vol = inflow_part + Pv + 0.0;
            st->inflow_volume = inflow_vol + st->inflow_volume;
            const double Qin = inflow_vol - Pv;
            const double Qout = qsum / 1000.0 / 3600.0 * ts * 60.0 + 0.0;
            const double fluxes = Pv + Qin - 0.0 - 0.0 - 0.0 - Qout - 0.0;
            const double S_GW = a.flag_gw ? st->gw_storage : 0.0;
            const double dS = (0.0 + storage + (a.ga ? uz : 0.0) + S_GW + 0.0) - st->S_prev;
            st->S_prev = st->S_prev + dS;
            st->volume_error = dS - fluxes;
            if (st->n_reports < a.n_rep_max) a.vol_err_hist[st->n_reports] = st->volume_error;
            st->n_reports += 1;
        }
    }
    // tmin_wq < 0, NaN or Inf raises at the end of the try block (:919-922):
    // the catch block then runs on the state of the accepted step (d_p = d_t)
    const bool late = !failed && a.flag_wq &&
                      (st->tmin_wq < 0 || isnan(st->tmin_wq) || isinf(st->tmin_wq));
    if (failed || late) {
        // ---------------- catch block (Main_While.m:926-993) ----------------
        if (late) dpm = dnm;
        if (st->inject_k == k) st->inject_k = -1;
        st->rejected += 1;
        double t_now = st->t - st->time_step_min;
        st->t_previous = t_now;
        const double wave_celerity = sqrt(9.81 * fmax(dtm, dpm));
        const double max_vel = st->max_vel_raster;
        st->catch_index += 1;
        const double factor = 1.0 / st->catch_index / a.factor_reduction;
        double nt = (a.adaptive == 1) ? factor * (a.alfa_min * a.R / wave_celerity)
                                      : factor * (a.alfa_min * a.R / (max_vel + wave_celerity));
        const double dt_wb = nt;
        nt = fmin(nt, dt_wb);
        if (st->catch_index > 1) st->max_dt = fmin(nt * 1.5, st->max_dt);
        nt = fmin(nt, st->max_dt);
        if (isinf(nt)) nt = st->min_dt;
        st->time_step_min = nt / 60.0;
        t_now = t_now + st->time_step_min;
        st->t = t_now;
        int cur = st->cursor;
        st->delta_p_agg = rain_depth(ser, a.n_rain, st->t_previous, t_now, &cur);
        st->cursor = cur;
        update_maps_inflow(a, &sst, st->t_previous, t_now);
    }
    st->done = !(st->t <= (routing_time + st->min_dt / 60.0)) ? 1 : 0;
    if (st->stop_accepted >= 0 && st->accepted >= st->stop_accepted) st->paused = 1;
    loc.map_n = sst.map_n;
    loc.map_cursor = sst.map_cursor;
    loc.map_err = sst.map_err;
    loc.map_dtot = sst.map_dtot;
    loc.map_k = sst.map_k;
    loc.inflow_cursor = sst.inflow_cursor;
#pragma unroll
    for (int j = 0; j < kMaxMapTerms; ++j) {
        loc.map_idx[j] = sst.map_idx[j];
        loc.map_w[j] = sst.map_w[j];
    }
#pragma unroll
    for (int g = 0; g < kMaxInflowGauges; ++g) loc.inflow_agg[g] = sst.inflow_agg[g];
    *a.st = loc;
}

template <typename T>
__global__ __launch_bounds__(kFinThreads) void finalize_kernel(FinArgs<T> a) {
    finalize_body<T>(a);
}

// one block per grid of an ensemble: block b finalises grid ids[b]
template <typename T>
__global__ __launch_bounds__(kFinThreads) void finalize_batch_kernel(const FinArgs<T> *args,
                                                                     const int *ids) {
    finalize_body<T>(args[ids[blockIdx.x]]);
}

template <typename T>
hipError_t launch_finalize(const FinArgs<T> &a, hipStream_t s) {
    hipLaunchKernelGGL((finalize_kernel<T>), dim3(1), dim3(kFinThreads), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_finalize_batch(const FinArgs<T> *args, const int *ids, int n_grids,
                                 hipStream_t s) {
    hipLaunchKernelGGL((finalize_batch_kernel<T>), dim3(n_grids), dim3(kFinThreads), 0, s, args,
                       ids);
    return hipGetLastError();
}
template hipError_t launch_finalize_batch<float>(const FinArgs<float> *, const int *, int,
                                                 hipStream_t);
template hipError_t launch_finalize_batch<double>(const FinArgs<double> *, const int *, int,
                                                  hipStream_t);

template hipError_t launch_finalize<float>(const FinArgs<float> &, hipStream_t);

// Surface (and soil) storage sums of the current state, FP64 per-block partials
// in a fixed order: mass_balance_check.m:62-64 outside the report cadence.
template <typename T>
__global__ __launch_bounds__(256) void storage_kernel(const T *d, const T *I, long long n,
                                                      double *part) {
    double sd = 0.0, si = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const double v = (double)d[c];
        if (!isnan(v)) sd += v;
        if (I) {
            const double u = (double)I[c];
            if (!isnan(u)) si += u;
        }
    }
    __shared__ double sh[2][4];
    sd = wave_sum(sd);
    si = wave_sum(si);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = sd;
        sh[1][w] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        part[2 * blockIdx.x + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

template <typename T>
hipError_t launch_storage(const T *d, const T *I, long long n, double *part, int blocks,
                          hipStream_t s) {
    hipLaunchKernelGGL((storage_kernel<T>), dim3(blocks), dim3(256), 0, s, d, I, n, part);
    return hipGetLastError();
}
template hipError_t launch_storage<float>(const float *, const float *, long long, double *, int,
                                          hipStream_t);
template hipError_t launch_storage<double>(const double *, const double *, long long, double *,
                                           int, hipStream_t);
template hipError_t launch_finalize<double>(const FinArgs<double> &, hipStream_t);

// Infiltration_Module.m:194-198 on the current head, in the reference's order
template <typename T>
__global__ __launch_bounds__(256) void vg_water_table_kernel(const double *__restrict__ h,
                                                             const double *__restrict__ elev,
                                                             const double *__restrict__ sd_map,
                                                             double sd_u, long long n, T *zw) {
    for (long long c = blockIdx.x * 256ll + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const double sd = sd_map ? sd_map[c] : sd_u;
        const double gw_depth = h[c] - (elev[c] - sd);
        const double zwt = fmax(sd - gw_depth, 0.0);
        zw[c] = copysign(T(zwt), zw[c]);
    }
}

template <typename T>
hipError_t launch_vg_water_table(const double *h, const double *elev, const double *soil_depth,
                                 double sd_u, long long n, T *zw, hipStream_t s) {
    const long long blocks = std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((vg_water_table_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, h, elev,
                       soil_depth, sd_u, n, zw);
    return hipGetLastError();
}
template hipError_t launch_vg_water_table<float>(const double *, const double *, const double *,
                                                 double, long long, float *, hipStream_t);
template hipError_t launch_vg_water_table<double>(const double *, const double *, const double *,
                                                  double, long long, double *, hipStream_t);

}  // namespace hp2d
