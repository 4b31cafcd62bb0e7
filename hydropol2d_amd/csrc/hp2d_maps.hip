// hp2d_maps.hip - map-stack snapshots of the device loop.
//
// save_output_maps.m:24-27 decides, after the dt refresh, whether the record
// state advanced (running_control.time_records < t_save); the finalize kernel
// takes that decision and raises st->snap_event.  This kernel, launched after
// finalize every step while map stacks are on, returns at once unless the
// event was raised; then it writes record st->snap_n - 1 of every field:
//   d, velocity and hazard_dv (snisb_velocity_hazard_snapshot, :832-848),
//   risk (Human_Instability_Module.m:8-25), I_t, GWdepth_save (:228-230),
//   Pol_Conc_Map, Pol_Mass_Map, Pol_Load_Map (:254-256).
// The velocities come from the step kernel, which stores them on the steps
// that can end in a record (t + max_dt reaches the next record time).
#include "hp2d_maps.hpp"

namespace hp2d {

template <typename T>
__device__ inline int snap_find_outlet(const long long *idx, int n, long long key) {
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

template <typename T>
__global__ __launch_bounds__(256) void snapshot_kernel(SnapArgs<T> a) {
    const SimState *st = a.st;
    if (!st->snap_event) return;
    const int slot = st->snap_n - 1;
    if (slot < 0 || slot >= a.slots) return;
    const int cur = (int)((st->k + 1) & 1);  // the buffers of the accepted step
    const int bp = (int)(st->wq_n & 1);
    const double dt_used_s = st->dt_used_min * 60.0, dt_new_s = st->time_step_min * 60.0;
    const long long off = (long long)slot * a.n;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < a.n; c += (long long)gridDim.x * 256) {
        const double d = (double)a.d[cur][c];
        const bool nan_cell = sizeof(T) == 8 ? isnan(a.elev[c]) : isnan((double)a.n2s[c]);
        // velocity and hazard snapshot
        const double dm = fmax(d / 1000.0, 0.0);
        double v = (double)a.vras[c];
        if (!isfinite(v)) v = 0.0;
        double hz = dm * v;
        if (dm < 0.01) hz = 0.0;
        if (!isfinite(hz)) hz = 0.0;
        a.stack[HP2D_MAP_DEPTH][off + c] = T(d);
        a.stack[HP2D_MAP_VELOCITY][off + c] = nan_cell ? T(NAN) : T(v);
        a.stack[HP2D_MAP_HAZARD][off + c] = nan_cell ? T(NAN) : T(hz);
        if (a.stack[HP2D_MAP_RISK]) {
            const double h = d / 1000.0, vt = (double)a.vtot[c];
            const double F_person = a.hi_weight * a.hi_g;
            const double F_buoy = (a.hi_w1 * a.hi_w2) * h * a.hi_ro * a.hi_g;
            const double avail = a.hi_mu * (fmax(F_person - F_buoy, 0.0));
            const double hydro = fmax(1.0 / 2.0 * (a.hi_Cd * a.hi_ro * a.hi_w1 * h * (vt * vt)), 0.0);
            double r = 0.0;
            if (avail > 0) r = hydro / avail;
            else if (h > 0) r = 1.0;
            r = fmin(fmax(r, 0.0), 1.0);
            a.stack[HP2D_MAP_RISK][off + c] = nan_cell ? T(NAN) : T(r);
        }
        if (a.stack[HP2D_MAP_INFILTRATION]) a.stack[HP2D_MAP_INFILTRATION][off + c] = a.I[cur][c];
        if (a.stack[HP2D_MAP_GW_DEPTH]) {
            const double sd = a.gw_sd ? a.gw_sd[c] : a.gw_sd_u;
            a.stack[HP2D_MAP_GW_DEPTH][off + c] = T(a.gw_h[c] - a.gw_elev[c] + sd);
        }
        if (a.stack[HP2D_MAP_POL_CONC]) {
            // I_tot_end_cell of the step (Local_Inertial_Model_D4.m:513-514) over the
            // refreshed dt: outflow pages x, y and the outlet's
            double q3 = 0.0;
            if (a.n_outlet > 0 && signbit((double)a.n2s[c])) {
                const int k = snap_find_outlet<T>(a.out_idx, a.n_outlet, c);
                if (k >= 0) q3 = (double)a.out_q[k];
            }
            const double pages = (double)a.qx[cur][c] + (double)a.qy[cur][c] + q3;
            const double itot = fabs(pages) * dt_used_s / 1000.0 * 1.0 / 3600.0 * (a.R * a.R);
            const double pc = (double)a.wqPc[c];
            a.stack[HP2D_MAP_POL_CONC][off + c] = T(pc);
            a.stack[HP2D_MAP_POL_MASS][off + c] = a.wqB[bp][c];
            a.stack[HP2D_MAP_POL_LOAD][off + c] = T(1.0 / 1000.0 * pc * itot / dt_new_s / (a.R * a.R));
        }
    }
}

template <typename T> hipError_t launch_snapshot(const SnapArgs<T> &a, hipStream_t s) {
    const long long blocks = std::min<long long>((a.n + 255) / 256, 2048);
    hipLaunchKernelGGL((snapshot_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template hipError_t launch_snapshot<float>(const SnapArgs<float> &, hipStream_t);
template hipError_t launch_snapshot<double>(const SnapArgs<double> &, hipStream_t);

}  // namespace hp2d
