// hp2d_capi.cpp - C ABI of include/hp2d.h: handles, drop-in per-call entry
// points and the device-resident step loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_forcing.hpp"
#include "hp2d_gw.hpp"
#include "hp2d_handle.hpp"
#include "hp2d_internal.hpp"
#include "hp2d_li.hpp"
#include "hp2d_maps.hpp"
#include "hp2d_sim.hpp"
#include "hp2d_simca.hpp"
#include "hp2d_simwq.hpp"

using namespace hp2d;

namespace {

// device map-stack records between host drains: a step takes one record at
// most, and a graph holds at most kSnapSlots steps while map stacks are on
constexpr int kSnapSlots = 16;

thread_local std::string g_last_error;

struct ErrSink {
    std::string *dst;
    void operator()(const std::string &m) const {
        if (dst) *dst = m;
        g_last_error = m;
    }
};

int default_tj(int nx) { return nx >= 4096 ? 64 : (nx >= 512 ? 32 : 16); }

template <typename T> struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        if (p) hipFree(p);
        p = nullptr;
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, count * sizeof(T));
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Sorted unique 0-based linear outlet indices from 1-based MATLAB rows/cols.
hp2d_status outlet_list(int ny, int nx, const int32_t *rows, const int32_t *cols, int n,
                        std::vector<long long> &out, std::string &err) {
    out.clear();
    for (int k = 0; k < n; ++k) {
        if (rows[k] < 1 || rows[k] > ny || cols[k] < 1 || cols[k] > nx) {
            err = "outlet index out of range";
            return HP2D_E_ARG;
        }
        out.push_back((long long)(rows[k] - 1) + (long long)(cols[k] - 1) * ny);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return HP2D_OK;
}

}  // namespace

// ============================================================================
// Handle (drop-in per-call boundaries)
// ============================================================================
struct hp2d_handle {
    int ny = 0, nx = 0, precision = HP2D_FP64;
    hipStream_t stream = nullptr;
    std::string err;
    size_t elem = 8;
    void *dev[20] = {nullptr};   // z (FP32: dzx), d, n2s, qx, qy, d_t, ox, oy, q3, hfx, hfy, hf3, qof,
                                 // ql, qr, qu, qd, itot, dzy (FP32)
    long long *out_idx = nullptr;
    void *out_n = nullptr;
    int out_cap = 0;
    int tj = 16, wpb = 4, rows = 1;
    void *ws = nullptr;          // module workspace (wash-off, groundwater)
    size_t ws_bytes = 0;
    void *ws2 = nullptr;         // second workspace (groundwater module around Boussinesq)
    size_t ws2_bytes = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};  // around the compute kernels of a timed drop-in
    bool ev_valid = false;
    HandleSlot slot[kHandleSlots];          // per-module resident state (subgrid tables)
};

static void set_err(hp2d_handle *h, const std::string &m) {
    if (h) h->err = m;
    g_last_error = m;
}

#define HTRY(h, expr)                                                                 \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            set_err(h, std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return HP2D_E_HIP;                                                        \
        }                                                                             \
    } while (0)

namespace hp2d {
// the thread's last error, for entry points defined in other translation units
void set_global_error(const std::string &m) { g_last_error = m; }
}  // namespace hp2d

extern "C" {

const char *hp2d_version(void) { return "hp2d-gfx950 0.1.0"; }

int32_t hp2d_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *hp2d_global_last_error(void) { return g_last_error.c_str(); }

// Largest grid: cell offsets are 32-bit, and a one-thread-per-cell launch of
// 256-thread blocks must stay under 2^32 work-items (a dispatch counts them in
// 32 bits)
constexpr double kMaxCells = 4294967296.0 - 65536.0;

hp2d_status hp2d_create(int32_t ny, int32_t nx, int32_t precision, hp2d_handle **out) {
    if (!out || ny < 1 || nx < 1 || (precision != HP2D_FP64 && precision != HP2D_FP32) ||
        (double)ny * (double)nx > kMaxCells) {
        g_last_error = "hp2d_create: bad arguments (grids are limited to 2^32 - 2^16 cells)";
        return HP2D_E_ARG;
    }
    if (hp2d_device_count() < 1) {
        g_last_error = "hp2d_create: no HIP device visible";
        return HP2D_E_NO_DEVICE;
    }
    hp2d_handle *h = new hp2d_handle();
    h->ny = ny;
    h->nx = nx;
    h->precision = precision;
    h->elem = precision == HP2D_FP64 ? 8 : 4;
    h->tj = default_tj(nx);
    h->rows = li_rows_ok(ny, 2) ? 2 : 1;
    const size_t n = (size_t)ny * nx;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    for (int i = 0; i < 19 && e == hipSuccess; ++i) e = hipMalloc(&h->dev[i], n * h->elem);
    if (e != hipSuccess) {
        g_last_error = std::string("hp2d_create: ") + hipGetErrorString(e);
        hp2d_destroy(h);
        return e == hipErrorOutOfMemory ? HP2D_E_NOMEM : HP2D_E_HIP;
    }
    *out = h;
    return HP2D_OK;
}

void hp2d_destroy(hp2d_handle *h) {
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (auto &p : h->dev)
        if (p) hipFree(p);
    if (h->out_idx) hipFree(h->out_idx);
    if (h->out_n) hipFree(h->out_n);
    if (h->ws) hipFree(h->ws);
    if (h->ws2) hipFree(h->ws2);
    for (auto &e : h->ev)
        if (e) hipEventDestroy(e);
    for (auto &sl : h->slot)
        if (sl.p && sl.del) sl.del(sl.p);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

const char *hp2d_last_error(const hp2d_handle *h) { return h ? h->err.c_str() : g_last_error.c_str(); }

hp2d_status hp2d_last_kernel_time(hp2d_handle *h, double *ms) {
    if (!h || !ms) return HP2D_E_ARG;
    if (!h->ev_valid) {
        set_err(h, "hp2d_last_kernel_time: no timed call on this handle yet");
        return HP2D_E_STATE;
    }
    HTRY(h, hipEventSynchronize(h->ev[1]));
    float t = 0.0f;
    HTRY(h, hipEventElapsedTime(&t, h->ev[0], h->ev[1]));
    *ms = (double)t;
    return HP2D_OK;
}

}  // extern "C"

hp2d_status hp2d_handle_shape(hp2d_handle *h, int32_t *ny, int32_t *nx, int32_t *precision) {
    if (!h) return HP2D_E_ARG;
    *ny = h->ny;
    *nx = h->nx;
    *precision = h->precision;
    return HP2D_OK;
}

hipStream_t hp2d_handle_stream(hp2d_handle *h) { return h->stream; }

HandleSlot *hp2d_handle_slot(hp2d_handle *h, int which) {
    return (which >= 0 && which < kHandleSlots) ? &h->slot[which] : nullptr;
}

hipError_t hp2d_handle_mark(hp2d_handle *h, int which) {
    if (!h->ev[which]) {
        hipError_t e = hipEventCreate(&h->ev[which]);
        if (e != hipSuccess) return e;
    }
    if (which == 0) h->ev_valid = false;
    hipError_t e = hipEventRecord(h->ev[which], h->stream);
    if (e == hipSuccess && which == 1) h->ev_valid = true;
    return e;
}

hp2d_status hp2d_handle_fail(hp2d_handle *h, const std::string &msg) {
    set_err(h, msg);
    return HP2D_E_HIP;
}

void *hp2d_handle_workspace(hp2d_handle *h, size_t bytes) {
    if (h->ws_bytes >= bytes) return h->ws;
    if (h->ws) {
        hipStreamSynchronize(h->stream);
        hipFree(h->ws);
        h->ws = nullptr;
        h->ws_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->ws, bytes);
    if (e != hipSuccess) {
        h->ws = nullptr;
        set_err(h, std::string("workspace hipMalloc: ") + hipGetErrorString(e));
        return nullptr;
    }
    h->ws_bytes = bytes;
    return h->ws;
}

void *hp2d_handle_workspace2(hp2d_handle *h, size_t bytes) {
    if (h->ws2_bytes >= bytes) return h->ws2;
    if (h->ws2) {
        hipStreamSynchronize(h->stream);
        hipFree(h->ws2);
        h->ws2 = nullptr;
        h->ws2_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->ws2, bytes);
    if (e != hipSuccess) {
        h->ws2 = nullptr;
        set_err(h, std::string("workspace hipMalloc: ") + hipGetErrorString(e));
        return nullptr;
    }
    h->ws2_bytes = bytes;
    return h->ws2;
}

// Elevation differences of the x and y faces a cell stores, for the FP32 step
// kernel: dzx(i,j) = z(i,j+1) - z(i,j) (0 in the last column), dzy(i,j) =
// z(i-1,j) - z(i,j) (0 in the first row), each taken in FP64 and rounded once,
// NaN when either cell lies outside the domain.
template <typename Z>
static void face_dz(int ny, int nx, const Z *z, std::vector<float> &dzx, std::vector<float> &dzy) {
    const size_t n = (size_t)ny * nx;
    dzx.assign(n, 0.0f);
    dzy.assign(n, 0.0f);
    for (int j = 0; j < nx; ++j)
        for (int i = 0; i < ny; ++i) {
            const size_t c = (size_t)i + (size_t)j * ny;
            if (j < nx - 1) dzx[c] = (float)((double)z[c + ny] - (double)z[c]);
            if (i > 0) dzy[c] = (float)((double)z[c - 1] - (double)z[c]);
            if (std::isnan((double)z[c])) dzx[c] = dzy[c] = NAN;
        }
}

template <typename T>
static hp2d_status li_dropin(hp2d_handle *h, int scheme, const T *z, const T *d_tot,
                             const T *rough, const T *rough2, T cell_area, T time_step_min,
                             T resolution, int outlet_type, T slope_outlet, const int32_t *rows,
                             const int32_t *cols, int n_out, const T *outflow_prev, int critical,
                             T *ql, T *qr, T *qu, T *qd, T *qof, T *d_t, T *itot, T *outflow,
                             T *Hf) {
    if (!h) return HP2D_E_ARG;
    if (scheme != 1) {
        set_err(h, "only flag_numerical_scheme = 1 (Bates) is implemented");
        return HP2D_E_UNSUPPORTED;
    }
    if (!z || !d_tot || !rough || !rough2 || !outflow_prev || (n_out > 0 && (!rows || !cols))) {
        set_err(h, "hp2d_local_inertial_d4: NULL input");
        return HP2D_E_ARG;
    }
    const int ny = h->ny, nx = h->nx;
    const size_t n = (size_t)ny * nx;
    std::vector<long long> oidx;
    std::string err;
    if (outlet_list(ny, nx, rows, cols, n_out, oidx, err) != HP2D_OK) {
        set_err(h, err);
        return HP2D_E_ARG;
    }
    // n^2 with the outlet flag in the sign bit; roughness of each outlet
    std::vector<T> n2s(rough2, rough2 + n);
    std::vector<T> on(oidx.size());
    for (size_t k = 0; k < oidx.size(); ++k) {
        n2s[oidx[k]] = std::copysign(std::fabs(n2s[oidx[k]]), T(-1));
        on[k] = rough[oidx[k]];
    }
    // the FP32 kernel reads the face elevation differences, and marks the cells
    // outside the domain by a NaN n^2
    std::vector<float> dzx, dzy;
    if (sizeof(T) == 4) {
        face_dz(ny, nx, z, dzx, dzy);
        for (size_t c = 0; c < n; ++c)
            if (std::isnan((double)z[c])) n2s[c] = T(NAN);
    }
    if ((int)oidx.size() > h->out_cap) {
        if (h->out_idx) hipFree(h->out_idx);
        if (h->out_n) hipFree(h->out_n);
        h->out_idx = nullptr;
        h->out_n = nullptr;
        h->out_cap = (int)oidx.size();
        HTRY(h, hipMalloc((void **)&h->out_idx, oidx.size() * sizeof(long long)));
        HTRY(h, hipMalloc(&h->out_n, oidx.size() * sizeof(T)));
    }
    hipStream_t s = h->stream;
    T **D = (T **)h->dev;
    if (sizeof(T) == 8) {
        HTRY(h, hipMemcpyAsync(D[0], z, n * sizeof(T), hipMemcpyHostToDevice, s));
    } else {
        HTRY(h, hipMemcpyAsync(D[0], dzx.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
        HTRY(h, hipMemcpyAsync(D[18], dzy.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    }
    HTRY(h, hipMemcpyAsync(D[1], d_tot, n * sizeof(T), hipMemcpyHostToDevice, s));
    HTRY(h, hipMemcpyAsync(D[2], n2s.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    HTRY(h, hipMemcpyAsync(D[3], outflow_prev, n * sizeof(T), hipMemcpyHostToDevice, s));
    HTRY(h, hipMemcpyAsync(D[4], outflow_prev + n, n * sizeof(T), hipMemcpyHostToDevice, s));
    if (!oidx.empty()) {
        HTRY(h, hipMemcpyAsync(h->out_idx, oidx.data(), oidx.size() * sizeof(long long),
                               hipMemcpyHostToDevice, s));
        HTRY(h, hipMemcpyAsync(h->out_n, on.data(), on.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    LiArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.ny = ny;
    a.nx = nx;
    a.tj = h->tj;
    a.rows = h->rows;
    a.n_slabs = li_n_slabs(ny, h->rows);
    a.waves_per_block = h->wpb;
    a.R = resolution;
    a.R2 = resolution * resolution;
    a.cell_area = cell_area;
    a.C_a = resolution * resolution;
    a.inv_R = T(1) / resolution;
    a.k_mmh_q = T((double)resolution / 3.6e6);
    a.k_q_mmh = T(3.6e6 / (double)cell_area);
    a.flag_critical = critical;
    a.outlet_type = outlet_type;
    a.sqrtS_type1 = std::sqrt(std::fabs(slope_outlet));
    a.dt_s = time_step_min * T(60);
    a.dt_h = time_step_min / T(60);
    a.z = D[0];
    a.dzx = D[0];
    a.dzy = D[18];
    a.d_src = D[1];
    a.n2s = D[2];
    a.qx_src = D[3];
    a.qy_src = D[4];
    a.d_dst = D[5];
    a.qx_dst = D[6];
    a.qy_dst = D[7];
    a.q3 = D[8];
    a.hfx = D[9];
    a.hfy = D[10];
    a.hf3 = D[11];
    a.outlet_flow = D[12];
    a.qleft = D[13];
    a.qright = D[14];
    a.qup = D[15];
    a.qdown = D[16];
    a.itot = D[17];
    a.out_idx = h->out_idx;
    a.out_n = (const T *)h->out_n;
    a.n_outlet = (int)oidx.size();
    HTRY(h, launch_li_step<T>(a, kModeDropIn, s));
    auto get = [&](T *dst, T *src) -> hipError_t {
        if (!dst) return hipSuccess;
        return hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s);
    };
    HTRY(h, get(d_t, D[5]));
    HTRY(h, get(ql, D[13]));
    HTRY(h, get(qr, D[14]));
    HTRY(h, get(qu, D[15]));
    HTRY(h, get(qd, D[16]));
    HTRY(h, get(qof, D[12]));
    HTRY(h, get(itot, D[17]));
    if (outflow) {
        HTRY(h, get(outflow, D[6]));
        HTRY(h, get(outflow + n, D[7]));
        HTRY(h, get(outflow + 2 * n, D[8]));
    }
    if (Hf) {
        HTRY(h, get(Hf, D[9]));
        HTRY(h, get(Hf + n, D[10]));
        HTRY(h, get(Hf + 2 * n, D[11]));
    }
    HTRY(h, hipStreamSynchronize(s));
    return HP2D_OK;
}

extern "C" {

hp2d_status hp2d_local_inertial_d4(hp2d_handle *h, int32_t flag_numerical_scheme,
                                   const double *z, const double *d_tot,
                                   const double *roughness_cell,
                                   const double *roughness_squared, double cell_area,
                                   double time_step_min, double resolution,
                                   int32_t outlet_type, double slope_outlet,
                                   const int32_t *row_outlet, const int32_t *col_outlet,
                                   int32_t n_outlet, const double *outflow_prev,
                                   int32_t flag_critical, double *qout_left, double *qout_right,
                                   double *qout_up, double *qout_down, double *outlet_flow,
                                   double *d_t, double *I_tot_end_cell, double *outflow,
                                   double *Hf) {
    if (!h || h->precision != HP2D_FP64) {
        set_err(h, "hp2d_local_inertial_d4 needs an HP2D_FP64 handle");
        return HP2D_E_ARG;
    }
    return li_dropin<double>(h, flag_numerical_scheme, z, d_tot, roughness_cell,
                             roughness_squared, cell_area, time_step_min, resolution, outlet_type,
                             slope_outlet, row_outlet, col_outlet, n_outlet, outflow_prev,
                             flag_critical, qout_left, qout_right, qout_up, qout_down, outlet_flow,
                             d_t, I_tot_end_cell, outflow, Hf);
}

hp2d_status hp2d_local_inertial_d4_f32(hp2d_handle *h, int32_t flag_numerical_scheme,
                                       const float *z, const float *d_tot,
                                       const float *roughness_cell,
                                       const float *roughness_squared, float cell_area,
                                       float time_step_min, float resolution,
                                       int32_t outlet_type, float slope_outlet,
                                       const int32_t *row_outlet, const int32_t *col_outlet,
                                       int32_t n_outlet, const float *outflow_prev,
                                       int32_t flag_critical, float *qout_left,
                                       float *qout_right, float *qout_up, float *qout_down,
                                       float *outlet_flow, float *d_t, float *I_tot_end_cell,
                                       float *outflow, float *Hf) {
    if (!h || h->precision != HP2D_FP32) {
        set_err(h, "hp2d_local_inertial_d4_f32 needs an HP2D_FP32 handle");
        return HP2D_E_ARG;
    }
    return li_dropin<float>(h, flag_numerical_scheme, z, d_tot, roughness_cell, roughness_squared,
                            cell_area, time_step_min, resolution, outlet_type, slope_outlet,
                            row_outlet, col_outlet, n_outlet, outflow_prev, flag_critical,
                            qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t,
                            I_tot_end_cell, outflow, Hf);
}

}  // extern "C"

// ============================================================================
// Device-resident simulation
// ============================================================================
struct hp2d_sim {
    hp2d_run_config cfg{};
    int ny = 0, nx = 0;
    size_t n = 0, elem = 8;
    hipStream_t stream = nullptr;
    std::string err;
    // device memory
    void *z = nullptr, *n2s = nullptr, *pattern = nullptr, *dmax = nullptr, *tpeak = nullptr;
    void *zy = nullptr;  // FP32: dzy (z holds dzx)
    void *d[2] = {nullptr, nullptr}, *qx[2] = {nullptr, nullptr}, *qy[2] = {nullptr, nullptr};
    void *I[2] = {nullptr, nullptr}, *Tc[2] = {nullptr, nullptr};
    uint8_t *soil = nullptr;
    GaClass *ga = nullptr;
    float *ga_f32 = nullptr;
    // van Genuchten infiltration (flag_infiltration = 2): class tables and the
    // depth to the water table per cell (sign bit = impervious)
    VgClass *vg = nullptr;
    float *vg_f32 = nullptr;
    void *vg_zwt = nullptr;
    double vg_l = 0.5;  // Mualem pore connectivity
    // CA routing (flags.flag_CA = 1): d_tot, the left / down / outlet outflows,
    // I_tot_end_cell and the optional h_0; the right / up outflows use qx / qy
    bool ca = false;
    void *ca_dtot = nullptr, *ca_ql = nullptr, *ca_qd = nullptr, *ca_qo = nullptr;
    void *ca_itot = nullptr, *ca_h0 = nullptr;
    int ca_blocks = 0;
    long long *out_idx = nullptr;
    void *out_n = nullptr, *out_q = nullptr, *out_d = nullptr;
    int n_outlet = 0;
    double *partials = nullptr;
    long long part_cap = 0;
    int n_partials = 0;
    SimState *st = nullptr;
    double *rain_t = nullptr, *rain_v = nullptr;
    int n_rain = 0;
    // rainfall maps and inflow hydrographs (update_spatial_BC.m:362-398, 469-518)
    // Human_Instability_Module max_risk and the map stacks (save_output_maps.m)
    void *max_risk = nullptr, *snap_vras = nullptr, *snap_vtot = nullptr;
    void *stack[HP2D_MAP_FIELDS] = {nullptr};
    double *snap_t = nullptr;
    int snap_slots = 0;
    std::vector<std::vector<double>> records[HP2D_MAP_FIELDS];
    std::vector<double> record_t;
    void *maps = nullptr;
    double *map_t = nullptr, *map_sum = nullptr;
    int n_maps = 0;
    double *inflow_t = nullptr, *inflow_v = nullptr, *gauge_cells = nullptr;
    uint8_t *gauge = nullptr;
    int n_inflow_times = 0, n_gauges = 0;
    double *hq = nullptr, *ht = nullptr, *hd = nullptr;
    int n_hydro_max = 0;
    double *dt_hist = nullptr;
    int n_dt_max = 4096;
    double *vol_err = nullptr;
    double *tune_sink = nullptr;
    int n_rep_max = 256;
    double sum_pattern = 0, n_valid = 0, drainage_area = 0, I_min = 0;
    int tj = 16, wpb = 4, rows = 1;  // step kernel launch shape (rows per lane)
    // graph
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    // timing of the last run_fixed
    double li_ms = 0, total_ms = 0;
    long long li_launches = 0;
    // water quality (build_up_wash_off in the loop)
    bool wq = false;
    void *wqB[2] = {nullptr, nullptr}, *wqBs = nullptr, *wqBx[2] = {nullptr, nullptr};
    void *wqOm = nullptr, *wqTW[2] = {nullptr, nullptr}, *wqPc = nullptr;
    void *wqC3 = nullptr, *wqC4 = nullptr;
    uint16_t *wq_cell = nullptr;
    double *wq_part = nullptr;
    WqCtl *wq_ctl = nullptr;
    int wq_blocks = 0, wq_scap = 2;
    double wq_min_Bt = 0, wq_Bmin = 0, wq_Bmax = 0;
    // groundwater (Groundwater_Module every gw_every_steps accepted steps)
    bool gw = false;
    hp2d_handle *gwh = nullptr;        // stream + workspace of the module's kernels
    double *gw_mem = nullptr;          // the module's device arrays (GwmArgs)
    GwmArgs gwa{};
    double *gw_col = nullptr;
    uint8_t *gw_dom = nullptr;
    hp2d_gw_params gwp{};
    hp2d_gw_state gws{};
    double gw_error = 0.0;
    long long gw_last = 0;             // accepted steps at the last module call
};

static void set_err(hp2d_sim *s, const std::string &m) {
    if (s) s->err = m;
    g_last_error = m;
}

#define STRY(s, expr)                                                                 \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            set_err(s, std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return HP2D_E_HIP;                                                        \
        }                                                                             \
    } while (0)

template <typename T> static LiArgs<T> sim_li_args(hp2d_sim *s) {
    LiArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    const hp2d_run_config &c = s->cfg;
    a.ny = s->ny;
    a.nx = s->nx;
    a.tj = s->tj;
    a.rows = s->rows;
    a.n_slabs = li_n_slabs(s->ny, s->rows);
    a.waves_per_block = s->wpb;
    a.R = T(c.resolution);
    a.R2 = T(c.resolution * c.resolution);
    a.cell_area = T(c.resolution * c.resolution);
    a.C_a = T(c.resolution * c.resolution);
    a.inv_R = T(1.0 / c.resolution);
    a.k_mmh_q = T(c.resolution / 3.6e6);
    a.k_q_mmh = T(3.6e6 / (c.resolution * c.resolution));
    a.flag_critical = c.flag_critical;
    a.outlet_type = c.outlet_type;
    a.adaptive = c.flag_adaptive_timestepping;
    a.sqrtS_type1 = T(std::sqrt(std::fabs(c.slope_outlet)));
    a.depth_tol_mm = T(c.depth_tolerance_mm);
    a.alfa_min = T(c.alfa_min);
    a.rain_maps = (const T *)s->maps;
    a.inflow_gauge = s->gauge;
    a.z = (const T *)s->z;
    a.dzx = (const T *)s->z;
    a.dzy = (const T *)s->zy;
    a.n2s = (const T *)s->n2s;
    for (int p = 0; p < 2; ++p) {
        a.d_buf[p] = (T *)s->d[p];
        a.qx_buf[p] = (T *)s->qx[p];
        a.qy_buf[p] = (T *)s->qy[p];
        a.I_buf[p] = (T *)s->I[p];
        a.T_buf[p] = (T *)s->Tc[p];
    }
    a.st = s->st;
    a.rain_pattern = (const T *)s->pattern;
    a.soil_class = s->soil;
    a.ga = s->ga;
    a.ga_f32 = s->ga_f32;
    a.vg = s->vg;
    a.vg_f32 = s->vg_f32;
    a.vg_zwt = (const T *)s->vg_zwt;
    a.vg_Ltop = T(std::fmax(c.vg_Ltop, 1e-6));
    a.vg_dhmax = T(c.vg_dh_max);
    a.vg_wK = T(std::fmin(std::fmax(c.vg_Ksurf_weight, 0.0), 1.0));
    a.vg_l = T(s->vg_l);
    a.n_soil_classes = (c.flag_infiltration == 1 || c.flag_infiltration == 2) ? c.n_soil_classes : 0;
    a.routing_time = c.routing_time_min;
    a.I_min = T(s->I_min);
    a.dmax_map = (T *)s->dmax;
    a.tpeak_map = (T *)s->tpeak;
    a.out_idx = s->out_idx;
    a.out_n = (const T *)s->out_n;
    a.n_outlet = s->n_outlet;
    a.out_q = (T *)s->out_q;
    a.out_d = (T *)s->out_d;
    a.partials = s->partials;
    a.part_cap = s->part_cap;
    a.track_dmax = s->wq ? 1 : 0;
    a.max_risk = (T *)s->max_risk;
    a.hi_weight = T(c.hi_weight_kg);
    a.hi_w1 = T(c.hi_width1_m);
    a.hi_w2 = T(c.hi_width2_m);
    a.hi_mu = T(c.hi_mu);
    a.hi_Cd = T(c.hi_Cd);
    a.hi_ro = T(c.hi_ro_water);
    a.hi_g = T(c.hi_gravity);
    a.snap_vras = (T *)s->snap_vras;
    a.snap_vtot = (T *)s->snap_vtot;
    a.rec_maps = c.record_time_maps_min;
    return a;
}

template <typename T> static WqSimArgs<T> sim_wq_args(hp2d_sim *s) {
    WqSimArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    const hp2d_run_config &c = s->cfg;
    a.ny = s->ny;
    a.nx = s->nx;
    a.n = (long long)s->n;
    a.A = c.resolution * c.resolution;
    a.min_Bt = s->wq_min_Bt;
    a.Bmin = s->wq_Bmin;
    a.Bmax = s->wq_Bmax;
    a.model = c.flag_wq_model;
    a.n_blocks = s->wq_blocks;
    a.s_cap = s->wq_scap;
    a.cell = s->wq_cell;
    a.out_idx = s->out_idx;
    a.out_q = (const T *)s->out_q;
    a.n_outlet = s->n_outlet;
    for (int p = 0; p < 2; ++p) {
        a.qx_buf[p] = (T *)s->qx[p];
        a.qy_buf[p] = (T *)s->qy[p];
        a.B[p] = (T *)s->wqB[p];
        a.Bx[p] = (T *)s->wqBx[p];
        a.TW[p] = (T *)s->wqTW[p];
    }
    a.C3 = (const T *)s->wqC3;
    a.C4 = (const T *)s->wqC4;
    a.n_classes = c.n_lulc_classes;
    a.Bs = (T *)s->wqBs;
    a.omass = (T *)s->wqOm;
    a.Pc = (T *)s->wqPc;
    a.part = s->wq_part;
    a.ctl = s->wq_ctl;
    a.st = s->st;
    return a;
}

template <typename T> static FinArgs<T> sim_fin_args(hp2d_sim *s) {
    FinArgs<T> f;
    std::memset(&f, 0, sizeof(f));
    const hp2d_run_config &c = s->cfg;
    f.st = s->st;
    f.partials = s->partials;
    f.part_cap = s->part_cap;
    // one partial per workgroup of the current launch shape
    f.n_partials = s->ca ? s->ca_blocks
                         : ((li_n_slabs(s->ny, s->rows) + s->wpb - 1) / s->wpb) *
                               ((s->nx + s->tj - 1) / s->tj);
    f.ca = s->ca ? 1 : 0;
    f.routing_time = c.routing_time_min;
    f.tsc = c.time_step_change_s;
    f.inc = c.time_step_increments_s;
    f.alfa_min = c.alfa_min;
    f.alfa_max = c.alfa_max;
    f.R = c.resolution;
    f.A = c.resolution * c.resolution;
    f.rec = c.record_time_hydrographs_min;
    f.report_pct = c.report_every_percent;
    f.factor_reduction = c.factor_reduction;
    f.drainage_area = s->drainage_area;
    f.sum_pattern = s->sum_pattern;
    f.n_valid = s->n_valid;
    f.flag_timestep = c.flag_timestep;
    f.adaptive = c.flag_adaptive_timestepping;
    f.ga = c.flag_infiltration == 1 || c.flag_infiltration == 2;  // I_t is kept
    f.flag_wq = s->wq ? 1 : 0;
    f.flag_gw = s->gw ? 1 : 0;
    f.has_pattern = s->pattern != nullptr;
    f.vmax_squared = s->elem == 4;
    f.n_rec = s->n_hydro_max;
    f.rain_t = s->rain_t;
    f.rain_v = s->rain_v;
    f.n_rain = s->n_rain;
    f.map_t = s->map_t;
    f.map_sum = s->map_sum;
    f.n_maps = s->n_maps;
    f.inflow_t = s->inflow_t;
    f.inflow_v = s->inflow_v;
    f.gauge_cells = s->gauge_cells;
    f.n_inflow_times = s->n_inflow_times;
    f.n_gauges = s->n_gauges;
    f.rec_maps = s->snap_slots > 0 ? c.record_time_maps_min : 0.0;
    f.snap_slots = s->snap_slots;
    f.snap_t = s->snap_t;
    f.out_q = (const T *)s->out_q;
    f.out_d = (const T *)s->out_d;
    f.n_outlet = s->n_outlet;
    f.hq = s->hq;
    f.ht = s->ht;
    f.hd = s->hd;
    f.n_hydro_max = s->n_hydro_max;
    f.dt_hist = s->dt_hist;
    f.n_dt_max = s->n_dt_max;
    f.vol_err_hist = s->vol_err;
    f.n_rep_max = s->n_rep_max;
    return f;
}

// the launches that follow the fused step: the wash-off (when enabled) and
// the dt / bookkeeping tail
template <typename T> static SnapArgs<T> sim_snap_args(hp2d_sim *s) {
    SnapArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    const hp2d_run_config &c = s->cfg;
    a.st = s->st;
    a.n = (long long)s->n;
    a.ny = s->ny;
    a.slots = s->snap_slots;
    for (int f = 0; f < HP2D_MAP_FIELDS; ++f) a.stack[f] = (T *)s->stack[f];
    for (int p = 0; p < 2; ++p) {
        a.d[p] = (const T *)s->d[p];
        a.I[p] = (const T *)s->I[p];
        a.wqB[p] = (const T *)s->wqB[p];
        a.qx[p] = (const T *)s->qx[p];
        a.qy[p] = (const T *)s->qy[p];
    }
    a.vras = (const T *)s->snap_vras;
    a.vtot = (const T *)s->snap_vtot;
    a.n2s = (const T *)s->n2s;
    a.elev = s->elem == 8 ? (const double *)s->z : nullptr;
    a.hi = c.flag_human_instability == 1;
    a.hi_weight = c.hi_weight_kg;
    a.hi_w1 = c.hi_width1_m;
    a.hi_w2 = c.hi_width2_m;
    a.hi_mu = c.hi_mu;
    a.hi_Cd = c.hi_Cd;
    a.hi_ro = c.hi_ro_water;
    a.hi_g = c.hi_gravity;
    if (s->gw) {
        a.gw_h = s->gwa.h;
        a.gw_elev = s->gwa.elev;
        a.gw_sd = s->gwa.soil_depth;
        a.gw_sd_u = s->gwa.soil_depth_u;
    }
    a.wqPc = (const T *)s->wqPc;
    a.out_idx = s->out_idx;
    a.out_q = (const T *)s->out_q;
    a.n_outlet = s->n_outlet;
    a.R = c.resolution;
    return a;
}

template <typename T> static hipError_t sim_after_li(hp2d_sim *s) {
    if (s->wq) {
        hipError_t e = launch_wq_step<T>(sim_wq_args<T>(s), s->stream);
        if (e != hipSuccess) return e;
    }
    hipError_t e = launch_finalize<T>(sim_fin_args<T>(s), s->stream);
    if (e != hipSuccess || s->snap_slots == 0) return e;
    return launch_snapshot<T>(sim_snap_args<T>(s), s->stream);
}

template <typename T> static CaSimArgs<T> sim_ca_args(hp2d_sim *s) {
    CaSimArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.li = sim_li_args<T>(s);
    a.dtot = (T *)s->ca_dtot;
    a.ql = (T *)s->ca_ql;
    a.qd = (T *)s->ca_qd;
    a.qo = (T *)s->ca_qo;
    a.itot = (T *)s->ca_itot;
    a.h0 = (const T *)s->ca_h0;
    a.blocks = s->ca_blocks;
    a.depth_tol_mm = T(s->cfg.depth_tolerance_mm);
    a.slope_outlet = T(s->cfg.slope_outlet);
    return a;
}

template <typename T> static hipError_t sim_step_launch(hp2d_sim *s, bool with_finalize) {
    hipError_t e;
    if (s->ca) {
        e = launch_ca_sim<T>(sim_ca_args<T>(s), s->stream);
    } else {
        LiArgs<T> a = sim_li_args<T>(s);
        e = launch_li_step<T>(a, kModeSim, s->stream);
    }
    if (e != hipSuccess || !with_finalize) return e;
    return sim_after_li<T>(s);
}

static hipError_t step_launch(hp2d_sim *s, bool with_finalize) {
    return s->elem == 8 ? sim_step_launch<double>(s, with_finalize)
                        : sim_step_launch<float>(s, with_finalize);
}

template <typename T> static std::vector<T> cast_vec(const double *src, size_t n, double shift = 0) {
    std::vector<T> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = T(src[i] - shift);
    return v;
}

extern "C" {

hp2d_status hp2d_sim_create(const hp2d_run_config *cfg, const hp2d_run_inputs *in,
                            hp2d_sim **out) {
    if (!cfg || !in || !out || cfg->ny < 1 || cfg->nx < 1 || !in->elevation || !in->roughness) {
        g_last_error = "hp2d_sim_create: bad arguments";
        return HP2D_E_ARG;
    }
    if ((double)cfg->ny * (double)cfg->nx > kMaxCells) {
        g_last_error = "hp2d_sim_create: grids are limited to 2^32 - 2^16 cells";
        return HP2D_E_ARG;
    }
    if (cfg->routing == HP2D_ROUTING_CA) {
        if (cfg->flag_waterquality == 1) {
            g_last_error = "hp2d_sim_create: CA routing in the loop runs without the wash-off";
            return HP2D_E_UNSUPPORTED;
        }
    } else if (cfg->routing != HP2D_ROUTING_INERTIAL || cfg->flag_numerical_scheme != 1) {
        g_last_error = "hp2d_sim_create: local-inertial routing with the Bates scheme, or CA routing";
        return HP2D_E_UNSUPPORTED;
    }
    if (cfg->flag_infiltration < 0 || cfg->flag_infiltration > 2) {
        g_last_error = "hp2d_sim_create: flag_infiltration must be 0, 1 (Green-Ampt) or 2 (van Genuchten)";
        return HP2D_E_ARG;
    }
    if (cfg->flag_groundwater == 1 && cfg->flag_infiltration == 0) {
        g_last_error = "hp2d_sim_create: the groundwater module needs the infiltration state I_t";
        return HP2D_E_UNSUPPORTED;
    }
    if (cfg->time_step_change_s <= 0 || cfg->time_step_increments_s <= 0 ||
        cfg->record_time_hydrographs_min <= 0 || cfg->routing_time_min <= 0 ||
        cfg->min_time_step_s <= 0) {
        g_last_error = "hp2d_sim_create: non-positive time control";
        return HP2D_E_ARG;
    }
    if (hp2d_device_count() < 1) {
        g_last_error = "hp2d_sim_create: no HIP device visible";
        return HP2D_E_NO_DEVICE;
    }
    hp2d_sim *s = new hp2d_sim();
    s->cfg = *cfg;
    s->ny = cfg->ny;
    s->nx = cfg->nx;
    s->n = (size_t)cfg->ny * cfg->nx;
    s->elem = cfg->precision == HP2D_FP64 ? 8 : 4;
    s->ca = cfg->routing == HP2D_ROUTING_CA;
    s->tj = default_tj(cfg->nx);
    s->rows = li_rows_ok(cfg->ny, 2) ? 2 : 1;
    const size_t n = s->n;
    const int ny = s->ny;
    const double R = cfg->resolution, A = R * R;
    auto fail = [&](hp2d_status st, const std::string &m) {
        g_last_error = m;
        hp2d_sim_destroy(s);
        return st;
    };
    // ---- host preparation ----
    std::vector<double> z(in->elevation, in->elevation + n), n2(n), d0(n, 0.0), I0(n, 0.0);
    std::vector<long long> oidx;
    double zmin = INFINITY;
    for (size_t c = 0; c < n; ++c) {
        const bool nan_cell = std::isnan(z[c]);
        if (!nan_cell) {
            zmin = std::min(zmin, z[c]);
            s->n_valid += 1;
            s->drainage_area += A;
        }
        n2[c] = in->roughness[c] * in->roughness[c];
        d0[c] = in->d0 ? in->d0[c] : 0.0;
        if (nan_cell) d0[c] = NAN;
        else if (!(d0[c] >= 0)) return fail(HP2D_E_ARG, "hp2d_sim_create: negative initial depth");
        if (in->outlet_mask && in->outlet_mask[c]) oidx.push_back((long long)c);
    }
    if (s->n_valid == 0) return fail(HP2D_E_ARG, "hp2d_sim_create: no valid cell");
    for (int i = 0; i < in->n_rain; ++i)
        if (!(in->rain_intensity_mmh[i] >= 0))
            return fail(HP2D_E_ARG, "hp2d_sim_create: negative rainfall intensity");
    // rainfall maps: their sums over the valid cells (the step's rain volume and
    // the report's P), strictly increasing stamps spaced at least max_dt / 3 so
    // that no step spans more than kMaxMapTerms of them
    std::vector<double> map_sum;
    if (cfg->n_rain_maps > 0) {
        if (!in->rain_maps_mmh || !in->rain_map_time_min)
            return fail(HP2D_E_ARG, "hp2d_sim_create: n_rain_maps > 0 without the maps");
        const int K = cfg->n_rain_maps;
        for (int k = 1; k < K; ++k)
            if (!((in->rain_map_time_min[k] - in->rain_map_time_min[k - 1]) * 60.0 >=
                  cfg->max_time_step_s / 3.0))
                return fail(HP2D_E_ARG, "hp2d_sim_create: rainfall map stamps must increase by at "
                                        "least max_time_step / 3");
        map_sum.assign(K, 0.0);
        for (int k = 0; k < K; ++k)
            for (size_t c = 0; c < n; ++c) {
                if (std::isnan(z[c])) continue;
                const double v = in->rain_maps_mmh[(size_t)k * n + c];
                if (!(v >= 0)) return fail(HP2D_E_ARG, "hp2d_sim_create: negative or NaN rainfall map value");
                map_sum[k] += v;
            }
    }
    // inflow hydrographs: the valid cells of each gauge
    std::vector<double> gauge_cells;
    if (cfg->flag_inflow == 1) {
        const int G = cfg->n_inflow_gauges;
        if (G < 1 || G > HP2D_MAX_INFLOW_GAUGES || !in->inflow_gauge || !in->inflow_time_min ||
            !in->inflow_rate_mmh || in->n_inflow_times < 1)
            return fail(HP2D_E_ARG, "hp2d_sim_create: inflow hydrographs incomplete (1 to 16 gauges)");
        gauge_cells.assign(G, 0.0);
        for (size_t c = 0; c < n; ++c) {
            const int g = in->inflow_gauge[c];
            if (g > G) return fail(HP2D_E_ARG, "hp2d_sim_create: inflow gauge index beyond n_inflow_gauges");
            if (g > 0 && !std::isnan(z[c])) gauge_cells[g - 1] += 1.0;
        }
    }
    if (in->rain_pattern)
        for (size_t c = 0; c < n; ++c) {
            if (std::isnan(z[c])) continue;
            if (!(in->rain_pattern[c] >= 0))
                return fail(HP2D_E_ARG, "hp2d_sim_create: negative rainfall pattern");
            s->sum_pattern += in->rain_pattern[c];
        }
    std::vector<double> n2s = n2;
    std::vector<double> on(oidx.size());
    for (size_t k = 0; k < oidx.size(); ++k) {
        n2s[oidx[k]] = std::copysign(std::fabs(n2s[oidx[k]]), -1.0);
        on[k] = in->roughness[oidx[k]];
    }
    // Green-Ampt classes (HydroPol2D_preprocessing.m:2249-2253)
    std::vector<GaClass> gac;
    const bool ga = cfg->flag_infiltration == 1;
    if (ga) {
        if (cfg->n_soil_classes < 1 || cfg->n_soil_classes > kMaxSoilClasses || !in->ksat_mmh ||
            !in->psi_mm || !in->dtheta ||
            !in->theta_sat_minus_r)
            return fail(HP2D_E_ARG, "hp2d_sim_create: Green-Ampt tables missing");
        for (int k = 0; k < cfg->n_soil_classes; ++k) {
            GaClass g;
            g.ksat = in->ksat_mmh[k];
            g.psi = in->psi_mm[k];
            g.dtheta = in->dtheta[k];
            const double kr = (1.0 / 75.0 * (std::sqrt(g.ksat / 25.4)));
            g.Tr = 4.5 / std::sqrt(g.ksat / 25.4);
            double Lu = 4.0 * std::sqrt(g.ksat / 25.4);
            Lu = Lu * 2.54 / 100.0;
            if (in->lu_m && !std::isnan(in->lu_m[k])) Lu = in->lu_m[k];
            g.k_out = in->theta_sat_minus_r[k] * kr * Lu * 1000.0;
            g.Lu_mm = Lu * 1000.0;
            gac.push_back(g);
        }
    }
    // van Genuchten classes and the depth to the water table
    // (Infiltration_Module.m:194-198: zwt = Soil_Depth - (h_t - (elevation - Soil_Depth)),
    // taken at max(zwt, 0); the head is gw_h0 until a groundwater update moves it)
    const bool vg = cfg->flag_infiltration == 2;
    const bool inf = ga || vg;
    std::vector<VgClass> vgc;
    std::vector<double> zw0;
    if (vg) {
        if (cfg->n_soil_classes < 1 || cfg->n_soil_classes > kMaxSoilClasses || !in->ksat_mmh ||
            !in->gw_theta_sat || !in->gw_theta_r || !in->gw_n_vg || !in->gw_alpha_vg ||
            !in->gw_h0 || !in->gw_soil_depth)
            return fail(HP2D_E_ARG, "hp2d_sim_create: van Genuchten tables, gw_h0 or gw_soil_depth missing");
        for (int k = 0; k < cfg->n_soil_classes; ++k) {
            VgClass g;
            g.ts = in->gw_theta_sat[k];
            g.tr = in->gw_theta_r[k];
            g.nv = in->gw_n_vg[k];
            g.alpha = in->gw_alpha_vg[k];
            g.ksat = in->ksat_mmh[k];
            if (!(g.nv > 1.0) || !(g.alpha > 0.0) || !(g.ksat >= 0.0))
                return fail(HP2D_E_ARG, "hp2d_sim_create: van Genuchten classes need n > 1, alpha > 0, ksat >= 0");
            vgc.push_back(g);
        }
        s->vg_l = in->gw_l_vg;
        zw0.resize(n);
        for (size_t c = 0; c < n; ++c) {
            const double sd = in->gw_soil_depth[c];
            const double gw_depth = in->gw_h0[c] - (z[c] - sd);
            const double zwt = std::fmax(sd - gw_depth, 0.0);
            zw0[c] = (in->impervious && in->impervious[c]) ? std::copysign(zwt, -1.0) : zwt;
        }
    }
    if (inf) {
        if (in->soil_class)
            for (size_t c = 0; c < n; ++c)
                if (in->soil_class[c] >= cfg->n_soil_classes)
                    return fail(HP2D_E_ARG, "hp2d_sim_create: soil class out of range");
        for (size_t c = 0; c < n; ++c) {
            I0[c] = in->I0_mm ? in->I0_mm[c] : in->I_min_mm;
            if (std::isnan(z[c])) I0[c] = NAN;
        }
    }
    // water quality: class tables, exponential build-up (HydroPol2D_preprocessing.m:2170-2174)
    // and initial_mass (:2319)
    s->wq = cfg->flag_waterquality == 1;
    std::vector<double> B0, wc3, wc4;
    std::vector<uint16_t> wcell;
    double initial_mass = 0.0;
    if (s->wq) {
        const int k = cfg->n_lulc_classes;
        if (k < 1 || k > 256 || !in->wq_C1 || !in->wq_C2 || !in->wq_C3 || !in->wq_C4)
            return fail(HP2D_E_ARG, "hp2d_sim_create: water-quality class tables missing");
        if (cfg->flag_wq_model != 0 && cfg->flag_wq_model != 1)
            return fail(HP2D_E_ARG, "hp2d_sim_create: flag_wq_model must be 0 or 1");
        B0.resize(n);
        wcell.resize(n);
        wc3.assign(in->wq_C3, in->wq_C3 + k);
        wc4.assign(in->wq_C4, in->wq_C4 + k);
        std::vector<uint8_t> is_out(n, 0);
        for (long long c : oidx) is_out[(size_t)c] = 1;
        for (size_t c = 0; c < n; ++c) {
            const int cls = in->lulc_class ? in->lulc_class[c] : 0;
            if (cls >= k) return fail(HP2D_E_ARG, "hp2d_sim_create: LULC class out of range");
            const bool nan_cell = std::isnan(z[c]);
            const double b0 = in->wq_C1[cls] * (1.0 - std::pow(2.718281828459045, -in->wq_C2[cls] * in->wq_ADD_days));
            B0[c] = b0 * A / 1e4 * (nan_cell ? 0.0 : 1.0);
            wcell[c] = (uint16_t)(cls | (nan_cell ? kWqNanBit : 0u) | (is_out[c] ? kWqOutletBit : 0u));
        }
        for (size_t c = 0; c < n; ++c)  // sum(sum(B_t(~isinf(B_t)))), column-major
            if (!std::isinf(B0[c])) initial_mass += B0[c];
    }
    // groundwater module: per-cell van Genuchten maps from the soil classes
    s->gw = cfg->flag_groundwater == 1;
    double S_GW = 0.0;
    std::vector<double> gw_host;  // 9 input maps, then h0
    if (s->gw) {
        const int k = cfg->n_soil_classes;
        if (!in->gw_h0 || !in->gw_soil_depth || !in->gw_theta_sat || !in->gw_theta_r ||
            !in->gw_n_vg || !in->gw_alpha_vg || cfg->gw_every_steps < 1 || ny < 2 || s->nx < 2)
            return fail(HP2D_E_ARG, "hp2d_sim_create: groundwater inputs missing");
        gw_host.resize(10 * n);
        for (size_t c = 0; c < n; ++c) {
            const int cls = in->soil_class ? in->soil_class[c] : 0;
            (void)k;
            gw_host[0 * n + c] = z[c];
            gw_host[1 * n + c] = in->gw_soil_depth[c];
            gw_host[2 * n + c] = in->gw_theta_sat[cls];
            gw_host[3 * n + c] = in->gw_theta_r[cls];
            gw_host[4 * n + c] = in->gw_n_vg[cls];
            gw_host[5 * n + c] = in->gw_alpha_vg[cls];
            gw_host[6 * n + c] = in->ksat_mmh[cls];
            gw_host[7 * n + c] = in->gw_ksat_mmh;
            gw_host[8 * n + c] = in->gw_Sy;
            gw_host[9 * n + c] = in->gw_h0[c];
        }
        // Main_While.m:258 S_GW = nansum(nansum(A .* Sy .* (h_t - (z - D))))
        for (int j = 0; j < s->nx; ++j) {
            double col = 0.0;
            for (int i = 0; i < ny; ++i) {
                const size_t c = (size_t)i + (size_t)j * ny;
                const double v = A * in->gw_Sy * (in->gw_h0[c] - (z[c] - in->gw_soil_depth[c]));
                if (!std::isnan(v)) col += v;
            }
            S_GW += col;
        }
    }
    // initial storage (Main_While.m:248-260), column-major nansum order
    double S_p = 0, S_UZ = 0;
    for (int j = 0; j < s->nx; ++j) {
        double cp = 0, cu = 0;
        for (int i = 0; i < ny; ++i) {
            const size_t c = (size_t)i + (size_t)j * ny;
            const double v = A * d0[c] / 1000.0;
            if (!std::isnan(v)) cp += v;
            if (inf) {
                const double u = A * I0[c] / 1000.0;
                if (!std::isnan(u)) cu += u;
            }
        }
        S_p += cp;
        S_UZ += cu;
    }
    // ---- device allocation ----
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    const size_t nb = n * s->elem;
    auto A_ = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    A_(&s->z, nb);
    if (s->elem == 4) A_(&s->zy, nb);
    if (cfg->n_rain_maps > 0) {
        s->n_maps = cfg->n_rain_maps;
        A_(&s->maps, nb * (size_t)s->n_maps);
        A_((void **)&s->map_t, (size_t)s->n_maps * 8);
        A_((void **)&s->map_sum, (size_t)s->n_maps * 8);
    }
    if (cfg->flag_human_instability == 1) A_(&s->max_risk, nb);
    if (cfg->record_time_maps_min > 0) {
        // one record at most per step, and at most kSnapSlots steps per graph
        s->snap_slots = kSnapSlots;
        A_(&s->snap_vras, nb);
        A_(&s->snap_vtot, nb);
        A_((void **)&s->snap_t, kSnapSlots * sizeof(double));
        const bool on[HP2D_MAP_FIELDS] = {true, true, true, cfg->flag_human_instability == 1,
                                          inf, s->gw, s->wq, s->wq, s->wq};
        for (int f = 0; f < HP2D_MAP_FIELDS; ++f)
            if (on[f]) A_(&s->stack[f], nb * kSnapSlots);
    }
    if (cfg->flag_inflow == 1) {
        s->n_gauges = cfg->n_inflow_gauges;
        s->n_inflow_times = in->n_inflow_times;
        A_((void **)&s->inflow_t, (size_t)s->n_inflow_times * 8);
        A_((void **)&s->inflow_v, (size_t)s->n_inflow_times * s->n_gauges * 8);
        A_((void **)&s->gauge_cells, (size_t)s->n_gauges * 8);
        A_((void **)&s->gauge, n);
    }
    A_(&s->n2s, nb);
    for (int p = 0; p < 2; ++p) {
        A_(&s->d[p], nb);
        A_(&s->qx[p], nb);
        A_(&s->qy[p], nb);
        if (inf) A_(&s->I[p], nb);
        if (ga) A_(&s->Tc[p], nb);
    }
    A_(&s->dmax, nb);
    if (in->rain_pattern) A_(&s->pattern, nb);
    if (inf && in->soil_class) A_((void **)&s->soil, n);
    if (vg) {
        A_((void **)&s->vg, vgc.size() * sizeof(VgClass));
        A_((void **)&s->vg_f32, vgc.size() * 8 * sizeof(float));
        A_(&s->vg_zwt, nb);
    }
    if (ga) A_((void **)&s->ga, gac.size() * sizeof(GaClass));
    if (ga) A_((void **)&s->ga_f32, gac.size() * 8 * sizeof(float));
    s->n_outlet = (int)oidx.size();
    if (s->n_outlet > 0) {
        A_((void **)&s->out_idx, oidx.size() * sizeof(long long));
        A_(&s->out_n, oidx.size() * s->elem);
        A_(&s->out_q, oidx.size() * s->elem);
        A_(&s->out_d, oidx.size() * s->elem);
    }
    const int n_slabs = (ny + kRowsOut - 1) / kRowsOut;
    s->n_partials = n_slabs * ((s->nx + 16 - 1) / 16);  // capacity for the smallest tj
    s->part_cap = s->n_partials;
    A_((void **)&s->partials, (size_t)s->n_partials * kPartCount * sizeof(double));
    if (s->ca) {
        // one lane per cell, in at most as many workgroups as there are partials
        s->ca_blocks = (int)std::min<long long>({(long long)((n + 255) / 256), (long long)s->n_partials, 8192LL});
        A_(&s->ca_dtot, nb);
        A_(&s->ca_ql, nb);
        A_(&s->ca_qd, nb);
        A_(&s->ca_qo, nb);
        A_(&s->ca_itot, nb);
        if (in->ca_h0_mm) A_(&s->ca_h0, nb);
    }
    A_((void **)&s->st, sizeof(SimState));
    s->n_rain = in->n_rain;
    A_((void **)&s->rain_t, std::max(1, in->n_rain) * sizeof(double));
    A_((void **)&s->rain_v, std::max(1, in->n_rain) * sizeof(double));
    s->n_hydro_max = (int)std::floor(cfg->routing_time_min / cfg->record_time_hydrographs_min + 1e-12) + 1;
    A_((void **)&s->hq, s->n_hydro_max * sizeof(double));
    A_((void **)&s->ht, s->n_hydro_max * sizeof(double));
    A_((void **)&s->hd, s->n_hydro_max * sizeof(double));
    A_((void **)&s->dt_hist, s->n_dt_max * sizeof(double));
    A_((void **)&s->vol_err, s->n_rep_max * sizeof(double));
    A_((void **)&s->tune_sink, sizeof(double));
    if (s->wq) {
        for (int p = 0; p < 2; ++p) {
            A_(&s->wqB[p], nb);
            A_(&s->wqTW[p], nb);
            if (cfg->flag_wq_model == 1) A_(&s->wqBx[p], nb);
        }
        A_(&s->wqBs, nb);
        A_(&s->wqOm, nb);
        A_(&s->wqPc, nb);
        A_(&s->wqC3, wc3.size() * s->elem);
        A_(&s->wqC4, wc4.size() * s->elem);
        A_((void **)&s->wq_cell, n * sizeof(uint16_t));
        s->wq_blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
        A_((void **)&s->wq_part, (size_t)s->wq_blocks * kWqFieldCount * sizeof(double));
        A_((void **)&s->wq_ctl, sizeof(WqCtl));
    }
    if (s->gw) {
        // 22 FP64 maps, then the soil-class tables (5 x kMaxSoilClasses)
        A_((void **)&s->gw_mem, (22 * n + 5 * (size_t)kMaxSoilClasses) * sizeof(double));
        A_((void **)&s->gw_col, 3 * (size_t)s->nx * sizeof(double));
        A_((void **)&s->gw_dom, n);
    }
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? HP2D_E_NOMEM : HP2D_E_HIP,
                    std::string("hp2d_sim_create: ") + hipGetErrorString(e));
    if (s->gw) {
        const hp2d_status hs = hp2d_create(ny, s->nx, HP2D_FP64, &s->gwh);
        if (hs != HP2D_OK) return fail(hs, "hp2d_sim_create: groundwater handle");
    }
    // ---- uploads ----
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    auto up_real = [&](void *dst, const std::vector<double> &v, double shift) {
        if (s->elem == 8) {
            std::vector<double> w(v.size());
            for (size_t i = 0; i < v.size(); ++i) w[i] = v[i] - shift;
            up(dst, w.data(), w.size() * 8);
        } else {
            std::vector<float> w = cast_vec<float>(v.data(), v.size(), shift);
            up(dst, w.data(), w.size() * 4);
        }
    };
    if (s->elem == 8) {
        up_real(s->z, z, 0.0);
        up_real(s->n2s, n2s, 0.0);
    } else {
        // FP32: the face elevation differences, and NaN n^2 outside the domain
        std::vector<float> dzx, dzy;
        face_dz(ny, s->nx, z.data(), dzx, dzy);
        up(s->z, dzx.data(), n * 4);
        up(s->zy, dzy.data(), n * 4);
        std::vector<double> n2m = n2s;
        for (size_t c = 0; c < n; ++c)
            if (std::isnan(z[c])) n2m[c] = NAN;
        up_real(s->n2s, n2m, 0.0);
    }
    std::vector<double> zeros(n, 0.0);
    for (int p = 0; p < 2; ++p) {
        up_real(s->d[p], d0, 0.0);
        up_real(s->qx[p], zeros, 0.0);
        up_real(s->qy[p], zeros, 0.0);
        if (inf) up_real(s->I[p], I0, 0.0);
        if (ga) up_real(s->Tc[p], zeros, 0.0);
    }
    up_real(s->dmax, d0, 0.0);
    if (s->ca) {
        // CA_States.I_tot_end_cell starts at 0; the outflows of a step are all written
        up_real(s->ca_itot, zeros, 0.0);
        up_real(s->ca_ql, zeros, 0.0);
        up_real(s->ca_qd, zeros, 0.0);
        up_real(s->ca_qo, zeros, 0.0);
        if (s->ca_h0) up_real(s->ca_h0, std::vector<double>(in->ca_h0_mm, in->ca_h0_mm + n), 0.0);
    }
    if (in->rain_pattern) up_real(s->pattern, std::vector<double>(in->rain_pattern, in->rain_pattern + n), 0.0);
    if (s->soil) up(s->soil, in->soil_class, n);
    if (ga) {
        up(s->ga, gac.data(), gac.size() * sizeof(GaClass));
        std::vector<float> gf(gac.size() * 8, 0.0f);
        for (size_t k = 0; k < gac.size(); ++k) {
            const GaClass &g = gac[k];
            const float v[8] = {(float)g.ksat, (float)g.psi, (float)g.dtheta, (float)g.Lu_mm,
                                (float)g.k_out, (float)g.Tr, 0.0f, 0.0f};
            std::memcpy(&gf[k * 8], v, sizeof(v));
        }
        up(s->ga_f32, gf.data(), gf.size() * sizeof(float));
    }
    if (vg) {
        up(s->vg, vgc.data(), vgc.size() * sizeof(VgClass));
        std::vector<float> vf(vgc.size() * 8, 0.0f);
        for (size_t k = 0; k < vgc.size(); ++k) {
            const VgClass &g = vgc[k];
            const double m = 1.0 - 1.0 / g.nv;
            const float v[8] = {(float)g.tr, (float)g.ts, (float)(1.0 / g.nv), (float)m,
                                (float)(1.0 / m), (float)(-1.0 / m), (float)(1.0 / g.alpha),
                                (float)(g.ksat / 1000.0 / 3600.0)};
            std::memcpy(&vf[k * 8], v, sizeof(v));
        }
        up(s->vg_f32, vf.data(), vf.size() * sizeof(float));
        up_real(s->vg_zwt, zw0, 0.0);
    }
    if (s->n_outlet > 0) {
        up(s->out_idx, oidx.data(), oidx.size() * sizeof(long long));
        up_real(s->out_n, on, 0.0);
        std::vector<double> zq(oidx.size(), 0.0);
        up_real(s->out_q, zq, 0.0);
        up_real(s->out_d, zq, 0.0);
    }
    if (in->n_rain > 0) {
        up(s->rain_t, in->rain_time_min, in->n_rain * sizeof(double));
        up(s->rain_v, in->rain_intensity_mmh, in->n_rain * sizeof(double));
    }
    if (s->n_maps > 0) {
        for (int k = 0; k < s->n_maps; ++k) {
            std::vector<double> m(in->rain_maps_mmh + (size_t)k * n, in->rain_maps_mmh + (size_t)(k + 1) * n);
            for (size_t c = 0; c < n; ++c)
                if (std::isnan(z[c])) m[c] = NAN;
            up_real((char *)s->maps + (size_t)k * nb, m, 0.0);
        }
        up(s->map_t, in->rain_map_time_min, (size_t)s->n_maps * 8);
        up(s->map_sum, map_sum.data(), (size_t)s->n_maps * 8);
    }
    if (s->n_gauges > 0) {
        up(s->inflow_t, in->inflow_time_min, (size_t)s->n_inflow_times * 8);
        up(s->inflow_v, in->inflow_rate_mmh, (size_t)s->n_inflow_times * s->n_gauges * 8);
        up(s->gauge_cells, gauge_cells.data(), (size_t)s->n_gauges * 8);
        up(s->gauge, in->inflow_gauge, n);
    }
    if (s->wq) {
        for (int p = 0; p < 2; ++p) {
            up_real(s->wqB[p], B0, 0.0);
            up_real(s->wqTW[p], zeros, 0.0);
        }
        up_real(s->wqPc, zeros, 0.0);
        up_real(s->wqC3, wc3, 0.0);
        up_real(s->wqC4, wc4, 0.0);
        up(s->wq_cell, wcell.data(), n * sizeof(uint16_t));
        s->wq_min_Bt = in->wq_min_Bt;
        s->wq_Bmin = in->wq_Bmin;
        s->wq_Bmax = in->wq_Bmax;
    }
    if (s->gw) {
        double *g = s->gw_mem;
        auto arr = [&](int k) { return g + (size_t)k * n; };
        up(g, gw_host.data(), 9 * n * sizeof(double));          // elev .. Sy
        up(arr(9), gw_host.data() + 9 * n, n * sizeof(double));  // h
        std::vector<double> pend(n);
        for (size_t c = 0; c < n; ++c) pend[c] = std::isnan(z[c]) ? NAN : 0.0;
        up(arr(14), pend.data(), n * sizeof(double));
        up(arr(15), pend.data(), n * sizeof(double));
        up(arr(16), pend.data(), n * sizeof(double));
        // the aquifer bed is static: set once here, so the module's recharge pass
        // does not rewrite it every call
        std::vector<double> zbed(n);
        for (size_t c = 0; c < n; ++c) zbed[c] = gw_host[0 * n + c] - gw_host[1 * n + c];
        up(arr(17), zbed.data(), n * sizeof(double));
        std::vector<uint8_t> dom(n);
        for (size_t c = 0; c < n; ++c) dom[c] = std::isnan(z[c]) ? 0 : 1;
        up(s->gw_dom, dom.data(), n);
        GwmArgs &a = s->gwa;
        a.ny = ny;
        a.nx = s->nx;
        a.n = (long long)n;
        a.elev = arr(0); a.soil_depth = arr(1); a.ts = arr(2); a.tr = arr(3); a.nv = arr(4);
        a.alpha = arr(5); a.ksat = arr(6);
        // a soil depth that is one value over the grid is passed as that value
        bool uniform_depth = true;
        for (size_t c = 1; c < n && uniform_depth; ++c)
            uniform_depth = std::memcmp(&in->gw_soil_depth[c], &in->gw_soil_depth[0], sizeof(double)) == 0;
        if (uniform_depth) {
            a.soil_depth = nullptr;
            a.soil_depth_u = in->gw_soil_depth[0];
        }
        if (s->soil) {
            // soil-class runs: the recharge and post passes read the class byte and
            // five per-class values instead of five FP64 maps
            double *tab = g + 22 * n;
            std::vector<double> t(5 * (size_t)kMaxSoilClasses, 0.0);
            for (int k = 0; k < cfg->n_soil_classes; ++k) {
                t[0 * kMaxSoilClasses + k] = in->gw_theta_sat[k];
                t[1 * kMaxSoilClasses + k] = in->gw_theta_r[k];
                t[2 * kMaxSoilClasses + k] = in->gw_n_vg[k];
                t[3 * kMaxSoilClasses + k] = in->gw_alpha_vg[k];
                t[4 * kMaxSoilClasses + k] = in->ksat_mmh[k];
            }
            up(tab, t.data(), t.size() * sizeof(double));
            a.soil_class = s->soil;
            a.cls_ts = tab;
            a.cls_tr = tab + kMaxSoilClasses;
            a.cls_nv = tab + 2 * kMaxSoilClasses;
            a.cls_alpha = tab + 3 * kMaxSoilClasses;
            a.cls_ksat = tab + 4 * kMaxSoilClasses;
        }
        // the aquifer's conductivity and specific yield are uniform in a run:
        // the kernels take the values instead of reading two FP64 maps per pass
        a.ksat_gw = nullptr; a.Sy = nullptr;
        a.ksat_gw_u = in->gw_ksat_mmh;
        a.Sy_u = in->gw_Sy;
        a.Kgw_u = in->gw_ksat_mmh / 1000.0 / 3600.0;
        // the cumulative recharge map is returned by the per-call entry point only
        a.h = arr(9); a.I = arr(10); a.d = arr(11); a.cum = nullptr; a.maxd = nullptr;
        a.statics_ready = 1;
        // the pending recharge / capillary maps are returned by the per-call entry
        // point but read by nothing in the loop: only the net exchange map is kept
        a.pn = arr(14); a.pr = nullptr; a.pc = nullptr; a.z_bed = arr(17); a.Kgw = nullptr;
        a.rfg = arr(19); a.q = arr(20); a.t2 = arr(21);
        hp2d_gw_params &p = s->gwp;
        p.ny = ny;
        p.nx = s->nx;
        p.resolution = R;
        p.routing_time_min = cfg->routing_time_min;
        p.flag_baseflow = 1;
        p.flag_async = 0;
        p.target_dt_min = 1440.0;
        p.min_dt_min = 1.0;
        p.max_head_change_m = 0.25;
        p.courant = cfg->gw_courant;
        p.l_vg = in->gw_l_vg;
    }
    std::vector<double> hz(s->n_hydro_max, 0.0);
    up(s->hq, hz.data(), hz.size() * 8);
    up(s->ht, hz.data(), hz.size() * 8);
    up(s->hd, hz.data(), hz.size() * 8);
    // ---- initial time control (Main_While.m:109-127, preprocessing.m:2469,2615,2618) ----
    SimState st;
    std::memset(&st, 0, sizeof(st));
    const double min_dt0 = cfg->min_time_step_s;
    st.k = 1;
    st.time_step_min = min_dt0 / 60.0;
    st.t = min_dt0 / 60.0;
    st.progress = (st.t / cfg->routing_time_min) * 100.0;
    st.t_previous = 0.0;
    st.cursor = 1;
    st.delta_p_agg = rain_depth(in->rain_time_min, in->rain_intensity_mmh, in->n_rain, 0.0, st.t,
                                &st.cursor);
    st.map_cursor = st.inflow_cursor = 1;
    st.map_last = 1;  // recording_parameters.last_record_maps (HydroPol2D_preprocessing.m:2396)
    if (s->n_maps > 0) {
        MapTerms mt;
        if (!map_terms(in->rain_map_time_min, s->n_maps, 0.0, st.t, &st.map_cursor, &mt))
            return fail(HP2D_E_ARG, "hp2d_sim_create: too many rainfall maps in the first step");
        st.map_n = mt.n;
        for (int j = 0; j < kMaxMapTerms; ++j) {
            st.map_idx[j] = j < mt.n ? mt.idx[j] : 0;
            st.map_w[j] = j < mt.n ? mt.w[j] : 0.0;
        }
        st.map_dtot = mt.dt_total;
        st.map_k = mt.k;
    }
    if (s->n_gauges > 0)
        inflow_depths(in->inflow_time_min, in->inflow_rate_mmh, in->n_inflow_times, s->n_gauges,
                      0.0, st.t, &st.inflow_cursor, st.inflow_agg);
    st.t_previous = 0.0;
    st.max_dt = std::min(cfg->record_time_hydrographs_min * 60.0, cfg->max_time_step_s);
    st.min_dt = std::min(cfg->record_time_hydrographs_min * 60.0, min_dt0);
    st.time_save_previous = 0.0;
    st.S_prev = 0.0 + S_p + S_UZ + S_GW;
    st.gw_storage = S_GW;
    st.wq_initial_mass = initial_mass;
    st.catch_index = 1;
    st.next_report = cfg->report_every_percent;
    st.last_record = 1;
    st.inject_k = -1;
    st.stop_accepted = -1;
    st.tmin_wq = INFINITY;
    st.done = !(st.t <= (cfg->routing_time_min + st.min_dt / 60.0)) ? 1 : 0;
    up(s->st, &st, sizeof(st));
    if (e != hipSuccess)
        return fail(HP2D_E_HIP, std::string("hp2d_sim_create upload: ") + hipGetErrorString(e));
    s->I_min = in->I_min_mm;
    s->err.clear();
    *out = s;
    return HP2D_OK;
}

void hp2d_sim_destroy(hp2d_sim *s) {
    if (!s) return;
    if (s->stream) hipStreamSynchronize(s->stream);
    if (s->graph) hipGraphExecDestroy(s->graph);
    void *ps[] = {s->z, s->zy, s->n2s, s->pattern, s->dmax, s->d[0], s->d[1], s->qx[0], s->qx[1],
                  s->qy[0], s->qy[1], s->I[0], s->I[1], s->Tc[0], s->Tc[1], s->soil, s->ga, s->ga_f32,
                  s->vg, s->vg_f32, s->vg_zwt, s->ca_dtot, s->ca_ql, s->ca_qd, s->ca_qo,
                  s->ca_itot, s->ca_h0,
                  s->out_idx, s->out_n, s->out_q, s->out_d, s->partials, s->st, s->rain_t,
                  s->rain_v, s->hq, s->ht, s->hd, s->dt_hist, s->vol_err, s->tune_sink,
                  s->tpeak, s->wqB[0], s->wqB[1], s->wqBs, s->wqBx[0], s->wqBx[1], s->wqOm,
                  s->wqTW[0], s->wqTW[1], s->wqPc, s->wqC3, s->wqC4, s->wq_cell, s->wq_part,
                  s->wq_ctl, s->gw_mem, s->gw_col, s->gw_dom, s->maps, s->map_t, s->map_sum,
                  s->inflow_t, s->inflow_v, s->gauge_cells, s->gauge, s->max_risk, s->snap_vras,
                  s->snap_vtot, s->snap_t, s->stack[0], s->stack[1], s->stack[2], s->stack[3],
                  s->stack[4], s->stack[5], s->stack[6], s->stack[7], s->stack[8]};
    for (void *p : ps)
        if (p) hipFree(p);
    if (s->gwh) hp2d_destroy(s->gwh);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
}

const char *hp2d_sim_last_error(const hp2d_sim *s) { return s ? s->err.c_str() : g_last_error.c_str(); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Running the loop
// ---------------------------------------------------------------------------
static hp2d_status read_state(hp2d_sim *s, SimState &st) {
    STRY(s, hipStreamSynchronize(s->stream));
    STRY(s, hipMemcpy(&st, s->st, sizeof(SimState), hipMemcpyDeviceToHost));
    if (st.map_err) {
        set_err(s, "rainfall maps: a step spans more than 4 map intervals");
        return HP2D_E_ARG;
    }
    return HP2D_OK;
}

static hp2d_status write_state(hp2d_sim *s, const SimState &st) {
    STRY(s, hipMemcpy(s->st, &st, sizeof(SimState), hipMemcpyHostToDevice));
    return HP2D_OK;
}

static hp2d_status ensure_graph(hp2d_sim *s, int steps) {
    if (s->graph && s->graph_steps == steps) return HP2D_OK;
    if (s->graph) {
        hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    STRY(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; ++i) {
        hipError_t e = step_launch(s, true);
        if (e != hipSuccess) {
            hipGraph_t dummy = nullptr;
            hipStreamEndCapture(s->stream, &dummy);
            if (dummy) hipGraphDestroy(dummy);
            set_err(s, std::string("graph capture: ") + hipGetErrorString(e));
            return HP2D_E_HIP;
        }
    }
    STRY(s, hipStreamEndCapture(s->stream, &g));
    hipError_t e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_err(s, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        s->graph = nullptr;
        return HP2D_E_HIP;
    }
    s->graph_steps = steps;
    return HP2D_OK;
}

// A step whose wash-off split exceeded the captured sub-steps left the hold
// flag set (every later launch was a no-op): run its remaining sub-steps, the
// fold (which clears the flag) and its finalize.
static hp2d_status sim_release_hold(hp2d_sim *s, SimState &st) {
    if (st.wq_err) {
        set_err(s, "build_up_wash_off: tmin_wq asks for more than 1e7 internal sub-steps "
                   "(the reference would not terminate)");
        return HP2D_E_ARG;
    }
    WqCtl ctl;
    STRY(s, hipMemcpy(&ctl, s->wq_ctl, sizeof(WqCtl), hipMemcpyDeviceToHost));
    hipError_t e;
    if (s->elem == 8) {
        e = launch_wq_finish<double>(sim_wq_args<double>(s), ctl.steps, s->stream);
        if (e == hipSuccess) e = launch_finalize<double>(sim_fin_args<double>(s), s->stream);
        if (e == hipSuccess && s->snap_slots > 0)
            e = launch_snapshot<double>(sim_snap_args<double>(s), s->stream);
    } else {
        e = launch_wq_finish<float>(sim_wq_args<float>(s), ctl.steps, s->stream);
        if (e == hipSuccess) e = launch_finalize<float>(sim_fin_args<float>(s), s->stream);
        if (e == hipSuccess && s->snap_slots > 0)
            e = launch_snapshot<float>(sim_snap_args<float>(s), s->stream);
    }
    STRY(s, e);
    return read_state(s, st);
}

// Groundwater_Module over the model time since its last call, on the current
// depth and infiltration buffers (an FP32 run's buffers are read as FP64 and
// rounded back by the module's kernels; the head and its ledgers stay FP64).
static hp2d_status sim_gw_update(hp2d_sim *s, SimState &st) {
    const int cur = (int)((st.k + 1) & 1);  // the buffers the next step reads
    GwmArgs a = s->gwa;
    if (s->elem == 8) {
        a.d = (double *)s->d[cur];
        a.I = (double *)s->I[cur];
    } else {  // the module reads and writes the FP32 state in place
        a.d32 = (float *)s->d[cur];
        a.I32 = (float *)s->I[cur];
        a.fast_boussinesq = s->cfg.gw_exact ? 0 : 1;
    }
    s->gwp.time_step_min = st.gw_elapsed_min;
    s->gwp.t_min = st.t;
    double err = 0.0, sgw = 0.0;
    int32_t ran = 0;
    const hp2d_status rc = gwm_run(s->gwh, &s->gwp, &s->gws, a, s->gw_col, s->gw_dom, &err, &ran, &sgw);
    if (rc != HP2D_OK) {
        set_err(s, std::string("groundwater module: ") + hp2d_last_error(s->gwh));
        return rc;
    }
    if (s->vg_zwt) {
        // the head moved: the van Genuchten closure's water-table depth follows
        const GwmArgs &g = s->gwa;
        const hipError_t e = s->elem == 8
            ? launch_vg_water_table<double>(g.h, g.elev, g.soil_depth, g.soil_depth_u,
                                            (long long)s->n, (double *)s->vg_zwt, s->stream)
            : launch_vg_water_table<float>(g.h, g.elev, g.soil_depth, g.soil_depth_u,
                                           (long long)s->n, (float *)s->vg_zwt, s->stream);
        STRY(s, e);
    }
    s->gw_error = err;
    s->gw_last = st.accepted;
    st.gw_elapsed_min = 0.0;
    st.gw_storage = sgw;
    return write_state(s, st);
}

// Move the records the device stack holds to the host (the Maps structure of
// save_output_maps.m, kept in host memory as the reference keeps it).
static hp2d_status sim_drain_maps(hp2d_sim *s, SimState &st) {
    if (s->snap_slots == 0 || st.snap_n == 0) return HP2D_OK;
    const int nr = st.snap_n;
    std::vector<double> tt(nr);
    STRY(s, hipMemcpy(tt.data(), s->snap_t, nr * sizeof(double), hipMemcpyDeviceToHost));
    for (int f = 0; f < HP2D_MAP_FIELDS; ++f) {
        if (!s->stack[f]) continue;
        for (int r = 0; r < nr; ++r) {
            std::vector<double> v(s->n);
            if (s->elem == 8) {
                STRY(s, hipMemcpy(v.data(), (char *)s->stack[f] + (size_t)r * s->n * 8, s->n * 8,
                                  hipMemcpyDeviceToHost));
            } else {
                std::vector<float> w(s->n);
                STRY(s, hipMemcpy(w.data(), (char *)s->stack[f] + (size_t)r * s->n * 4, s->n * 4,
                                  hipMemcpyDeviceToHost));
                for (size_t c = 0; c < s->n; ++c) v[c] = w[c];
            }
            s->records[f].push_back(std::move(v));
        }
    }
    for (int r = 0; r < nr; ++r) s->record_t.push_back(tt[r]);
    st.snap_n = 0;
    return write_state(s, st);
}

// Advance by n_accept accepted steps (0: to the end of the run) in captured
// graphs of `batch` steps, pausing for the groundwater calls and releasing
// wash-off holds between graphs.
static hp2d_status sim_advance(hp2d_sim *s, int64_t n_accept, int batch, int32_t *done) {
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    const int64_t target = n_accept > 0 ? st.accepted + n_accept : -1;
    const int every = s->gw ? s->cfg.gw_every_steps : 0;
    if (every > 0) batch = std::min(batch, every);
    if (s->snap_slots > 0) batch = std::min(batch, s->snap_slots);
    const bool use_graph = batch >= 8 || (batch >= 1 && (s->wq || s->gw || s->snap_slots > 0));
    if (use_graph && (rc = ensure_graph(s, batch))) return rc;
    // every launch of a graph accepts, rejects or skips (paused / held) each of
    // its steps, and a release always completes the held step; a run longer
    // than this bound is a failure of the time control, not an endless loop
    const int64_t cap = (n_accept > 0 ? n_accept : (int64_t)1 << 40) * 64 + (1 << 20);
    int64_t launches = 0;
    while (!st.done && (target < 0 || st.accepted < target)) {
        if (++launches > cap) {
            set_err(s, "sim: the step loop made no progress");
            return HP2D_E_HIP;
        }
        if (every > 0 && st.accepted - s->gw_last >= every) {
            if ((rc = sim_gw_update(s, st))) return rc;
        }
        int64_t stop = target;
        if (every > 0) {
            const int64_t g = s->gw_last + every;
            stop = stop < 0 ? g : std::min(stop, g);
        }
        st.stop_accepted = stop;
        st.paused = 0;
        if ((rc = write_state(s, st))) return rc;
        if (use_graph) {
            STRY(s, hipGraphLaunch(s->graph, s->stream));
        } else {
            for (int i = 0; i < batch; ++i) STRY(s, step_launch(s, true));
        }
        if ((rc = read_state(s, st))) return rc;
        if (st.hold && (rc = sim_release_hold(s, st))) return rc;
        if ((rc = sim_drain_maps(s, st))) return rc;
    }
    st.stop_accepted = -1;
    st.paused = 0;
    if ((rc = write_state(s, st))) return rc;
    if (done) *done = st.done;
    return HP2D_OK;
}

extern "C" {

hp2d_status hp2d_sim_run(hp2d_sim *s, int64_t max_steps, int32_t *done) {
    if (!s || max_steps < 0) return HP2D_E_ARG;
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    if (st.done) {
        if (done) *done = 1;
        return HP2D_OK;
    }
    if (s->wq || s->gw || s->snap_slots > 0) {
        const int batch = (int)std::min<int64_t>(64, max_steps > 0 ? max_steps : 64);
        if ((rc = sim_advance(s, max_steps, batch, done))) return rc;
        if ((rc = read_state(s, st))) return rc;
        if (st.neg_events > 0 && s->err.empty())
            s->err = "negative depth below -1e-8 mm met before clean-up (Hydrological_Model.m:45)";
        return HP2D_OK;
    }
    st.stop_accepted = max_steps > 0 ? st.accepted + max_steps : -1;
    st.paused = 0;
    if ((rc = write_state(s, st))) return rc;
    const int batch = (int)std::min<int64_t>(64, max_steps > 0 ? max_steps : 64);
    const bool use_graph = batch >= 8;
    if (use_graph && (rc = ensure_graph(s, batch))) return rc;
    // every launch either accepts or rejects a step; rejected steps only come
    // from injected failures, so this bound always ends the loop.
    const int64_t cap = (max_steps > 0 ? max_steps : (int64_t)1 << 40) + 1024;
    int64_t launched = 0;
    while (launched < cap) {
        if (use_graph) {
            STRY(s, hipGraphLaunch(s->graph, s->stream));
        } else {
            for (int i = 0; i < batch; ++i) STRY(s, step_launch(s, true));
        }
        launched += batch;
        if ((rc = read_state(s, st))) return rc;
        if (st.done || st.paused) break;
    }
    st.stop_accepted = -1;
    st.paused = 0;
    if ((rc = write_state(s, st))) return rc;
    if (st.neg_events > 0 && s->err.empty())
        s->err = "negative depth below -1e-8 mm met before clean-up (Hydrological_Model.m:45)";
    if (done) *done = st.done;
    return HP2D_OK;
}

hp2d_status hp2d_sim_run_fixed(hp2d_sim *s, int64_t n_steps, int32_t graph_steps) {
    if (!s || n_steps < 0) return HP2D_E_ARG;
    hipEvent_t e0, e1, a0, a1;
    STRY(s, hipEventCreate(&e0));
    STRY(s, hipEventCreate(&e1));
    STRY(s, hipEventCreate(&a0));
    STRY(s, hipEventCreate(&a1));
    s->li_ms = 0;
    s->li_launches = 0;
    if (graph_steps > 0) {
        // the batch sim_advance uses (a groundwater call cuts every N steps), so
        // that the graph built here is the one it launches
        if (s->gw) graph_steps = std::min<int32_t>(graph_steps, s->cfg.gw_every_steps);
        if (s->snap_slots > 0) graph_steps = std::min<int32_t>(graph_steps, s->snap_slots);
        hp2d_status rc = ensure_graph(s, graph_steps);
        if (rc) return rc;
    }
    STRY(s, hipEventRecord(e0, s->stream));
    if (graph_steps > 0 && (s->wq || s->gw || s->snap_slots > 0)) {
        hp2d_status rc = sim_advance(s, n_steps, graph_steps, nullptr);
        if (rc) return rc;
    } else if (graph_steps > 0) {
        int64_t done_steps = 0;
        while (done_steps + graph_steps <= n_steps) {
            STRY(s, hipGraphLaunch(s->graph, s->stream));
            done_steps += graph_steps;
        }
        for (; done_steps < n_steps; ++done_steps) STRY(s, step_launch(s, true));
    } else {
        for (int64_t i = 0; i < n_steps; ++i) {
            if (s->gw) {  // a groundwater call opens the step that follows every N accepted ones
                SimState st;
                hp2d_status rc = read_state(s, st);
                if (!rc && st.accepted - s->gw_last >= s->cfg.gw_every_steps) rc = sim_gw_update(s, st);
                if (rc) return rc;
            }
            STRY(s, hipEventRecord(a0, s->stream));
            hipError_t e = step_launch(s, false);  // the routing step (LI or the CA passes)
            STRY(s, e);
            STRY(s, hipEventRecord(a1, s->stream));
            if (s->elem == 8) e = sim_after_li<double>(s);
            else e = sim_after_li<float>(s);
            STRY(s, e);
            STRY(s, hipEventSynchronize(a1));
            if (s->wq || s->snap_slots > 0) {  // a held wash-off split completes before the
                SimState st;                    // next launch; map records move to the host
                hp2d_status rc = read_state(s, st);
                if (!rc && st.hold) rc = sim_release_hold(s, st);
                if (!rc) rc = sim_drain_maps(s, st);
                if (rc) return rc;
            }
            float ms = 0;
            STRY(s, hipEventElapsedTime(&ms, a0, a1));
            s->li_ms += ms;
            s->li_launches += 1;
        }
    }
    STRY(s, hipEventRecord(e1, s->stream));
    STRY(s, hipEventSynchronize(e1));
    float tot = 0;
    STRY(s, hipEventElapsedTime(&tot, e0, e1));
    s->total_ms = tot;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(a0);
    hipEventDestroy(a1);
    return HP2D_OK;
}

hp2d_status hp2d_sim_kernel_time(hp2d_sim *s, double *li_ms, double *total_ms,
                                 int64_t *launches) {
    if (!s) return HP2D_E_ARG;
    if (li_ms) *li_ms = s->li_launches ? s->li_ms / (double)s->li_launches : 0.0;
    if (total_ms) *total_ms = s->total_ms;
    if (launches) *launches = s->li_launches;
    return HP2D_OK;
}

hp2d_status hp2d_sim_report(hp2d_sim *s, hp2d_run_report *rep) {
    if (!s || !rep) return HP2D_E_ARG;
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    rep->steps = st.accepted;
    rep->rejected_steps = st.rejected;
    rep->t_min = st.t;
    rep->time_step_s = st.time_step_min * 60.0;
    rep->max_velocity = st.max_velocity;
    rep->volume_error_m3 = st.volume_error;
    rep->n_reports = st.n_reports;
    rep->outflow_volume_m3 = st.outflow_volume;
    rep->inflow_volume_m3 = st.inflow_volume;
    rep->outlet_runoff_volume_mm = st.runoff_mm;
    rep->cum_rain_m3 = st.cum_rain;
    rep->cum_inflow_m3 = st.cum_inflow;
    rep->n_map_records = (int32_t)s->record_t.size();
    rep->cum_infiltration_m3 = st.cum_inf;
    rep->cum_outflow_m3 = st.cum_out;
    {
        // surface storage of the current state, FP64 (mass_balance_check.m:62)
        const int cur = (int)((st.k + 1) & 1);
        const int blocks = 1024;
        double *part = nullptr;
        STRY(s, hipMalloc((void **)&part, 2 * blocks * sizeof(double)));
        hipError_t e = s->elem == 8
                           ? launch_storage<double>((const double *)s->d[cur], nullptr,
                                                    (long long)s->n, part, blocks, s->stream)
                           : launch_storage<float>((const float *)s->d[cur], nullptr,
                                                   (long long)s->n, part, blocks, s->stream);
        std::vector<double> hp(2 * blocks);
        if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        hipFree(part);
        STRY(s, e);
        double sum = 0.0;
        for (int b = 0; b < blocks; ++b) sum += hp[2 * b];
        const double A = s->cfg.resolution * s->cfg.resolution;
        rep->storage_m3 = sum * (A / 1000.0);
        if (s->wq) {  // sum of the current B_t
            const void *Bc = s->wqB[st.wq_n & 1];
            STRY(s, hipMalloc((void **)&part, 2 * blocks * sizeof(double)));
            e = s->elem == 8 ? launch_storage<double>((const double *)Bc, nullptr, (long long)s->n,
                                                      part, blocks, s->stream)
                             : launch_storage<float>((const float *)Bc, nullptr, (long long)s->n,
                                                     part, blocks, s->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            hipFree(part);
            STRY(s, e);
            double sb = 0.0;
            for (int b = 0; b < blocks; ++b) sb += hp[2 * b];
            rep->wq_mass_kg = sb;
        }
    }
    rep->catch_index = st.catch_index;
    if (s->wq) {
        rep->wq_mass_lost_kg = st.wq_mass_lost;
        rep->wq_out_conc_mgl = st.wq_out_conc;
        rep->wq_tmin_s = st.tmin_wq;
        rep->wq_mass_outlet_mg = st.wq_mass_outlet;
        rep->wq_vol_outlet = st.wq_vol_outlet;
        rep->wq_substep_steps = st.wq_split_steps;
    }
    if (s->gw) {
        rep->gw_storage_m3 = st.gw_storage;
        rep->gw_total_recharge_m3 = s->gws.total_recharge_m3;
        rep->gw_total_exfiltration_m3 = s->gws.total_exfiltration_m3;
        rep->gw_error_m3 = s->gw_error;
        rep->gw_updates = s->gws.n_updates;
    }
    return HP2D_OK;
}


static hp2d_status fetch_real(hp2d_sim *s, double *dst, const void *src, double shift) {
    if (!dst || !src) return HP2D_OK;
    if (s->elem == 8) {
        STRY(s, hipMemcpy(dst, src, s->n * 8, hipMemcpyDeviceToHost));
        if (shift != 0)
            for (size_t i = 0; i < s->n; ++i) dst[i] += shift;
    } else {
        std::vector<float> tmp(s->n);
        STRY(s, hipMemcpy(tmp.data(), src, s->n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < s->n; ++i) dst[i] = (double)tmp[i] + shift;
    }
    return HP2D_OK;
}

hp2d_status hp2d_sim_get_state(hp2d_sim *s, double *d_t, double *outflow, double *Hf,
                               double *I_t, double *dmax, double *B_t, double *gw_h) {
    if (!s) return HP2D_E_ARG;
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    if (Hf || (B_t && !s->wq) || (gw_h && !s->gw)) {
        set_err(s, "Hf is not kept by the device loop; B_t / gw_h need water quality / groundwater");
        return HP2D_E_UNSUPPORTED;
    }
    if (B_t && (rc = fetch_real(s, B_t, s->wqB[st.wq_n & 1], 0.0))) return rc;
    if (gw_h) STRY(s, hipMemcpy(gw_h, s->gwa.h, s->n * 8, hipMemcpyDeviceToHost));
    const int cur = (int)((st.k + 1) & 1);  // buffer holding the last accepted state
    if ((rc = fetch_real(s, d_t, s->d[cur], 0.0))) return rc;
    if (outflow) {
        if ((rc = fetch_real(s, outflow, s->qx[cur], 0.0))) return rc;
        if ((rc = fetch_real(s, outflow + s->n, s->qy[cur], 0.0))) return rc;
        std::vector<double> q3(s->n, 0.0);
        if (s->n_outlet > 0) {
            std::vector<long long> idx(s->n_outlet);
            STRY(s, hipMemcpy(idx.data(), s->out_idx, idx.size() * 8, hipMemcpyDeviceToHost));
            std::vector<double> q(s->n_outlet);
            if (s->elem == 8) {
                STRY(s, hipMemcpy(q.data(), s->out_q, q.size() * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> qf(s->n_outlet);
                STRY(s, hipMemcpy(qf.data(), s->out_q, qf.size() * 4, hipMemcpyDeviceToHost));
                for (size_t k = 0; k < qf.size(); ++k) q[k] = qf[k];
            }
            for (size_t k = 0; k < idx.size(); ++k) q3[idx[k]] = q[k];
        }
        std::memcpy(outflow + 2 * s->n, q3.data(), s->n * 8);
    }
    if (I_t) {
        if (!s->I[cur]) {
            set_err(s, "no infiltration state in this run");
            return HP2D_E_STATE;
        }
        if ((rc = fetch_real(s, I_t, s->I[cur], 0.0))) return rc;
    }
    if ((rc = fetch_real(s, dmax, s->dmax, 0.0))) return rc;
    return HP2D_OK;
}

hp2d_status hp2d_sim_get_wq_maps(hp2d_sim *s, double *P_conc, double *Tot_Washed) {
    if (!s) return HP2D_E_ARG;
    if (!s->wq) {
        set_err(s, "hp2d_sim_get_wq_maps: water quality is off for this run");
        return HP2D_E_UNSUPPORTED;
    }
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    if (P_conc && (rc = fetch_real(s, P_conc, s->wqPc, 0.0))) return rc;
    if (Tot_Washed && (rc = fetch_real(s, Tot_Washed, s->wqTW[st.wq_n & 1], 0.0))) return rc;
    return HP2D_OK;
}

hp2d_status hp2d_sim_get_hydrograph(hp2d_sim *s, double *q_m3s, double *time_min,
                                    double *depth_outlet_mm, int32_t n_max, int32_t *n_out) {
    if (!s) return HP2D_E_ARG;
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    const int n = std::min(n_max, s->n_hydro_max);
    if (q_m3s) STRY(s, hipMemcpy(q_m3s, s->hq, n * 8, hipMemcpyDeviceToHost));
    if (time_min) STRY(s, hipMemcpy(time_min, s->ht, n * 8, hipMemcpyDeviceToHost));
    if (depth_outlet_mm) STRY(s, hipMemcpy(depth_outlet_mm, s->hd, n * 8, hipMemcpyDeviceToHost));
    if (n_out) *n_out = std::min(st.n_hydro, n);
    return HP2D_OK;
}

hp2d_status hp2d_sim_get_dt_history(hp2d_sim *s, double *dt_s, int32_t n_max, int32_t *n_out) {
    if (!s || !dt_s) return HP2D_E_ARG;
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    std::vector<double> ring(s->n_dt_max);
    STRY(s, hipMemcpy(ring.data(), s->dt_hist, ring.size() * 8, hipMemcpyDeviceToHost));
    const long long total = st.accepted;
    const long long avail = std::min<long long>(total, s->n_dt_max);
    const long long take = std::min<long long>(avail, n_max);
    for (long long i = 0; i < take; ++i) {
        const long long step = total - take + i;
        dt_s[i] = ring[step % s->n_dt_max];
    }
    if (n_out) *n_out = (int32_t)take;
    return HP2D_OK;
}

hp2d_status hp2d_sim_get_max_risk(hp2d_sim *s, double *max_risk) {
    if (!s || !max_risk) return HP2D_E_ARG;
    if (!s->max_risk) {
        set_err(s, "flag_human_instability is off in this run");
        return HP2D_E_STATE;
    }
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    return fetch_real(s, max_risk, s->max_risk, 0.0);
}

hp2d_status hp2d_sim_get_map_records(hp2d_sim *s, int32_t field, int32_t first, int32_t count,
                                     double *out, double *t_min, int32_t *n_records) {
    if (!s || field < 0 || field >= HP2D_MAP_FIELDS || first < 0 || count < 0) return HP2D_E_ARG;
    const int32_t total = (int32_t)s->record_t.size();
    if (n_records) *n_records = total;
    if (count == 0) return HP2D_OK;
    if (first + count > total) {
        set_err(s, "map records: the range passes the records taken");
        return HP2D_E_ARG;
    }
    if (out) {
        if (s->records[field].empty()) {
            set_err(s, "map records: this field is not recorded in this run");
            return HP2D_E_STATE;
        }
        for (int32_t r = 0; r < count; ++r)
            std::memcpy(out + (size_t)r * s->n, s->records[field][first + r].data(), s->n * 8);
    }
    if (t_min)
        for (int32_t r = 0; r < count; ++r) t_min[r] = s->record_t[first + r];
    return HP2D_OK;
}

hp2d_status hp2d_debug_inject_failure(hp2d_sim *s, int64_t step_k, int32_t kind) {
    if (!s || kind != 1) return HP2D_E_ARG;
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    st.inject_k = step_k;
    return write_state(s, st);
}

}  // extern "C"

namespace hp2d {
hipError_t launch_bates(const double *q_p, double dt, const double *Hf, const double *S,
                        const double *n_sq, long long n_cells, int n_pages, double *q,
                        hipStream_t s);
}

extern "C" hp2d_status hp2d_inertial_solver(hp2d_handle *h, int32_t flag_numerical_scheme,
                                            const double *q_p, double dt, const double *Hf,
                                            const double *S, const double *n_sq, int32_t n_pages,
                                            double *q_out) {
    if (!h || !q_p || !Hf || !S || !n_sq || !q_out || n_pages < 1) return HP2D_E_ARG;
    if (flag_numerical_scheme != 1) {
        set_err(h, "only flag_numerical_scheme = 1 (Bates) is implemented");
        return HP2D_E_UNSUPPORTED;
    }
    const size_t n = (size_t)h->ny * h->nx, np = n * n_pages;
    double *dq = nullptr, *dH = nullptr, *dS = nullptr, *dn = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc((void **)&dq, np * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&dH, np * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&dS, np * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&dn, n * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, np * 8);
    if (e == hipSuccess) e = hipMemcpyAsync(dq, q_p, np * 8, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dH, Hf, np * 8, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dS, S, np * 8, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dn, n_sq, n * 8, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hp2d::launch_bates(dq, dt, dH, dS, dn, (long long)n, n_pages, dout, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(q_out, dout, np * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    for (double *p : {dq, dH, dS, dn, dout})
        if (p) hipFree(p);
    if (e != hipSuccess) {
        set_err(h, std::string("hp2d_inertial_solver: ") + hipGetErrorString(e));
        return HP2D_E_HIP;
    }
    return HP2D_OK;
}

// Entry points whose kernels live in their own translation units are
// declared in include/hp2d.h and defined there (hp2d_ca.hip, hp2d_wq.hip,
// hp2d_gw.hip).

// ---------------------------------------------------------------------------
// Launch-shape autotuner
// ---------------------------------------------------------------------------
namespace {
struct TuneKey {
    int dev, ny, nx, elem;
    bool operator<(const TuneKey &o) const {
        return std::tie(dev, ny, nx, elem) < std::tie(o.dev, o.ny, o.nx, o.elem);
    }
};
std::mutex g_tune_mu;
struct Shape {
    int tj, wpb, rows;
};
std::map<TuneKey, Shape> g_tune_cache;
}  // namespace

extern "C" hp2d_status hp2d_sim_autotune(hp2d_sim *s, int32_t *best_tj, int32_t *best_wpb,
                                         double *best_ms) {
    if (!s) return HP2D_E_ARG;
    if (s->ca) {
        set_err(s, "hp2d_sim_autotune: CA runs have one launch shape (one lane per cell)");
        return HP2D_E_UNSUPPORTED;
    }
    int dev = 0;
    STRY(s, hipGetDevice(&dev));
    const TuneKey key{dev, s->ny, s->nx, (int)s->elem};
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            s->tj = it->second.tj;
            s->wpb = it->second.wpb;
            s->rows = it->second.rows;
            if (s->graph) {
                hipGraphExecDestroy(s->graph);
                s->graph = nullptr;
            }
            if (best_tj) *best_tj = s->tj;
            if (best_wpb) *best_wpb = s->wpb;
            if (best_ms) *best_ms = -1.0;
            return HP2D_OK;
        }
    }
    SimState st;
    hp2d_status rc = read_state(s, st);
    if (rc) return rc;
    if (st.done) {
        set_err(s, "autotune needs a sim that has not finished");
        return HP2D_E_STATE;
    }
    // The trials run the run's next step as it will run: the step kernel reads
    // the current parity of the state and writes the other one, and the maps it
    // raises (depth, time-to-peak and risk maxima) take the values the real
    // next step gives them (a maximum is idempotent), so repeating it changes
    // nothing the next step does not; the finalize kernel runs dry (its
    // reduction only, into a sink), so no clock, ledger or dt moves.
    const long long saved_stop = st.stop_accepted;
    const int saved_paused = st.paused;
    st.stop_accepted = -1;
    st.paused = 0;
    if ((rc = write_state(s, st))) return rc;
    hipEvent_t e0, e1;
    STRY(s, hipEventCreate(&e0));
    STRY(s, hipEventCreate(&e1));
    // the candidate shapes, each timed in three interleaved passes (so a
    // slow moment of the device does not fall on one shape only), ranked by
    // the median of its passes
    std::vector<Shape> shapes;
    for (int rows : {1, 2}) {
        if (!li_rows_ok(s->ny, rows)) continue;
        for (int tj : {16, 32, 64, 128}) {
            if (tj > 16 && (s->nx + tj - 1) / tj < 2) continue;
            for (int wpb : {1, 2, 4}) shapes.push_back({tj, wpb, rows});
        }
    }
    // time the fused kernel plus the reduction of its partials
    auto tune_step = [&]() -> hipError_t {
        hipError_t e = step_launch(s, false);
        if (e != hipSuccess) return e;
        if (s->elem == 8) {
            FinArgs<double> f = sim_fin_args<double>(s);
            f.dry_run = 1;
            f.partials_out = s->tune_sink;
            return launch_finalize<double>(f, s->stream);
        }
        FinArgs<float> f = sim_fin_args<float>(s);
        f.dry_run = 1;
        f.partials_out = s->tune_sink;
        return launch_finalize<float>(f, s->stream);
    };
    constexpr int kPasses = 3, kReps = 6;
    std::vector<std::vector<double>> per(shapes.size());
    for (int pass = 0; pass < kPasses; ++pass) {
        for (size_t k = 0; k < shapes.size(); ++k) {
            s->tj = shapes[k].tj;
            s->wpb = shapes[k].wpb;
            s->rows = shapes[k].rows;
            for (int i = 0; i < 2; ++i) STRY(s, tune_step());
            STRY(s, hipEventRecord(e0, s->stream));
            for (int i = 0; i < kReps; ++i) STRY(s, tune_step());
            STRY(s, hipEventRecord(e1, s->stream));
            STRY(s, hipEventSynchronize(e1));
            float ms = 0;
            STRY(s, hipEventElapsedTime(&ms, e0, e1));
            per[k].push_back(ms / kReps);
        }
    }
    double best = 1e30;
    Shape bs = shapes.empty() ? Shape{s->tj, s->wpb, s->rows} : shapes[0];
    for (size_t k = 0; k < shapes.size(); ++k) {
        std::sort(per[k].begin(), per[k].end());
        const double med = per[k][kPasses / 2];
        if (med < best) {
            best = med;
            bs = shapes[k];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    s->tj = bs.tj;
    s->wpb = bs.wpb;
    s->rows = bs.rows;
    if (s->graph) {
        hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    st.stop_accepted = saved_stop;
    st.paused = saved_paused;
    if ((rc = write_state(s, st))) return rc;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune_cache[key] = bs;
    }
    if (best_tj) *best_tj = bs.tj;
    if (best_wpb) *best_wpb = bs.wpb;
    if (best_ms) *best_ms = best;
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_sim_launch_shape(hp2d_sim *s, int32_t *tj, int32_t *wpb, int32_t *rows) {
    if (!s) return HP2D_E_ARG;
    if (tj) *tj = s->tj;
    if (wpb) *wpb = s->wpb;
    if (rows) *rows = s->rows;
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_sim_set_launch_shape(hp2d_sim *s, int32_t tj, int32_t wpb, int32_t rows) {
    if (!s) return HP2D_E_ARG;
    if (s->ca) {
        set_err(s, "hp2d_sim_set_launch_shape: CA runs have one launch shape (one lane per cell)");
        return HP2D_E_UNSUPPORTED;
    }
    if (tj < 16 || tj > 4096 || (wpb != 1 && wpb != 2 && wpb != 4) || !li_rows_ok(s->ny, rows)) {
        set_err(s, "hp2d_sim_set_launch_shape: tj must be 16..4096, wpb 1, 2 or 4, rows 1 (or 2 for an even ny)");
        return HP2D_E_ARG;
    }
    s->tj = tj;
    s->wpb = wpb;
    s->rows = rows;
    if (s->graph) {  // captured launches carry the old shape
        hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    return HP2D_OK;
}

// ============================================================================
// Time to peak (derived output: no reference counterpart)
// ============================================================================
extern "C" {

hp2d_status hp2d_sim_record_time_to_peak(hp2d_sim *s) {
    if (!s) return HP2D_E_ARG;
    if (s->tpeak) return HP2D_OK;
    STRY(s, hipMalloc(&s->tpeak, s->n * s->elem));
    STRY(s, hipMemsetAsync(s->tpeak, 0, s->n * s->elem, s->stream));
    if (s->graph) {  // captured launches carry the old arguments
        hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
        s->graph_steps = 0;
    }
    return HP2D_OK;
}

hp2d_status hp2d_sim_get_time_to_peak(hp2d_sim *s, double *tpeak_min) {
    if (!s || !tpeak_min) return HP2D_E_ARG;
    if (!s->tpeak) {
        set_err(s, "hp2d_sim_record_time_to_peak was not called");
        return HP2D_E_ARG;
    }
    STRY(s, hipStreamSynchronize(s->stream));
    return fetch_real(s, tpeak_min, s->tpeak, 0.0);
}

}  // extern "C"

// ============================================================================
// Ensemble: many independent runs advanced by one batched launch per phase
// ============================================================================
struct hp2d_ensemble {
    std::vector<hp2d_sim *> m;
    hipStream_t stream = nullptr;
    std::string err;
    size_t elem = 4;
    int wpb = 4, critical = 0;
    void *li_args = nullptr, *fin_args = nullptr;  // per-member argument tables
    int *tile_grid = nullptr, *first_tile = nullptr, *active_ids = nullptr;
    int *done = nullptr;
    SimState **states = nullptr;
    std::vector<int> active;
    int n_tiles = 0;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    double last_ms = 0;
};

static void set_err(hp2d_ensemble *e, const std::string &m) {
    if (e) e->err = m;
    g_last_error = m;
}

#define ETRY(e, expr)                                                                 \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            set_err(e, std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return HP2D_E_HIP;                                                        \
        }                                                                             \
    } while (0)

namespace {

__global__ void gather_done_kernel(SimState *const *states, int n, int *done) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // (a member paused at its groundwater call is not done: the host makes the
    // call and releases it before the next chunk)
    if (i < n) done[i] = states[i]->done ? 1 : 0;
}

int member_tiles(const hp2d_sim *s) {
    return ((li_n_slabs(s->ny, s->rows) + s->wpb - 1) / s->wpb) *
           ((s->nx + s->tj - 1) / s->tj);
}

// tile table of the active members: tiles of one member are contiguous
hp2d_status build_tiles(hp2d_ensemble *e) {
    std::vector<int> tg, first(e->m.size(), 0);
    for (int id : e->active) {
        first[id] = (int)tg.size();
        tg.insert(tg.end(), member_tiles(e->m[id]), id);
    }
    e->n_tiles = (int)tg.size();
    if (!tg.empty())
        ETRY(e, hipMemcpyAsync(e->tile_grid, tg.data(), tg.size() * 4, hipMemcpyHostToDevice,
                               e->stream));
    ETRY(e, hipMemcpyAsync(e->first_tile, first.data(), first.size() * 4, hipMemcpyHostToDevice,
                           e->stream));
    if (!e->active.empty())
        ETRY(e, hipMemcpyAsync(e->active_ids, e->active.data(), e->active.size() * 4,
                               hipMemcpyHostToDevice, e->stream));
    ETRY(e, hipStreamSynchronize(e->stream));
    if (e->graph) {
        hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
        e->graph_steps = 0;
    }
    return HP2D_OK;
}

// the wash-off of each active member that has one, after the batched step
// (HydroPol2D_Main_While.m:683-695): the member's own kernels and arguments
template <typename T> hipError_t ensemble_wq(hp2d_ensemble *e) {
    for (int id : e->active) {
        hp2d_sim *m = e->m[id];
        if (!m->wq) continue;
        const hipError_t r = launch_wq_step<T>(sim_wq_args<T>(m), e->stream);
        if (r != hipSuccess) return r;
    }
    return hipSuccess;
}

hipError_t ensemble_step(hp2d_ensemble *e) {
    hipError_t r;
    const int na = (int)e->active.size();
    if (e->elem == 4) {
        r = launch_li_step_batch<float>((const LiArgs<float> *)e->li_args, e->tile_grid,
                                        e->first_tile, e->n_tiles, e->wpb, e->critical, e->stream);
        if (r == hipSuccess) r = ensemble_wq<float>(e);
        if (r == hipSuccess)
            r = launch_finalize_batch<float>((const FinArgs<float> *)e->fin_args, e->active_ids, na,
                                             e->stream);
    } else {
        r = launch_li_step_batch<double>((const LiArgs<double> *)e->li_args, e->tile_grid,
                                         e->first_tile, e->n_tiles, e->wpb, e->critical, e->stream);
        if (r == hipSuccess) r = ensemble_wq<double>(e);
        if (r == hipSuccess)
            r = launch_finalize_batch<double>((const FinArgs<double> *)e->fin_args, e->active_ids,
                                              na, e->stream);
    }
    return r;
}

hp2d_status ensemble_graph(hp2d_ensemble *e, int steps) {
    if (e->graph && e->graph_steps == steps) return HP2D_OK;
    if (e->graph) {
        hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    ETRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; ++i) {
        hipError_t r = ensemble_step(e);
        if (r != hipSuccess) {
            hipGraph_t dummy = nullptr;
            hipStreamEndCapture(e->stream, &dummy);
            if (dummy) hipGraphDestroy(dummy);
            set_err(e, std::string("ensemble graph capture: ") + hipGetErrorString(r));
            return HP2D_E_HIP;
        }
    }
    ETRY(e, hipStreamEndCapture(e->stream, &g));
    hipError_t r = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (r != hipSuccess) {
        e->graph = nullptr;
        set_err(e, std::string("hipGraphInstantiate: ") + hipGetErrorString(r));
        return HP2D_E_HIP;
    }
    e->graph_steps = steps;
    return HP2D_OK;
}

}  // namespace

extern "C" {

void hp2d_ensemble_destroy(hp2d_ensemble *e) {
    if (!e) return;
    if (e->stream) hipStreamSynchronize(e->stream);
    if (e->graph) hipGraphExecDestroy(e->graph);
    void *ps[] = {e->li_args, e->fin_args, e->tile_grid, e->first_tile, e->active_ids, e->done,
                  e->states};
    for (void *p : ps)
        if (p) hipFree(p);
    for (hp2d_sim *s : e->m) hp2d_sim_destroy(s);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

const char *hp2d_ensemble_last_error(const hp2d_ensemble *e) {
    return e ? e->err.c_str() : g_last_error.c_str();
}

hp2d_status hp2d_ensemble_create(int32_t n_members, const hp2d_run_config *cfgs,
                                 const hp2d_run_inputs *inputs, int32_t cols_per_wave,
                                 hp2d_ensemble **out) {
    if (!out || n_members < 1 || !cfgs || !inputs) {
        g_last_error = "hp2d_ensemble_create: bad arguments";
        return HP2D_E_ARG;
    }
    if (cols_per_wave != 0 && cols_per_wave != 16 && cols_per_wave != 32 && cols_per_wave != 64 &&
        cols_per_wave != 128) {
        g_last_error = "hp2d_ensemble_create: cols_per_wave must be 0, 16, 32, 64 or 128";
        return HP2D_E_ARG;
    }
    for (int i = 0; i < n_members; ++i) {
        if (cfgs[i].record_time_maps_min > 0 || cfgs[i].routing != HP2D_ROUTING_INERTIAL) {
            g_last_error = "hp2d_ensemble_create: members run the local-inertial step (no CA "
                           "routing or map stacks in the batched launch)";
            return HP2D_E_UNSUPPORTED;
        }
    }
    for (int i = 1; i < n_members; ++i) {
        if (cfgs[i].precision != cfgs[0].precision ||
            cfgs[i].flag_critical != cfgs[0].flag_critical) {
            g_last_error = "hp2d_ensemble_create: members must share precision and flag_critical";
            return HP2D_E_ARG;
        }
    }
    hp2d_ensemble *e = new hp2d_ensemble();
    e->elem = cfgs[0].precision == HP2D_FP64 ? 8 : 4;
    e->critical = cfgs[0].flag_critical;
    for (int i = 0; i < n_members; ++i) {
        hp2d_sim *s = nullptr;
        hp2d_status rc = hp2d_sim_create(&cfgs[i], &inputs[i], &s);
        if (rc != HP2D_OK) {
            e->err = "member " + std::to_string(i) + ": " + g_last_error;
            g_last_error = e->err;
            hp2d_ensemble_destroy(e);
            return rc;
        }
        s->wpb = e->wpb;
        s->rows = 1;  // the batched kernel holds one row per lane
        if (cols_per_wave) s->tj = cols_per_wave;
        e->m.push_back(s);
    }
    const size_t nm = e->m.size();
    size_t tiles = 0;
    for (hp2d_sim *s : e->m) {
        const int nt = member_tiles(s);
        if ((long long)nt > s->part_cap) {
            set_err(e, "member partials smaller than its tile count");
            hp2d_ensemble_destroy(e);
            return HP2D_E_ARG;
        }
        tiles += nt;
    }
    const size_t la = e->elem == 4 ? sizeof(LiArgs<float>) : sizeof(LiArgs<double>);
    const size_t fa = e->elem == 4 ? sizeof(FinArgs<float>) : sizeof(FinArgs<double>);
    std::vector<char> hl(la * nm), hf(fa * nm);
    std::vector<SimState *> hs(nm);
    for (size_t i = 0; i < nm; ++i) {
        if (e->elem == 4) {
            LiArgs<float> a = sim_li_args<float>(e->m[i]);
            FinArgs<float> f = sim_fin_args<float>(e->m[i]);
            std::memcpy(&hl[i * la], &a, la);
            std::memcpy(&hf[i * fa], &f, fa);
        } else {
            LiArgs<double> a = sim_li_args<double>(e->m[i]);
            FinArgs<double> f = sim_fin_args<double>(e->m[i]);
            std::memcpy(&hl[i * la], &a, la);
            std::memcpy(&hf[i * fa], &f, fa);
        }
        hs[i] = e->m[i]->st;
    }
    hipError_t r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (r == hipSuccess) r = hipMalloc(&e->li_args, la * nm);
    if (r == hipSuccess) r = hipMalloc(&e->fin_args, fa * nm);
    if (r == hipSuccess) r = hipMalloc((void **)&e->tile_grid, std::max<size_t>(tiles, 1) * 4);
    if (r == hipSuccess) r = hipMalloc((void **)&e->first_tile, nm * 4);
    if (r == hipSuccess) r = hipMalloc((void **)&e->active_ids, nm * 4);
    if (r == hipSuccess) r = hipMalloc((void **)&e->done, nm * 4);
    if (r == hipSuccess) r = hipMalloc((void **)&e->states, nm * sizeof(SimState *));
    // member streams may still be uploading
    for (size_t i = 0; i < nm && r == hipSuccess; ++i) r = hipStreamSynchronize(e->m[i]->stream);
    if (r == hipSuccess) r = hipMemcpy(e->li_args, hl.data(), la * nm, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemcpy(e->fin_args, hf.data(), fa * nm, hipMemcpyHostToDevice);
    if (r == hipSuccess)
        r = hipMemcpy(e->states, hs.data(), nm * sizeof(SimState *), hipMemcpyHostToDevice);
    if (r != hipSuccess) {
        set_err(e, std::string("hp2d_ensemble_create: ") + hipGetErrorString(r));
        hp2d_ensemble_destroy(e);
        return r == hipErrorOutOfMemory ? HP2D_E_NOMEM : HP2D_E_HIP;
    }
    for (size_t i = 0; i < nm; ++i) e->active.push_back((int)i);
    hp2d_status rc = build_tiles(e);
    if (rc != HP2D_OK) {
        hp2d_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return HP2D_OK;
}

hp2d_status hp2d_ensemble_run(hp2d_ensemble *e, int64_t max_steps, int32_t graph_steps,
                              int32_t *n_done) {
    if (!e || max_steps < 0 || graph_steps < 0) return HP2D_E_ARG;
    hipEvent_t e0, e1;
    ETRY(e, hipEventCreate(&e0));
    ETRY(e, hipEventCreate(&e1));
    ETRY(e, hipEventRecord(e0, e->stream));
    int64_t steps = 0;
    std::vector<int> done(e->m.size(), 0);
    // drop members that finished before this call
    {
        hipLaunchKernelGGL(gather_done_kernel, dim3(((int)e->m.size() + 255) / 256), dim3(256), 0,
                           e->stream, (SimState *const *)e->states, (int)e->m.size(), e->done);
        ETRY(e, hipGetLastError());
        ETRY(e, hipMemcpyAsync(done.data(), e->done, done.size() * 4, hipMemcpyDeviceToHost,
                               e->stream));
        ETRY(e, hipStreamSynchronize(e->stream));
        std::vector<int> still;
        for (int id : e->active)
            if (!done[id]) still.push_back(id);
        if (still.size() != e->active.size()) {
            e->active.swap(still);
            hp2d_status rc = build_tiles(e);
            if (rc) return rc;
        }
    }
    // members with the groundwater module pause at their next call (every N
    // accepted steps of their own); a chunk is at most the shortest such cadence
    int gw_min = 0;
    for (hp2d_sim *m : e->m)
        if (m->gw) gw_min = gw_min == 0 ? m->cfg.gw_every_steps : std::min(gw_min, m->cfg.gw_every_steps);
    if (gw_min > 0 && graph_steps > gw_min) graph_steps = gw_min;
    const int64_t eager_chunk = gw_min > 0 ? std::min(64, gw_min) : 64;
    while (steps < max_steps && !e->active.empty()) {
        // groundwater calls that are due, then each such member's pause point
        for (int id : e->active) {
            hp2d_sim *m = e->m[id];
            if (!m->gw) continue;
            SimState st;
            hp2d_status rc = read_state(m, st);
            if (!rc && st.accepted - m->gw_last >= m->cfg.gw_every_steps) rc = sim_gw_update(m, st);
            if (rc) {
                set_err(e, "member " + std::to_string(id) + ": " + m->err);
                return rc;
            }
            st.stop_accepted = m->gw_last + m->cfg.gw_every_steps;
            st.paused = 0;
            if ((rc = write_state(m, st))) return rc;
            ETRY(e, hipStreamSynchronize(m->stream));
        }
        const int64_t chunk = graph_steps > 0 ? std::min<int64_t>(graph_steps, max_steps - steps)
                                              : std::min<int64_t>(eager_chunk, max_steps - steps);
        if (graph_steps > 0 && chunk == graph_steps) {
            hp2d_status rc = ensemble_graph(e, graph_steps);
            if (rc) return rc;
            ETRY(e, hipGraphLaunch(e->graph, e->stream));
        } else {
            for (int64_t k = 0; k < chunk; ++k) ETRY(e, ensemble_step(e));
        }
        steps += chunk;
        // a wash-off split longer than the captured sub-steps held its member:
        // the host completes the held step (as the single-run loop does)
        bool any_wq = false;
        for (int id : e->active) any_wq = any_wq || e->m[id]->wq;
        if (any_wq) {
            ETRY(e, hipStreamSynchronize(e->stream));
            for (int id : e->active) {
                hp2d_sim *m = e->m[id];
                if (!m->wq) continue;
                SimState st;
                hp2d_status rc = read_state(m, st);
                if (!rc && st.hold) rc = sim_release_hold(m, st);
                if (rc) {
                    set_err(e, "member " + std::to_string(id) + ": " + m->err);
                    return rc;
                }
            }
        }
        hipLaunchKernelGGL(gather_done_kernel, dim3(((int)e->m.size() + 255) / 256), dim3(256), 0,
                           e->stream, (SimState *const *)e->states, (int)e->m.size(), e->done);
        ETRY(e, hipGetLastError());
        ETRY(e, hipMemcpyAsync(done.data(), e->done, done.size() * 4, hipMemcpyDeviceToHost,
                               e->stream));
        ETRY(e, hipStreamSynchronize(e->stream));
        std::vector<int> still;
        for (int id : e->active)
            if (!done[id]) still.push_back(id);
        if (still.size() != e->active.size()) {
            e->active.swap(still);
            if (!e->active.empty()) {
                hp2d_status rc = build_tiles(e);
                if (rc) return rc;
            }
        }
    }
    ETRY(e, hipEventRecord(e1, e->stream));
    ETRY(e, hipEventSynchronize(e1));
    float ms = 0;
    ETRY(e, hipEventElapsedTime(&ms, e0, e1));
    e->last_ms = ms;
    for (hp2d_sim *m : e->m) {  // no pause point outlives the call
        if (!m->gw) continue;
        SimState st;
        hp2d_status rc = read_state(m, st);
        if (rc) return rc;
        st.stop_accepted = -1;
        st.paused = 0;
        if ((rc = write_state(m, st))) return rc;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (n_done) *n_done = (int32_t)(e->m.size() - e->active.size());
    return HP2D_OK;
}

hp2d_status hp2d_ensemble_member(hp2d_ensemble *e, int32_t i, hp2d_sim **sim) {
    if (!e || !sim || i < 0 || i >= (int32_t)e->m.size()) return HP2D_E_ARG;
    ETRY(e, hipStreamSynchronize(e->stream));
    *sim = e->m[i];
    return HP2D_OK;
}

hp2d_status hp2d_ensemble_record_time_to_peak(hp2d_ensemble *e) {
    if (!e) return HP2D_E_ARG;
    const size_t la = e->elem == 4 ? sizeof(LiArgs<float>) : sizeof(LiArgs<double>);
    std::vector<char> hl(la * e->m.size());
    for (size_t i = 0; i < e->m.size(); ++i) {
        hp2d_status rc = hp2d_sim_record_time_to_peak(e->m[i]);
        if (rc) {
            set_err(e, e->m[i]->err);
            return rc;
        }
        ETRY(e, hipStreamSynchronize(e->m[i]->stream));
        if (e->elem == 4) {
            LiArgs<float> a = sim_li_args<float>(e->m[i]);
            std::memcpy(&hl[i * la], &a, la);
        } else {
            LiArgs<double> a = sim_li_args<double>(e->m[i]);
            std::memcpy(&hl[i * la], &a, la);
        }
    }
    ETRY(e, hipMemcpy(e->li_args, hl.data(), hl.size(), hipMemcpyHostToDevice));
    if (e->graph) {
        hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
        e->graph_steps = 0;
    }
    return HP2D_OK;
}

hp2d_status hp2d_ensemble_elapsed(hp2d_ensemble *e, double *ms) {
    if (!e || !ms) return HP2D_E_ARG;
    *ms = e->last_ms;
    return HP2D_OK;
}

}  // extern "C"
