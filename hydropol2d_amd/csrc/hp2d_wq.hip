// hp2d_wq.hip - pollutant build-up / wash-off for gfx950: the four-direction
// build_up_wash_off.m:11-183 and the eight-direction build_up_wash_off_8D.m:1-143.
//
// One wash-off pass is two elementwise kernels over the column-major grid:
// wq_rates_kernel evaluates the five (nine) outflow rates of every cell (power
// law on the face discharge), wq_net_kernel gathers the four (eight)
// neighbours' rates into the net rate dW (and, inside a sub-step, advances B).
// The eight-direction routine differs in more than its pages: its dW is
// inflow minus outflow, its tmin_wq bounds the cells that gain mass (10 s
// when none does), and Tot_Washed accumulates max(B_begin - B_t, 0).
//
// The scalars the reference branches on (min dW, min B_t_mid, the tmin_wq
// candidate) are reduced per 256-thread block into a partials array that the
// host folds with NaN-omitting min, so one host synchronisation decides the
// sub-stepping and the sub-steps themselves run back to back on the stream.
//
// Four-direction face rates below zero (the inflow faces of the signed
// local-inertial outputs) carry no wash-off: q < 0 is taken as 0 before the
// power law (NaN stays NaN).  For the non-negative cellular-automata outputs
// this is the reference expression unchanged.  tot_q_out keeps the signed sum
// (:37).  The eight-direction pages (non-negative by construction) enter the
// power law as given.
#include <cmath>
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_handle.hpp"
#include "hp2d_internal.hpp"

namespace hp2d {

constexpr int kWqBlock = 256;
// block partial fields
enum WqPart { kWqMinDW = 0, kWqTminCand = 1, kWqMinMid = 2, kWqParts = 3 };

struct WqArgs {
    int ny, nx;
    long long n;
    double A, min_Bt, Bmin, Bmax;
    int model;
    const double *C3, *C4, *q5;   // q5: pages left, right, up, down, outlet(, NE, SE, SW, NW) [mm/h]
    const uint8_t *nan_mask;      // page 1 of idx_nan_5, or NULL
    double *B, *Bx;               // B_t [kg], B_t_extra [kg]
    double *W, *totW, *totQ, *dW; // W: 5 (9) pages [kg/h]
    double *part;                 // [block][kWqParts]
};

template <int NW>
__device__ __forceinline__ void block_min_store(double v, double *dst) {
    __shared__ double sh[NW];
    v = wave_fmin(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int k = 1; k < NW; ++k) m = fmin(m, sh[k]);
        *dst = m;
    }
}

// W_out_t (:24-33, :91-101), idx_bt (:17, :102-108), idx_nan_5 (:41-43, first pass only)
template <bool D8>
__global__ __launch_bounds__(kWqBlock) void wq_rates_kernel(WqArgs a, int first_pass) {
    constexpr int NP = D8 ? 9 : 5;
    const long long c = blockIdx.x * (long long)kWqBlock + threadIdx.x;
    if (c >= a.n) return;
    const double B = a.B[c];
    double f_bt = B;
    if (a.model == 1) {
        const double Br = a.Bmin / 1000.0 * a.A, Bm = a.Bmax / 1000.0 * a.A;
        double bx = a.Bx[c];
        if (B * 1000.0 / a.A <= a.Bmin) bx = 0.0;
        if (B * 1000.0 / a.A >= a.Bmax) bx = Bm;
        a.Bx[c] = bx;
        f_bt = 1.0 + fmax(bx - Br, 0.0);
    }
    const bool low = (B * 1000.0 / a.A < a.min_Bt) || isnan(B) || isinf(B);
    const bool masked = first_pass && a.nan_mask && a.nan_mask[c];
    double tot = 0.0, tq = 0.0;
    for (int p = 0; p < NP; ++p) {
        const double qs = a.q5[p * a.n + c];
        const double q = (!D8 && qs < 0.0) ? 0.0 : qs;
        double w = a.C3[c] * pow(q / 1000.0 / 3600.0 * a.A, a.C4[c]) * f_bt;
        if (low || masked) w = 0.0;
        a.W[p * a.n + c] = w;
        tot += w;
        tq += qs;
    }
    a.totW[c] = tot;
    if (first_pass) a.totQ[c] = masked ? 0.0 : tq;
}

// dW (:45-59, :110-120); with dt_s > 0 the sub-step mass update (:121-125).
// The block partial kWqMinDW holds min(dW) (four directions) or min(-dW)
// (eight directions), so that "min >= 0" is the large-timestep test of both.
template <bool D8>
__global__ __launch_bounds__(kWqBlock) void wq_net_kernel(WqArgs a, double dt_s, double *omass,
                                                          double *olmass) {
    const long long c = blockIdx.x * (long long)kWqBlock + threadIdx.x;
    double v = NAN;
    if (c < a.n) {
        const int ny = a.ny;
        const int i = (int)(c % ny), j = (int)(c / ny);
        const double in_l = j > 0 ? a.W[1 * a.n + c - ny] : 0.0;
        const double in_r = j < a.nx - 1 ? a.W[0 * a.n + c + ny] : 0.0;
        const double in_u = i > 0 ? a.W[3 * a.n + c - 1] : 0.0;
        const double in_d = i < ny - 1 ? a.W[2 * a.n + c + 1] : 0.0;
        double dW;
        if (D8) {  // build_up_wash_off_8D.m:34-47
            const long long n = a.n;
            const int nx = a.nx;
            const double in_ne = (i > 0 && j < nx - 1) ? a.W[7 * n + c - 1 + ny] : 0.0;
            const double in_se = (i < ny - 1 && j < nx - 1) ? a.W[8 * n + c + 1 + ny] : 0.0;
            const double in_sw = (i < ny - 1 && j > 0) ? a.W[5 * n + c + 1 - ny] : 0.0;
            const double in_nw = (i > 0 && j > 0) ? a.W[6 * n + c - 1 - ny] : 0.0;
            dW = (in_l + in_r + in_u + in_d + in_ne + in_se + in_sw + in_nw) - a.totW[c];
        } else {
            dW = a.totW[c] - (in_l + in_r + in_u + in_d);
        }
        a.dW[c] = dW;
        if (dt_s > 0.0) {
            omass[c] = omass[c] + a.totW[c] * dt_s / 3600.0;
            olmass[c] = olmass[c] + a.W[4 * a.n + c] * dt_s / 3600.0;
            a.B[c] = D8 ? a.B[c] + dW * dt_s / 3600.0 : a.B[c] - dW * dt_s / 3600.0;
        }
        v = D8 ? -dW : dW;
    }
    block_min_store<kWqBlock / 64>(v, &a.part[blockIdx.x * kWqParts + kWqMinDW]);
}

// B_t_mid = B_t - dW*time_step/60 (:79) and the tmin_wq candidate of the
// current B_t (:61-76): min over cells of m/(|min(dW,0)| + 1e-16); eight
// directions: B_t + dW*time_step/60 and m/(|max(dW,0)| + 1e-16) (8D.m:48-61)
template <bool D8>
__global__ __launch_bounds__(kWqBlock) void wq_mid_kernel(WqArgs a, double time_step_min,
                                                          double *Bmid) {
    const long long c = blockIdx.x * (long long)kWqBlock + threadIdx.x;
    double vm = NAN, vt = NAN;
    if (c < a.n) {
        const double B = a.B[c], dW = a.dW[c];
        vm = D8 ? B + dW * time_step_min / 60.0 : B - dW * time_step_min / 60.0;
        Bmid[c] = vm;
        const double mc = (B * 1000.0 / a.A < a.min_Bt) ? INFINITY : B;
        const double d = D8 ? (dW < 0.0 ? 0.0 : dW) : (dW > 0.0 ? 0.0 : dW);
        vt = mc / (fabs(d) + 1e-16);
    }
    block_min_store<kWqBlock / 64>(vm, &a.part[blockIdx.x * kWqParts + kWqMinMid]);
    __syncthreads();
    block_min_store<kWqBlock / 64>(vt, &a.part[blockIdx.x * kWqParts + kWqTminCand]);
}

// B_t(B_t<0) = 0 (:137-140) and the closing tmin_wq candidate (:144-159)
template <bool D8>
__global__ __launch_bounds__(kWqBlock) void wq_close_kernel(WqArgs a) {
    const long long c = blockIdx.x * (long long)kWqBlock + threadIdx.x;
    double vt = NAN;
    if (c < a.n) {
        double B = a.B[c];
        if (B < 0.0) B = 0.0;
        a.B[c] = B;
        const double mc = (B * 1000.0 / a.A < a.min_Bt) ? INFINITY : B;
        const double dW = a.dW[c];
        const double d = D8 ? (dW < 0.0 ? 0.0 : dW) : (dW > 0.0 ? 0.0 : dW);
        vt = mc / (fabs(d) + 1e-16);
    }
    block_min_store<kWqBlock / 64>(vt, &a.part[blockIdx.x * kWqParts + kWqTminCand]);
}

}  // namespace hp2d

namespace hp2d {
namespace {

// one wash-off call on host arrays; pages[p] is flow page p [mm/h]
template <bool D8>
hp2d_status washoff_call(const char *name, hp2d_handle *h, const double *C_3, const double *C_4,
                         const double *const *pages, const double *B_t, double time_step_min,
                         double cell_area, const uint8_t *nan_mask, int32_t flag_wq_model,
                         const double *Tot_Washed, double Bmin, double Bmax, double min_Bt,
                         double *B_t_out, double *P_conc, double *Out_Conc, double *tmin_wq,
                         double *tot_W_out, double *mass_lost, double *Tot_Washed_out) {
    constexpr int NP = D8 ? 9 : 5;
    bool pages_ok = true;
    for (int p = 0; p < NP; ++p) pages_ok = pages_ok && pages[p];
    if (!h || !C_3 || !C_4 || !pages_ok || !B_t || !B_t_out || !P_conc || !Out_Conc || !tmin_wq ||
        !mass_lost || !Tot_Washed_out || !(time_step_min > 0.0) || !(cell_area > 0.0))
        return HP2D_E_ARG;
    const double *outlet_flow = pages[4];
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const size_t n = (size_t)ny * nx;
    const int blocks = (int)((n + kWqBlock - 1) / kWqBlock);
    hipStream_t s = hp2d_handle_stream(h);

    // workspace: C3 C4 q[NP] B Bx W[NP] totW totQ dW Bmid omass olmass | partials | mask
    const size_t n_arrays = 10 + 2 * NP + 1;
    const size_t part_off = n_arrays * n * sizeof(double);
    const size_t mask_off = part_off + (size_t)blocks * kWqParts * sizeof(double);
    char *ws = (char *)hp2d_handle_workspace(h, mask_off + n + 256);
    if (!ws) return HP2D_E_NOMEM;
    double *p = (double *)ws;
    auto take = [&](size_t k) {
        double *r = p;
        p += k * n;
        return r;
    };
    WqArgs a;
    a.ny = ny;
    a.nx = nx;
    a.n = (long long)n;
    a.A = cell_area;
    a.min_Bt = min_Bt;
    a.Bmin = Bmin;
    a.Bmax = Bmax;
    a.model = flag_wq_model;
    double *dC3 = take(1), *dC4 = take(1), *dq5 = take(NP), *dB = take(1), *dBx = take(1);
    a.C3 = dC3;
    a.C4 = dC4;
    a.q5 = dq5;
    a.B = dB;
    a.Bx = dBx;
    a.W = take(NP);
    a.totW = take(1);
    a.totQ = take(1);
    a.dW = take(1);
    double *Bmid = take(1), *omass = take(1), *olmass = take(1);
    a.part = (double *)(ws + part_off);
    uint8_t *dmask = nan_mask ? (uint8_t *)(ws + mask_off) : nullptr;
    a.nan_mask = dmask;

    hipError_t e = hipSuccess;
    auto H2D = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    auto D2H = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    auto launched = [&]() {
        if (e == hipSuccess) e = hipGetLastError();
    };
    const size_t nb = n * sizeof(double);
    H2D(dC3, C_3, nb);
    H2D(dC4, C_4, nb);
    for (int p = 0; p < NP; ++p) H2D(dq5 + p * n, pages[p], nb);
    H2D(dB, B_t, nb);
    H2D(dBx, B_t, nb);  // B_t_extra = B_t (:19)
    if (dmask) H2D(dmask, nan_mask, n);

    std::vector<double> hp((size_t)blocks * kWqParts);
    auto fold = [&](int field) {
        double m = NAN;
        for (int b = 0; b < blocks; ++b) m = std::fmin(m, hp[(size_t)b * kWqParts + field]);
        return m;
    };
    auto sync_parts = [&]() {
        D2H(hp.data(), a.part, hp.size() * sizeof(double));
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    };
    const double large_timestep = D8 ? 10.0 : time_step_min * 60.0;  // 8D.m:3
    auto tmin_from_parts = [&]() {
        return fold(kWqMinDW) >= 0 ? large_timestep : 3600.0 * fold(kWqTminCand);
    };

    // first pass (:15-79)
    if (e == hipSuccess) {
        hipLaunchKernelGGL(wq_rates_kernel<D8>, dim3(blocks), dim3(kWqBlock), 0, s, a, 1);
        launched();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(wq_net_kernel<D8>, dim3(blocks), dim3(kWqBlock), 0, s, a, 0.0, omass, olmass);
        launched();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(wq_mid_kernel<D8>, dim3(blocks), dim3(kWqBlock), 0, s, a, time_step_min, Bmid);
        launched();
    }
    std::vector<double> totQ(n), totW(n), Wout(n), B(n);
    D2H(totQ.data(), a.totQ, nb);
    sync_parts();
    const double tmin = tmin_from_parts();
    const bool substep = e == hipSuccess && (fold(kWqMinMid) < 0 || tmin < time_step_min * 60.0);
    if (substep) {
        // internal sub-steps (:80-132); rates are held, B_t changes
        const double n_dec = time_step_min * 60.0 / tmin;
        const double n_floor = std::floor(n_dec);
        if (!(n_floor < 1e7)) {
            return hp2d_handle_fail(h, std::string(name) + ": tmin_wq gives " +
                                           std::to_string(n_dec) + " sub-steps");
        }
        const long long steps = (long long)n_floor + 1;
        if (e == hipSuccess) e = hipMemsetAsync(omass, 0, 2 * nb, s);  // omass, olmass adjacent
        for (long long k = 1; k <= steps && e == hipSuccess; ++k) {
            const double dt = (k == steps) ? (n_dec - n_floor) * tmin
                                           : time_step_min * 60.0 / (double)steps;
            hipLaunchKernelGGL(wq_rates_kernel<D8>, dim3(blocks), dim3(kWqBlock), 0, s, a, 0);
            hipLaunchKernelGGL(wq_net_kernel<D8>, dim3(blocks), dim3(kWqBlock), 0, s, a, dt, omass,
                               olmass);
            launched();
        }
        D2H(totW.data(), omass, nb);
        D2H(Wout.data(), olmass, nb);
    } else {
        D2H(totW.data(), a.totW, nb);
        D2H(Wout.data(), a.W + 4 * n, nb);
        a.B = Bmid;  // B_t = B_t_mid (:134)
    }
    D2H(B.data(), a.B, nb);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(wq_close_kernel<D8>, dim3(blocks), dim3(kWqBlock), 0, s, a);
        launched();
    }
    sync_parts();
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string(name) + ": " + hipGetErrorString(e));
    const double tmin_end = tmin_from_parts();

    if (substep) {
        for (size_t c = 0; c < n; ++c) {
            totW[c] = totW[c] / (time_step_min / 60.0);  // :128
            Wout[c] = Wout[c] / (time_step_min / 60.0);  // :130
        }
    }
    // sum(sum(B_t(B_t<0))): the negatives in column-major order (:134-140)
    double lost = 0.0;
    for (size_t c = 0; c < n; ++c)
        if (B[c] < 0) lost += B[c];
    *mass_lost = lost + *mass_lost;
    for (size_t c = 0; c < n; ++c) {
        double b = B[c] < 0 ? 0.0 : B[c];
        if (std::fabs(b) < 1e-12) b = 0.0;  // :171
        B[c] = b;
    }
    double sum_wo = 0.0, sum_qo = 0.0;  // sum(sum(.)): column sums, then their sum
    for (int j = 0; j < nx; ++j) {
        double cw = 0.0, cq = 0.0;
        for (int i = 0; i < ny; ++i) {
            cw += Wout[(size_t)i + (size_t)j * ny];
            cq += outlet_flow[(size_t)i + (size_t)j * ny];
        }
        sum_wo += cw;
        sum_qo += cq;
    }
    for (size_t c = 0; c < n; ++c) {
        const double tq = std::fmax(totQ[c], 10.0);  // :175
        P_conc[c] = std::fmax(1e6 * (totW[c]) / (tq * cell_area), 0.0);
        const double tw0 = Tot_Washed ? Tot_Washed[c] : 0.0;
        Tot_Washed_out[c] = D8 ? (std::fmax(B_t[c] - B[c], 0.0)) + tw0  // 8D.m:139
                               : totW[c] * time_step_min / 60.0 + tw0;
        if (tot_W_out) tot_W_out[c] = totW[c];
    }
    for (size_t c = 0; c < n; ++c) B_t_out[c] = B[c];  // after the reads of B_t (may alias)
    *Out_Conc = std::fmax(1000.0 * sum_wo / (sum_qo / 1000.0 * cell_area), 0.0);
    *tmin_wq = tmin_end;
    return HP2D_OK;
}

}  // namespace
}  // namespace hp2d

extern "C" hp2d_status hp2d_build_up_wash_off(
    hp2d_handle *h, const double *C_3, const double *C_4, const double *qout_left,
    const double *qout_right, const double *qout_up, const double *qout_down,
    const double *outlet_flow, const double *B_t, double time_step_min, double cell_area,
    const uint8_t *nan_mask, int32_t flag_wq_model, const double *Tot_Washed, double Bmin,
    double Bmax, double min_Bt, double *B_t_out, double *P_conc, double *Out_Conc,
    double *tmin_wq, double *tot_W_out, double *mass_lost, double *Tot_Washed_out) {
    const double *pages[5] = {qout_left, qout_right, qout_up, qout_down, outlet_flow};
    return hp2d::washoff_call<false>("hp2d_build_up_wash_off", h, C_3, C_4, pages, B_t,
                                     time_step_min, cell_area, nan_mask, flag_wq_model, Tot_Washed,
                                     Bmin, Bmax, min_Bt, B_t_out, P_conc, Out_Conc, tmin_wq,
                                     tot_W_out, mass_lost, Tot_Washed_out);
}

extern "C" hp2d_status hp2d_build_up_wash_off_8d(
    hp2d_handle *h, const double *C_3, const double *C_4, const double *outflow, const double *B_t,
    double time_step_min, double cell_area, const uint8_t *nan_mask, int32_t flag_wq_model,
    const double *Tot_Washed, double Bmin, double Bmax, double min_Bt, double *B_t_out,
    double *P_conc, double *Out_Conc, double *tmin_wq, double *tot_W_out, double *mass_lost,
    double *Tot_Washed_out) {
    int32_t ny = 0, nx = 0, prec = 0;
    if (!h || !outflow) return HP2D_E_ARG;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    const size_t n = (size_t)ny * nx;
    const double *pages[9];
    for (int p = 0; p < 9; ++p) pages[p] = outflow + p * n;
    return hp2d::washoff_call<true>("hp2d_build_up_wash_off_8d", h, C_3, C_4, pages, B_t,
                                    time_step_min, cell_area, nan_mask, flag_wq_model, Tot_Washed,
                                    Bmin, Bmax, min_Bt, B_t_out, P_conc, Out_Conc, tmin_wq,
                                    tot_W_out, mass_lost, Tot_Washed_out);
}
