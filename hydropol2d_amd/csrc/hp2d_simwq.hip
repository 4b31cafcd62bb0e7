// hp2d_simwq.hip - build_up_wash_off (build_up_wash_off.m:11-183) inside the
// device-resident step loop, after the local-inertial update of
// HydroPol2D_Main_While.m:683-689.
//
// The five face pages of the reference (left, right, up, down, outlet) are
// not stored: a cell's pages are its own face fluxes and its neighbours'
// (qout_left = -q_x of the left cell, qout_down = -q_y of the cell below,
// Local_Inertial_Model_D4.m:347-350), and a neighbour's rate towards the cell
// is recomputed from the neighbour's mass, class and the shared face flux, so
// one grid-stride pass yields dW without a W buffer.
//
// Per step: wq_first_kernel (rates, dW, B_t_mid, tentative outputs of the
// unsplit case, block partials) -> wq_ctl_kernel (one block: the tmin_wq and
// split decision, the closing scalars of the unsplit case) -> s_cap guarded
// wq_sub_kernel launches (the internal sub-steps, :80-132; the last one
// closes the call: clamp, tmin_wq, P_conc, Tot_Washed) -> wq_fold_kernel (one
// block: the closing scalars of a split call).  A split longer than s_cap
// raises SimState::hold, which makes every later kernel of the captured step
// sequence a no-op until the host has run the remaining sub-steps.
//
// Face rates below zero carry no wash-off (q < 0 is taken as 0 before the
// power law, as in the per-call drop-in); tot_q_out keeps the signed sum.
#include "hp2d_simwq.hpp"

namespace hp2d {

constexpr int kWqThreads = 256;
constexpr int kWqFoldThreads = 1024;

// pow(0, c) as pow gives it: +0 for c > 0, 1 for c == 0, +inf for c < 0.  The
// FP64 power law takes this case without a call (a face without flow in the
// page's direction, the common case there)
template <typename T> __device__ __forceinline__ T wq_pow0(T c) {
    return isnan(c) ? c : (c > T(0) ? T(0) : (c == T(0) ? T(1) : T(INFINITY)));
}
__device__ __forceinline__ double wq_pow(double x, double c) { return x == 0.0 ? wq_pow0(c) : pow(x, c); }
// FP32: exp2(c log2 x) without a branch.  The zero base needs no case of its
// own: log2(0) = -inf, so c > 0 gives exp2(-inf) = +0 and c < 0 gives +inf, as
// pow does; only c = 0 (pow(x, 0) = 1 for every x, NaN included) is selected.
__device__ __forceinline__ float wq_pow(float x, float c) {
    const float r = __builtin_amdgcn_exp2f(c * __builtin_amdgcn_logf(x));
    return c == 0.0f ? 1.0f : r;
}

// FP64 keeps the reference's operation order; FP32 multiplies by hardware
// reciprocals (the precision policy of the FP32 step kernel)
__device__ __forceinline__ double wq_div(double x, double y) { return x / y; }
__device__ __forceinline__ float wq_div(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }
__device__ __forceinline__ double wq_gm2(double B, double A) { return B * 1000.0 / A; }  // [g/m2]
__device__ __forceinline__ float wq_gm2(float B, float A) { return B * (1000.0f * __builtin_amdgcn_rcpf(A)); }
__device__ __forceinline__ double wq_qA(double q, double A) { return q / 1000.0 / 3600.0 * A; }  // [m3/s]
__device__ __forceinline__ float wq_qA(float q, float A) { return q * (A * 2.77777778e-7f); }

// class tables of a launch, staged in LDS
template <typename T> struct WqLds {
    T C3[256], C4[256];
};

template <typename T>
__device__ __forceinline__ void wq_stage_classes(const WqSimArgs<T> &a, WqLds<T> &lds) {
    for (int k = threadIdx.x; k < 256; k += kWqThreads) {
        lds.C3[k] = a.C3[k < a.n_classes ? k : 0];
        lds.C4[k] = a.C4[k < a.n_classes ? k : 0];
    }
    __syncthreads();
}

// wash-off parameters of one cell for the current pass
template <typename T> struct WqCell {
    T B, f, C3, C4;
    bool zero;  // idx_bt (or the idx_nan_5 mask of the first pass): no flux
};

// the parameters of a cell from its flag word fl, its mass B and (rating
// model) the entry B_t_extra bx; bx_out receives the updated B_t_extra
template <typename T>
__device__ __forceinline__ WqCell<T> wq_cell_of(const WqSimArgs<T> &a, const WqLds<T> &lds, unsigned fl,
                                                T B, T bx, bool first, T &bx_out) {
    WqCell<T> w;
    const int cls = (int)(fl & 0xffu);
    w.B = B;
    w.C3 = lds.C3[cls];
    w.C4 = lds.C4[cls];
    T f = B;
    bx_out = bx;
    if (a.model == 1) {  // rating curve (:23-28, :89-94)
        const T Br = T(a.Bmin / 1000.0 * a.A), Bm = T(a.Bmax / 1000.0 * a.A);
        if (wq_gm2(B, T(a.A)) <= T(a.Bmin)) bx = T(0);
        if (wq_gm2(B, T(a.A)) >= T(a.Bmax)) bx = Bm;
        bx_out = bx;
        f = T(1) + dmax(bx - Br, T(0));
    }
    w.f = f;
    const bool low = (wq_gm2(B, T(a.A)) < T(a.min_Bt)) || isnan(B) || isinf(B);
    w.zero = low || (first && (fl & kWqNanBit));
    return w;
}

template <typename T>
__device__ __forceinline__ WqCell<T> wq_cell(const WqSimArgs<T> &a, const WqLds<T> &lds, long long c,
                                             const T *Bin, const T *Bx_in, bool first, T *Bx_out) {
    const T B = Bin[c];
    const T bx_in = (a.model == 1 && !first) ? Bx_in[c] : B;
    T bx;
    const WqCell<T> w = wq_cell_of(a, lds, (unsigned)a.cell[c], B, bx_in, first, bx);
    if (a.model == 1 && Bx_out) Bx_out[c] = bx;
    return w;
}

// W of one page: C3 (q/1000/3600*A)^C4 f(B) [kg/h]
template <typename T>
__device__ __forceinline__ T wq_rate(const WqCell<T> &w, T q, T A) {
    const T qq = q < T(0) ? T(0) : q;  // NaN stays NaN
    const T r = w.C3 * wq_pow(wq_qA(qq, A), w.C4) * w.f;
    return w.zero ? T(0) : r;
}

// a neighbour's page towards the cell.  FP64: the neighbour's parameters are
// only read when the face carries flow in that direction (or C4 = 0 makes a
// zero flux count), which skips the software pow of the dry faces.  FP32: the
// rate is evaluated on every face without a branch (the hardware exp2/log2
// cost less than the divergent branch); a dry face gives C3 * 0 * f = 0 as the
// skipped case does, f being finite wherever the cell is not masked.
template <typename T>
__device__ __forceinline__ T wq_inflow(const WqSimArgs<T> &a, const WqLds<T> &lds, long long nb,
                                       T q, T A, const T *Bin, const T *Bx_in, bool first) {
    if constexpr (sizeof(T) == 8) {
        const T qq = q < T(0) ? T(0) : q;
        if (qq == T(0) && wq_pow0(lds.C4[a.cell[nb] & 0xffu]) == T(0)) return T(0);
    }
    return wq_rate(wq_cell(a, lds, nb, Bin, Bx_in, first, (T *)nullptr), q, A);
}

template <typename T>
__device__ __forceinline__ T outlet_q(const WqSimArgs<T> &a, long long c, unsigned fl) {
    if (!(fl & kWqOutletBit)) return T(0);
    int lo = 0, hi = a.n_outlet - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const long long v = a.out_idx[mid];
        if (v == c) return a.out_q[mid];
        if (v < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return T(0);
}

// the pass is skipped when the run is over / paused, when the step was
// rejected before the wash-off (injected failure), or on hold
__device__ __forceinline__ bool wq_skip(const SimState *st, bool force) {
    return st->done || st->paused || st->inject_k == st->k || (st->hold && !force);
}

// fixed-order block reduction of up to 8 fields (mins or sums)
template <int NF>
__device__ __forceinline__ void block_store(double (&v)[NF], const bool (&is_min)[NF],
                                            double *part, int nb) {
    __shared__ double sh[NF][kWqThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        v[f] = is_min[f] ? wave_fmin(v[f]) : wave_sum(v[f]);
        if (lane == 0) sh[f][w] = v[f];
    }
    __syncthreads();
    if (threadIdx.x < NF) {
        const int f = threadIdx.x;
        double acc = sh[f][0];
        for (int k = 1; k < kWqThreads / kWave; ++k) acc = is_min[f] ? fmin(acc, sh[f][k]) : acc + sh[f][k];
        part[(long long)f * nb + blockIdx.x] = acc;
    }
}

// rates of one cell against its current B: the five outflow pages, their
// sum, the signed flux sum and the net rate dW (:24-59, :89-120)
template <typename T> struct WqRates {
    T W4, totW, totQ, dW;
    WqCell<T> own;
};

template <typename T>
__device__ __forceinline__ WqRates<T> wq_rates(const WqSimArgs<T> &a, const WqLds<T> &lds, int i,
                                               int j, long long c, const T *qx, const T *qy,
                                               const T *Bin, const T *Bx_in, T *Bx_out, bool first) {
    const int ny = a.ny, nx = a.nx;
    const T A = T(a.A);
    WqRates<T> r;
    const WqCell<T> w = wq_cell(a, lds, c, Bin, Bx_in, first, Bx_out);
    r.own = w;
    const T qr = qx[c], qu = qy[c];
    const T ql = j > 0 ? -qx[c - ny] : T(0);
    const T qd = i < ny - 1 ? -qy[c + 1] : T(0);
    const T qo = outlet_q(a, c, (unsigned)a.cell[c]);
    const T W0 = wq_rate(w, ql, A), W1 = wq_rate(w, qr, A), W2 = wq_rate(w, qu, A);
    const T W3 = wq_rate(w, qd, A), W4 = wq_rate(w, qo, A);
    r.W4 = W4;
    r.totW = W0 + W1 + W2 + W3 + W4;
    r.totQ = (a.cell[c] & kWqNanBit) ? T(0) : ql + qr + qu + qd + qo;
    // the neighbours' pages towards this cell (:51-55)
    T in_l = T(0), in_r = T(0), in_u = T(0), in_d = T(0);
    if (j > 0) in_l = wq_inflow(a, lds, c - ny, -ql, A, Bin, Bx_in, first);
    if (j < nx - 1) in_r = wq_inflow(a, lds, c + ny, -qr, A, Bin, Bx_in, first);
    if (i > 0) in_u = wq_inflow(a, lds, c - 1, -qu, A, Bin, Bx_in, first);
    if (i < ny - 1) in_d = wq_inflow(a, lds, c + 1, -qd, A, Bin, Bx_in, first);
    r.dW = r.totW - (in_l + in_r + in_u + in_d);
    return r;
}

// Loads of one cell of the first pass: its flag word and mass and its four
// neighbours' (left, right, up, down; a neighbour outside the grid reads the
// cell itself and is masked), its two face fluxes and the neighbours' shared
// ones, and its entry Tot_Washed.  The FP32 pass fetches them one tile ahead.
template <typename T> struct WqIn {
    T B, Bn[4], qr, qu, qxl, qyd, tw;
    unsigned fl, fln[4];
};

template <typename T>
__device__ __forceinline__ void wq_load(const WqSimArgs<T> &a, int i, int j, long long c, const T *qx,
                                        const T *qy, const T *Bin, const T *TWin, WqIn<T> &in) {
    const int ny = a.ny, nx = a.nx;
    const long long nb[4] = {j > 0 ? c - ny : c, j < nx - 1 ? c + ny : c, i > 0 ? c - 1 : c,
                             i < ny - 1 ? c + 1 : c};
    in.fl = a.cell[c];
    in.B = Bin[c];
    in.qr = qx[c];
    in.qu = qy[c];
    in.qxl = qx[nb[0]];
    in.qyd = qy[nb[3]];
    in.tw = TWin[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        in.fln[k] = a.cell[nb[k]];
        in.Bn[k] = Bin[nb[k]];
    }
}

// wq_rates of the first pass from preloaded inputs (the same arithmetic)
template <typename T>
__device__ __forceinline__ WqRates<T> wq_rates_in(const WqSimArgs<T> &a, const WqLds<T> &lds, int i,
                                                  int j, long long c, const WqIn<T> &in, T *Bx_out) {
    const int ny = a.ny, nx = a.nx;
    const T A = T(a.A);
    WqRates<T> r;
    T bx;
    const WqCell<T> w = wq_cell_of(a, lds, in.fl, in.B, in.B, true, bx);
    if (a.model == 1 && Bx_out) Bx_out[c] = bx;
    r.own = w;
    const T qr = in.qr, qu = in.qu;
    const T ql = j > 0 ? -in.qxl : T(0);
    const T qd = i < ny - 1 ? -in.qyd : T(0);
    const T qo = outlet_q(a, c, in.fl);
    const T W0 = wq_rate(w, ql, A), W1 = wq_rate(w, qr, A), W2 = wq_rate(w, qu, A);
    const T W3 = wq_rate(w, qd, A), W4 = wq_rate(w, qo, A);
    r.W4 = W4;
    r.totW = W0 + W1 + W2 + W3 + W4;
    r.totQ = (in.fl & kWqNanBit) ? T(0) : ql + qr + qu + qd + qo;
    const T qn[4] = {-ql, -qr, -qu, -qd};
    const bool has[4] = {j > 0, j < nx - 1, i > 0, i < ny - 1};
    T inr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T unused;
        const T v = wq_rate(wq_cell_of(a, lds, in.fln[k], in.Bn[k], in.Bn[k], true, unused), qn[k], A);
        inr[k] = has[k] ? v : T(0);
    }
    r.dW = r.totW - (inr[0] + inr[1] + inr[2] + inr[3]);
    return r;
}

// ---------------------------------------------------------------------------
// first pass (:13-79): rates of the entry B_t, dW, B_t_mid, tmin candidates;
// writes the unsplit outcome (B_t, Tot_Washed, P_conc), overwritten by the
// last sub-step when the call splits
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kWqThreads) void wq_first_kernel(WqSimArgs<T> a) {
    const SimState *st = a.st;
    if (wq_skip(st, false)) return;
    __shared__ WqLds<T> lds;
    wq_stage_classes(a, lds);
    const int qp = (int)(st->k & 1);          // fluxes just written by the step
    const int bp = (int)(st->wq_n & 1);       // entry B_t of this call
    const T *qx = a.qx_buf[qp], *qy = a.qy_buf[qp];
    const T *Bin = a.B[bp];
    T *Bout = a.B[bp ^ 1];
    const T *TWin = a.TW[bp];
    T *TWout = a.TW[bp ^ 1];
    const double dt_min = st->time_step_min;
    const T A = T(a.A);
    // minima in the working precision (the conversion to FP64 is exact and
    // monotone, so the minimum converts unchanged), sums in FP64
    T mn[4] = {T(NAN), T(NAN), T(NAN), T(NAN)};
    double s_wout = 0.0, s_b = 0.0;
    // tiles of 256 rows of one column: tile t is (row block t % n_rb, column t / n_rb)
    // (fewer than 2^32 cells, so fewer than 2^24 tiles)
    const unsigned n_rb = (unsigned)(a.ny + kWqThreads - 1) / kWqThreads;
    const unsigned n_tiles = n_rb * (unsigned)a.nx;
    T *Bx_out = a.model == 1 ? a.Bx[0] : (T *)nullptr;
    auto cell_of_tile = [&](unsigned t, int &i, int &j) {
        j = (int)(t / n_rb);
        i = (int)(t % n_rb) * kWqThreads + (int)threadIdx.x;
        return (long long)j * a.ny + i;
    };
    // FP32: the loads of tile t + gridDim.x are issued before tile t is
    // computed (the FP64 pass is bound by its software pow, not by latency)
    constexpr bool kPrefetch = sizeof(T) == 4;
    WqIn<T> nxt;
    if (kPrefetch && blockIdx.x < n_tiles) {
        int i0, j0;
        const long long c0 = cell_of_tile(blockIdx.x, i0, j0);
        if (i0 < a.ny) wq_load(a, i0, j0, c0, qx, qy, Bin, TWin, nxt);
    }
    for (unsigned t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        int i, j;
        const long long c = cell_of_tile(t, i, j);
        WqRates<T> r;
        T tw_in;
        if constexpr (kPrefetch) {
            const WqIn<T> cur = nxt;
            const unsigned tn = t + gridDim.x;
            if (tn < n_tiles) {
                int in_i, in_j;
                const long long cn = cell_of_tile(tn, in_i, in_j);
                if (in_i < a.ny) wq_load(a, in_i, in_j, cn, qx, qy, Bin, TWin, nxt);
            }
            if (i >= a.ny) continue;
            r = wq_rates_in(a, lds, i, j, c, cur, Bx_out);
            tw_in = cur.tw;
        } else {
            if (i >= a.ny) continue;
            r = wq_rates(a, lds, i, j, c, qx, qy, Bin, (const T *)nullptr, Bx_out, true);
            tw_in = TWin[c];
        }
        const T B = r.own.B, dW = r.dW, totW = r.totW;
        const T Bmid = B - wq_div(dW * T(dt_min), T(60));
        // tmin candidates (:61-76, :144-159): m_calc / (|min(dW,0)| + 1e-16)
        const T d0 = dW > T(0) ? T(0) : dW;
        const T mc0 = (wq_gm2(B, A) < T(a.min_Bt)) ? T(INFINITY) : B;
        const T mce = (wq_gm2(Bmid, A) < T(a.min_Bt)) ? T(INFINITY) : Bmid;
        const T cand0 = wq_div(mc0, fabs(d0) + T(1e-16));
        const T cande = wq_div(mce, fabs(d0) + T(1e-16));
        const T Bfin = fabs(Bmid) < T(1e-12) ? T(0) : Bmid;  // :171
        Bout[c] = Bfin;
        TWout[c] = wq_div(totW * T(dt_min), T(60)) + tw_in;
        const T tq = dmax(r.totQ, T(10));
        a.Pc[c] = dmax(wq_div(T(1e6) * totW, tq * A), T(0));
        mn[kWqMinDW] = dmin(mn[kWqMinDW], dW);
        mn[kWqMinMid] = dmin(mn[kWqMinMid], Bmid);
        mn[kWqCand0] = dmin(mn[kWqCand0], cand0);
        mn[kWqCandEnd] = dmin(mn[kWqCandEnd], cande);
        s_wout += (double)r.W4;
        if (!isinf(Bfin)) s_b += (double)Bfin;
    }
    double v[6] = {(double)mn[0], (double)mn[1], (double)mn[2], (double)mn[3], s_wout, s_b};
    const bool is_min[6] = {true, true, true, true, false, false};
    block_store<6>(v, is_min, a.part, a.n_blocks);
}

// one block: folds the first pass; decides the split; closes an unsplit call.
// Each thread issues all of its loads (NF fields x 4 partials) before it
// combines any of them, so the fold costs one memory latency, not 4*NF.
template <int NF>
__device__ __forceinline__ void fold_fields(const double *part, int nb, const bool (&is_min)[NF],
                                            double *out) {
    constexpr int kU = 4;
    __shared__ double sh[NF][kWqFoldThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    double acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = is_min[f] ? NAN : 0.0;
    for (int b0 = threadIdx.x; b0 < nb; b0 += kU * kWqFoldThreads) {
        double x[NF][kU];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int b = b0 + u * kWqFoldThreads;
                x[f][u] = b < nb ? part[(long long)f * nb + b] : (is_min[f] ? NAN : 0.0);
            }
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int u = 0; u < kU; ++u) acc[f] = is_min[f] ? fmin(acc[f], x[f][u]) : acc[f] + x[f][u];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const double v = is_min[f] ? wave_fmin(acc[f]) : wave_sum(acc[f]);
        if (lane == 0) sh[f][w] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int f = 0; f < NF; ++f) {
            double r = sh[f][0];
            for (int k = 1; k < kWqFoldThreads / kWave; ++k)
                r = is_min[f] ? fmin(r, sh[f][k]) : r + sh[f][k];
            out[f] = r;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWqFoldThreads) void wq_ctl_kernel(WqSimArgs<T> a) {
    SimState *st = a.st;
    if (wq_skip(st, false)) return;
    double r[6];
    const bool is_min[6] = {true, true, true, true, false, false};
    fold_fields<6>(a.part, a.n_blocks, is_min, r);
    if (threadIdx.x != 0) return;
    const double dt_s = st->time_step_min * 60.0;  // large_timestep
    const double tmin0 = r[kWqMinDW] >= 0 ? dt_s : 3600.0 * r[kWqCand0];
    WqCtl *ctl = a.ctl;
    ctl->tmin0 = tmin0;
    ctl->dt_s = dt_s;
    const bool split = r[kWqMinMid] < 0 || tmin0 < dt_s;
    ctl->split = split ? 1 : 0;
    if (!split) {
        // B_t = B_t_mid: nothing negative to clamp, mass_lost += 0
        st->tmin_wq = r[kWqMinDW] >= 0 ? dt_s : 3600.0 * r[kWqCandEnd];
        st->wq_sum_wout = r[kWqSumWout];
        st->wq_sum_b = r[kWqSumB];
        ctl->steps = 0;
        return;
    }
    const double n_dec = dt_s / tmin0;
    const double n_floor = floor(n_dec);
    ctl->n_dec = n_dec;
    ctl->n_floor = n_floor;
    st->wq_split_steps += 1;
    if (!(n_floor < 1e7)) {  // tmin_wq = 0 would loop for ever in the reference
        st->wq_err = 1;
        st->hold = 1;
        ctl->steps = 0;
        return;
    }
    ctl->steps = (long long)n_floor + 1;
    if (ctl->steps > a.s_cap) st->hold = 1;
}

// ---------------------------------------------------------------------------
// sub-step k of a split call (:85-125); the last one closes the call
// (:128-140 clamp and mass_lost, :144-159 tmin_wq, :171, :175-182)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kWqThreads) void wq_sub_kernel(WqSimArgs<T> a, int k, int force) {
    const SimState *st = a.st;
    if (wq_skip(st, force != 0)) return;
    const WqCtl *ctl = a.ctl;
    if (!ctl->split) return;
    const long long steps = ctl->steps;
    if (k > steps) return;
    __shared__ WqLds<T> lds;
    wq_stage_classes(a, lds);
    const int qp = (int)(st->k & 1), bp = (int)(st->wq_n & 1);
    const T *qx = a.qx_buf[qp], *qy = a.qy_buf[qp];
    // ping-pong so that the last sub-step lands in the call's output buffer
    T *dst_final = a.B[bp ^ 1];
    auto buf = [&](long long s) { return ((steps - s) % 2 == 0) ? dst_final : a.Bs; };
    const T *Bin = (k == 1) ? a.B[bp] : buf(k - 1);
    T *Bout = buf(k);
    const T *Bx_in = a.model == 1 ? a.Bx[(k - 1) & 1] : nullptr;
    T *Bx_out = a.model == 1 ? a.Bx[k & 1] : nullptr;
    const double dt = (k == steps) ? (ctl->n_dec - ctl->n_floor) * ctl->tmin0
                                   : ctl->dt_s / (double)steps;
    const bool last = (k == steps);
    const double dt_min = st->time_step_min;
    const T A = T(a.A), dtT = T(dt);
    const T *TWin = a.TW[bp];
    T *TWout = a.TW[bp ^ 1];
    T mn_dw = T(NAN), mn_cand = T(NAN);       // minima in the working precision
    double v[5] = {NAN, NAN, 0.0, 0.0, 0.0};  // minDW, candEnd, sumNeg, outMass, sumB
    const bool is_min[5] = {true, true, false, false, false};
    // tiles of 256 rows of one column: tile t is (row block t % n_rb, column t / n_rb)
    const unsigned n_rb = (unsigned)(a.ny + kWqThreads - 1) / kWqThreads;
    const unsigned n_tiles = n_rb * (unsigned)a.nx;
    for (unsigned t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int j = (int)(t / n_rb), i = (int)(t % n_rb) * kWqThreads + (int)threadIdx.x;
        if (i >= a.ny) continue;
        const long long c = (long long)j * a.ny + i;
        const WqRates<T> r = wq_rates(a, lds, i, j, c, qx, qy, Bin, Bx_in, Bx_out, false);
        const T dW = r.dW, totW = r.totW;
        const T om = (k == 1 ? T(0) : a.omass[c]) + wq_div(totW * dtT, T(3600));
        T Bn = r.own.B - wq_div(dW * dtT, T(3600));
        v[3] += (double)wq_div(r.W4 * dtT, T(3600));
        if (!last) {
            a.omass[c] = om;
            Bout[c] = Bn;
            continue;
        }
        if (Bn < T(0)) v[2] += (double)Bn;
        if (Bn < T(0)) Bn = T(0);
        const T d0 = dW > T(0) ? T(0) : dW;
        const T mce = (wq_gm2(Bn, A) < T(a.min_Bt)) ? T(INFINITY) : Bn;
        mn_dw = dmin(mn_dw, dW);
        mn_cand = dmin(mn_cand, wq_div(mce, fabs(d0) + T(1e-16)));
        if (fabs(Bn) < T(1e-12)) Bn = T(0);
        Bout[c] = Bn;
        if (!isinf(Bn)) v[4] += (double)Bn;
        const T totWa = wq_div(om, wq_div(T(dt_min), T(60)));
        TWout[c] = wq_div(totWa * T(dt_min), T(60)) + TWin[c];
        a.Pc[c] = dmax(wq_div(T(1e6) * totWa, dmax(r.totQ, T(10)) * A), T(0));
    }
    v[0] = (double)mn_dw;
    v[1] = (double)mn_cand;
    // partials: the outlet mass accumulates over the sub-steps in this block's slot
    __shared__ double prev;
    if (threadIdx.x == 0) prev = (k == 1) ? 0.0 : a.part[(long long)kWqOutMass * a.n_blocks + blockIdx.x];
    __syncthreads();
    const double prev_out = prev;
    __syncthreads();
    double u[5] = {v[0], v[1], v[2], v[3], v[4]};
    const bool m[5] = {is_min[0], is_min[1], is_min[2], is_min[3], is_min[4]};
    // store as fields kWqMinDW, kWqCandEnd, kWqSumNeg, kWqOutMass, kWqSumB
    __shared__ double sh[5][kWqThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        u[f] = m[f] ? wave_fmin(u[f]) : wave_sum(u[f]);
        if (lane == 0) sh[f][wv] = u[f];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int f = threadIdx.x;
        double acc = sh[f][0];
        for (int q = 1; q < kWqThreads / kWave; ++q) acc = m[f] ? fmin(acc, sh[f][q]) : acc + sh[f][q];
        const int field[5] = {kWqMinDW, kWqCandEnd, kWqSumNeg, kWqOutMass, kWqSumB};
        if (f == 3) acc = prev_out + acc;
        if (f == 3 || last) a.part[(long long)field[f] * a.n_blocks + blockIdx.x] = acc;
    }
}

// one block: the closing scalars of a split call
template <typename T>
__global__ __launch_bounds__(kWqFoldThreads) void wq_fold_kernel(WqSimArgs<T> a, int force) {
    SimState *st = a.st;
    if (wq_skip(st, force != 0)) return;
    const WqCtl *ctl = a.ctl;
    if (!ctl->split) return;
    double r[8];
    const bool is_min[8] = {true, false, false, true, false, false, false, false};
    // fields kWqMinDW .. kWqOutMass are read; kWqMinMid, kWqCand0, kWqSumWout are stale
    fold_fields<8>(a.part, a.n_blocks, is_min, r);
    if (threadIdx.x != 0) return;
    const double dt_s = ctl->dt_s;
    st->tmin_wq = r[kWqMinDW] >= 0 ? dt_s : 3600.0 * r[kWqCandEnd];
    st->wq_mass_lost = r[kWqSumNeg] + st->wq_mass_lost;
    st->wq_sum_wout = r[kWqOutMass] / (st->time_step_min / 60.0);
    st->wq_sum_b = r[kWqSumB];
    if (force) st->hold = 0;
}

template <typename T> hipError_t launch_wq_step(const WqSimArgs<T> &a, hipStream_t s) {
    hipLaunchKernelGGL((wq_first_kernel<T>), dim3(a.n_blocks), dim3(kWqThreads), 0, s, a);
    hipLaunchKernelGGL((wq_ctl_kernel<T>), dim3(1), dim3(kWqFoldThreads), 0, s, a);
    for (int k = 1; k <= a.s_cap; ++k)
        hipLaunchKernelGGL((wq_sub_kernel<T>), dim3(a.n_blocks), dim3(kWqThreads), 0, s, a, k, 0);
    hipLaunchKernelGGL((wq_fold_kernel<T>), dim3(1), dim3(kWqFoldThreads), 0, s, a, 0);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_wq_finish(const WqSimArgs<T> &a, long long steps, hipStream_t s) {
    for (long long k = 1; k <= steps; ++k)
        hipLaunchKernelGGL((wq_sub_kernel<T>), dim3(a.n_blocks), dim3(kWqThreads), 0, s, a,
                           (int)k, 1);
    hipLaunchKernelGGL((wq_fold_kernel<T>), dim3(1), dim3(kWqFoldThreads), 0, s, a, 1);
    return hipGetLastError();
}

template hipError_t launch_wq_step<float>(const WqSimArgs<float> &, hipStream_t);
template hipError_t launch_wq_step<double>(const WqSimArgs<double> &, hipStream_t);
template hipError_t launch_wq_finish<float>(const WqSimArgs<float> &, long long, hipStream_t);
template hipError_t launch_wq_finish<double>(const WqSimArgs<double> &, long long, hipStream_t);

}  // namespace hp2d
