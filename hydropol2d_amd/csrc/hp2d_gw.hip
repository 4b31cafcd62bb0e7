// hp2d_gw.hip - explicit Boussinesq groundwater solver
// (Boussinesq_2D_explicit.m:59-195) for gfx950.
//
// One sub-step of the reference's predictor-corrector is five elementwise
// kernels over the column-major grid plus two one-thread control kernels:
//   ctl_begin  while-test, n_steps = ceil(T/dt), dt = min(dt_GW, dt)
//   flux       Darcy face fluxes of h_t (donor-capped, left/upper donor)
//   predict    h_p = max(h_t - dt/Sy div F1, z0), Dirichlet
//   flux(avg)  fluxes of h_p, averaged with F1 in place
//   update     h = h_t + dt/Sy (-div F + R), Dirichlet, >= z0, exfiltration,
//              <= surface, and the column sums of the closing mass balance
//   velocity   Darcy velocities of the new head, the catchment mask, and the
//              next Courant limit as an atomic minimum over positive doubles
//   ctl_end    kk / trim / t bookkeeping of the reference loop
// The loop control lives on the device, so the host only launches batches of
// sub-steps and reads the done flag between batches; kernels of a sub-step
// past the end return at once.
// The closing mass-balance sums are per-column FP64 sums in row order, taken
// by the last sub-step's update pass and finished on the host in column
// order: the same association as the reference's nansum(nansum(.)).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_handle.hpp"
#include "hp2d_internal.hpp"
#include "hp2d_colred.hpp"
#include "hp2d_gw.hpp"

namespace hp2d {

// (row, column) of a linear cell index; 32-bit arithmetic below 2^32 cells
__device__ __forceinline__ void gw_rc(long long c, int ny, int &i, int &j) {
    if (c < 0x100000000LL) {
        const unsigned u = (unsigned)c;
        i = (int)(u % (unsigned)ny);
        j = (int)(u / (unsigned)ny);
    } else {
        i = (int)(c % ny);
        j = (int)(c / ny);
    }
}

__device__ __forceinline__ double signum(double F) {
    return isnan(F) ? NAN : (double)((F > 0.0) - (F < 0.0));
}

__global__ void gw_ctl_init(GwCtl *ctl, double T) {
    ctl->t = 0.0;
    ctl->T = T;
    ctl->dt = T;
    ctl->dt_used = T;
    ctl->n_steps = 1.0;
    ctl->dtgw_bits = kInfBits;
    ctl->done = 0;
    ctl->nsub = 0;
    ctl->last = 0;
}

__global__ void gw_ctl_begin(GwCtl *ctl) {
    if (ctl->done) return;
    if (!(ctl->t < ctl->T)) {
        ctl->done = 1;
        return;
    }
    ctl->n_steps = ceil(ctl->T / ctl->dt);
    const double dt_gw = __longlong_as_double((long long)ctl->dtgw_bits);
    ctl->dt = fmin(dt_gw, ctl->dt);
    ctl->dt_used = ctl->dt;
    ctl->dtgw_bits = kInfBits;
    // whether gw_ctl_end will find the loop over after this sub-step (the same
    // trim and clock arithmetic): the next Courant limit is then not needed
    const int kk = 2;
    const double dt_next = ((double)kk == ctl->n_steps - 1.0) ? ctl->T - kk * ctl->dt : ctl->dt;
    ctl->last = !(ctl->t + dt_next < ctl->T);
}

__global__ void gw_ctl_end(GwCtl *ctl) {
    if (ctl->done) return;
    ctl->nsub += 1;
    const int kk = 2;  // reset to 1 at the top of every pass, then incremented
    if ((double)kk == ctl->n_steps - 1.0) ctl->dt = ctl->T - kk * ctl->dt;
    ctl->t = ctl->t + ctl->dt;
    // the loop test of the next pass, taken now so that the host sees the end
    // without launching a pass that would only find it
    if (!(ctl->t < ctl->T)) ctl->done = 1;
}

// compute_fluxes_conservative; with avg, F := 0.5*(F_prev + F)
__global__ __launch_bounds__(kGwBlock) void gw_flux_kernel(GwArgs a, const double *__restrict__ h,
                                                           double *Fx, double *Fy, int avg) {
    if (a.ctl->done) return;
    const long long c = xcd_logical_tile() * (long long)kGwBlock + threadIdx.x;
    if (c >= a.n) return;
    const int ny = a.ny;
    int i, j;
    gw_rc(c, ny, i, j);
    const double dt = a.ctl->dt_used;
    const double H = fmax(h[c] - a.z0[c], 0.0);
    double fx = 0.0, fy = 0.0;
    if (j < a.nx - 1) {
        const long long r = c + ny;
        const double Hr = fmax(h[r] - a.z0[r], 0.0);
        const double Hx = 0.5 * (H + Hr);
        const double Kx = 0.5 * (gw_K(a, c) + gw_K(a, r));
        const double dHdx = (h[r] - h[c]) / a.dx;
        double F = -Kx * Hx * dHdx;
        const double cap = gw_Sy(a, c) * H / dt;
        F = signum(F) * fmin(fabs(F), cap);
        if (isnan(Hr)) F = 0.0;
        fx = F;
    }
    if (i < ny - 1) {
        const long long d = c + 1;
        const double Hd = fmax(h[d] - a.z0[d], 0.0);
        const double Hy = 0.5 * (H + Hd);
        const double Ky = 0.5 * (gw_K(a, c) + gw_K(a, d));
        const double dHdy = (h[d] - h[c]) / a.dy;
        double F = -Ky * Hy * dHdy;
        const double cap = gw_Sy(a, c) * H / dt;
        F = signum(F) * fmin(fabs(F), cap);
        if (isnan(Hd)) F = 0.0;
        fy = F;
    }
    if (avg) {
        fx = 0.5 * (Fx[c] + fx);
        fy = 0.5 * (Fy[c] + fy);
    }
    Fx[c] = fx;
    Fy[c] = fy;
}

// compute_flux_divergence at cell (i, j)
__device__ __forceinline__ double gw_div(const GwArgs &a, const double *Fx, const double *Fy,
                                         long long c, int i, int j) {
    double div_x, div_y;
    if (j == 0)
        div_x = Fx[c] / a.dx;
    else if (j == a.nx - 1)
        div_x = -Fx[c - a.ny] / a.dx;
    else
        div_x = (Fx[c] - Fx[c - a.ny]) / a.dx;
    if (i == 0)
        div_y = Fy[c] / a.dy;
    else if (i == a.ny - 1)
        div_y = -Fy[c - 1] / a.dy;
    else
        div_y = (Fy[c] - Fy[c - 1]) / a.dy;
    return div_x + div_y;
}

// predictor (:113-118)
__global__ __launch_bounds__(kGwBlock) void gw_predict_kernel(GwArgs a, const double *h_t,
                                                              const double *Fx, const double *Fy,
                                                              double *h_p) {
    if (a.ctl->done) return;
    const long long c = xcd_logical_tile() * (long long)kGwBlock + threadIdx.x;
    if (c >= a.n) return;
    int i, j;
    gw_rc(c, a.ny, i, j);
    const double dt = a.ctl->dt_used;
    double v = h_t[c] + dt / gw_Sy(a, c) * (-gw_div(a, Fx, Fy, c, i, j));
    if (a.dirichlet && a.dirichlet[c]) v = a.h_dir[c];
    h_p[c] = fmax(v, a.z0[c]);
}

// corrector update, Dirichlet, bed, exfiltration and surface clamp (:128-148),
// one block per column.  The block also leaves the column's sums of the
// closing mass balance (R*dt, q_exf*dt, Sy*max(h0 - z0, 0), Sy*max(h - z0, 0)
// with h masked outside the catchment, as the velocity pass masks it): the
// last sub-step's sums are the call's.
__global__ __launch_bounds__(kColThreads) void gw_update_kernel(GwArgs a, const double *h_t,
                                                                const double *Fx, const double *Fy,
                                                                double *h_out, double *q_exf,
                                                                const double *h0, double *col) {
    if (a.ctl->done) return;
    const double dt = a.ctl->dt_used;
    // columns in XCD-contiguous runs: a column's neighbours, whose heads and
    // fluxes it reads, sit in the same L2
    const int j = xcd_logical_tile();
    const bool is_max[4] = {false, false, false, false};
    column_reduce<4>(a.ny, j, is_max, [&](long long c, double (&s)[4]) {
        const int i = (int)(c - (long long)j * a.ny);
        double v = h_t[c] + dt / gw_Sy(a, c) * (-gw_div(a, Fx, Fy, c, i, j) + a.R[c]);
        if (a.dirichlet && a.dirichlet[c]) v = a.h_dir[c];
        v = fmax(v, a.z0[c]);
        const double hs = a.z0[c] + (a.h_soil ? a.h_soil[c] : a.h_soil_u);
        const bool out = a.catchment && !a.catchment[c];
        double q = fmax(0.0, v - hs) * gw_Sy(a, c) / dt;
        if (out) q = NAN;
        q_exf[c] = q;
        const double hn = fmin(v, hs);
        h_out[c] = hn;
        s[0] = a.R[c] * dt;
        s[1] = q * dt;
        s[2] = gw_Sy(a, c) * fmax(h0[c] - a.z0[c], 0.0);
        s[3] = gw_Sy(a, c) * fmax((out ? NAN : hn) - a.z0[c], 0.0);
    }, col);
}

// ---- FP32 runs: one sub-step in two passes ----
// The predictor and the corrector march along the columns: one wave owns 62
// rows (lanes 1..62; lanes 0 and 63 are the one-row halo) and keeps the
// column to its left in registers, so each array is read once and no flux
// array is stored.  The heads, the saturated thicknesses and the head
// differences across a face are taken in FP64; the face flux, its donor cap
// and the divergence in FP32 with reciprocal multiplies (an FP32 run's
// precision; a call moves the head by far less than the head's FP32
// resolution, so the head and its update stay FP64).  Same fluxes, limiter,
// divergence, bed, exfiltration and surface clamp as the four passes above,
// which the FP64 runs keep in the reference's operation order.
constexpr int kGwRowsOut = kWave - 2;  // rows finalised per wave

__device__ __forceinline__ double gw_lane_up(double v) { return __shfl_up(v, 1); }
__device__ __forceinline__ double gw_lane_down(double v) { return __shfl_down(v, 1); }
__device__ __forceinline__ float gw_lane_up(float v) { return __shfl_up(v, 1); }
__device__ __forceinline__ float gw_lane_down(float v) { return __shfl_down(v, 1); }

// flux across the face from cell A (donor side: the left / upper cell) to B
__device__ __forceinline__ float gw_face32(double hA, float HA, double hB, float HB, float K,
                                           float Sy_A, float inv_d, float inv_dt) {
    const float Hx = 0.5f * (HA + HB);
    const float F = -K * Hx * ((float)(hB - hA) * inv_d);
    const float cap = Sy_A * HA * inv_dt;
    float Fl = copysignf(fminf(fabsf(F), cap), F);
    if (F == 0.0f) Fl = 0.0f;
    if (isnan(F)) Fl = F;  // signum(NaN) * ... of the reference
    return Fl;
}

// one head's values in a column: h, H = max(h - z0, 0)
struct GwCol {
    double h;
    float H;
};

// the head this sub-step starts from: the entry head, then the two masked-head
// buffers in turn (a sub-step never writes the head it reads)
__device__ __forceinline__ const double *gw_fast_head(const GwArgs &a, const double *h0,
                                                      const double *hA, const double *hB) {
    const int k = a.ctl->nsub;
    return k == 0 ? h0 : ((k & 1) ? hA : hB);
}

// CORR = false: the predictor (:113-118), h_p from the fluxes of h_t.
// CORR = true: the corrector (:119-148) from the averaged fluxes of h_t and
// h_p; the masked head goes to the buffer the next sub-step reads (hA after an
// even count, hB after an odd one), the unmasked head to h_out for the next
// Courant limit, and each workgroup leaves its sums of the closing mass
// balance (R*dt, q_exf*dt, Sy*max(h0 - z0, 0), Sy*max(h - z0, 0)) in parts.
template <bool CORR>
__global__ __launch_bounds__(256) void gw_march_kernel(GwArgs a, const double *h0, double *hA,
                                                       double *hB, double *h_p, double *h_out,
                                                       double *q_exf, double *parts, int tj) {
    if (a.ctl->done) return;
    const int ny = a.ny, nx = a.nx;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int n_slabs = (ny + kGwRowsOut - 1) / kGwRowsOut;
    const int n_bx = (n_slabs + 3) / 4;
    const int logical = xcd_logical_tile();
    const int bx = logical % n_bx, by = logical / n_bx;
    const int slab = bx * 4 + wv;
    const int r = slab * kGwRowsOut - 1 + lane;
    const bool row_in = slab < n_slabs && r >= 0 && r < ny;
    const bool own = row_in && lane >= 1 && lane <= kGwRowsOut;
    const int c0 = by * tj, c1 = min(c0 + tj, nx);
    const double dt = a.ctl->dt_used, inv_dt64 = 1.0 / dt;
    const float inv_dt = (float)inv_dt64, inv_dx = (float)(1.0 / a.dx), inv_dy = (float)(1.0 / a.dy);
    const double *h_t = gw_fast_head(a, h0, hA, hB);
    double *h_mask = (a.ctl->nsub & 1) ? hB : hA;

    // one column's loads of the lane's row (zero thickness outside the grid):
    // the heads with their saturated thickness, the conductivity and specific
    // yield, and (corrector) the recharge and the entry head of the balance
    struct Col {
        GwCol t, p;
        double z0, R, h0;
        float K, Sy;
    };
    auto load = [&](int j) {
        Col v;
        v.t = GwCol{0.0, 0.0f};
        v.p = GwCol{0.0, 0.0f};
        v.z0 = v.R = v.h0 = 0.0;
        v.K = v.Sy = 0.0f;
        if (row_in && j >= 0 && j < nx) {
            const long long c = (long long)j * ny + r;
            v.z0 = a.z0[c];
            v.t.h = h_t[c];
            v.t.H = (float)fmax(v.t.h - v.z0, 0.0);
            if constexpr (CORR) {
                v.p.h = h_p[c];
                v.p.H = (float)fmax(v.p.h - v.z0, 0.0);
                v.R = a.R[c];
                v.h0 = h0[c];
            }
            v.K = (float)gw_K(a, c);
            v.Sy = (float)gw_Sy(a, c);
        }
        return v;
    };
    // the flux of the x face between columns j and j + 1 (0 past the grid edge)
    auto fx = [&](const GwCol &A, const GwCol &B, float KA, float KB, float SyA, int j) {
        return (j >= 0 && j < nx - 1) ? gw_face32(A.h, A.H, B.h, B.H, 0.5f * (KA + KB), SyA, inv_dx, inv_dt)
                                      : 0.0f;
    };
    // the flux of the y face between this row and the next one down (every
    // lane takes part in the shuffles)
    auto fy = [&](const GwCol &A, float KA, float SyA) {
        const double hB = gw_lane_down(A.h);
        const float HB = gw_lane_down(A.H), KB = gw_lane_down(KA);
        const float F = gw_face32(A.h, A.H, hB, HB, 0.5f * (KA + KB), SyA, inv_dy, inv_dt);
        return r < ny - 1 ? F : 0.0f;
    };

    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (slab < n_slabs) {
        // loads run two columns ahead of the compute
        Col prev = load(c0 - 1), cur = load(c0), nxt = load(c0 + 1);
        // the x faces between c0 - 1 and c0
        float fl1 = fx(prev.t, cur.t, prev.K, cur.K, prev.Sy, c0 - 1);
        float fl2 = 0.0f;
        if constexpr (CORR) fl2 = fx(prev.p, cur.p, prev.K, cur.K, prev.Sy, c0 - 1);
        for (int j = c0; j < c1; ++j) {
            const Col nxt2 = load(j + 2 < c1 + 1 ? j + 2 : nx);  // past the chunk: nothing
            const float fr1 = fx(cur.t, nxt.t, cur.K, nxt.K, cur.Sy, j);
            const float fd1 = fy(cur.t, cur.K, cur.Sy);
            const float up1 = gw_lane_up(fd1);
            const float fu1 = r > 0 ? up1 : 0.0f;
            float fr = fr1, fl = fl1, fd = fd1, fu = fu1;
            float fr2 = 0.0f;
            if constexpr (CORR) {
                fr2 = fx(cur.p, nxt.p, cur.K, nxt.K, cur.Sy, j);
                const float fd2 = fy(cur.p, cur.K, cur.Sy);
                const float up2 = gw_lane_up(fd2);
                const float fu2 = r > 0 ? up2 : 0.0f;
                fr = 0.5f * (fr1 + fr2);
                fl = 0.5f * (fl1 + fl2);
                fd = 0.5f * (fd1 + fd2);
                fu = 0.5f * (fu1 + fu2);
            }
            if (own) {
                const long long c = (long long)j * ny + r;
                const float div = (fr - fl) * inv_dx + (fd - fu) * inv_dy;
                const double Sy = gw_Sy(a, c);
                if constexpr (!CORR) {
                    double v = cur.t.h + dt / Sy * (-(double)div);
                    if (a.dirichlet && a.dirichlet[c]) v = a.h_dir[c];
                    h_p[c] = fmax(v, cur.z0);
                } else {
                    double v = cur.t.h + dt / Sy * (-(double)div + cur.R);
                    if (a.dirichlet && a.dirichlet[c]) v = a.h_dir[c];
                    v = fmax(v, cur.z0);
                    const double hs = cur.z0 + (a.h_soil ? a.h_soil[c] : a.h_soil_u);
                    const bool out = a.catchment && !a.catchment[c];
                    double q = fmax(0.0, v - hs) * Sy * inv_dt64;
                    if (out) q = NAN;
                    q_exf[c] = q;
                    const double hn = fmin(v, hs);
                    h_out[c] = hn;
                    h_mask[c] = out ? NAN : hn;
                    const double v0 = cur.R * dt, v1 = q * dt;
                    const double v2 = Sy * fmax(cur.h0 - cur.z0, 0.0);
                    const double v3 = Sy * fmax((out ? NAN : hn) - cur.z0, 0.0);
                    if (!isnan(v0)) s0 += v0;
                    if (!isnan(v1)) s1 += v1;
                    if (!isnan(v2)) s2 += v2;
                    if (!isnan(v3)) s3 += v3;
                }
            }
            // shift the march
            cur = nxt;
            nxt = nxt2;
            fl1 = fr1;
            if constexpr (CORR) fl2 = fr2;
        }
    }
    if constexpr (CORR) {
        // workgroup sums in a fixed order (the waves in turn)
        __shared__ double sh[4][4];
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        s3 = wave_sum(s3);
        if (lane == 0) {
            sh[wv][0] = s0;
            sh[wv][1] = s1;
            sh[wv][2] = s2;
            sh[wv][3] = s3;
        }
        __syncthreads();
        if (threadIdx.x < 4)
            parts[4LL * logical + threadIdx.x] =
                sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
    }
}

// workgroups of the marching passes (column chunks of kGwMarchCols)
constexpr int kGwMarchCols = 32;
inline int gw_march_blocks(int ny, int nx) {
    const int n_slabs = (ny + kGwRowsOut - 1) / kGwRowsOut;
    return ((n_slabs + 3) / 4) * ((nx + kGwMarchCols - 1) / kGwMarchCols);
}

// compute_boussinesq_velocities (+ mask) and the Courant candidates of
// compute_stable_dt; h_masked (if given) receives h with NaN outside
__global__ __launch_bounds__(kGwBlock) void gw_velocity_kernel(GwArgs a, const double *h,
                                                               double *ux, double *uy,
                                                               double *h_masked, int mask_on,
                                                               int check_done) {
    if (check_done && a.ctl->done) return;
    // the fast sub-step leaves the masked head itself: after the call's last
    // sub-step nothing here is needed
    if (check_done && !h_masked && !ux && !uy && a.ctl->last) return;
    double best = INFINITY;
    const int ny = a.ny;
    // grid-stride over the cells: one atomic per workgroup of a bounded grid
    for (long long c = blockIdx.x * (long long)kGwBlock + threadIdx.x; c < a.n;
         c += (long long)gridDim.x * kGwBlock) {
        int i, j;
        gw_rc(c, ny, i, j);
        double vx = NAN, vy = NAN;
        if (j < a.nx - 1) vx = -(0.5 * (gw_K(a, c) + gw_K(a, c + ny))) * ((h[c + ny] - h[c]) / a.dx);
        if (i < ny - 1) vy = -(0.5 * (gw_K(a, c) + gw_K(a, c + 1))) * ((h[c + 1] - h[c]) / a.dy);
        const bool out = mask_on && a.catchment && !a.catchment[c];
        if (out) {
            vx = NAN;
            vy = NAN;
        }
        if (ux) ux[c] = vx;  // the Darcy velocities, when the caller returns them
        if (uy) uy[c] = vy;
        if (h_masked) h_masked[c] = out ? NAN : h[c];
        const double cx = a.courant * a.dx / fabs(vx), cy = a.courant * a.dy / fabs(vy);
        if (isfinite(cx) && cx > 0.0) best = fmin(best, cx);
        if (isfinite(cy) && cy > 0.0) best = fmin(best, cy);
    }
    best = wave_fmin(best);
    __shared__ double sh[kGwBlock / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        for (int k = 1; k < kGwBlock / 64; ++k) m = fmin(m, sh[k]);
        if (m < INFINITY)
            atomicMin(&a.ctl->dtgw_bits, (unsigned long long)__double_as_longlong(m));
    }
}

// per-column nansum of R*dt, q_exf*dt, Sy*max(h0 - z0, 0), Sy*max(h - z0, 0)
__global__ __launch_bounds__(kColThreads) void gw_colsum_kernel(GwArgs a, const double *q_exf,
                                                                const double *h0, const double *h,
                                                                double *out) {
    const double dt = a.ctl->dt_used;
    const bool is_max[4] = {false, false, false, false};
    column_reduce<4>(a.ny, blockIdx.x, is_max, [&](long long c, double (&v)[4]) {
        v[0] = a.R[c] * dt;
        v[1] = q_exf[c] * dt;
        v[2] = gw_Sy(a, c) * fmax(h0[c] - a.z0[c], 0.0);
        v[3] = gw_Sy(a, c) * fmax(h[c] - a.z0[c], 0.0);
    }, out);
}

hp2d_status gw_run(hp2d_handle *h, const GwArgs &a_in, double dt, const double *dh0,
                   const GwScratch &w, double *error, int32_t *n_substeps,
                   const double **h_final) {
    const int ny = a_in.ny, nx = a_in.nx;
    const size_t n = (size_t)ny * nx, nb = n * sizeof(double);
    const int blocks = (int)((n + kGwBlock - 1) / kGwBlock);
    hipStream_t s = hp2d_handle_stream(h);
    GwArgs a = a_in;
    a.ctl = w.ctl;
    // the first sub-step reads the entry head in place; the velocity pass of
    // each sub-step leaves the (masked) head of the next in w.ht
    hipError_t e = hipSuccess;
    const dim3 g(blocks), b(kGwBlock), gv(std::min(blocks, 2048));
    {
        hipLaunchKernelGGL(gw_ctl_init, dim3(1), dim3(1), 0, s, w.ctl, dt);
        // velocities of the entry head (not masked, Boussinesq_2D_explicit.m:91)
        hipLaunchKernelGGL(gw_velocity_kernel, gv, b, 0, s, a, dh0, w.ux, w.uy, (double *)nullptr, 0,
                           0);
        e = hipGetLastError();
    }
    bool first = true;
    GwCtl hc{};
    const int kMaxSub = 1000000;
    int batch = 1;
    while (e == hipSuccess) {
        for (int k = 0; k < batch; ++k) {
            const double *ht = first ? dh0 : (const double *)w.ht;
            first = false;
            hipLaunchKernelGGL(gw_ctl_begin, dim3(1), dim3(1), 0, s, w.ctl);
            if (a.fast) {
                // masked heads alternate between w.ht and w.Fx (unused here);
                // the workgroup sums of the balance go to w.Fy
                const dim3 gm(gw_march_blocks(ny, nx));
                hipLaunchKernelGGL(gw_march_kernel<false>, gm, dim3(256), 0, s, a, dh0, w.ht, w.Fx,
                                   w.hp, w.tmp, w.q, w.Fy, kGwMarchCols);
                hipLaunchKernelGGL(gw_march_kernel<true>, gm, dim3(256), 0, s, a, dh0, w.ht, w.Fx,
                                   w.hp, w.tmp, w.q, w.Fy, kGwMarchCols);
                hipLaunchKernelGGL(gw_velocity_kernel, gv, b, 0, s, a, (const double *)w.tmp, w.ux,
                                   w.uy, (double *)nullptr, 1, 1);
            } else {
                hipLaunchKernelGGL(gw_flux_kernel, g, b, 0, s, a, ht, w.Fx, w.Fy, 0);
                hipLaunchKernelGGL(gw_predict_kernel, g, b, 0, s, a, ht, (const double *)w.Fx,
                                   (const double *)w.Fy, w.hp);
                hipLaunchKernelGGL(gw_flux_kernel, g, b, 0, s, a, (const double *)w.hp, w.Fx, w.Fy, 1);
                hipLaunchKernelGGL(gw_update_kernel, dim3(nx), dim3(kColThreads), 0, s, a, ht,
                                   (const double *)w.Fx, (const double *)w.Fy, w.tmp, w.q, dh0, w.col);
                hipLaunchKernelGGL(gw_velocity_kernel, gv, b, 0, s, a, (const double *)w.tmp, w.ux,
                                   w.uy, w.ht, 1, 1);
            }
            hipLaunchKernelGGL(gw_ctl_end, dim3(1), dim3(1), 0, s, w.ctl);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&hc, w.ctl, sizeof(GwCtl), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess || hc.done) break;
        if (hc.nsub > kMaxSub) return hp2d_handle_fail(h, "Boussinesq: more than 1e6 sub-steps");
        batch = batch < 64 ? batch * 2 : 64;
    }
    if (e == hipSuccess && hc.nsub == 0) {  // no sub-step ran: the entry head, and its sums
        e = hipMemcpyAsync(w.ht, dh0, nb, hipMemcpyDeviceToDevice, s);
        hipLaunchKernelGGL(gw_colsum_kernel, dim3(nx), dim3(kColThreads), 0, s, a, (const double *)w.q, dh0,
                           (const double *)w.ht, w.col);
        e = hipGetLastError();
    }
    // the closing sums: per column in column order, or (fast sub-steps) per
    // workgroup of the marching corrector
    const bool fast_sums = a.fast && hc.nsub > 0;
    const size_t n_parts = fast_sums ? (size_t)gw_march_blocks(ny, nx) : (size_t)nx;
    std::vector<double> col(4 * n_parts);
    if (e == hipSuccess)
        e = hipMemcpyAsync(col.data(), fast_sums ? w.Fy : w.col, col.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hp2d_handle_fail(h, std::string("Boussinesq: ") + hipGetErrorString(e));
    double sums[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t j = 0; j < n_parts; ++j)
        for (int k = 0; k < 4; ++k) sums[k] += col[4 * j + k];
    const double A = a.dx * a.dy;
    const double recharge_mass = sums[0] * A, exfil_mass = sums[1] * A;
    const double storage_prev = sums[2] * A, storage_curr = sums[3] * A;
    *error = recharge_mass - exfil_mass - (storage_curr - storage_prev);
    if (n_substeps) *n_substeps = hc.nsub;
    // the final head: w.ht, or w.Fx after an even count of fast sub-steps
    if (h_final) *h_final = (a.fast && hc.nsub > 0 && hc.nsub % 2 == 0) ? w.Fx : w.ht;
    return HP2D_OK;
}

size_t gw_scratch_bytes(int ny, int nx) {
    const size_t n = (size_t)ny * nx;
    return 8 * n * sizeof(double) + 4 * (size_t)nx * sizeof(double) + 256;
}

GwScratch gw_scratch(char *base, int ny, int nx) {
    const size_t n = (size_t)ny * nx;
    GwScratch w;
    double *p = (double *)base;
    w.ht = p; p += n;
    w.hp = p; p += n;
    w.tmp = p; p += n;
    w.Fx = p; p += n;
    w.Fy = p; p += n;
    w.ux = p; p += n;
    w.uy = p; p += n;
    w.q = p; p += n;
    w.col = p;
    w.ctl = (GwCtl *)(base + 8 * n * sizeof(double) + 4 * (size_t)nx * sizeof(double));
    return w;
}

}  // namespace hp2d

extern "C" hp2d_status hp2d_boussinesq_2d(hp2d_handle *h, double dt, double dx, double dy,
                                          const double *h_in, const double *z0, const double *Sy,
                                          const double *R, const double *K, double courant,
                                          const double *h_soil, const uint8_t *catchment_mask,
                                          const uint8_t *dirichlet_mask, const double *h_dirichlet,
                                          double *h_t, double *u_x, double *u_y, double *q_exf,
                                          double *q_river, double *error, int32_t *n_substeps) {
    using namespace hp2d;
    if (!h || !h_in || !z0 || !Sy || !R || !K || !h_soil || !h_t || !q_exf || !error ||
        !(dt > 0.0) || !(dx > 0.0) || !(dy > 0.0) || !(courant > 0.0) ||
        (dirichlet_mask && !h_dirichlet))
        return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_status st = hp2d_handle_shape(h, &ny, &nx, &prec);
    if (st != HP2D_OK) return st;
    if (ny < 2 || nx < 2) return HP2D_E_ARG;  // the reference indexes column/row 2
    const size_t n = (size_t)ny * nx, nb = n * sizeof(double);
    hipStream_t s = hp2d_handle_stream(h);
    // workspace: inputs z0 Sy R K h_soil h_dir h0 | masks | Boussinesq scratch
    const size_t in_bytes = 7 * nb;
    const size_t mask_off = in_bytes;
    const size_t scr_off = ((mask_off + 2 * n + 255) / 256) * 256;
    char *ws = (char *)hp2d_handle_workspace(h, scr_off + gw_scratch_bytes(ny, nx) + 256);
    if (!ws) return HP2D_E_NOMEM;
    double *p = (double *)ws;
    double *dz0 = p, *dSy = p + n, *dR = p + 2 * n, *dK = p + 3 * n, *dhs = p + 4 * n,
           *dhd = p + 5 * n, *dh0 = p + 6 * n;
    uint8_t *dcm = catchment_mask ? (uint8_t *)(ws + mask_off) : nullptr;
    uint8_t *ddm = dirichlet_mask ? (uint8_t *)(ws + mask_off + n) : nullptr;
    GwScratch w = gw_scratch(ws + scr_off, ny, nx);
    if (!u_x) w.ux = nullptr;  // velocities only where the caller returns them
    if (!u_y) w.uy = nullptr;
    hipError_t e = hipSuccess;
    auto H2D = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    auto D2H = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    H2D(dz0, z0, nb);
    H2D(dSy, Sy, nb);
    H2D(dR, R, nb);
    H2D(dK, K, nb);
    H2D(dhs, h_soil, nb);
    H2D(dh0, h_in, nb);
    if (ddm) {
        H2D(dhd, h_dirichlet, nb);
        H2D(ddm, dirichlet_mask, n);
    }
    if (dcm) H2D(dcm, catchment_mask, n);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_boussinesq_2d: ") + hipGetErrorString(e));
    GwArgs a;
    a.ny = ny;
    a.nx = nx;
    a.n = (long long)n;
    a.dx = dx;
    a.dy = dy;
    a.courant = courant;
    a.z0 = dz0;
    a.Sy = dSy;
    a.R = dR;
    a.K = dK;
    a.h_soil = dhs;
    a.h_dir = dhd;
    a.catchment = dcm;
    a.dirichlet = ddm;
    a.ctl = w.ctl;
    a.fast = prec == HP2D_FP32;  // an FP32 handle: the FP32 runs' two-pass sub-steps
    const double *h_fin = w.ht;
    st = gw_run(h, a, dt, dh0, w, error, n_substeps, &h_fin);
    if (st != HP2D_OK) return st;
    D2H(h_t, h_fin, nb);
    D2H(q_exf, w.q, nb);
    if (u_x) D2H(u_x, w.ux, nb);
    if (u_y) D2H(u_y, w.uy, nb);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_boussinesq_2d: ") + hipGetErrorString(e));
    if (q_river)
        for (size_t c = 0; c < n; ++c) q_river[c] = 0.0;
    return HP2D_OK;
}
