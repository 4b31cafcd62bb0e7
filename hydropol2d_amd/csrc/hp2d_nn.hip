// hp2d_nn.hip - BF16 MFMA kernels of the flood surrogate (forward, backward,
// optimiser) for gfx950.
//
// Activations are NHWC bf16 (channels innermost, a multiple of 32); 3x3
// weights are [Cout][3][3][Cin] bf16, i.e. a Cout x 9*Cin matrix whose k index
// is (kh*3 + kw)*Cin + ci.  Every 3x3 convolution is an implicit GEMM on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//   forward      C[pixel][cout]   = im2col(x)[pixel][k] * W^T[k][cout]
//   data grad    the forward GEMM on the output gradient with the weights
//                flipped in space and transposed in channels
//   weight grad  C[cout][k]       = g^T[cout][pixel] * im2col(x)[pixel][k],
//                split over pixel ranges, FP32 atomics into the gradient
// A workgroup of 4 waves computes a TM x TN tile (2 x 2 waves, each wave
// (TM/2) x (TN/2) from 16x16 MFMA blocks); each K step of 32 stages a TM x 32
// and a TN x 32 bf16 tile in LDS (rows padded against bank conflicts), from
// which every lane reads its 8-element operand with one 16-byte LDS read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "hp2d.h"

namespace hp2d {
void set_global_error(const std::string &m);
}

namespace hp2d {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kKStep = 32;
constexpr int kLdsRow = kKStep + 8;  // 40 bf16 = 80 B per staged row

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                      // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

enum { kFwd = 0, kWgrad = 1 };

struct ConvArgs {
    int N, H, W, Cin, Cout;
    const uint16_t *x;      // [N,H,W,Cin]
    const uint16_t *w;      // forward: [Cout][9*Cin]
    const uint16_t *g;      // wgrad: output gradient [N,H,W,Cout]
    const float *bias;      // forward, may be NULL
    const uint16_t *skip;   // forward: added after the activation, may be NULL
    uint16_t *y;            // forward: activation output (relu(conv+b) or conv+b)
    uint16_t *y_skip;       // forward: y + skip (when skip)
    float *gw;              // wgrad: [Cout][9*Cin] FP32, accumulated
    int relu;
    long long pix_per_block;  // wgrad split
};

// One 16-byte chunk (8 channels) of the input pixel under tap (kh, kw) of
// output pixel p, or zeros outside the image.
__device__ __forceinline__ uint4 load_tap(const ConvArgs &a, const uint16_t *src, int C, long long p,
                                          int kh, int kw, int c0) {
    const long long hw = (long long)a.H * a.W;
    const int n = (int)(p / hw);
    const int rem = (int)(p - (long long)n * hw);
    const int h = rem / a.W + kh - 1, w = rem % a.W + kw - 1;
    if (h < 0 || h >= a.H || w < 0 || w >= a.W) return make_uint4(0, 0, 0, 0);
    const uint16_t *q = src + (((long long)n * a.H + h) * a.W + w) * C + c0;
    return *reinterpret_cast<const uint4 *>(q);
}

template <int TM, int TN>
__device__ __forceinline__ void mfma_tile(const uint16_t (*As)[kLdsRow], const uint16_t (*Bs)[kLdsRow],
                                          f32x4 (&acc)[TM / 32][TN / 32], int wm, int wn, int lane) {
    const int r = lane & 15, kc = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < TM / 32; ++i) {
        const uint4 av = *reinterpret_cast<const uint4 *>(&As[wm * (TM / 2) + i * 16 + r][kc]);
        const bf16x8 A = __builtin_bit_cast(bf16x8, av);
#pragma unroll
        for (int j = 0; j < TN / 32; ++j) {
            const uint4 bv = *reinterpret_cast<const uint4 *>(&Bs[wn * (TN / 2) + j * 16 + r][kc]);
            const bf16x8 B = __builtin_bit_cast(bf16x8, bv);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[i][j], 0, 0, 0);
        }
    }
}

// Forward (and data-gradient) convolution: M = pixels, N = Cout, K = 9*Cin.
template <int TM, int TN>
__global__ __launch_bounds__(kBlock) void conv3x3_fwd_kernel(ConvArgs a) {
    __shared__ uint16_t As[TM][kLdsRow];
    __shared__ uint16_t Bs[TN][kLdsRow];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long long P = (long long)a.N * a.H * a.W;
    const long long p0 = (long long)blockIdx.x * TM;
    const int n0 = blockIdx.y * TN;
    f32x4 acc[TM / 32][TN / 32];
#pragma unroll
    for (int i = 0; i < TM / 32; ++i)
#pragma unroll
        for (int j = 0; j < TN / 32; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = 9 * a.Cin;
    for (int k0 = 0; k0 < K; k0 += kKStep) {
        const int tap = k0 / a.Cin, c0 = k0 - tap * a.Cin;
        const int kh = tap / 3, kw = tap % 3;
        // A: TM pixels x 32 channels (4 chunks of 8 per pixel)
        for (int e = tid; e < TM * 4; e += kBlock) {
            const int row = e >> 2, ch = (e & 3) * 8;
            const long long p = p0 + row;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (p < P) v = load_tap(a, a.x, a.Cin, p, kh, kw, c0 + ch);
            *reinterpret_cast<uint4 *>(&As[row][ch]) = v;
        }
        // B: TN output channels x 32 k
        for (int e = tid; e < TN * 4; e += kBlock) {
            const int row = e >> 2, ch = (e & 3) * 8;
            const int n = n0 + row;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (n < a.Cout) v = *reinterpret_cast<const uint4 *>(a.w + (long long)n * K + k0 + ch);
            *reinterpret_cast<uint4 *>(&Bs[row][ch]) = v;
        }
        __syncthreads();
        mfma_tile<TM, TN>(As, Bs, acc, wm, wn, lane);
        __syncthreads();
    }
    // epilogue: lane holds column n = lane%16 and rows 4*(lane/16)+q of each 16x16 block
    const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < TM / 32; ++i)
#pragma unroll
        for (int j = 0; j < TN / 32; ++j) {
            const int n = n0 + wn * (TN / 2) + j * 16 + col;
            if (n >= a.Cout) continue;
            const float b = a.bias ? a.bias[n] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const long long p = p0 + wm * (TM / 2) + i * 16 + rq + q;
                if (p >= P) continue;
                float v = acc[i][j][q] + b;
                if (a.relu) v = fmaxf(v, 0.f);
                const uint16_t yb = f2bf(v);
                a.y[p * a.Cout + n] = yb;
                if (a.skip) a.y_skip[p * a.Cout + n] = f2bf(bf2f(yb) + bf2f(a.skip[p * a.Cout + n]));
            }
        }
}

// Weight gradient: M = Cout, N = 9*Cin (a block covers one tap and TN
// channels), K = pixels of this block's range; FP32 atomics into a.gw.
template <int TM, int TN>
__global__ __launch_bounds__(kBlock) void conv3x3_wgrad_kernel(ConvArgs a) {
    __shared__ uint16_t As[TM][kLdsRow];  // [cout][pixel]
    __shared__ uint16_t Bs[TN][kLdsRow];  // [channel][pixel]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long long P = (long long)a.N * a.H * a.W;
    const int m0 = blockIdx.x * TM;                       // cout block
    const int nb = blockIdx.y;                            // (tap, channel block)
    const int cblocks = a.Cin / TN;
    const int tap = nb / cblocks, c0 = (nb % cblocks) * TN;
    const int kh = tap / 3, kw = tap % 3;
    const long long pbeg = (long long)blockIdx.z * a.pix_per_block;
    const long long pend = pbeg + a.pix_per_block < P ? pbeg + a.pix_per_block : P;
    f32x4 acc[TM / 32][TN / 32];
#pragma unroll
    for (int i = 0; i < TM / 32; ++i)
#pragma unroll
        for (int j = 0; j < TN / 32; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long k0 = pbeg; k0 < pend; k0 += kKStep) {
        // A^T: 32 pixels x TM output channels of g, transposed into As[cout][pixel]
        for (int e = tid; e < kKStep * (TM / 8); e += kBlock) {
            const int px = e / (TM / 8), ch = (e % (TM / 8)) * 8;
            const long long p = k0 + px;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (p < pend) v = *reinterpret_cast<const uint4 *>(a.g + p * a.Cout + m0 + ch);
            const uint16_t *s = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
            for (int q = 0; q < 8; ++q) As[ch + q][px] = s[q];
        }
        // B^T: 32 pixels x TN input channels under the tap, into Bs[channel][pixel]
        for (int e = tid; e < kKStep * (TN / 8); e += kBlock) {
            const int px = e / (TN / 8), ch = (e % (TN / 8)) * 8;
            const long long p = k0 + px;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (p < pend) v = load_tap(a, a.x, a.Cin, p, kh, kw, c0 + ch);
            const uint16_t *s = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
            for (int q = 0; q < 8; ++q) Bs[ch + q][px] = s[q];
        }
        __syncthreads();
        mfma_tile<TM, TN>(As, Bs, acc, wm, wn, lane);
        __syncthreads();
    }
    const int col = lane & 15, rq = (lane >> 4) * 4;
    const int K = 9 * a.Cin;
#pragma unroll
    for (int i = 0; i < TM / 32; ++i)
#pragma unroll
        for (int j = 0; j < TN / 32; ++j) {
            const int kk = tap * a.Cin + c0 + wn * (TN / 2) + j * 16 + col;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm * (TM / 2) + i * 16 + rq + q;
                atomicAdd(&a.gw[(long long)m * K + kk], acc[i][j][q]);
            }
        }
}

// grid-stride loop bounds of the element-wise kernels
__device__ __forceinline__ long long gs_first() { return blockIdx.x * (long long)blockDim.x + threadIdx.x; }
__device__ __forceinline__ long long gs_stride() { return (long long)gridDim.x * blockDim.x; }

// W'[ci][kh][kw][co] = W[co][2-kh][2-kw][ci]: the data-gradient weights
__global__ void flip_transpose_kernel(const uint16_t *w, int Cin, int Cout, uint16_t *wt) {
    const long long total = 9LL * Cin * Cout;
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        const int ci = (int)(e % Cin);
        const int tap = (int)((e / Cin) % 9);
        const int co = (int)(e / (9LL * Cin));
        const int kh = tap / 3, kw = tap % 3;
        wt[((long long)ci * 9 + (2 - kh) * 3 + (2 - kw)) * Cout + co] = w[e];
    }
}

// ---- 8-channel vectors: the element-wise passes move 16 bytes per thread ----
// (every tensor here has a channel count that is a multiple of 32)
__device__ __forceinline__ void unpack8(const uint4 u, float (&f)[8]) {
    const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[2 * k] = __uint_as_float(w[k] << 16);
        f[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
    unsigned w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (unsigned)f2bf(f[2 * k]) | ((unsigned)f2bf(f[2 * k + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// bias gradient: gb[c] += sum over pixels of g[p][c].  A thread sums one
// 8-channel vector of a pixel over a strided set of pixels; the block folds
// the vectors of each channel group in LDS and adds one value per channel.
__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t *g, long long P, int C, float *gb) {
    const int V = C / 8;                       // vectors per pixel (C <= 2048)
    const int per = 256 / V;                   // pixels a block covers per pass
    const int v = threadIdx.x % V, sub = threadIdx.x / V;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (sub < per) {
        const uint4 *g4 = reinterpret_cast<const uint4 *>(g);
        for (long long p = blockIdx.x * (long long)per + sub; p < P; p += (long long)gridDim.x * per) {
            float f[8];
            unpack8(g4[p * V + v], f);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += f[k];
        }
    }
    __shared__ float red[256 * 8];
#pragma unroll
    for (int k = 0; k < 8; ++k) red[threadIdx.x * 8 + k] = s[k];
    __syncthreads();
    for (int c = threadIdx.x; c < C; c += 256) {
        const int vv = c / 8, k = c % 8;
        float t = 0.f;
        for (int q = 0; q < per; ++q) t += red[(q * V + vv) * 8 + k];
        atomicAdd(&gb[c], t);
    }
}

// g *= (y > 0): the ReLU backward on the stored activation
__global__ void relu_mask_kernel(uint16_t *g, const uint16_t *y, long long n) {
    const long long n8 = n / 8;
    uint4 *g4 = reinterpret_cast<uint4 *>(g);
    const uint4 *y4 = reinterpret_cast<const uint4 *>(y);
    for (long long i = gs_first(); i < n8; i += gs_stride()) {
        float fy[8];
        unpack8(y4[i], fy);
        uint4 u = g4[i];
        unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!(fy[2 * k] > 0.f)) w[k] &= 0xffff0000u;
            if (!(fy[2 * k + 1] > 0.f)) w[k] &= 0x0000ffffu;
        }
        g4[i] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    for (long long i = n8 * 8 + gs_first(); i < n; i += gs_stride())
        if (!(bf2f(y[i]) > 0.f)) g[i] = 0;
}

// out = a + b (bf16, for gradients meeting at a skip connection)
__global__ void add_kernel(const uint16_t *a, const uint16_t *b, uint16_t *out, long long n) {
    const long long n8 = n / 8;
    const uint4 *a4 = reinterpret_cast<const uint4 *>(a), *b4 = reinterpret_cast<const uint4 *>(b);
    uint4 *o4 = reinterpret_cast<uint4 *>(out);
    for (long long i = gs_first(); i < n8; i += gs_stride()) {
        float fa[8], fb[8];
        unpack8(a4[i], fa);
        unpack8(b4[i], fb);
#pragma unroll
        for (int k = 0; k < 8; ++k) fa[k] = fa[k] + fb[k];
        o4[i] = pack8(fa);
    }
    for (long long i = n8 * 8 + gs_first(); i < n; i += gs_stride()) out[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// Vector e of an N x H x W x (8V) tensor as (row q = n*H + h, column w, vector v)
// with 32-bit arithmetic once the row is known.
__device__ __forceinline__ void pix_of(long long e, int W, int V, long long &q, int &w, int &v) {
    const long long WV = (long long)W * V;
    q = e / WV;
    const int r = (int)(e - q * WV);
    w = r / V;
    v = r - w * V;
}

// 2x2 average pooling (H, W even; output H/2 x W/2) and its backward
__global__ void avgpool2_kernel(const uint16_t *x, int N, int H, int W, int C, uint16_t *y) {
    const int Ho = H / 2, Wo = W / 2, V = C / 8;
    const long long total = (long long)N * Ho * Wo * V;
    const uint4 *x4 = reinterpret_cast<const uint4 *>(x);
    uint4 *y4 = reinterpret_cast<uint4 *>(y);
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        long long q;
        int wo, v;
        pix_of(e, Wo, V, q, wo, v);  // q = n*Ho + ho
        const long long r0 = 2 * q;  // = n*H + 2*ho
        const long long i00 = (r0 * W + 2 * wo) * V + v, i10 = i00 + (long long)W * V;
        float a[8], b[8], c[8], d[8];
        unpack8(x4[i00], a);
        unpack8(x4[i00 + V], b);
        unpack8(x4[i10], c);
        unpack8(x4[i10 + V], d);
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = 0.25f * (a[k] + b[k] + c[k] + d[k]);
        y4[e] = pack8(a);
    }
}
__global__ void avgpool2_bwd_kernel(const uint16_t *gy, int N, int H, int W, int C, uint16_t *gx) {
    const int V = C / 8;
    const long long total = (long long)N * H * W * V;
    const uint4 *g4 = reinterpret_cast<const uint4 *>(gy);
    uint4 *o4 = reinterpret_cast<uint4 *>(gx);
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        long long q;
        int w, v;
        pix_of(e, W, V, q, w, v);  // q = n*H + h
        const long long qh = q / 2;  // = n*(H/2) + h/2 (H even)
        float f[8];
        unpack8(g4[(qh * (W / 2) + w / 2) * V + v], f);
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = 0.25f * f[k];
        o4[e] = pack8(f);
    }
}

// nearest 2x upsampling (output H, W) and its backward (sum over the 2x2 block)
__global__ void upsample2_kernel(const uint16_t *x, int N, int H, int W, int C, uint16_t *y) {
    const int V = C / 8;
    const long long total = (long long)N * H * W * V;
    const uint4 *x4 = reinterpret_cast<const uint4 *>(x);
    uint4 *y4 = reinterpret_cast<uint4 *>(y);
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        long long q;
        int w, v;
        pix_of(e, W, V, q, w, v);  // q = n*H + h
        y4[e] = x4[((q / 2) * (W / 2) + w / 2) * V + v];
    }
}
__global__ void upsample2_bwd_kernel(const uint16_t *gy, int N, int H, int W, int C, uint16_t *gx) {
    const int Ho = H / 2, Wo = W / 2, V = C / 8;
    const long long total = (long long)N * Ho * Wo * V;
    const uint4 *g4 = reinterpret_cast<const uint4 *>(gy);
    uint4 *o4 = reinterpret_cast<uint4 *>(gx);
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        long long q;
        int wo, v;
        pix_of(e, Wo, V, q, wo, v);  // q = n*Ho + ho
        const long long i00 = (2 * q * W + 2 * wo) * V + v, i10 = i00 + (long long)W * V;
        float a[8], b[8], c[8], d[8];
        unpack8(g4[i00], a);
        unpack8(g4[i00 + V], b);
        unpack8(g4[i10], c);
        unpack8(g4[i10 + V], d);
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = a[k] + b[k] + c[k] + d[k];
        o4[e] = pack8(a);
    }
}

// 1x1 head: out[p][o] = b[o] + sum_c x[p][c] w[o][c] (FP32 out, O outputs);
// channels read as 16-byte vectors
__global__ void head_fwd_kernel(const uint16_t *x, long long P, int C, const float *w,
                                const float *b, int O, float *out) {
    const int V = C / 8;
    const uint4 *x4 = reinterpret_cast<const uint4 *>(x);
    for (long long p = gs_first(); p < P; p += gs_stride()) {
        float s[4] = {0.f, 0.f, 0.f, 0.f};  // O <= 4
        for (int v = 0; v < V; ++v) {
            float f[8];
            unpack8(x4[p * V + v], f);
            for (int o = 0; o < O; ++o)
#pragma unroll
                for (int k = 0; k < 8; ++k) s[o] += f[k] * w[o * C + v * 8 + k];
        }
        for (int o = 0; o < O; ++o) out[p * O + o] = b[o] + s[o];
    }
}
// its backward: gx[p][c] = sum_o gy[p][o] w[o][c] (bf16); gw, gb accumulated
// in FP32 (grid-stride over pixels, per-thread sums, one LDS fold per block)
template <int C, int O>
__global__ __launch_bounds__(256) void head_bwd_kernel(const uint16_t *x, long long P,
                                                       const float *w, const float *gy,
                                                       uint16_t *gx, float *gw, float *gb) {
    float aw[O][C], ab[O];
#pragma unroll
    for (int o = 0; o < O; ++o) {
        ab[o] = 0.f;
#pragma unroll
        for (int c = 0; c < C; ++c) aw[o][c] = 0.f;
    }
    for (long long p = blockIdx.x * 256LL + threadIdx.x; p < P; p += (long long)gridDim.x * 256) {
        float go[O];
#pragma unroll
        for (int o = 0; o < O; ++o) {
            go[o] = gy[p * O + o];
            ab[o] += go[o];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float xc = bf2f(x[p * C + c]);
            float s = 0.f;
#pragma unroll
            for (int o = 0; o < O; ++o) {
                s += go[o] * w[o * C + c];
                aw[o][c] += go[o] * xc;
            }
            gx[p * C + c] = f2bf(s);
        }
    }
    __shared__ float red[O * C + O];
    for (int i = threadIdx.x; i < O * C + O; i += 256) red[i] = 0.f;
    __syncthreads();
#pragma unroll
    for (int o = 0; o < O; ++o) {
        float v = ab[o];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) atomicAdd(&red[O * C + o], v);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float u = aw[o][c];
            for (int off = 32; off > 0; off >>= 1) u += __shfl_xor(u, off);
            if ((threadIdx.x & 63) == 0) atomicAdd(&red[o * C + c], u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < O * C; i += 256) atomicAdd(&gw[i], red[i]);
    if (threadIdx.x < O) atomicAdd(&gb[threadIdx.x], red[O * C + threadIdx.x]);
}

// masked mean squared error over the O output maps; gy = d loss / d out.  The
// loss is folded per block (one atomic add per block of a capped grid).
__global__ __launch_bounds__(256) void mse_kernel(const float *out, const float *target,
                                                  const uint8_t *mask, long long P, int O,
                                                  float inv_count, float *gy, float *loss) {
    float l = 0.f;
    for (long long p = gs_first(); p < P; p += gs_stride()) {
        const bool on = !mask || mask[p];
        for (int o = 0; o < O; ++o) {
            const float d = on ? out[p * O + o] - target[p * O + o] : 0.f;
            gy[p * O + o] = 2.f * d * inv_count;
            l += d * d;
        }
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) * inv_count);
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the bf16 copy
__global__ void adam_kernel(float *wt, float *m, float *v, const float *g, long long n, float lr,
                            float b1, float b2, float eps, float c1, float c2, uint16_t *wb) {
    for (long long i = gs_first(); i < n; i += gs_stride()) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float w = wt[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        wt[i] = w;
        if (wb) wb[i] = f2bf(w);
    }
}

__global__ void f32_to_bf16_kernel(const float *x, long long n, uint16_t *y) {
    for (long long i = gs_first(); i < n; i += gs_stride()) y[i] = f2bf(x[i]);
}

// Grid of an element-wise launch: one thread per element up to kMaxGridBlocks
// blocks, the kernels' grid-stride loops take the rest.  A dispatch counts its
// work-items in 32 bits, so an uncapped grid over more than 2^32 elements
// would silently cover only part of them.
constexpr long long kMaxGridBlocks = 1 << 16;
inline dim3 grid1(long long n, int b = 256) {
    long long blocks = (n + b - 1) / b;
    if (blocks > kMaxGridBlocks) blocks = kMaxGridBlocks;
    return dim3((unsigned)(blocks > 0 ? blocks : 1));
}

}  // namespace nn
}  // namespace hp2d

using namespace hp2d::nn;

extern "C" {

hp2d_status hp2d_nn_conv3x3_fwd(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                                const uint16_t *x, const uint16_t *w, const float *bias,
                                int32_t relu, const uint16_t *skip, uint16_t *y, uint16_t *y_skip,
                                void *stream) {
    if (N < 1 || H < 1 || W < 1 || Cin % 32 || Cout % 32 || !x || !w || !y || (skip && !y_skip))
        return HP2D_E_ARG;
    ConvArgs a{};
    a.N = N; a.H = H; a.W = W; a.Cin = Cin; a.Cout = Cout;
    a.x = x; a.w = w; a.bias = bias; a.skip = skip; a.y = y; a.y_skip = y_skip; a.relu = relu;
    const long long P = (long long)N * H * W;
    hipStream_t s = (hipStream_t)stream;
    if (Cout % 64 == 0)
        hipLaunchKernelGGL((conv3x3_fwd_kernel<64, 64>), dim3((unsigned)((P + 63) / 64), Cout / 64),
                           dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((conv3x3_fwd_kernel<64, 32>), dim3((unsigned)((P + 63) / 64), Cout / 32),
                           dim3(kBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_conv3x3_dgrad(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                                  const uint16_t *gy, const uint16_t *w, uint16_t *w_flip,
                                  uint16_t *gx, void *stream) {
    if (!gy || !w || !w_flip || !gx || Cin % 32 || Cout % 32) return HP2D_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(flip_transpose_kernel, grid1(9LL * Cin * Cout), dim3(256), 0, s, w, Cin, Cout,
                       w_flip);
    if (hipGetLastError() != hipSuccess) return HP2D_E_HIP;
    return hp2d_nn_conv3x3_fwd(N, H, W, Cout, Cin, gy, w_flip, nullptr, 0, nullptr, gx, nullptr,
                               stream);
}

}  // extern "C"

// Device scratch of the allocation-free data gradient: the flipped, transposed
// weights of one layer.  Every GPU holds its own copy of the array until the
// process ends; its device address differs per GPU and is taken with
// hipGetSymbolAddress on the first call on each device (the array's name is a
// host-side stand-in and never reaches a kernel).
static constexpr long long kNoAllocFlipMax = 9LL * 512 * 256;  // bf16 elements, 2.25 MiB
static __device__ __align__(256) uint16_t scratch[kNoAllocFlipMax];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
uint16_t *scratch_ptr[kMaxDevices] = {};

// the current device's address of `scratch`, or nullptr with the error set
uint16_t *scratch_on_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        hp2d::set_global_error("hp2d_nn_conv3x3_dgrad_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            hp2d::set_global_error(std::string("hp2d_nn_conv3x3_dgrad_noalloc: "
                                               "hipGetSymbolAddress: ") + hipGetErrorString(e));
            return nullptr;
        }
        scratch_ptr[dev] = static_cast<uint16_t *>(p);
    }
    return scratch_ptr[dev];
}
}  // namespace

extern "C" {

hp2d_status hp2d_nn_conv3x3_dgrad_noalloc(int32_t N, int32_t H, int32_t W, int32_t Cin,
                                          int32_t Cout, const uint16_t *gy, const uint16_t *w,
                                          uint16_t *gx, void *stream) {
    if (!gy || !w || !gx || N < 1 || H < 1 || W < 1 || Cin < 32 || Cout < 32 || Cin % 32 ||
        Cout % 32) {
        hp2d::set_global_error("hp2d_nn_conv3x3_dgrad_noalloc: bad arguments");
        return HP2D_E_ARG;
    }
    if (9LL * Cin * Cout > kNoAllocFlipMax) {
        hp2d::set_global_error("hp2d_nn_conv3x3_dgrad_noalloc: 9*Cin*Cout = " +
                               std::to_string(9LL * Cin * Cout) + " exceeds the " +
                               std::to_string(kNoAllocFlipMax) +
                               " bf16 weights of the device scratch (Cin*Cout <= 131072)");
        return HP2D_E_ARG;
    }
    uint16_t *w_flip = scratch_on_current_device();
    if (!w_flip) return HP2D_E_HIP;
    // the same kernels as hp2d_nn_conv3x3_dgrad, with the workspace in scratch;
    // the flip overwrites every element the forward GEMM then reads
    return hp2d_nn_conv3x3_dgrad(N, H, W, Cin, Cout, gy, w, w_flip, gx, stream);
}

hp2d_status hp2d_nn_conv3x3_wgrad(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                                  const uint16_t *x, const uint16_t *gy, float *gw, float *gb,
                                  void *stream) {
    if (!x || !gy || !gw || Cin % 32 || Cout % 32) return HP2D_E_ARG;
    ConvArgs a{};
    a.N = N; a.H = H; a.W = W; a.Cin = Cin; a.Cout = Cout;
    a.x = x; a.g = gy; a.gw = gw;
    const long long P = (long long)N * H * W;
    // split the pixel reduction so that the grid has a few thousand workgroups
    const int TN = Cin % 64 == 0 ? 64 : 32, TM = Cout % 64 == 0 ? 64 : 32;
    const long long tiles = (long long)(Cout / TM) * 9 * (Cin / TN);
    long long split = (4096 + tiles - 1) / tiles;
    long long ppb = (P + split - 1) / split;
    ppb = ((ppb + kKStep - 1) / kKStep) * kKStep;
    if (ppb < kKStep) ppb = kKStep;
    a.pix_per_block = ppb;
    const unsigned gz = (unsigned)((P + ppb - 1) / ppb);
    const dim3 g((unsigned)(Cout / TM), (unsigned)(9 * (Cin / TN)), gz);
    hipStream_t s = (hipStream_t)stream;
    if (TM == 64 && TN == 64)
        hipLaunchKernelGGL((conv3x3_wgrad_kernel<64, 64>), g, dim3(kBlock), 0, s, a);
    else if (TM == 64)
        hipLaunchKernelGGL((conv3x3_wgrad_kernel<64, 32>), g, dim3(kBlock), 0, s, a);
    else if (TN == 64)
        hipLaunchKernelGGL((conv3x3_wgrad_kernel<32, 64>), g, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((conv3x3_wgrad_kernel<32, 32>), g, dim3(kBlock), 0, s, a);
    if (hipGetLastError() != hipSuccess) return HP2D_E_HIP;
    if (gb) {
        if (Cout / 8 > 256) return HP2D_E_UNSUPPORTED;
        hipLaunchKernelGGL(bias_grad_kernel, dim3(1024), dim3(256), 0, s, gy, P, Cout, gb);
        if (hipGetLastError() != hipSuccess) return HP2D_E_HIP;
    }
    return HP2D_OK;
}

hp2d_status hp2d_nn_relu_mask(uint16_t *g, const uint16_t *y, int64_t n, void *stream) {
    hipLaunchKernelGGL(relu_mask_kernel, grid1((n + 7) / 8), dim3(256), 0, (hipStream_t)stream, g, y,
                       (long long)n);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_add(const uint16_t *a, const uint16_t *b, uint16_t *out, int64_t n,
                        void *stream) {
    hipLaunchKernelGGL(add_kernel, grid1((n + 7) / 8), dim3(256), 0, (hipStream_t)stream, a, b, out,
                       (long long)n);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_avgpool2(const uint16_t *x, int32_t N, int32_t H, int32_t W, int32_t C,
                             uint16_t *y, int32_t backward, void *stream) {
    if (H % 2 || W % 2) return HP2D_E_ARG;
    if (C % 8) return HP2D_E_ARG;
    const long long out_n = (backward ? (long long)N * H * W * C : (long long)N * (H / 2) * (W / 2) * C) / 8;
    if (backward)
        hipLaunchKernelGGL(avgpool2_bwd_kernel, grid1(out_n), dim3(256), 0, (hipStream_t)stream, x, N,
                           H, W, C, y);
    else
        hipLaunchKernelGGL(avgpool2_kernel, grid1(out_n), dim3(256), 0, (hipStream_t)stream, x, N, H,
                           W, C, y);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_upsample2(const uint16_t *x, int32_t N, int32_t H, int32_t W, int32_t C,
                              uint16_t *y, int32_t backward, void *stream) {
    if (H % 2 || W % 2) return HP2D_E_ARG;
    if (C % 8) return HP2D_E_ARG;
    const long long out_n = (backward ? (long long)N * (H / 2) * (W / 2) * C : (long long)N * H * W * C) / 8;
    if (backward)
        hipLaunchKernelGGL(upsample2_bwd_kernel, grid1(out_n), dim3(256), 0, (hipStream_t)stream, x,
                           N, H, W, C, y);
    else
        hipLaunchKernelGGL(upsample2_kernel, grid1(out_n), dim3(256), 0, (hipStream_t)stream, x, N, H,
                           W, C, y);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_head(const uint16_t *x, int64_t P, int32_t C, const float *w, const float *b,
                         int32_t O, float *out, const float *gy, uint16_t *gx, float *gw, float *gb,
                         void *stream) {
    if (!x || !w || (!gy && (!b || !out)) || (gy && (!gx || !gw || !gb))) return HP2D_E_ARG;
    if (C % 8 || O > 4) return HP2D_E_UNSUPPORTED;
    if (!gy) {
        hipLaunchKernelGGL(head_fwd_kernel, grid1(P), dim3(256), 0, (hipStream_t)stream, x,
                           (long long)P, C, w, b, O, out);
    } else {
        if (C != 32 || O != 2) return HP2D_E_UNSUPPORTED;  // the surrogate's head
        const long long blocks = P / 256 < 1024 ? (P + 255) / 256 : 1024;
        hipLaunchKernelGGL((head_bwd_kernel<32, 2>), dim3((unsigned)blocks), dim3(256), 0,
                           (hipStream_t)stream, x, (long long)P, w, gy, gx, gw, gb);
    }
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_mse(const float *out, const float *target, const uint8_t *mask, int64_t P,
                        int32_t O, float inv_count, float *gy, float *loss, void *stream) {
    const long long blocks = std::min<long long>((P + 255) / 256, 1024);
    hipLaunchKernelGGL(mse_kernel, dim3((unsigned)std::max<long long>(blocks, 1)), dim3(256), 0,
                       (hipStream_t)stream, out, target, mask,
                       (long long)P, O, inv_count, gy, loss);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_adam(float *w, float *m, float *v, const float *g, int64_t n, float lr,
                         float beta1, float beta2, float eps, int32_t step, uint16_t *w_bf16,
                         void *stream) {
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, grid1(n), dim3(256), 0, (hipStream_t)stream, w, m, v, g,
                       (long long)n, lr, beta1, beta2, eps, c1, c2, w_bf16);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_to_bf16(const float *x, int64_t n, uint16_t *y, void *stream) {
    hipLaunchKernelGGL(f32_to_bf16_kernel, grid1(n), dim3(256), 0, (hipStream_t)stream, x,
                       (long long)n, y);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

}  // extern "C"
