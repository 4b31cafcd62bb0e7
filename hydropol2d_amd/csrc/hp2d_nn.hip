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

// bias gradient: gb[c] += sum over pixels of g[p][c]
__global__ void bias_grad_kernel(const uint16_t *g, long long P, int C, float *gb) {
    const int c = threadIdx.x % C;
    const int lanes_per_c = blockDim.x / C;
    const int sub = threadIdx.x / C;
    float s = 0.f;
    for (long long p = blockIdx.x * (long long)lanes_per_c + sub; p < P;
         p += (long long)gridDim.x * lanes_per_c)
        s += bf2f(g[p * C + c]);
    atomicAdd(&gb[c], s);
}

// g *= (y > 0): the ReLU backward on the stored activation
__global__ void relu_mask_kernel(uint16_t *g, const uint16_t *y, long long n) {
    for (long long i = gs_first(); i < n; i += gs_stride())
        if (!(bf2f(y[i]) > 0.f)) g[i] = 0;
}

// out = a + b (bf16, for gradients meeting at a skip connection)
__global__ void add_kernel(const uint16_t *a, const uint16_t *b, uint16_t *out, long long n) {
    for (long long i = gs_first(); i < n; i += gs_stride()) out[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// 2x2 average pooling (H, W even) and its backward
__global__ void avgpool2_kernel(const uint16_t *x, int N, int H, int W, int C, uint16_t *y) {
    const int Ho = H / 2, Wo = W / 2;
    const long long total = (long long)N * Ho * Wo * C;
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        const int c = (int)(e % C);
        long long r = e / C;
        const int wo = (int)(r % Wo);
        r /= Wo;
        const int ho = (int)(r % Ho);
        const int n = (int)(r / Ho);
        auto at = [&](int h, int w) { return bf2f(x[(((long long)n * H + h) * W + w) * C + c]); };
        const float s = at(2 * ho, 2 * wo) + at(2 * ho, 2 * wo + 1) + at(2 * ho + 1, 2 * wo) +
                        at(2 * ho + 1, 2 * wo + 1);
        y[e] = f2bf(0.25f * s);
    }
}
__global__ void avgpool2_bwd_kernel(const uint16_t *gy, int N, int H, int W, int C, uint16_t *gx) {
    const long long total = (long long)N * H * W * C;
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        const int c = (int)(e % C);
        long long r = e / C;
        const int w = (int)(r % W);
        r /= W;
        const int h = (int)(r % H);
        const int n = (int)(r / H);
        gx[e] = f2bf(0.25f * bf2f(gy[(((long long)n * (H / 2) + h / 2) * (W / 2) + w / 2) * C + c]));
    }
}

// nearest 2x upsampling (output H, W) and its backward (sum over the 2x2 block)
__global__ void upsample2_kernel(const uint16_t *x, int N, int H, int W, int C, uint16_t *y) {
    const long long total = (long long)N * H * W * C;
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        const int c = (int)(e % C);
        long long r = e / C;
        const int w = (int)(r % W);
        r /= W;
        const int h = (int)(r % H);
        const int n = (int)(r / H);
        y[e] = x[(((long long)n * (H / 2) + h / 2) * (W / 2) + w / 2) * C + c];
    }
}
__global__ void upsample2_bwd_kernel(const uint16_t *gy, int N, int H, int W, int C, uint16_t *gx) {
    const int Ho = H / 2, Wo = W / 2;
    const long long total = (long long)N * Ho * Wo * C;
    for (long long e = gs_first(); e < total; e += gs_stride()) {
        const int c = (int)(e % C);
        long long r = e / C;
        const int wo = (int)(r % Wo);
        r /= Wo;
        const int ho = (int)(r % Ho);
        const int n = (int)(r / Ho);
        auto at = [&](int h, int w) { return bf2f(gy[(((long long)n * H + h) * W + w) * C + c]); };
        gx[e] = f2bf(at(2 * ho, 2 * wo) + at(2 * ho, 2 * wo + 1) + at(2 * ho + 1, 2 * wo) +
                     at(2 * ho + 1, 2 * wo + 1));
    }
}

// 1x1 head: out[p][o] = b[o] + sum_c x[p][c] w[o][c] (FP32 out, O outputs)
__global__ void head_fwd_kernel(const uint16_t *x, long long P, int C, const float *w,
                                const float *b, int O, float *out) {
    for (long long p = gs_first(); p < P; p += gs_stride()) {
        for (int o = 0; o < O; ++o) {
            float s = b[o];
            for (int c = 0; c < C; ++c) s += bf2f(x[p * C + c]) * w[o * C + c];
            out[p * O + o] = s;
        }
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

// masked mean squared error over the O output maps; gy = d loss / d out
__global__ void mse_kernel(const float *out, const float *target, const uint8_t *mask, long long P,
                           int O, float inv_count, float *gy, float *loss) {
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
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l * inv_count);
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
        const int per = 256 / Cout > 0 ? 256 / Cout : 1;
        hipLaunchKernelGGL(bias_grad_kernel, dim3(1024), dim3(per * Cout), 0, s, gy, P, Cout, gb);
        if (hipGetLastError() != hipSuccess) return HP2D_E_HIP;
    }
    return HP2D_OK;
}

hp2d_status hp2d_nn_relu_mask(uint16_t *g, const uint16_t *y, int64_t n, void *stream) {
    hipLaunchKernelGGL(relu_mask_kernel, grid1(n), dim3(256), 0, (hipStream_t)stream, g, y,
                       (long long)n);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_add(const uint16_t *a, const uint16_t *b, uint16_t *out, int64_t n,
                        void *stream) {
    hipLaunchKernelGGL(add_kernel, grid1(n), dim3(256), 0, (hipStream_t)stream, a, b, out,
                       (long long)n);
    return hipGetLastError() == hipSuccess ? HP2D_OK : HP2D_E_HIP;
}

hp2d_status hp2d_nn_avgpool2(const uint16_t *x, int32_t N, int32_t H, int32_t W, int32_t C,
                             uint16_t *y, int32_t backward, void *stream) {
    if (H % 2 || W % 2) return HP2D_E_ARG;
    const long long out_n = backward ? (long long)N * H * W * C : (long long)N * (H / 2) * (W / 2) * C;
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
    const long long out_n = backward ? (long long)N * (H / 2) * (W / 2) * C : (long long)N * H * W * C;
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
    hipLaunchKernelGGL(mse_kernel, grid1(P), dim3(256), 0, (hipStream_t)stream, out, target, mask,
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
