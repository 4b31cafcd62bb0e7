/*
 * hp2d_nn.c - FP64-accumulating CPU reference of the surrogate's convolution
 * layers (TEST INFRASTRUCTURE ONLY).
 *
 * The surrogate has no counterpart in the reference; these loops are the
 * textbook definitions of a 3x3, stride-1, zero-padded 2-D convolution and of
 * its gradients, on NHWC activations and [Cout][3][3][Cin] weights:
 *   y[n,h,w,o] = b[o] + sum_{kh,kw,i} x[n,h+kh-1,w+kw-1,i] * W[o,kh,kw,i]
 *   dx[n,h,w,i] = sum_{kh,kw,o} gy[n,h-kh+1,w-kw+1,o] * W[o,kh,kw,i]
 *   dW[o,kh,kw,i] = sum_{n,h,w} gy[n,h,w,o] * x[n,h+kh-1,w+kw-1,i]
 *   db[o] = sum_{n,h,w} gy[n,h,w,o]
 * They are pinned in tests/test_oracle_nn.py by central finite differences of
 * the forward loop and by the 1x1-kernel case that reduces to a matrix product.
 */
#include <stdint.h>
#include <stdlib.h>

#define XI(n, h, w, c, H, W, C) ((((int64_t)(n) * (H) + (h)) * (W) + (w)) * (C) + (c))
#define WI(o, kh, kw, i, Cin) ((((int64_t)(o) * 3 + (kh)) * 3 + (kw)) * (Cin) + (i))

void oracle_conv3x3_fwd(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                        const float *x, const float *w, const float *b, int32_t relu, float *y) {
    for (int n = 0; n < N; ++n)
        for (int h = 0; h < H; ++h)
            for (int ww = 0; ww < W; ++ww)
                for (int o = 0; o < Cout; ++o) {
                    double acc = b ? b[o] : 0.0;
                    for (int kh = 0; kh < 3; ++kh) {
                        const int hh = h + kh - 1;
                        if (hh < 0 || hh >= H) continue;
                        for (int kw = 0; kw < 3; ++kw) {
                            const int wc = ww + kw - 1;
                            if (wc < 0 || wc >= W) continue;
                            const float *xp = x + XI(n, hh, wc, 0, H, W, Cin);
                            const float *wp = w + WI(o, kh, kw, 0, Cin);
                            for (int i = 0; i < Cin; ++i) acc += (double)xp[i] * (double)wp[i];
                        }
                    }
                    if (relu && acc < 0) acc = 0;
                    y[XI(n, h, ww, o, H, W, Cout)] = (float)acc;
                }
}

/* gy is the gradient with respect to the convolution output before any
 * activation; any of gx, gw, gb may be NULL */
void oracle_conv3x3_bwd(int32_t N, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
                        const float *x, const float *w, const float *gy, float *gx, float *gw,
                        float *gb) {
    if (gx) {
        for (int n = 0; n < N; ++n)
            for (int h = 0; h < H; ++h)
                for (int ww = 0; ww < W; ++ww)
                    for (int i = 0; i < Cin; ++i) {
                        double acc = 0.0;
                        for (int kh = 0; kh < 3; ++kh) {
                            const int ho = h - kh + 1;
                            if (ho < 0 || ho >= H) continue;
                            for (int kw = 0; kw < 3; ++kw) {
                                const int wo = ww - kw + 1;
                                if (wo < 0 || wo >= W) continue;
                                for (int o = 0; o < Cout; ++o)
                                    acc += (double)gy[XI(n, ho, wo, o, H, W, Cout)] *
                                           (double)w[WI(o, kh, kw, i, Cin)];
                            }
                        }
                        gx[XI(n, h, ww, i, H, W, Cin)] = (float)acc;
                    }
    }
    if (gw) {
        for (int o = 0; o < Cout; ++o)
            for (int kh = 0; kh < 3; ++kh)
                for (int kw = 0; kw < 3; ++kw)
                    for (int i = 0; i < Cin; ++i) {
                        double acc = 0.0;
                        for (int n = 0; n < N; ++n)
                            for (int h = 0; h < H; ++h) {
                                const int hh = h + kh - 1;
                                if (hh < 0 || hh >= H) continue;
                                for (int ww = 0; ww < W; ++ww) {
                                    const int wc = ww + kw - 1;
                                    if (wc < 0 || wc >= W) continue;
                                    acc += (double)gy[XI(n, h, ww, o, H, W, Cout)] *
                                           (double)x[XI(n, hh, wc, i, H, W, Cin)];
                                }
                            }
                        gw[WI(o, kh, kw, i, Cin)] = (float)acc;
                    }
    }
    if (gb) {
        for (int o = 0; o < Cout; ++o) {
            double acc = 0.0;
            for (int64_t p = 0; p < (int64_t)N * H * W; ++p) acc += (double)gy[p * Cout + o];
            gb[o] = (float)acc;
        }
    }
}
