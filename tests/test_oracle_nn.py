"""Pins of the CPU convolution reference used to check the surrogate's kernels.

The reference is pinned without restating its loops: a centre-only kernel
reduces the convolution to a matrix product, a one-tap kernel to a shift,
and the gradients are checked through the adjoint identities
<gy, conv(dx, W)> = <gx, dx> and <gy, conv(x, dW)> = <gW, dW>, which a
transposed, flipped or mis-offset gradient loop fails.
"""
import numpy as np
import pytest

from oracle import oracle


def test_centre_tap_is_a_matrix_product():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 7, 6)).astype(np.float32)
    w = np.zeros((4, 3, 3, 6), dtype=np.float32)
    w[:, 1, 1, :] = rng.normal(size=(4, 6))
    b = rng.normal(size=4).astype(np.float32)
    y = oracle.conv3x3(x, w, b)
    ref = x.astype(np.float64) @ w[:, 1, 1, :].T.astype(np.float64) + b
    np.testing.assert_allclose(y, ref, rtol=1e-6, atol=1e-6)
    yr = oracle.conv3x3(x, w, b, relu=True)
    np.testing.assert_allclose(yr, np.maximum(ref, 0), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("kh,kw", [(0, 1), (2, 0), (1, 2)])
def test_single_tap_is_a_zero_padded_shift(kh, kw):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 6, 5, 3)).astype(np.float32)
    w = np.zeros((3, 3, 3, 3), dtype=np.float32)
    w[:, kh, kw, :] = np.eye(3)
    y = oracle.conv3x3(x, w)
    ref = np.zeros_like(x)
    dh, dw = kh - 1, kw - 1
    H, W = 6, 5
    for h in range(H):
        for c in range(W):
            if 0 <= h + dh < H and 0 <= c + dw < W:
                ref[0, h, c] = x[0, h + dh, c + dw]
    np.testing.assert_array_equal(y, ref)


def test_gradients_satisfy_the_adjoint_identities():
    rng = np.random.default_rng(2)
    N, H, W, Cin, Cout = 2, 6, 7, 5, 4
    x = rng.normal(size=(N, H, W, Cin)).astype(np.float32)
    w = rng.normal(size=(Cout, 3, 3, Cin)).astype(np.float32)
    gy = rng.normal(size=(N, H, W, Cout)).astype(np.float32)
    gx, gw, gb = oracle.conv3x3_grad(x, w, gy)
    dx = rng.normal(size=x.shape).astype(np.float32)
    dw = rng.normal(size=w.shape).astype(np.float32)
    lhs_x = np.sum(gy.astype(np.float64) * oracle.conv3x3(dx, w))
    np.testing.assert_allclose(lhs_x, np.sum(gx.astype(np.float64) * dx), rtol=1e-5)
    lhs_w = np.sum(gy.astype(np.float64) * oracle.conv3x3(x, dw))
    np.testing.assert_allclose(lhs_w, np.sum(gw.astype(np.float64) * dw), rtol=1e-5)
    np.testing.assert_allclose(gb, gy.sum(axis=(0, 1, 2)), rtol=1e-5)
