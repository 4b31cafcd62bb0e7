"""CPU reference of the surrogate network built on the oracle's convolution,
with bf16 rounding at the points where the GPU stores bf16 (test infrastructure).

Layer order and shapes follow hydropol2d_amd/surrogate.py; the convolution
itself is oracle.conv3x3 / oracle.conv3x3_grad (FP64 accumulation), so a
mismatch in the GPU's im2col, weight flip, MFMA lane layout or gradient
routing shows up as a difference here.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle

ENC = [(32, 32), (32, 64), (64, 128), (128, 256)]
DEC = [(256, 128), (128, 64), (64, 32)]


def bf16(x):
    """Round to bf16 (nearest even) and return as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


# where the GPU stores bf16; a test of this module swaps in the identity
ROUND = bf16


def w4(wmat, ci):
    """[Cout][9*Cin] -> [Cout][3][3][Cin]."""
    return wmat.reshape(wmat.shape[0], 3, 3, ci)


def pool(x):
    return ROUND(0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2]))


def pool_bwd(g):
    return ROUND(0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2))


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(g):
    return ROUND(g[:, 0::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 0::2] + g[:, 1::2, 1::2])


def conv(x, wmat, b, relu=True, mask=None):
    """ReLU(conv + b), or (conv + b) * mask with a frozen activation pattern."""
    ci = x.shape[-1]
    if mask is None:
        return ROUND(oracle.conv3x3(x, w4(wmat, ci), b, relu=relu))
    return ROUND(oracle.conv3x3(x, w4(wmat, ci), b, relu=False) * mask)


def forward(params, head_w, head_b, x, masks=None):
    """params: list of 7 (w [Cout][9Cin] bf16-valued, b) pairs. Returns (out, cache).
    With `masks` (the "mask" entries of a previous cache) the ReLU pattern is
    frozen, which makes the network linear in any one layer's parameters."""
    c = {"x": x}
    m = masks or {}
    cur = x
    for lv in range(4):
        c[f"e{lv}"] = conv(cur, *params[lv], mask=m.get(f"e{lv}"))
        c.setdefault("mask", {})[f"e{lv}"] = (c[f"e{lv}"] > 0).astype(np.float32)
        if lv < 3:
            c[f"p{lv}"] = pool(c[f"e{lv}"])
            cur = c[f"p{lv}"]
    cur = c["e3"]
    for k in range(3):
        lv = 2 - k
        c[f"u{lv}"] = up(cur)
        c[f"d{lv}r"] = conv(c[f"u{lv}"], *params[4 + k], mask=m.get(f"d{lv}r"))
        c["mask"][f"d{lv}r"] = (c[f"d{lv}r"] > 0).astype(np.float32)
        c[f"d{lv}"] = ROUND(c[f"d{lv}r"] + c[f"e{lv}"])
        cur = c[f"d{lv}"]
    out = c["d0"].astype(np.float64) @ head_w.T.astype(np.float64) + head_b
    return out.astype(np.float32), c


def backward(params, head_w, c, g_out):
    """Parameter gradients (list of (gw [Cout][9Cin], gb)) and the head's (gw, gb)."""
    grads = [None] * 7
    P = g_out.shape[0] * g_out.shape[1] * g_out.shape[2]
    go = g_out.reshape(P, -1).astype(np.float64)
    d0 = c["d0"].reshape(P, -1).astype(np.float64)
    head_gw = (go.T @ d0).astype(np.float32)
    head_gb = go.sum(axis=0).astype(np.float32)
    g_d = ROUND((go @ head_w.astype(np.float64)).reshape(c["d0"].shape))
    skip = {}
    for lv in range(3):
        li = 6 - lv
        skip[lv] = g_d
        g_c = np.where(c[f"d{lv}r"] > 0, g_d, 0).astype(np.float32)
        ci = c[f"u{lv}"].shape[-1]
        gx, gw, gb = oracle.conv3x3_grad(c[f"u{lv}"], w4(params[li][0], ci), g_c)
        grads[li] = (gw.reshape(gw.shape[0], -1), gb)
        g_d = up_bwd(ROUND(gx))
    g_e = g_d
    for lv in range(3, -1, -1):
        if lv < 3:
            g_e = ROUND(g_e + skip[lv])
        g_c = np.where(c[f"e{lv}"] > 0, g_e, 0).astype(np.float32)
        x_in = c["x"] if lv == 0 else c[f"p{lv - 1}"]
        ci = x_in.shape[-1]
        gx, gw, gb = oracle.conv3x3_grad(x_in, w4(params[lv][0], ci), g_c)
        grads[lv] = (gw.reshape(gw.shape[0], -1), gb)
        if lv > 0:
            g_e = pool_bwd(ROUND(gx))
    return grads, (head_gw, head_gb)
