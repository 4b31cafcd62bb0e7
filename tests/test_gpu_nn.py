"""GPU checks of the surrogate's BF16 MFMA kernels against the CPU oracle.

Tolerances: outputs stored in bf16 may differ from the FP64-accumulated
reference by one bf16 rounding (2^-8 relative) plus FP32 accumulation order;
FP32 weight gradients (atomics in any order) to 1e-3 relative; whole-network
parameter gradients, which pass through a dozen bf16 roundings, to 3e-2 in
relative Frobenius norm per tensor.
"""
import numpy as np
import pytest

import nn_cases
from oracle import oracle

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _dev_bf16(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda().to(torch.bfloat16)


def _dev_f32(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _np(t):
    return t.float().cpu().numpy()


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _lib():
    from hydropol2d_amd import _lib as L
    return L.load()


def _close_bf16(got, ref, name=""):
    scale = np.max(np.abs(ref)) + 1e-30
    err = np.abs(got - ref)
    bound = 2.0 ** -7 * np.abs(ref) + 1e-3 * scale
    assert (err <= bound).all(), (name, float(np.max(err / (bound + 1e-30))))


@pytest.mark.parametrize("N,H,W,Cin,Cout", [
    (2, 16, 16, 32, 32), (1, 24, 40, 64, 32), (2, 8, 8, 32, 128), (1, 13, 7, 96, 64),
])
def test_conv_forward_matches_oracle(N, H, W, Cin, Cout):
    torch = _torch()
    rng = np.random.default_rng(N * 100 + Cin)
    x = nn_cases.bf16(rng.normal(size=(N, H, W, Cin)))
    w = nn_cases.bf16(rng.normal(0, 0.1, size=(Cout, 9 * Cin)))
    b = rng.normal(size=Cout).astype(np.float32)
    skip = nn_cases.bf16(rng.normal(size=(N, H, W, Cout)))
    dx, dw, db, ds = _dev_bf16(x), _dev_bf16(w), _dev_f32(b), _dev_bf16(skip)
    y = torch.empty((N, H, W, Cout), dtype=torch.bfloat16, device="cuda")
    ys = torch.empty_like(y)
    st = _lib().hp2d_nn_conv3x3_fwd(N, H, W, Cin, Cout, _ptr(dx), _ptr(dw), _ptr(db), 1, _ptr(ds),
                                    _ptr(y), _ptr(ys), None)
    assert st == 0
    torch.cuda.synchronize()
    ref = oracle.conv3x3(x, nn_cases.w4(w, Cin), b, relu=True)
    _close_bf16(_np(y), ref, "y")
    _close_bf16(_np(ys), _np(y) + skip, "y+skip")


@pytest.mark.parametrize("N,H,W,Cin,Cout", [(2, 16, 16, 32, 64), (1, 12, 20, 64, 32)])
def test_conv_gradients_match_oracle(N, H, W, Cin, Cout):
    torch = _torch()
    rng = np.random.default_rng(7)
    x = nn_cases.bf16(rng.normal(size=(N, H, W, Cin)))
    w = nn_cases.bf16(rng.normal(0, 0.1, size=(Cout, 9 * Cin)))
    g = nn_cases.bf16(rng.normal(size=(N, H, W, Cout)))
    dx, dw, dg = _dev_bf16(x), _dev_bf16(w), _dev_bf16(g)
    gx = torch.empty((N, H, W, Cin), dtype=torch.bfloat16, device="cuda")
    wflip = torch.empty((Cin, 9 * Cout), dtype=torch.bfloat16, device="cuda")
    gw = torch.zeros((Cout, 9 * Cin), dtype=torch.float32, device="cuda")
    gb = torch.zeros(Cout, dtype=torch.float32, device="cuda")
    L = _lib()
    assert L.hp2d_nn_conv3x3_dgrad(N, H, W, Cin, Cout, _ptr(dg), _ptr(dw), _ptr(wflip), _ptr(gx),
                                   None) == 0
    assert L.hp2d_nn_conv3x3_wgrad(N, H, W, Cin, Cout, _ptr(dx), _ptr(dg), _ptr(gw), _ptr(gb),
                                   None) == 0
    torch.cuda.synchronize()
    rgx, rgw, rgb = oracle.conv3x3_grad(x, nn_cases.w4(w, Cin), g)
    _close_bf16(_np(gx), rgx, "gx")
    np.testing.assert_allclose(_np(gw), rgw.reshape(Cout, -1), rtol=1e-3,
                               atol=1e-4 * np.max(np.abs(rgw)))
    np.testing.assert_allclose(_np(gb), rgb, rtol=1e-3, atol=1e-4 * np.max(np.abs(rgb)))


def _small_net(seed=0, H=32, W=32, B=2):
    from hydropol2d_amd.surrogate import UNetSurrogate
    net = UNetSurrogate(H, W, batch=B, seed=seed)
    rng = np.random.default_rng(seed + 1)
    x = np.zeros((B, H, W, 32), dtype=np.float32)
    x[..., :6] = rng.normal(size=(B, H, W, 6))
    x = nn_cases.bf16(x)
    return net, x, rng


def test_network_forward_and_gradients_match_cpu_reference():
    torch = _torch()
    net, x, rng = _small_net()
    B, H, W = x.shape[:3]
    params = [(nn_cases.bf16(_np(p["w"])), _np(p["b"])) for p in net.convs]
    hw, hb = _np(net.head["w"]), _np(net.head["b"])
    out = _np(net.forward(_dev_bf16(x)))
    ref_out, cache = nn_cases.forward(params, hw, hb, x)
    err = np.linalg.norm(out - ref_out) / np.linalg.norm(ref_out)
    assert err < 1e-2, err
    g_out = rng.normal(size=out.shape).astype(np.float32) / out.size
    net.zero_grad()
    net.backward(_dev_f32(g_out), B)
    torch.cuda.synchronize()
    grads, (hgw, hgb) = nn_cases.backward(params, hw, cache, g_out)
    rel = lambda a, b: np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
    assert rel(_np(net.head["gw"]), hgw) < 3e-2
    assert rel(_np(net.head["gb"]), hgb) < 3e-2
    # the reference rounds to bf16 where the GPU stores bf16, but an FP32 MFMA
    # sum and an FP64 sum round the same value to different bf16 neighbours now
    # and then (2^-8 each); the flips compound along the backward chain, to a
    # few percent at the first encoder layers, which sit under all seven others
    for li, (gw, gb) in enumerate(grads):
        assert rel(_np(net.convs[li]["gw"]), gw) < 5e-2, (li, rel(_np(net.convs[li]["gw"]), gw))
        assert rel(_np(net.convs[li]["gb"]), gb) < 5e-2, (li, rel(_np(net.convs[li]["gb"]), gb))


def test_network_256_forward_and_gradients_match_cpu_reference():
    """The surrogate network at its benchmark size (256 x 256 input stacks,
    32 -> 256 channels over four levels): one forward and one backward pass of
    one sample against the CPU reference network (FP64 convolutions with bf16
    rounding where the GPU stores bf16)."""
    torch = _torch()
    print("network 256 x 256 forward + backward against the CPU reference", flush=True)
    net, x, rng = _small_net(seed=3, H=256, W=256, B=1)
    params = [(nn_cases.bf16(_np(p["w"])), _np(p["b"])) for p in net.convs]
    hw, hb = _np(net.head["w"]), _np(net.head["b"])
    out = _np(net.forward(_dev_bf16(x)))
    ref_out, cache = nn_cases.forward(params, hw, hb, x)
    assert np.linalg.norm(out - ref_out) / np.linalg.norm(ref_out) < 1e-2
    g_out = rng.normal(size=out.shape).astype(np.float32) / out.size
    net.zero_grad()
    net.backward(_dev_f32(g_out), 1)
    torch.cuda.synchronize()
    grads, (hgw, hgb) = nn_cases.backward(params, hw, cache, g_out)
    rel = lambda a, b: np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
    assert rel(_np(net.head["gw"]), hgw) < 3e-2
    assert rel(_np(net.head["gb"]), hgb) < 3e-2
    for li, (gw, gb) in enumerate(grads):
        assert rel(_np(net.convs[li]["gw"]), gw) < 5e-2, (li, rel(_np(net.convs[li]["gw"]), gw))
        assert rel(_np(net.convs[li]["gb"]), gb) < 5e-2, (li, rel(_np(net.convs[li]["gb"]), gb))


def test_training_fits_a_smooth_map():
    # a learnable target (a smoothed, nonlinear function of the input maps):
    # 150 Adam steps cut the loss by 10x
    torch = _torch()
    net, x, rng = _small_net(seed=3, H=32, W=32, B=8)
    xs = np.zeros((8, 32, 32, 32), dtype=np.float32)
    xs[..., :6] = rng.normal(size=(8, 32, 32, 6))
    xs = nn_cases.bf16(xs)
    t = np.stack([np.tanh(xs[..., 0] + 0.5 * xs[..., 1]), 0.3 * xs[..., 2] ** 2], axis=-1)
    dx, dt = _dev_bf16(xs), _dev_f32(t)
    losses = [float(net.train_step(dx, dt, lr=3e-3).item()) for _ in range(150)]
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
    assert np.isfinite(losses).all()


def test_elementwise_kernels_cover_more_than_2_32_elements():
    # a dispatch counts its work-items in 32 bits: the element-wise kernels run
    # a capped grid with grid-stride loops, so the last elements of a tensor of
    # more than 2^32 elements (a 4352-member stack of 256 x 256 x 32 inputs is
    # 9.1e9) are converted too (17 GiB + 8.6 GiB of device memory)
    torch = _torch()
    n = (1 << 32) + 4096
    x = torch.full((n,), 1.5, dtype=torch.float32, device="cuda")
    x[-4096:] = 2.5
    x[(1 << 31):(1 << 31) + 8] = -3.0
    y = torch.zeros((n,), dtype=torch.bfloat16, device="cuda")
    assert _lib().hp2d_nn_to_bf16(_ptr(x), n, _ptr(y), None) == 0
    torch.cuda.synchronize()
    del x
    assert bool((y[-4096:] == 2.5).all())
    assert bool((y[(1 << 31):(1 << 31) + 8] == -3.0).all())
    assert bool((y[:4096] == 1.5).all()) and float(y[(1 << 32) - 1]) == 1.5
    del y
    torch.cuda.empty_cache()
