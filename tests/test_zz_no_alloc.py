"""The allocation-free data gradient (hp2d_nn_conv3x3_dgrad_noalloc) against
hp2d_nn_conv3x3_dgrad with a caller-allocated workspace, through the C ABI:
bit-identical results on two calls in a row, the second smaller than the
first (the per-GPU scratch keeps the first call's contents), and a layer
beyond the scratch's maximum refused with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _case(torch, rng, N, H, W, Cin, Cout):
    gy = torch.from_numpy(rng.normal(size=(N, H, W, Cout)).astype(np.float32)).cuda().to(
        torch.bfloat16)
    w = torch.from_numpy(rng.normal(size=(Cout, 9 * Cin)).astype(np.float32) * 0.05).cuda().to(
        torch.bfloat16)
    return gy, w


@pytest.mark.parametrize("shapes", [
    # (N, H, W, Cin, Cout) of two calls in a row, the second smaller; the first
    # is the largest layer the scratch holds (9 * 256 * 512 weights)
    ((1, 16, 24, 256, 512), (2, 8, 8, 32, 64)),
    ((2, 32, 16, 64, 128), (1, 8, 16, 32, 32)),
])
def test_noalloc_dgrad_matches_the_workspace_entry_point(shapes):
    import torch
    from hydropol2d_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(5)
    for N, H, W, Cin, Cout in shapes:
        print(f"dgrad noalloc N={N} H={H} W={W} Cin={Cin} Cout={Cout}", flush=True)
        gy, w = _case(torch, rng, N, H, W, Cin, Cout)
        w_flip = torch.empty(9 * Cin * Cout, dtype=torch.bfloat16, device="cuda")
        gx_ref = torch.full((N, H, W, Cin), float("nan"), dtype=torch.bfloat16, device="cuda")
        gx = torch.full((N, H, W, Cin), float("nan"), dtype=torch.bfloat16, device="cuda")
        assert L.hp2d_nn_conv3x3_dgrad(N, H, W, Cin, Cout, _ptr(gy), _ptr(w), _ptr(w_flip),
                                       _ptr(gx_ref), None) == 0
        assert L.hp2d_nn_conv3x3_dgrad_noalloc(N, H, W, Cin, Cout, _ptr(gy), _ptr(w), _ptr(gx),
                                               None) == 0
        torch.cuda.synchronize()
        assert torch.isfinite(gx_ref.float()).all()
        assert torch.equal(gx.view(torch.int16), gx_ref.view(torch.int16))


def test_noalloc_dgrad_refuses_a_layer_beyond_the_scratch():
    import torch
    from hydropol2d_amd import _lib
    L = _lib.load()
    t = torch.zeros(32, dtype=torch.bfloat16, device="cuda")
    # 9 * 512 * 512 weights exceed the 9 * 512 * 256 of the scratch
    rc = L.hp2d_nn_conv3x3_dgrad_noalloc(1, 1, 1, 512, 512, _ptr(t), _ptr(t), _ptr(t), None)
    assert rc != 0
    assert b"exceeds" in L.hp2d_global_last_error()
