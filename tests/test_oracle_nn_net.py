"""The CPU reference network used to check the surrogate's GPU gradients,
pinned by central finite differences of its own forward pass (rounding off).

With the ReLU pattern frozen at the base point the network is linear in any
one layer's parameters, so the directional differences are exact up to FP32
rounding and pin the gradient routing (skips, pooling, upsampling, the flipped
data-gradient weights) without the noise of crossing activation kinks."""
import numpy as np

import nn_cases


def test_reference_network_gradients_by_finite_differences(monkeypatch):
    monkeypatch.setattr(nn_cases, "ROUND", lambda a: np.asarray(a, dtype=np.float32))
    rng = np.random.default_rng(0)
    B, H, W = 1, 16, 16
    params = []
    for ci, co in nn_cases.ENC + nn_cases.DEC:
        params.append((rng.normal(0, np.sqrt(2.0 / (9 * ci)), (co, 9 * ci)).astype(np.float32),
                       rng.normal(0, 0.1, co).astype(np.float32)))
    hw = rng.normal(0, 0.2, (2, 32)).astype(np.float32)
    hb = np.zeros(2, dtype=np.float32)
    x = np.zeros((B, H, W, 32), dtype=np.float32)
    x[..., :6] = rng.normal(size=(B, H, W, 6))
    out, cache = nn_cases.forward(params, hw, hb, x)
    g_out = rng.normal(size=out.shape).astype(np.float32)
    grads, (hgw, hgb) = nn_cases.backward(params, hw, cache, g_out)

    masks = cache["mask"]

    def loss(ps, hw_):
        o, _ = nn_cases.forward(ps, hw_, hb, x, masks=masks)
        return float(np.sum(o.astype(np.float64) * g_out))

    # directional derivatives along a random full-tensor direction per layer
    h = 1e-2
    for li in range(7):
        D = rng.normal(size=params[li][0].shape).astype(np.float32)
        if li == 0:
            D.reshape(D.shape[0], 9, 32)[:, :, 6:] = 0   # padded input channels
        vals = []
        for sgn in (1, -1):
            ps = [(w.copy(), b) for w, b in params]
            ps[li] = (ps[li][0] + sgn * h * D, ps[li][1])
            vals.append(loss(ps, hw))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(np.sum(grads[li][0].astype(np.float64) * D))
        assert abs(fd - an) <= 1e-3 * max(abs(fd), abs(an)), (li, fd, an)
        Db = rng.normal(size=params[li][1].shape).astype(np.float32)
        vals = []
        for sgn in (1, -1):
            ps = [(w, b.copy()) for w, b in params]
            ps[li] = (ps[li][0], ps[li][1] + sgn * h * Db)
            vals.append(loss(ps, hw))
        fd = (vals[0] - vals[1]) / (2 * h)
        an = float(np.sum(grads[li][1].astype(np.float64) * Db))
        assert abs(fd - an) <= 1e-3 * max(abs(fd), abs(an)), (li, "bias", fd, an)
    # head weight
    for _ in range(2):
        idx = (rng.integers(2), rng.integers(32))
        vals = []
        for sgn in (1, -1):
            h2 = hw.copy()
            h2[idx] += sgn * h
            vals.append(loss(params, h2))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - hgw[idx]) <= 1e-3 * max(abs(fd), abs(hgw[idx])), (idx, fd, hgw[idx])
