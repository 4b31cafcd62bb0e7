"""Flood surrogate: a 4-level convolutional encoder-decoder on the gfx950 BF16
MFMA kernels of libhp2d (hp2d_nn_*), trained on the solver's own peak-depth
and time-to-peak maps and judged by the solver.

The network has no counterpart in the reference.  Inputs per member are six
maps [elevation, Manning n, Green-Ampt Ks, initial soil moisture, total
rainfall depth, peak rainfall intensity] and two derived from them: the D8
upslope contributing area (log-scaled) and the kinematic normal depth of the
peak excess rainfall over that area, (n q / sqrt(S))^(3/5) with
q = (i_peak - Ks) A_up / dx and S the steepest D8 slope; normalised and
zero-padded to 32 channels.  Outputs are [peak depth, time to peak].  The
contributing area is a non-local function of the DEM that a 4-level network
of 3x3 convolutions sees only over its receptive field; handing it over as a
map, with the steady-state depth it implies, lets the network place and
size the channels of a 256 x 256 member and learn the corrections (finite
rain duration, storage, infiltration) on top.

    E0 conv 32->32  @H     relu                    (skip to D0)
    E1 conv 32->64  @H/2   relu  after avg-pool    (skip to D1)
    E2 conv 64->128 @H/4   relu  after avg-pool    (skip to D2)
    E3 conv 128->256 @H/8  relu  after avg-pool
    D2 conv 256->128 @H/4  relu  after 2x upsample, + E2
    D1 conv 128->64  @H/2  relu  after 2x upsample, + E1
    D0 conv 64->32   @H    relu  after 2x upsample, + E0
    head 1x1 32->2 (FP32)

Activations and 3x3 weights are bf16, accumulation and the optimiser state
FP32 (Adam).  PyTorch only allocates device memory and supplies the stream;
every layer, gradient and update runs in libhp2d.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

ENC = [(32, 32), (32, 64), (64, 128), (128, 256)]
DEC = [(256, 128), (128, 64), (64, 32)]
N_IN, N_PAD, N_OUT = 8, 32, 2


def _torch():
    import torch
    return torch


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class UNetSurrogate:
    def __init__(self, height: int = 256, width: int = 256, batch: int = 64, seed: int = 0):
        torch = _torch()
        if height % 8 or width % 8:
            raise ValueError("height and width must be multiples of 8")
        self.L = _lib.load()
        _lib.require_device()
        self.H, self.W, self.B = height, width, batch
        self.dev = torch.device("cuda")
        rng = np.random.default_rng(seed)
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.convs = []
        fan_in_real = [N_IN] + [c for c, _ in ENC[1:]] + [c for c, _ in DEC]
        for li, (ci, co) in enumerate(ENC + DEC):
            std = np.sqrt(2.0 / (9 * fan_in_real[li]))
            w = rng.normal(0.0, std, (co, 9 * ci)).astype(np.float32)
            p = {"ci": ci, "co": co,
                 "w": torch.from_numpy(w).to(self.dev), "b": torch.zeros(co, **f32)}
            for k in ("w", "b"):
                p["m" + k] = torch.zeros_like(p[k])
                p["v" + k] = torch.zeros_like(p[k])
                p["g" + k] = torch.zeros_like(p[k])
            p["wb"] = torch.empty(co, 9 * ci, dtype=torch.bfloat16, device=self.dev)
            p["wflip"] = torch.empty(ci, 9 * co, dtype=torch.bfloat16, device=self.dev)
            self._chk(self.L.hp2d_nn_to_bf16(_p(p["w"]), p["w"].numel(), _p(p["wb"]), self._s()))
            self.convs.append(p)
        hw = rng.normal(0.0, np.sqrt(1.0 / 32), (N_OUT, 32)).astype(np.float32)
        self.head = {"w": torch.from_numpy(hw).to(self.dev), "b": torch.zeros(N_OUT, **f32)}
        for k in ("w", "b"):
            self.head["m" + k] = torch.zeros_like(self.head[k])
            self.head["v" + k] = torch.zeros_like(self.head[k])
            self.head["g" + k] = torch.zeros_like(self.head[k])
        self.step_count = 0
        self._alloc(batch)

    # ------------------------------------------------------------------ memory
    def _alloc(self, batch):
        torch = _torch()
        bf = dict(dtype=torch.bfloat16, device=self.dev)
        H, W = self.H, self.W
        shape = lambda lv, c: (batch, H >> lv, W >> lv, c)
        A = {}
        for lv, (_, co) in enumerate(ENC):
            A[f"e{lv}"] = torch.empty(shape(lv, co), **bf)
            if lv < 3:
                A[f"p{lv}"] = torch.empty(shape(lv + 1, co), **bf)
        for k, (ci, co) in enumerate(DEC):
            lv = 2 - k
            A[f"u{lv}"] = torch.empty(shape(lv, ci), **bf)
            A[f"d{lv}r"] = torch.empty(shape(lv, co), **bf)
            A[f"d{lv}"] = torch.empty(shape(lv, co), **bf)
        A["out"] = torch.empty((batch, H, W, N_OUT), dtype=torch.float32, device=self.dev)
        self.A, self.batch_cap = A, batch
        self.G = {}      # gradient buffers, allocated on first use
        self._x = None   # the input of the last forward pass

    def reserve(self, batch):
        """Grow the activation buffers to hold `batch` members per forward pass
        (inference in large batches after training in small ones)."""
        if batch > self.batch_cap:
            self._alloc(batch)

    def _g(self, name, shape):
        torch = _torch()
        t = self.G.get(name)
        if t is None or tuple(t.shape) != tuple(shape):
            t = torch.empty(shape, dtype=torch.bfloat16, device=self.dev)
            self.G[name] = t
        return t

    def _s(self):
        return C.c_void_p(_torch().cuda.current_stream().cuda_stream)

    @staticmethod
    def _chk(st):
        if st != 0:
            raise _lib.HP2DError(st, "surrogate kernel launch failed")

    # ----------------------------------------------------------------- forward
    def _conv(self, li, x, y, n, lv, skip=None, y_skip=None):
        p = self.convs[li]
        self._chk(self.L.hp2d_nn_conv3x3_fwd(n, self.H >> lv, self.W >> lv, p["ci"], p["co"], _p(x),
                                             _p(p["wb"]), _p(p["b"]), 1, _p(skip), _p(y),
                                             _p(y_skip), self._s()))

    def forward(self, x, n=None):
        """x: [n, H, W, 32] bf16 device tensor (first 6 channels used) -> [n, H, W, 2] FP32."""
        n = x.shape[0] if n is None else n
        if n > self.batch_cap:
            raise ValueError("batch larger than the allocated capacity")
        A, L, s = self.A, self.L, self._s()
        self._x = x
        cur = x
        for lv in range(4):
            self._conv(lv, cur, A[f"e{lv}"], n, lv)
            if lv < 3:
                co = ENC[lv][1]
                self._chk(L.hp2d_nn_avgpool2(_p(A[f"e{lv}"]), n, self.H >> lv, self.W >> lv, co,
                                             _p(A[f"p{lv}"]), 0, s))
                cur = A[f"p{lv}"]
        cur = A["e3"]
        for k, (ci, co) in enumerate(DEC):
            lv = 2 - k
            self._chk(L.hp2d_nn_upsample2(_p(cur), n, self.H >> lv, self.W >> lv, ci, _p(A[f"u{lv}"]),
                                          0, s))
            self._conv(4 + k, A[f"u{lv}"], A[f"d{lv}r"], n, lv, skip=A[f"e{lv}"],
                       y_skip=A[f"d{lv}"])
            cur = A[f"d{lv}"]
        P = n * self.H * self.W
        self._chk(L.hp2d_nn_head(_p(A["d0"]), P, 32, _p(self.head["w"]), _p(self.head["b"]), N_OUT,
                                 _p(A["out"]), None, None, None, None, s))
        return A["out"][:n]

    # ---------------------------------------------------------------- backward
    def backward(self, g_out, n):
        """Accumulate parameter gradients for d loss / d out = g_out [n,H,W,2] (FP32)."""
        A, L, s = self.A, self.L, self._s()
        H, W = self.H, self.W
        P = n * H * W
        g_d = self._g("g_d0", (n, H, W, 32))
        self._chk(L.hp2d_nn_head(_p(A["d0"]), P, 32, _p(self.head["w"]), None, N_OUT, None,
                                 _p(g_out), _p(g_d), _p(self.head["gw"]), _p(self.head["gb"]), s))
        skip_grads = {}
        for lv in range(3):  # D0, D1, D2: the reverse of the forward order
            ci, co = DEC[2 - lv]
            li = 6 - lv
            h, w = H >> lv, W >> lv
            # d = relu(conv(u)) + e: the skip receives g_d unchanged
            skip_grads[lv] = g_d
            g_c = self._g(f"g_c{li}", (n, h, w, co))
            g_c.copy_(g_d)
            self._chk(L.hp2d_nn_relu_mask(_p(g_c), _p(A[f"d{lv}r"]), g_c.numel(), s))
            p = self.convs[li]
            self._chk(L.hp2d_nn_conv3x3_wgrad(n, h, w, ci, co, _p(A[f"u{lv}"]), _p(g_c), _p(p["gw"]),
                                              _p(p["gb"]), s))
            g_u = self._g(f"g_u{lv}", (n, h, w, ci))
            self._chk(L.hp2d_nn_conv3x3_dgrad(n, h, w, ci, co, _p(g_c), _p(p["wb"]), _p(p["wflip"]),
                                              _p(g_u), s))
            g_d = self._g(f"g_up{lv}", (n, h // 2, w // 2, ci))
            self._chk(L.hp2d_nn_upsample2(_p(g_u), n, h, w, ci, _p(g_d), 1, s))
        # g_d now holds d loss / d e3
        g_e = g_d
        for lv in range(3, -1, -1):
            ci, co = ENC[lv]
            h, w = H >> lv, W >> lv
            if lv < 3:
                # e_lv feeds the pool (gradient g_pool) and the decoder skip
                g_sum = self._g(f"g_e{lv}", (n, h, w, co))
                self._chk(L.hp2d_nn_add(_p(g_e), _p(skip_grads[lv]), _p(g_sum), g_sum.numel(), s))
                g_e = g_sum
            g_c = g_e
            self._chk(L.hp2d_nn_relu_mask(_p(g_c), _p(A[f"e{lv}"]), g_c.numel(), s))
            p = self.convs[lv]
            x_in = self._x if lv == 0 else A[f"p{lv - 1}"]
            self._chk(L.hp2d_nn_conv3x3_wgrad(n, h, w, ci, co, _p(x_in), _p(g_c), _p(p["gw"]),
                                              _p(p["gb"]), s))
            if lv == 0:
                break
            g_p = self._g(f"g_p{lv - 1}", (n, h, w, ci))
            self._chk(L.hp2d_nn_conv3x3_dgrad(n, h, w, ci, co, _p(g_c), _p(p["wb"]), _p(p["wflip"]),
                                              _p(g_p), s))
            g_e = self._g(f"g_pool{lv - 1}", (n, 2 * h, 2 * w, ci))
            self._chk(L.hp2d_nn_avgpool2(_p(g_p), n, 2 * h, 2 * w, ci, _p(g_e), 1, s))

    def zero_grad(self):
        for p in self.convs + [self.head]:
            p["gw"].zero_()
            p["gb"].zero_()

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        L, s, t = self.L, self._s(), self.step_count
        for p in self.convs:
            self._chk(L.hp2d_nn_adam(_p(p["w"]), _p(p["mw"]), _p(p["vw"]), _p(p["gw"]),
                                     p["w"].numel(), lr, beta1, beta2, eps, t, _p(p["wb"]), s))
            self._chk(L.hp2d_nn_adam(_p(p["b"]), _p(p["mb"]), _p(p["vb"]), _p(p["gb"]),
                                     p["b"].numel(), lr, beta1, beta2, eps, t, None, s))
        h = self.head
        for k in ("w", "b"):
            self._chk(L.hp2d_nn_adam(_p(h[k]), _p(h["m" + k]), _p(h["v" + k]), _p(h["g" + k]),
                                     h[k].numel(), lr, beta1, beta2, eps, t, None, s))

    def train_step(self, x, target, mask=None, lr=1e-3):
        """One Adam step on a batch: x [n,H,W,32] bf16, target [n,H,W,2] FP32,
        mask [n,H,W] uint8 or None.  Returns the loss (FP32 device scalar)."""
        torch = _torch()
        n = x.shape[0]
        out = self.forward(x, n)
        g_out = torch.empty((n, self.H, self.W, N_OUT), dtype=torch.float32, device=self.dev)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        P = n * self.H * self.W
        count = float(P * N_OUT) if mask is None else float(max(int(mask.sum().item()), 1) * N_OUT)
        self._chk(self.L.hp2d_nn_mse(_p(out), _p(target), _p(mask), P, N_OUT, 1.0 / count, _p(g_out),
                                     _p(loss), self._s()))
        self.zero_grad()
        self.backward(g_out, n)
        self.adam_step(lr)
        return loss

    def predict(self, x):
        """Inference in chunks of the allocated batch; returns a new FP32 tensor."""
        torch = _torch()
        out = torch.empty((x.shape[0], self.H, self.W, N_OUT), dtype=torch.float32, device=self.dev)
        for i in range(0, x.shape[0], self.batch_cap):
            xb = x[i:i + self.batch_cap]
            out[i:i + xb.shape[0]].copy_(self.forward(xb, xb.shape[0]))
        return out

    def to_bf16(self, x32):
        """FP32 device tensor -> bf16 device tensor through the library kernel."""
        torch = _torch()
        y = torch.empty(x32.shape, dtype=torch.bfloat16, device=self.dev)
        self._chk(self.L.hp2d_nn_to_bf16(_p(x32), x32.numel(), _p(y), self._s()))
        return y


# ---------------------------------------------------------------------------
# Training data from the solver, and the solver as the surrogate's judge
# ---------------------------------------------------------------------------
THETA_S = 0.4


def member_features(spec):
    """The input maps of one member [H, W, 8] (FP32, normalised):
    elevation (minus its mean, / 10 m), Manning n / 0.02, Ks / 5 mm/h, initial
    soil moisture theta_s - dtheta, total rainfall depth / 25 mm, peak intensity
    / 50 mm/h, log10 of the D8 upslope contributing cells / 4, kinematic normal
    depth of the peak excess rainfall [dm] (slopes floored at 1e-3)."""
    from .synthetic import d8_slope, flow_accumulation
    z = np.asarray(spec.elevation, dtype=np.float64)
    H, W = z.shape
    ga = spec.infiltration
    ks = float(ga.ksat_mmh[0]) if ga is not None else 0.0
    dth = float(ga.dtheta[0]) if ga is not None else 0.0
    t = np.asarray(spec.rain_time_min, dtype=np.float64)
    v = np.asarray(spec.rain_intensity_mmh, dtype=np.float64)
    t0 = np.concatenate([[0.0], t[:-1]])
    dur = np.minimum(t, spec.routing_time_min) - np.minimum(t0, spec.routing_time_min)
    total = float(np.sum(v * np.maximum(dur, 0.0)) / 60.0)
    peak = float(v.max())
    f = np.empty((H, W, N_IN), dtype=np.float32)
    f[..., 0] = (z - z.mean()) / 10.0
    f[..., 1] = np.asarray(spec.roughness) / 0.02
    f[..., 2] = ks / 5.0
    f[..., 3] = (THETA_S - dth) / THETA_S
    f[..., 4] = total / 25.0
    f[..., 5] = peak / 50.0
    area = np.maximum(flow_accumulation(z), 1.0)
    f[..., 6] = np.log10(area) / 4.0
    dx = float(spec.resolution)
    q = max(peak - ks, 0.0) / 3.6e6 * area * dx  # [m2/s] per unit width
    slope = np.maximum(np.nan_to_num(d8_slope(z, dx), nan=0.0), 1e-3)
    h_kin = (np.asarray(spec.roughness, dtype=np.float64) * q / np.sqrt(slope)) ** 0.6
    f[..., 7] = h_kin * 10.0
    return f


# target units: peak depth in cm, time to peak in units of 10 min, so that
# the squared error the network is trained on is mostly the depth's (the judge
# measures peak-depth RMSE) while the time to peak still carries weight
DEPTH_SCALE = 100.0
TPEAK_SCALE = 6.0


def solve_members(specs, graph_steps=50, chunk=4096, log=None):
    """Run the members as batched ensembles of up to `chunk` members; returns
    FP32 arrays (peak depth [m], time to peak [h]) of shape [n, H, W] and the
    wall time.  `log` (optional) receives a line after each batch."""
    import time
    from .simulation import Ensemble
    t0 = time.perf_counter()
    peaks, tpk = [], []
    for c0 in range(0, len(specs), chunk):
        with Ensemble(specs[c0:c0 + chunk]) as ens:
            ens.record_time_to_peak()
            ens.run(graph_steps=graph_steps)
            for i in range(min(chunk, len(specs) - c0)):
                m = ens.member(i)
                peaks.append(m.state()["dmax"] / 1000.0)
                tpk.append(m.time_to_peak() / 60.0)
        if log:
            log(f"solved {len(peaks)} of {len(specs)} members ({time.perf_counter() - t0:.1f} s)")
    return (np.asarray(peaks, dtype=np.float32), np.asarray(tpk, dtype=np.float32),
            time.perf_counter() - t0)


def stack_inputs(net, specs, workers=8, chunk=512):
    """Device bf16 input tensor [n, H, W, 32] of the members (N_IN maps, zero padding).
    The host maps of the members are built on `workers` threads (the library's
    host routines release the interpreter lock), `chunk` members at a time, so
    that only the bf16 stack is held whole (a 20 000-member stack at 256 x 256
    is 84 GB in bf16, twice that in FP32)."""
    from concurrent.futures import ThreadPoolExecutor
    torch = _torch()
    n = len(specs)
    out = torch.empty((n, net.H, net.W, N_PAD), dtype=torch.bfloat16, device=net.dev)
    x = np.zeros((min(chunk, n), net.H, net.W, N_PAD), dtype=np.float32)
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        for c0 in range(0, n, chunk):
            m = min(chunk, n - c0)

            def fill(i):
                x[i, :, :, :N_IN] = member_features(specs[c0 + i])

            list(pool.map(fill, range(m)))
            xd = torch.from_numpy(x[:m]).to(net.dev)
            net._chk(net.L.hp2d_nn_to_bf16(_p(xd), xd.numel(), _p(out[c0:c0 + m]), net._s()))
    return out


def stack_targets(net, peak_m, tpeak_h):
    torch = _torch()
    t = np.stack([peak_m * DEPTH_SCALE, tpeak_h * TPEAK_SCALE], axis=-1).astype(np.float32)
    return torch.from_numpy(t).to(net.dev)


def train(net, x, y, epochs=10, batch=32, lr=1e-3, seed=0, log=None, lr_final=None):
    """Mini-batch Adam over (x, y) device tensors; returns the loss of each epoch.
    With lr_final the learning rate follows a cosine from lr to lr_final."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    hist = []
    lr0 = lr
    for ep in range(epochs):
        if lr_final is not None and epochs > 1:
            lr = lr_final + 0.5 * (lr0 - lr_final) * (1 + np.cos(np.pi * ep / (epochs - 1)))
        order = rng.permutation(n)
        tot, cnt = 0.0, 0
        for i in range(0, n - batch + 1, batch):
            idx = torch.from_numpy(order[i:i + batch]).to(net.dev)
            loss = net.train_step(x.index_select(0, idx), y.index_select(0, idx), lr=lr)
            tot += float(loss.item())
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.5f}")
    return hist


def judged_predictions(net, specs, judge_fraction=0.05, tol_m=0.02, seed=0, batch=1024):
    """Surrogate peak-depth maps for all members, judged by the solver.

    A `judge_fraction` sample of the members is re-run by the plain solver;
    if the surrogate's peak-depth RMSE on that sample exceeds `tol_m` the
    error indicator trips and every member falls back to the solver.  Returns
    a dict with the per-member peak depth [m], the judge's RMSE, whether the
    fallback fired, the network's own peak-depth [m] and time-to-peak [h] maps
    and the wall times of each part."""
    import time
    torch = _torch()
    n = len(specs)
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    net.reserve(min(batch, n))  # inference in batches of up to `batch` members
    x = stack_inputs(net, specs)
    pred = net.predict(x)
    torch.cuda.synchronize()
    peak = (pred[..., 0] / DEPTH_SCALE).clamp(min=0.0).cpu().numpy()
    tpeak = (pred[..., 1] / TPEAK_SCALE).clamp(min=0.0).cpu().numpy()
    t_inf = time.perf_counter() - t0
    net_peak = peak.copy()
    k = max(1, int(round(judge_fraction * n)))
    judged = np.sort(rng.choice(n, size=k, replace=False))
    jp, _, t_judge = solve_members([specs[i] for i in judged])
    rmse = float(np.sqrt(np.mean((peak[judged] - jp) ** 2)))
    fallback = rmse > tol_m
    t_fb = 0.0
    if fallback:
        full, _, t_fb = solve_members(specs)
        peak = full
    else:
        peak[judged] = jp
    return {"peak_m": peak, "net_peak_m": net_peak, "net_tpeak_h": tpeak,
            "judge_rmse_m": rmse, "fallback": fallback, "judged": judged,
            "t_inference_s": t_inf, "t_judge_s": t_judge, "t_fallback_s": t_fb,
            "t_total_s": t_inf + t_judge + t_fb}
