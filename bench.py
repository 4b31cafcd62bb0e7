#!/usr/bin/env python3
"""Throughput of the HydroPol2D explicit surface-water step on one MI355X.

A "step" is one iteration of the device-resident Main_While loop on the
4096 x 4096 full-physics workload (sink-filled fractal DEM, 8 LULC classes, 5 Green-Ampt
soil classes, Gaussian-random-field rainfall on a triangular hyetograph,
critical-depth outlets on one edge, one pollutant with build-up / wash-off,
the groundwater module every 10 surface steps): the fused rainfall +
Green-Ampt + local-inertial kernel, the wash-off passes, the one-block
dt/bookkeeping kernel, and every 10th step the groundwater module (recharge
bucket + Boussinesq).  Inputs are resident in HBM before timing; the model
first spins up untimed for 10 min of model time; steps run as captured
hipGraphs, built before the timed window.

Prints one JSON line (value = cell-updates/s of the whole job), with the
roofline of the dominant kernel, that of the whole step, and the FP64 CPU
oracle on one thread and on all host cores.  Beside it: the batched-ensemble
study (members/hour against the plain per-member loop, --ensemble-members)
the surrogate studies (members/hour of the surrogate with its 5 % solver
judge, --surrogate-train; its skill against the training-mean map and its
time-to-peak error on wetter members, --surrogate-wet-train) and the D8
drop-in kernels (--d8-size, --d8-only).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=400)
    p.add_argument("--warmup", type=int, default=40)
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    p.add_argument("--graph-steps", type=int, default=50)
    p.add_argument("--spinup-min", type=float, default=10.0,
                   help="untimed model minutes run before the timed window")
    p.add_argument("--spinup-max-steps", type=int, default=60000,
                   help="cap on the spin-up steps (bounds the run's wall time)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=11,
                   help="steps of the CPU oracle sample (11 includes one groundwater call)")
    p.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    p.add_argument("--launch-shape", default="", metavar="TJ,WPB,ROWS",
                   help="run the step kernel at this launch shape (columns per wave task, waves "
                        "per block, rows per lane) instead of the tuned one")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="threads of the all-core CPU leg (0: OMP_NUM_THREADS, else every core)")
    p.add_argument("--ensemble-members", type=int, default=512,
                   help="members of the batched-ensemble measurement (0 skips it)")
    p.add_argument("--ensemble-max-log2", type=int, default=10,
                   help="largest member side 2^k (sides log-uniform from 2^7; 10 = the 1024^2 of "
                        "the 512-member configuration)")
    p.add_argument("--ensemble-plain-sample", type=int, default=64, metavar="K",
                   help="time the plain per-member solver on every (members/K)-th member only and "
                        "scale by cell-steps (0: every member)")
    p.add_argument("--surrogate-train", type=int, default=1536, metavar="N",
                   help="the surrogate's members/hour: train on N solved 256 x 256 members of "
                        "the dry family, judge 256 more (0 skips it; about 1 min at N = 1536)")
    p.add_argument("--surrogate-wet-train", type=int, default=1024, metavar="N",
                   help="the surrogate's skill: train on N members of the wet family (120 mm/h "
                        "for 45 min), evaluate 256 more (0 skips it; about 1 min at N = 1024)")
    p.add_argument("--surrogate-epochs", type=int, default=60,
                   help="training epochs at batch 64")
    p.add_argument("--d8-size", type=int, default=4096, metavar="N",
                   help="the D8 drop-in study on an N x N grid (0 skips it)")
    p.add_argument("--nn-only", action="store_true",
                   help="run only the surrogate network's kernel study (training steps at batch "
                        "32 and inference at batch 256 on 256 x 256 stacks) and print its JSON line")
    p.add_argument("--variants-only", action="store_true",
                   help="time the loop's other paths on the full-physics workload: the van Genuchten "
                        "closure in place of Green-Ampt, and CA routing (without the wash-off)")
    p.add_argument("--d8-only", action="store_true",
                   help="run only the D8 drop-in study and print its JSON line")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="only run 5 device copies of MIB MiB (the known-size kernel that "
                        "scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return p.parse_args()


PROFILE_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "li_step_kernel_4096_fp32.md")


def measured_traffic():
    """HBM bytes per li_step_kernel launch from the committed counter summary
    (a line `traffic_bytes_per_launch = N`), or None when it is absent."""
    try:
        with open(PROFILE_SUMMARY) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1].split()[0])
    except (OSError, ValueError, IndexError):
        return None
    return None


def calibrate_copy(mib):
    import torch
    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"[bench] copy calibration: 5 copies of {n * 4} bytes", flush=True)


def ensemble_throughput(n_members, max_log2=9, plain_sample=0):
    """Members/hour of the batched ensemble against the plain per-member GPU
    solver (the same device loop, one member after another), on C4-style
    members: grid sides log-uniform in [128, 2^max_log2], dx = 5 m, perturbed
    n, Ks and rainfall, 2 h simulated each.  With plain_sample = K the plain
    solver runs every (n/K)-th member only, and its time for all members is
    that sample's time scaled by the cell-steps of all members over the
    sample's (both counted from the batched run)."""
    from hydropol2d_amd import synthetic
    from hydropol2d_amd.simulation import Ensemble, Simulation
    specs = synthetic.ensemble_c4(n_members, seed=1000, min_log2=7, max_log2=max_log2,
                                  routing_time_min=120.0)
    sizes = [sp.shape[0] * sp.shape[1] for sp in specs]
    with Ensemble(specs) as ens:
        t0 = time.perf_counter()
        ens.run(graph_steps=50)
        t_batch = time.perf_counter() - t0
        peaks = [ens.member(i).state()["dmax"] for i in range(len(specs))]
        steps = [ens.member(i).report()["steps"] for i in range(len(specs))]
    cell_steps = [c * k for c, k in zip(sizes, steps)]
    stride = max(1, n_members // plain_sample) if plain_sample > 0 else 1
    sample = list(range(0, n_members, stride))
    t_sample = 0.0
    max_diff = 0.0
    print(f"[bench] ensemble: batched run of {n_members} members in {t_batch:.2f} s",
          file=sys.stderr, flush=True)
    for k, i in enumerate(sample):
        with Simulation(specs[i]) as sim:
            t0 = time.perf_counter()
            sim.run()
            t_sample += time.perf_counter() - t0
            max_diff = max(max_diff,
                           float(np.nanmax(np.abs(sim.state()["dmax"] - peaks[i]))) / 1000)
        if (k + 1) % 16 == 0:
            print(f"[bench] ensemble: {k + 1} of {len(sample)} members run alone ({t_sample:.1f} s)",
                  file=sys.stderr, flush=True)
    t_plain = t_sample * sum(cell_steps) / sum(cell_steps[i] for i in sample)
    print(f"[bench] ensemble {n_members} members: batched {t_batch:.2f} s, plain {t_plain:.2f} s "
          f"({len(sample)} members timed)", file=sys.stderr, flush=True)
    return {"members": n_members, "cells": int(sum(sizes)), "member_steps": int(sum(steps)),
            "cell_updates": int(sum(cell_steps)),
            "cell_updates_per_s_batched": sum(cell_steps) / t_batch,
            "members_per_hour_batched": n_members / t_batch * 3600.0,
            "members_per_hour_plain": n_members / t_plain * 3600.0,
            "speedup": t_plain / t_batch, "plain_members_timed": len(sample),
            "max_peak_depth_diff_m": max_diff,
            "workload": f"C4-style: sides 2^U(7,{max_log2}), dx=5 m, fractal DEM on a 1 % tilt, "
                        "n/Ks/rain x e^U(ln .5, ln 1.5) around C1, 2 h each, FP32"}


def surrogate_study(n_train, epochs, n_eval=256, n_plain=8, wet=False, batch=64):
    """The surrogate with its 5 % solver judge on 256 x 256 C4-style members
    (2 h each, FP32).  The training data are the solver's own peak-depth and
    time-to-peak maps of n_train members; the n_eval held-out members give the
    peak-depth RMSE against the solver (and against the training-mean map, the
    forecast without the network), the time-to-peak RMSE on their wet cells
    (peak depth >= 1 cm) and the judge's verdict at 2 cm.  The dry family (50
    mm/h for 30 min around the C1 values) sets the members/hour against the
    plain per-member solver and the batched ensemble; the wet family (120 mm/h
    for 45 min, RMS peak depth about 11 cm) measures the network's skill."""
    import torch
    from hydropol2d_amd import surrogate as S
    from hydropol2d_amd import synthetic
    from hydropol2d_amd.simulation import Simulation
    tag = "wet" if wet else "dry"

    def log(msg):
        print(f"[bench] surrogate ({tag}): {msg}", file=sys.stderr, flush=True)

    rain = dict(rain_mmh=120.0, rain_duration_min=45.0) if wet else {}
    specs = synthetic.ensemble_c4(n_train + n_eval, seed=6000 if wet else 5000, min_log2=8, max_log2=8,
                                  routing_time_min=120.0, vary_soil_moisture=True, filled=True, **rain)
    pk, tp, t_data = S.solve_members(specs, log=log)
    net = S.UNetSurrogate(256, 256, batch=batch, seed=1)
    t0 = time.perf_counter()
    x = S.stack_inputs(net, specs[:n_train])
    y = S.stack_targets(net, pk[:n_train], tp[:n_train])
    log(f"input maps of {n_train} members in {time.perf_counter() - t0:.1f} s")
    S.train(net, x, y, epochs=epochs, batch=batch, lr=1e-3, lr_final=5e-5, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del x, y
    res = S.judged_predictions(net, specs[n_train:], judge_fraction=0.05, tol_m=0.02)
    ref, ref_tp = pk[n_train:], tp[n_train:]
    rmse = float(np.sqrt(np.mean((res["peak_m"] - ref) ** 2)))
    unjudged = np.setdiff1d(np.arange(n_eval), res["judged"])
    rmse_net = float(np.sqrt(np.mean((res["net_peak_m"][unjudged] - ref[unjudged]) ** 2)))
    rmse_mean = float(np.sqrt(np.mean((pk[:n_train].mean(axis=0)[None] - ref) ** 2)))
    wetc = ref >= 0.01
    ttp_net = float(np.sqrt(np.mean((res["net_tpeak_h"][wetc] - ref_tp[wetc]) ** 2)) * 60.0)
    ttp_mean = float(np.sqrt(np.mean((np.broadcast_to(tp[:n_train].mean(axis=0)[None], ref_tp.shape)[wetc]
                                      - ref_tp[wetc]) ** 2)) * 60.0)
    t_pred = res["t_inference_s"] + res["t_judge_s"] + res["t_fallback_s"]
    out = {
        "members_train": n_train, "members_eval": n_eval, "epochs": epochs, "batch": batch,
        "rms_peak_depth_m": float(np.sqrt(np.mean(ref ** 2))),
        "peak_depth_rmse_m": rmse, "surrogate_rmse_m_unjudged": rmse_net,
        "rmse_mean_map_m": rmse_mean, "rmse_zero_map_m": float(np.sqrt(np.mean(ref ** 2))),
        "skill_vs_mean_map": 1.0 - rmse_net / rmse_mean,
        "time_to_peak_rmse_min_wet_cells": ttp_net, "time_to_peak_rmse_min_mean_map": ttp_mean,
        "judge_rmse_m": res["judge_rmse_m"], "judge_members": int(len(res["judged"])),
        "fallback": bool(res["fallback"]),
        "s_inference": res["t_inference_s"], "s_judge": res["t_judge_s"],
        "s_fallback": res["t_fallback_s"], "s_training_data": t_data, "s_training": t_train,
    }
    log(f"data {t_data:.1f} s, train {t_train:.1f} s, {n_eval} members in {t_pred:.3f} s; "
        f"RMSE {rmse_net * 100:.2f} cm (mean map {rmse_mean * 100:.2f} cm), time to peak "
        f"{ttp_net:.1f} min (mean map {ttp_mean:.1f} min), judge {res['judge_rmse_m'] * 100:.2f} cm"
        f"{', fallback' if res['fallback'] else ''}")
    if wet:
        out["workload"] = ("256x256 C4-style members, dx=5 m, sink-filled fractal DEM on a 1 % tilt, "
                           "n/Ks x e^U(ln .5, ln 1.5), rain 120 mm/h x e^U(ln .5, ln 1.5) for 45 min, "
                           "soil moisture deficit U(0.15, 0.35), 2 h each, FP32")
        return out
    t_plain = 0.0
    for sp in specs[n_train:n_train + n_plain]:
        with Simulation(sp) as sim:
            t1 = time.perf_counter()
            sim.run()
            t_plain += time.perf_counter() - t1
    per_plain = t_plain / n_plain
    per_batch = t_data / len(specs)
    study = 20000  # members of a probabilistic study (C5 scale)
    t_study = t_data + t_train + study / n_eval * t_pred
    out.update({
        "members_per_hour_surrogate": n_eval / t_pred * 3600.0,
        "members_per_hour_batched": 3600.0 / per_batch,
        "members_per_hour_plain": 3600.0 / per_plain,
        "speedup_vs_plain": per_plain * n_eval / t_pred,
        "speedup_vs_batched": per_batch * n_eval / t_pred,
        "speedup_vs_plain_20000_member_study_incl_training": study * per_plain / t_study,
        "workload": "256x256 C4-style members, dx=5 m, sink-filled fractal DEM on a 1 % tilt, "
                    "n/Ks/rain x e^U(ln .5, ln 1.5), soil moisture deficit U(0.15, 0.35), 2 h "
                    "each, FP32; judge = 5 % of the members re-run by the solver, fallback above "
                    "2 cm",
    })
    return out


def nn_study(train_batch=32, infer_batch=256, reps=10):
    """The surrogate network's kernels on 256 x 256 input stacks (random
    inputs and targets: the timing does not depend on them): training steps
    (forward, MSE, backward, Adam) at batch train_batch and inference at
    batch infer_batch, as FLOP rates of the 3x3 convolutions (forward, data
    gradient and weight gradient of the training step each count one forward's
    FLOPs) against the dense BF16 MFMA peak."""
    import torch
    from hydropol2d_amd import surrogate as S
    H = W = 256
    net = S.UNetSurrogate(H, W, batch=max(train_batch, infer_batch), seed=1)
    rng = np.random.default_rng(0)
    x32 = torch.from_numpy(rng.normal(size=(train_batch, H, W, 32)).astype(np.float32)).cuda()
    x = net.to_bf16(x32)
    y = torch.from_numpy(rng.normal(size=(train_batch, H, W, 2)).astype(np.float32)).cuda()
    flop_fwd = 0
    for lv, (ci, co) in enumerate(S.ENC):
        flop_fwd += 2 * (H >> lv) * (W >> lv) * co * 9 * ci
    for k, (ci, co) in enumerate(S.DEC):
        flop_fwd += 2 * (H >> (2 - k)) * (W >> (2 - k)) * co * 9 * ci
    for _ in range(2):
        net.train_step(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.train_step(x, y)
    torch.cuda.synchronize()
    t_train = (time.perf_counter() - t0) / reps
    xb = x.repeat(infer_batch // train_batch, 1, 1, 1)
    net.predict(xb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.predict(xb)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    peak = 2.5e15
    train_rate = 3 * flop_fwd * train_batch / t_train
    inf_rate = flop_fwd * infer_batch / t_inf
    print(f"[bench] nn: train step {t_train * 1e3:.2f} ms ({train_rate / 1e12:.1f} TFLOP/s), "
          f"inference {t_inf * 1e3:.2f} ms ({inf_rate / 1e12:.1f} TFLOP/s)", file=sys.stderr,
          flush=True)
    return {"grid": [H, W], "conv_gflop_per_sample_forward": flop_fwd / 1e9,
            "train_batch": train_batch, "ms_per_train_step": t_train * 1e3,
            "train_tflops": train_rate / 1e12, "train_frac_of_bf16_peak": train_rate / peak,
            "infer_batch": infer_batch, "ms_per_inference_batch": t_inf * 1e3,
            "infer_tflops": inf_rate / 1e12, "infer_frac_of_bf16_peak": inf_rate / peak,
            "peak_tflops": peak / 1e12, "dtype": "bf16 (fp32 accumulation)"}


# algorithmic HBM bytes per cell of the D8 drop-in kernels (FP64): the inputs
# read once, the outputs written once
D8_BYTES = {
    # z, d_tot, d_p, roughness, 9 previous pages + outlet byte | 9 pages, 9 Hf, d_t, I_tot
    "bates8d_kernel": (13 * 8 + 1, 20 * 8),
    # z, d_tot, roughness, h_0, I_prev + outlet byte | 9 pages, d_t, I_tot
    "ca8d_kernel": (5 * 8 + 1, 11 * 8),
    # 9 pages, d_t | qin, d_t
    "gather_kernel": (10 * 8, 2 * 8),
    # z, d_tot, roughness, the two momentum pages | five outflow pages, d_t, I_tot, 3 Hf, C_a
    "full momentum step (4 kernels)": (5 * 8, 11 * 8),
    # d_tot, two previous flux pages | 3 outflow pages, d_t, I_tot, 3 Hf, C_a, eta, V
    # (the table reads are counted apart: see subgrid_table_bytes)
    "subgrid step (all kernels)": (3 * 8, 11 * 8),
}


def d8_study(n, calls=4):
    """Kernel time of the routing drop-ins beside the main step (the D8
    routines, the gather, the full-momentum step on an n x n grid, the subgrid
    step on n/2 x n/2 coarse cells): HIP events around the kernels on the
    handle's stream, host<->device copies excluded, as HBM throughput of their
    algorithmic bytes."""
    import ctypes as C

    from hydropol2d_amd import _lib
    from hydropol2d_amd import local_inertial as li
    rng = np.random.default_rng(3)
    z = np.asfortranarray(rng.random((n, n)) * 0.5 + np.linspace(0.0, 20.0, n)[None, :])
    d = np.asfortranarray(rng.random((n, n)) * 200.0)
    res = (None,) * 12
    L = _lib.load()
    h = li._Handle.get(n, n, 0)
    t = C.c_double()
    out = {}

    def timed(fn):
        ms = []
        for _ in range(calls):
            r = fn()
            _lib.check(L.hp2d_last_kernel_time(h, C.byref(t)), handle=h)
            ms.append(t.value)
        return r, float(np.median(ms[1:]))

    g, ms_b = timed(lambda: li.Bates_Inertial_8D(*res, 0, z, d, d, 0.03, 4.0, 0.002, 2.0, None, 2,
                                                   0.0, [n], [n], 1.0, None, None, 1))
    q = g[11]
    del g
    _, ms_c = timed(lambda: li.CA_Routing_8D(*res, 0, z, d, 0.03, 4.0, 0.002, None, 2.0,
                                             np.zeros((n, n)), None, 2, 0.0, [n], [n], None, 1,
                                             1.0))
    _, ms_g = timed(lambda: li.outflows_to_inflows(q, 0.002, d))
    del q
    from hydropol2d_amd import subgrid as sgm
    _, ms_f = timed(lambda: li.Full_Momentum_Model_D4(1, *res, 0, z, d, None, 0.03, None, 4.0,
                                                      0.002, 2.0, None, 2, 0.0, [n], [n], 1.0,
                                                      None, None, 0, 0, None, None, None, None,
                                                      None, None, None, None, None, None, 0, 0,
                                                      None, None))
    # subgrid: an n/2 x n/2 coarse grid of 2 x 2 fine cells (10 levels)
    nc = n // 2

    class _R:
        def __init__(self, Z, cs):
            self.Z, self.cellsize = Z, cs
    T, _ = sgm.Subgrid_Properties_Lookup(_R(z, 2.0), _R(np.full((n, n), 0.03), 2.0),
                                         _R(np.zeros((nc, nc)), 4.0), 4.0, nr_levels=10)
    hs = li._Handle.get(nc, nc, 0)
    dc = np.asfortranarray(d[::2, ::2] * 4.0)
    ms_s = []
    prev = None
    for _ in range(calls):
        g = sgm.Local_Inertial_Model_D4_Subgrid(1, *res, 0, None, dc, None, None, None, 16.0,
                                                0.002, 4.0, None, 2, 0.0, [nc], [nc], 1.0, prev,
                                                None, 0, None, None, None, None, None, None, None,
                                                None, None, None, 0, T)
        prev = g[7]
        _lib.check(L.hp2d_last_kernel_time(hs, C.byref(t)), handle=hs)
        ms_s.append(t.value)
    ms_s = float(np.median(ms_s[1:]))
    del T
    for name, ms in (("bates8d_kernel", ms_b), ("ca8d_kernel", ms_c), ("gather_kernel", ms_g),
                     ("full momentum step (4 kernels)", ms_f)):
        rd, wr = D8_BYTES[name]
        gbs = (rd + wr) * n * n / (ms / 1000.0) / 1e9
        out[name] = {"avg_kernel_ms": ms, "bytes_per_cell": rd + wr, "achieved": gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "cell_updates_per_s": n * n / (ms / 1000.0)}
        print(f"[bench] d8 {name}: {ms:.4f} ms, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    rd, wr = D8_BYTES["subgrid step (all kernels)"]
    out["subgrid step (all kernels)"] = {
        "avg_kernel_ms": ms_s, "grid": [nc, nc], "nr_levels": 10, "bytes_per_cell": rd + wr,
        "achieved": (rd + wr) * nc * nc / (ms_s / 1000.0) / 1e9, "unit": "GB/s",
        "cell_updates_per_s": nc * nc / (ms_s / 1000.0),
        "note": "HIP events around the whole step, which waits once per internal sub-step "
                "for its negative-volume flag; table reads not in bytes_per_cell"}
    print(f"[bench] subgrid step {nc}x{nc}: {ms_s:.4f} ms", file=sys.stderr, flush=True)
    return {"grid": [n, n], "dtype": "f64", "calls": calls, "kernels": out}


def step_bytes_per_cell(spec, elem, li_bytes):
    """Algorithmic bytes per cell of one whole step of the full-physics loop
    (FP32 state; the groundwater module keeps its head and ledgers in FP64),
    kernel by kernel; each distinct per-cell array read once and written once
    per launch, class tables and per-block partials left out."""
    parts = {"li_step_kernel": li_bytes}
    if spec.water_quality is not None:
        # wq_first_kernel: read the LULC/flag word 2, B 4, q_x 4, q_y 4, Tot_Washed 4;
        # write B 4, Tot_Washed 4, P_conc 4 (the internal sub-steps are guarded no-ops
        # unless a split occurs)
        parts["wq_first_kernel"] = 30
    if spec.groundwater is not None:
        every = spec.groundwater.every_steps
        per_call = {
            # bucket: read elevation 8, head 8, soil class 1, I 4, pending net / recharge /
            # capillary 24, cumulative recharge 8; write I 4, the three pending 24,
            # cumulative 8, recharge rate 8, q_exf 8
            "gwm_recharge_kernel": 105,
            # Boussinesq predictor faces: read h 8, z_bed 8; write Fx 8, Fy 8
            "gw_flux_kernel (predictor)": 32,
            # h_p = h + dt/Sy (-div F): read h, Fx, Fy, z_bed; write h_p
            "gw_predict_kernel": 40,
            # corrector faces averaged with the predictor's: read h_p, z_bed, Fx, Fy; write Fx, Fy
            "gw_flux_kernel (corrector)": 48,
            # head update + exfiltration: read h, Fx, Fy, R, z_bed; write q_exf, h
            "gw_update_kernel": 56,
            # Courant dt of the entry head (read h) and of the new head (read h and the
            # catchment byte, write the masked head)
            "gw_velocity_kernel x2": 25,
            # takes over head and q_exf, returns exfiltration to the surface, resets the
            # pending sums, excess: read elevation 8, h 8, q 8, d 4, class 1, I 4; write h 8,
            # q 8, d 4, pending 24, excess 8
            "gwm_post_kernel": 85,
            # unsaturated-zone excess back to the surface: read I 4, excess 8, d 4; write I, d
            "gwm_excess_kernel": 24,
        }
        parts["groundwater module per call"] = sum(per_call.values())
        parts["groundwater module per step"] = sum(per_call.values()) / every
    total = sum(v for k, v in parts.items() if k != "groundwater module per call")
    return total, parts


def algorithmic_bytes_per_cell(spec, elem):
    """Distinct per-cell arrays read once and written once per step."""
    # d, n^2 (outlet flag in the sign), q_x, q_y and the elevation: FP64 reads z,
    # FP32 the two face elevation differences (dz_x, dz_y)
    reads = 4 * elem + (elem if elem == 8 else 2 * elem)
    writes = 3 * elem           # d, q_x, q_y
    if spec.rain_pattern is not None:
        reads += elem           # rainfall multiplier map
    if spec.infiltration is not None:
        reads += 2 * elem + (1 if spec.infiltration.soil_class is not None else 0)  # I, T, class
        writes += 2 * elem      # I, T
    reads += elem               # dmax (save_output_maps.m:588-600); it is stored only where
    return reads, writes        # the maximum rises, so its write is not counted


def variant_study(n=4096, steps=400, spin_steps=2000, graph_steps=50):
    """The full-physics workload (FP32) with the van Genuchten Infiltration_Module
    in place of Green-Ampt (the aquifer's soil tables, Ltop 5 cm), and with CA
    routing in place of the local-inertial step (the loop runs CA without the
    wash-off).  After a spin-up, `steps` steps are timed as graphs; the routing
    kernels' own time is the average over 20 eager steps timed with HIP events."""
    from hydropol2d_amd import synthetic
    from hydropol2d_amd.runspec import VanGenuchten
    from hydropol2d_amd.simulation import Simulation
    import torch
    out = {}
    for name in ("vg", "ca"):
        spec = synthetic.full_physics_c2(n, n, 2.0, precision="fp32", routing_time_min=1e6)
        if name == "vg":
            g, ga = spec.groundwater, spec.infiltration
            spec.infiltration = VanGenuchten(
                ksat_mmh=ga.ksat_mmh, theta_sat=g.theta_sat, theta_r=g.theta_r, n_vg=g.n_vg,
                alpha_vg=g.alpha_vg, h0=g.h0, soil_depth=g.soil_depth, soil_class=ga.soil_class,
                l_vg=g.l_vg)
        else:
            spec.routing = "ca"
            spec.water_quality = None
        with Simulation(spec) as sim:
            sim.run_fixed(spin_steps, graph_steps=graph_steps)
            if name == "vg":  # the step kernel's launch shape, tuned as the default run's is
                sim.autotune()
            torch.cuda.synchronize() if torch.cuda.is_available() else None
            t0 = time.perf_counter()
            sim.run_fixed(steps, graph_steps=graph_steps)
            wall = time.perf_counter() - t0
            rep = sim.report()
            sim.run_fixed(20, graph_steps=0)
            k_ms, _, launches = sim.kernel_time()
            sim_shape = sim.launch_shape() if name == "vg" else "one lane per cell"
        ms = wall / steps * 1e3
        out[name] = {"ms_per_step": ms, "G_cell_updates_per_s": n * n / (ms * 1e-3) / 1e9,
                     "launch": sim_shape,
                     "routing_kernel_ms": k_ms, "routing_launches_timed": launches,
                     "t_min": rep["t_min"],
                     "workload": f"{n}x{n} full-physics FP32, " +
                                 ("van Genuchten infiltration" if name == "vg" else
                                  "CA routing, no wash-off")}
        print(f"[bench] variant {name}: {ms:.3f} ms/step, routing kernels "
              f"{out[name]['routing_kernel_ms']:.3f} ms", file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.nn_only:
        print(json.dumps({"nn": nn_study()}), flush=True)
        return
    if args.d8_only:
        print(json.dumps({"d8": d8_study(args.d8_size or 4096)}), flush=True)
        return
    if args.variants_only:
        print(json.dumps({"variants": variant_study(args.size, args.steps)}), flush=True)
        return
    from hydropol2d_amd import synthetic
    from hydropol2d_amd.simulation import Simulation

    n = args.size
    t0 = time.time()
    spec = synthetic.full_physics_c2(n, n, 2.0, precision=args.precision, routing_time_min=1e6)
    print(f"[bench] built {n}x{n} workload in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    elem = 4 if args.precision == "fp32" else 8
    cells = n * n
    sim = Simulation(spec)
    # warm-up (also lets water start moving: rainfall begins at t = 0)
    sim.run_fixed(args.warmup, graph_steps=min(args.graph_steps, max(args.warmup, 1)))
    print("[bench] warm-up done", file=sys.stderr, flush=True)
    # untimed spin-up: the rain fills the domain and water starts to move; the
    # same graph size as the timed window, so its graph is built here
    ts = time.perf_counter()
    spin = 0
    while spin < args.spinup_max_steps:
        chunk = min(2000, args.spinup_max_steps - spin)
        sim.run_fixed(chunk, graph_steps=args.graph_steps)
        spin += chunk
        t_model = sim.report()["t_min"]
        print(f"[bench] spin-up {spin} steps, t = {t_model:.3f} min, "
              f"{time.perf_counter() - ts:.1f} s", file=sys.stderr, flush=True)
        if t_model >= args.spinup_min:
            break
    # the launch shape is tuned on the wet domain of the timed window (the
    # tuner drops the captured graph); the next graph's worth of steps builds
    # the timed window's graph again, so that the timed interval holds
    # launches only
    if args.launch_shape:
        tj, wpb, rows = (int(v) for v in args.launch_shape.split(","))
        sim.set_launch_shape(tj, wpb, rows)
        tune = sim.launch_shape()
    else:
        tune = sim.launch_shape() if args.no_tune else sim.autotune()
    print(f"[bench] launch shape {tune}", file=sys.stderr, flush=True)
    sim.run_fixed(args.graph_steps, graph_steps=args.graph_steps)
    t_start_min = sim.report()["t_min"]
    # timed region: K whole steps through captured graphs; run_fixed (like the
    # autotuner before it) returns only after synchronising the simulation's
    # stream, so the device is idle on both sides of the wall-clock interval
    t1 = time.perf_counter()
    sim.run_fixed(args.steps, graph_steps=args.graph_steps)
    t2 = time.perf_counter()
    _, graph_ms, _ = sim.kernel_time()
    elapsed_s = t2 - t1
    ms_per_step = elapsed_s * 1000.0 / args.steps
    value = cells * args.steps / elapsed_s
    # dominant kernel: average duration from HIP events around each launch on the sim stream
    sim.run_fixed(min(args.steps, 50), graph_steps=0)
    li_ms, _, launches = sim.kernel_time()
    rd, wr = algorithmic_bytes_per_cell(spec, elem)
    bytes_per_launch = (rd + wr) * cells
    achieved = bytes_per_launch / (li_ms / 1000.0) / 1e9
    step_b, step_parts = step_bytes_per_cell(spec, elem, rd + wr)
    step_gbs = step_b * cells / (ms_per_step / 1000.0) / 1e9
    rep = sim.report()
    sim.close()
    print(f"[bench] {args.steps} steps, {ms_per_step:.4f} ms/step, LI kernel {li_ms:.4f} ms "
          f"({achieved:.0f} GB/s), window t={t_start_min:.3f}..{rep['t_min']:.3f} min",
          file=sys.stderr, flush=True)

    cpu = cpu1 = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        cpu_spec = synthetic.full_physics_c2(n, n, 2.0, precision="fp64", routing_time_min=1e6)
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", 0) or 0) or \
            (os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
        legs = {}
        for mode, nthreads in (("st", 1), ("omp", threads)):
            oracle.use_openmp(mode == "omp")
            tc = time.perf_counter()
            oracle.run(cpu_spec, max_steps=args.cpu_steps, n_dt_max=16)
            legs[mode] = (time.perf_counter() - tc, nthreads)
            print(f"[bench] cpu oracle ({nthreads} thread(s)): "
                  f"{cells * args.cpu_steps / legs[mode][0]:.3e} cell-updates/s",
                  file=sys.stderr, flush=True)
        oracle.use_openmp(False)
        sample = f"first {args.cpu_steps} steps (one groundwater call) of the same {n}x{n} workload"
        cpu = {"value": cells * args.cpu_steps / legs["omp"][0], "unit": "cell-updates/s",
               "cores": legs["omp"][1], "kind": "port", "nproc": os.cpu_count(),
               "sample": f"{sample}, FP64 C oracle with OpenMP on its element-wise loops "
                         f"({legs['omp'][1]} threads, reductions serial), {legs['omp'][0]:.1f} s"}
        cpu1 = {"value": cells * args.cpu_steps / legs["st"][0], "unit": "cell-updates/s",
                "cores": 1, "kind": "port",
                "sample": f"{sample}, FP64 C oracle, one thread, {legs['st'][0]:.1f} s"}

    ens = (ensemble_throughput(args.ensemble_members, args.ensemble_max_log2,
                               args.ensemble_plain_sample)
           if args.ensemble_members > 0 else None)
    sur = (surrogate_study(args.surrogate_train, args.surrogate_epochs)
           if args.surrogate_train > 0 else None)
    sur_wet = (surrogate_study(args.surrogate_wet_train, args.surrogate_epochs, wet=True)
               if args.surrogate_wet_train > 0 else None)
    d8 = d8_study(args.d8_size) if args.d8_size > 0 else None

    line = {
        "metric": "cell-updates/s (% HBM roofline)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": f"full-physics single domain {n}x{n} dx=2 m: sink-filled fractal DEM, "
                               "8 LULC Manning classes, 5 Green-Ampt soil classes, GRF rainfall "
                               "on a 2 h triangular hyetograph, critical-depth edge outlets, one "
                               "pollutant (build-up / wash-off), groundwater module (Sy=0.2, "
                               "K=1e-5 m/s) every 10 steps; fused rain+GA+local-inertial step, "
                               "wash-off passes, dt control; timed window after "
                               f"{t_start_min:.1f} min of model time",
                   "cells": cells, "precision": args.precision,
                   "launch": {"cols_per_wave": tune["tj"], "waves_per_block": tune["waves_per_block"],
                              "rows_per_lane": tune["rows"]},
                   "bytes_per_cell_update": rd + wr},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "algorithmic_bytes": bytes_per_launch,
                     "kernel": "li_step_kernel", "avg_kernel_ms": li_ms,
                     "launches_timed": launches, "graph_event_ms": graph_ms},
        "step_roofline": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "bytes_per_cell_step": step_b, "per_kernel_bytes_per_cell": step_parts,
                          "ms_per_step": ms_per_step},
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "ensemble": ens,
        "surrogate": sur,
        "surrogate_wet": sur_wet,
        "d8": d8,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
