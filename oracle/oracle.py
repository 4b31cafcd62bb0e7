"""ctypes wrapper of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from hydropol2d_amd import abi
from hydropol2d_amd.runspec import RunSpec

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libhp2d_oracle.so")
_SO_OMP = os.path.join(_HERE, "_build", "libhp2d_oracle_omp.so")
# HP2D_ORACLE_SO names another build of the same sources (the sanitizer build
# of the memory-safety test); it replaces the single-threaded library
_SO_OVERRIDE = os.environ.get("HP2D_ORACLE_SO")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO) or not os.path.exists(_SO_OMP):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def use_openmp(on: bool = True):
    """Route the following calls to the OpenMP build of the oracle (the
    all-core CPU baseline: its element-wise loops run on OMP_NUM_THREADS
    threads, with results bit-identical to the single-threaded build), or
    back to the single-threaded one."""
    global _mode
    _mode = "omp" if on else "st"


_mode = "st"
_libs = {}


class DtParams(C.Structure):
    _fields_ = [
        ("resolution", C.c_double), ("alfa_min", C.c_double), ("alfa_max", C.c_double),
        ("max_time_step", C.c_double), ("min_time_step", C.c_double),
        ("time_step_increments", C.c_double), ("depth_tolerance_mm", C.c_double),
        ("flag_timestep", C.c_int32), ("flag_adaptive_timestepping", C.c_int32),
        ("flag_waterquality", C.c_int32), ("tmin_wq", C.c_double), ("nt_raw", C.c_double),
        ("total_velocity", _dp),
    ]


class RunOut(C.Structure):
    _fields_ = [
        ("d_t", _dp), ("outflow", _dp), ("I_t", _dp), ("dmax", _dp), ("B_t", _dp),
        ("gw_h", _dp), ("hydro_q", _dp), ("hydro_t", _dp), ("hydro_depth", _dp),
        ("n_hydro_max", C.c_int32), ("n_hydro", C.c_int32), ("dt_hist", _dp),
        ("n_dt_max", C.c_int32), ("max_steps", C.c_int64), ("inject_fail_step", C.c_int64),
        ("rep", abi.RunReport), ("volume_error_hist", _dp), ("n_report_max", C.c_int32),
        ("P_conc", _dp), ("Tot_Washed", _dp),
        ("dt_follow", _dp), ("n_follow", C.c_int64), ("n_dt_ties", C.c_int64),
        ("n_dt_mismatch", C.c_int64), ("stop_after_rollback", C.c_int32),
        ("rollback_dt_s", C.c_double), ("rollback_max_dt_s", C.c_double),
        ("dt_force", C.c_int32), ("max_risk", _dp), ("rec", _dp), ("rec_t", _dp),
        ("n_rec_max", C.c_int32), ("n_rec", C.c_int32),
    ]


def lib():
    if _mode not in _libs:
        build()
        _libs[_mode] = _load(_SO_OMP if _mode == "omp" else (_SO_OVERRIDE or _SO))
    return _libs[_mode]


def _load(path):
    if True:
        L = C.CDLL(path)
        L.oracle_inertial_solver.argtypes = [C.c_int32, C.c_int64, C.c_int32, _dp, C.c_double,
                                             _dp, _dp, _dp, _dp]
        L.oracle_li_d4.argtypes = [C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp,
                                   C.c_double, C.c_double, C.c_double, C.c_int32, C.c_double,
                                   _i64p, C.c_int32, _dp, C.c_int32] + [_dp] * 9
        L.oracle_li_d4.restype = C.c_int32
        L.oracle_refresh_dt.argtypes = [C.c_int32, C.c_int32, _dp, _dp, _dp,
                                        C.POINTER(DtParams), _dp, _dp]
        L.oracle_refresh_dt.restype = C.c_double
        L.oracle_refresh_dt_ca.argtypes = [C.c_int32, C.c_int32, _dp, _dp,
                                           C.POINTER(DtParams), _dp, _dp]
        L.oracle_refresh_dt_ca.restype = C.c_double
        L.oracle_rain_avg.argtypes = [_dp, _dp, C.c_int32, C.c_double, C.c_double,
                                      C.POINTER(C.c_int32)]
        L.oracle_rain_avg.restype = C.c_double
        L.oracle_run.argtypes = [C.POINTER(abi.RunConfig), C.POINTER(abi.RunInputs),
                                 C.POINTER(RunOut)]
        L.oracle_run.restype = C.c_int32
        L.oracle_ga_params.argtypes = [C.c_double] * 5 + [_dp]
        L.oracle_ga_params.restype = None
        L.oracle_ga_cells.argtypes = [C.c_int64, C.c_int32, C.c_double, C.c_double, _dp, _dp,
                                      _dp, _dp, _dp, _dp]
        L.oracle_ga_cells.restype = None
        L.oracle_idw.argtypes = [C.c_int32, _dp, _dp, _dp, C.c_int32, _dp, C.c_int32, _dp,
                                 C.c_double, C.c_double, _dp]
        L.oracle_idw.restype = None
        if hasattr(L, "oracle_ca_routing"):
            L.oracle_ca_routing.argtypes = [C.c_int32, C.c_int32, _dp, _dp, _dp, C.c_double,
                                            C.c_double, _dp, C.c_double, _dp, C.c_int32,
                                            C.c_double, _i64p, C.c_int32, C.c_int32] + [_dp] * 7
            L.oracle_ca_routing.restype = C.c_int32
        L.oracle_washoff.argtypes = [C.c_int32, C.c_int32] + [_dp] * 8 + [C.c_double, C.c_double,
                                     C.POINTER(C.c_uint8), C.c_int32, _dp, _dp, C.c_double,
                                     C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.oracle_washoff.restype = C.c_int32
        _u8 = C.POINTER(C.c_uint8)
        L.oracle_full_momentum.argtypes = ([C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp,
                                            C.c_double, C.c_double, _u8, C.c_int32, C.c_double,
                                            _i64p, C.c_int32, C.c_double, _dp, C.c_int32, _u8,
                                            C.c_int32, C.POINTER(_u8), _dp, _dp, _dp, _dp, _dp])
        L.oracle_full_momentum.restype = C.c_int32
        _tp = C.POINTER(abi.SubgridTablesC)
        L.oracle_subgrid_build.argtypes = [_dp, C.c_int32, C.c_int32, C.c_double, _dp, C.c_int32,
                                           C.c_double, C.c_int32, C.c_int32, C.c_int32, _tp]
        L.oracle_subgrid_build.restype = C.c_int32
        L.oracle_subgrid_queries.argtypes = [_tp, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp]
        L.oracle_subgrid_queries.restype = None
        L.oracle_subgrid_face_state.argtypes = [C.POINTER(abi.SubgridFaces), C.c_int64, C.c_int32,
                                                _dp, _dp, _dp, _dp]
        L.oracle_subgrid_face_state.restype = None
        L.oracle_subgrid_volume.argtypes = [_tp, C.c_int32, C.c_int32, _dp, _dp, C.c_double,
                                            _dp, _dp, _dp, _dp, _dp]
        L.oracle_subgrid_volume.restype = None
        L.oracle_li_subgrid.argtypes = ([C.c_int32, C.c_int32, _tp, _dp, C.c_double, C.c_double,
                                         C.c_int32, C.c_double, _i64p, C.POINTER(C.c_int32),
                                         C.c_int32, _dp, C.c_int32, _u8] + [_dp] * 7
                                        + [C.POINTER(C.c_int32)])
        L.oracle_li_subgrid.restype = C.c_int32
        L.oracle_bates_8d.argtypes = ([C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, C.c_double,
                                       C.c_double, C.c_double, C.c_int32, C.c_double, _i64p,
                                       C.c_int32, _dp, C.c_int32, _dp, _dp, _dp, _dp])
        L.oracle_bates_8d.restype = C.c_int32
        L.oracle_ca_8d.argtypes = ([C.c_int32, C.c_int32, _dp, _dp, _dp, C.c_double, C.c_double,
                                    _dp, C.c_double, _dp, C.c_int32, C.c_double, _i64p,
                                    C.c_int32, C.c_int32, C.c_double, _dp, _dp, _dp])
        L.oracle_ca_8d.restype = C.c_int32
        L.oracle_gather.argtypes = [C.c_int32, C.c_int32, _dp, C.c_int32, C.c_double, _dp, _dp]
        L.oracle_gather.restype = None
        L.oracle_washoff_8d.argtypes = ([C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, C.c_double,
                                         C.c_double, _u8, C.c_int32, _dp, _dp, C.c_double,
                                         C.c_double, C.c_double, _dp, _dp, _dp, _dp])
        L.oracle_washoff_8d.restype = C.c_int32
        L.oracle_boussinesq_2d.argtypes = ([C.c_int32, C.c_int32, C.c_double, C.c_double,
                                            C.c_double] + [_dp] * 5 + [C.c_double, _dp, _u8, _u8,
                                            _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_int32)])
        L.oracle_boussinesq_2d.restype = C.c_int32
        L.oracle_gw_recharge.argtypes = ([C.c_int64, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp,
                                          _dp, _dp, C.c_double, _dp, _dp, _dp, _dp])
        L.oracle_gw_recharge.restype = C.c_int32
        _fp = C.POINTER(C.c_float)
        L.oracle_conv3x3_fwd.argtypes = [C.c_int32] * 5 + [_fp, _fp, _fp, C.c_int32, _fp]
        L.oracle_conv3x3_fwd.restype = None
        L.oracle_conv3x3_bwd.argtypes = [C.c_int32] * 5 + [_fp] * 6
        L.oracle_conv3x3_bwd.restype = None
        L.oracle_infiltration.argtypes = ([C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double]
                                          + [_dp] * 8 + [_u8, C.c_double, C.c_double, C.c_double,
                                                         C.c_double] + [_dp] * 7)
        L.oracle_infiltration.restype = C.c_int32
        L.oracle_groundwater_module.argtypes = [C.POINTER(GwmParams), C.POINTER(GwmState)] + \
            [_dp] * 6 + [_dp, C.POINTER(C.c_int32), _dp]
        L.oracle_groundwater_module.restype = C.c_int32
    return L


class GwmParams(C.Structure):
    _fields_ = [("ny", C.c_int32), ("nx", C.c_int32), ("time_step_min", C.c_double),
                ("resolution", C.c_double), ("t_min", C.c_double), ("routing_time_min", C.c_double),
                ("flag_baseflow", C.c_int32), ("flag_async", C.c_int32),
                ("target_dt_min", C.c_double), ("min_dt_min", C.c_double),
                ("max_head_change_m", C.c_double), ("courant", C.c_double), ("l_vg", C.c_double)] + \
               [(k, C.POINTER(C.c_double)) for k in ("elevation", "soil_depth", "theta_sat", "theta_r",
                                                     "n_vg", "alpha_vg", "ksat_mmh", "ksat_gw_mmh",
                                                     "Sy")] + [("domain", C.POINTER(C.c_uint8))]


class GwmState(C.Structure):
    _fields_ = [("pending_net", C.POINTER(C.c_double)), ("pending_recharge", C.POINTER(C.c_double)),
                ("pending_capillary", C.POINTER(C.c_double)), ("elapsed_s", C.c_double),
                ("total_recharge_m3", C.c_double), ("total_capillary_m3", C.c_double),
                ("total_exfiltration_m3", C.c_double), ("last_dt_s", C.c_double),
                ("n_updates", C.c_int32)]


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else _dp()


def outlet_linear(row_outlet, col_outlet, ny):
    """sub2ind(...)-1 of 1-based MATLAB rows/cols (unique, find() order)."""
    r = np.asarray(row_outlet, dtype=np.int64).ravel()
    c = np.asarray(col_outlet, dtype=np.int64).ravel()
    return np.ascontiguousarray((r - 1) + (c - 1) * ny, dtype=np.int64)


def inertial_solver(q_p, dt, Hf, S, n_sq):
    q_p, Hf, S, n_sq = _f(q_p), _f(Hf), _f(S), _f(n_sq)
    n_cells = n_sq.size
    n_pages = q_p.size // n_cells
    q = np.zeros_like(q_p)
    lib().oracle_inertial_solver(1, n_cells, n_pages, _p(q_p), dt, _p(Hf), _p(S), _p(n_sq), _p(q))
    return q


def li_d4(z, d_tot, roughness_cell, roughness_squared, cell_area, time_step_min, resolution,
          outlet_type, slope_outlet, row_outlet, col_outlet, outflow_prev, flag_critical=0):
    """Local_Inertial_Model_D4 (subgrid off) on the oracle; returns a dict of outputs."""
    z = _f(z)
    ny, nx = z.shape
    d_tot = _f(d_tot)
    rc, rs = _f(roughness_cell), _f(roughness_squared)
    of_prev = _f(outflow_prev).reshape(ny, nx, 3, order="F")
    oidx = outlet_linear(row_outlet, col_outlet, ny)
    outs = {k: np.zeros((ny, nx), order="F") for k in
            ("qout_left", "qout_right", "qout_up", "qout_down", "outlet_flow", "d_t",
             "I_tot_end_cell")}
    outs["outflow"] = np.zeros((ny, nx, 3), order="F")
    outs["Hf"] = np.zeros((ny, nx, 3), order="F")
    lib().oracle_li_d4(ny, nx, 1, _p(z), _p(d_tot), _p(rc), _p(rs), cell_area, time_step_min,
                       resolution, outlet_type, slope_outlet, oidx.ctypes.data_as(_i64p),
                       oidx.size, _p(of_prev), flag_critical,
                       *[_p(outs[k]) for k in ("qout_left", "qout_right", "qout_up", "qout_down",
                                               "outlet_flow", "d_t", "I_tot_end_cell",
                                               "outflow", "Hf")])
    return outs


def ca_routing(z, d_tot, roughness, cell_area, time_step_min, h_0, resolution,
               I_tot_end_cell_prev, outlet_type, slope_outlet, row_outlet, col_outlet,
               flag_critical=0):
    z = _f(z)
    ny, nx = z.shape
    oidx = outlet_linear(row_outlet, col_outlet, ny)
    h0 = _f(h_0) if h_0 is not None else None
    outs = {k: np.zeros((ny, nx), order="F") for k in
            ("qout_left", "qout_right", "qout_up", "qout_down", "outlet_flow", "d_t",
             "I_tot_end_cell")}
    lib().oracle_ca_routing(ny, nx, _p(z), _p(_f(d_tot)), _p(_f(roughness)), cell_area,
                            time_step_min, _p(h0), resolution, _p(_f(I_tot_end_cell_prev)),
                            outlet_type, slope_outlet, oidx.ctypes.data_as(_i64p), oidx.size,
                            flag_critical,
                            *[_p(outs[k]) for k in ("qout_left", "qout_right", "qout_up",
                                                    "qout_down", "outlet_flow", "d_t",
                                                    "I_tot_end_cell")])
    return outs


def refresh_dt(qout_right, qout_up, Hf, resolution, alfa_min=0.7, alfa_max=0.7,
               max_time_step=10.0, min_time_step=0.01, time_step_increments=0.001,
               depth_tolerance_mm=1.0, flag_timestep=1, flag_adaptive_timestepping=1,
               flag_waterquality=0, tmin_wq=np.inf):
    qr, qu, Hf = _f(qout_right), _f(qout_up), _f(Hf)
    ny, nx = qr.shape
    p = DtParams(resolution, alfa_min, alfa_max, max_time_step, min_time_step,
                 time_step_increments, depth_tolerance_mm, flag_timestep,
                 flag_adaptive_timestepping, flag_waterquality, tmin_wq)
    vmax = C.c_double()
    raster = np.zeros((ny, nx), order="F")
    tcr = lib().oracle_refresh_dt(ny, nx, _p(qr), _p(qu), _p(Hf), C.byref(p), C.byref(vmax),
                                  _p(raster))
    return tcr, vmax.value, raster


def refresh_dt_ca(q4, d_t, resolution, alfa_min=0.7, alfa_max=0.7, max_time_step=10.0,
                  min_time_step=0.01, time_step_increments=0.001, depth_tolerance_mm=1.0,
                  flag_timestep=1, flag_adaptive_timestepping=1):
    """refreshing_timestep with flags.flag_CA = 1; q4: ny x nx x 4 (left, right, up,
    down) [mm/h], d_t [mm].  Returns (tcr [s], max_velocity, velocity_raster)."""
    q = np.asfortranarray(np.asarray(q4, dtype=np.float64))
    d = _f(d_t)
    ny, nx = d.shape
    p = DtParams(resolution, alfa_min, alfa_max, max_time_step, min_time_step,
                 time_step_increments, depth_tolerance_mm, flag_timestep,
                 flag_adaptive_timestepping, 0, np.inf)
    vmax = C.c_double()
    raster = np.zeros((ny, nx), order="F")
    tcr = lib().oracle_refresh_dt_ca(ny, nx, _p(q), _p(d), C.byref(p), C.byref(vmax), _p(raster))
    return tcr, vmax.value, raster


def rain_avg(times, values, t0, t1, cursor=1):
    times, values = _f(times).ravel(), _f(values).ravel()
    cur = C.c_int32(cursor)
    v = lib().oracle_rain_avg(_p(times), _p(values), min(times.size, values.size), t0, t1,
                              C.byref(cur))
    return v, cur.value


def ga_params(ksat_mmh, psi_mm, dtheta, theta_sat_minus_r, lu_m=np.nan):
    """Derived Green-Ampt parameters of one soil class: dict of ksat, psi,
    dtheta, Lu_mm, k_out, Tr (HydroPol2D_preprocessing.m:2249-2253)."""
    out = np.zeros(6)
    lib().oracle_ga_params(ksat_mmh, psi_mm, dtheta, theta_sat_minus_r, lu_m, _p(out))
    return dict(zip(("ksat", "psi", "dtheta", "Lu_mm", "k_out", "Tr"), out.tolist()))


def ga_step(params, first_step, dt_h, d, I_p, T, I_min_mm=1e-3):
    """One legacy explicit Green-Ampt update of the cells d (post-rain depth,
    mm), I_p (mm), T (recovery clock, h) with one class's parameters (the
    dict of ga_params).  Returns (d_after, I_t, T, f)."""
    d = np.array(d, dtype=np.float64, ndmin=1)
    I_p = np.ascontiguousarray(np.broadcast_to(np.asarray(I_p, dtype=np.float64), d.shape))
    T = np.array(np.broadcast_to(np.asarray(T, dtype=np.float64), d.shape))
    I_t = np.zeros_like(d)
    f = np.zeros_like(d)
    p6 = np.array([params[k] for k in ("ksat", "psi", "dtheta", "Lu_mm", "k_out", "Tr")])
    lib().oracle_ga_cells(d.size, int(bool(first_step)), dt_h, I_min_mm, _p(p6), _p(d), _p(I_p),
                          _p(I_t), _p(T), _p(f))
    return d, I_t, T, f


def run(spec: RunSpec, max_steps: int = 0, inject_fail_step: int = 0, n_dt_max: int = 200000,
        n_report_max: int = 64, dt_follow=None, stop_after_rollback: bool = False,
        dt_force: bool = False, n_map_records: int = 0):
    """Full Main_While run on the oracle.  Returns a dict of final state + ledgers.

    dt_follow: the accepted-step dt sequence [s] of another implementation; the
    oracle adopts it only at floating-point ties of the dt snap (see
    oracle_run_out) and reports n_dt_ties / n_dt_mismatch; with dt_force it
    adopts every one of them (an FP32 run compared on its own time levels)."""
    ny, nx = spec.shape
    cfg, ins, keep = spec.to_c()
    cfg.precision = abi.HP2D_FP64
    n_h = spec.n_hydro_records()
    res = {
        "d_t": np.zeros((ny, nx), order="F"),
        "outflow": np.zeros((ny, nx, 3), order="F"),
        "I_t": np.zeros((ny, nx), order="F"),
        "dmax": np.zeros((ny, nx), order="F"),
        "B_t": np.zeros((ny, nx), order="F"),
        "P_conc": np.zeros((ny, nx), order="F"),
        "Tot_Washed": np.zeros((ny, nx), order="F"),
        "gw_h": np.zeros((ny, nx), order="F"),
        "hydro_q": np.zeros(n_h),
        "hydro_t": np.zeros(n_h),
        "hydro_depth": np.zeros(n_h),
        "dt_hist": np.zeros(n_dt_max),
        "volume_error_hist": np.zeros(n_report_max),
    }
    out = RunOut()
    out.d_t, out.outflow, out.I_t, out.dmax = (_p(res[k]) for k in ("d_t", "outflow", "I_t", "dmax"))
    out.B_t, out.P_conc, out.Tot_Washed, out.gw_h = (_p(res[k]) for k in
                                                     ("B_t", "P_conc", "Tot_Washed", "gw_h"))
    out.hydro_q, out.hydro_t, out.hydro_depth = (_p(res[k]) for k in
                                                 ("hydro_q", "hydro_t", "hydro_depth"))
    out.n_hydro_max = n_h
    out.dt_hist = _p(res["dt_hist"])
    out.n_dt_max = n_dt_max
    out.max_steps = max_steps
    out.inject_fail_step = inject_fail_step
    out.volume_error_hist = _p(res["volume_error_hist"])
    out.n_report_max = n_report_max
    if dt_follow is not None:
        follow = np.ascontiguousarray(dt_follow, dtype=np.float64)
        out.dt_follow = _p(follow)
        out.n_follow = follow.size
    res["max_risk"] = np.zeros((ny, nx), order="F")
    out.max_risk = _p(res["max_risk"])
    if n_map_records > 0:
        recbuf = np.zeros(n_map_records * len(abi.MAP_FIELDS) * ny * nx)
        rect = np.zeros(n_map_records)
        out.rec, out.rec_t, out.n_rec_max = _p(recbuf), _p(rect), n_map_records
    out.stop_after_rollback = int(stop_after_rollback)
    out.dt_force = int(dt_force)
    rc = lib().oracle_run(C.byref(cfg), C.byref(ins), C.byref(out))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed with status {rc}")
    rep = out.rep.as_dict()
    res["report"] = rep
    if n_map_records > 0:
        nr = out.n_rec
        R = recbuf.reshape(n_map_records, len(abi.MAP_FIELDS), nx, ny)[:nr]
        res["map_records"] = {name: np.asfortranarray(R[:, f].transpose(2, 1, 0))
                              for f, name in enumerate(abi.MAP_FIELDS)}
        res["map_record_t"] = rect[:nr]
    res["rollback_dt_s"] = out.rollback_dt_s
    res["rollback_max_dt_s"] = out.rollback_max_dt_s
    res["n_hydro"] = out.n_hydro
    res["n_dt_ties"] = out.n_dt_ties
    res["n_dt_mismatch"] = out.n_dt_mismatch
    res["dt_hist"] = res["dt_hist"][: min(rep["steps"], n_dt_max)]
    res["volume_error_hist"] = res["volume_error_hist"][: min(rep["n_reports"], n_report_max)]
    del keep
    return res


def washoff(C3, C4, qout_left, qout_right, qout_up, qout_down, outlet_flow, B_t, time_step_min,
            cell_area, nan_mask=None, flag_wq_model=0, mass_lost=0.0, Tot_Washed=None, Bmin=0.0,
            Bmax=0.0, min_Bt=0.0):
    """build_up_wash_off on the oracle; returns a dict (B_t, P_conc, Out_Conc, tmin_wq,
    tot_W_out, mass_lost, Tot_Washed)."""
    B = _f(B_t).copy(order="F")
    if B.ndim == 1:
        B = B.reshape(1, -1, order="F")
    ny, nx = B.shape
    sh = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))
    C3, C4 = sh(C3), sh(C4)
    qs = [sh(q) for q in (qout_left, qout_right, qout_up, qout_down, outlet_flow)]
    tw = sh(0.0 if Tot_Washed is None else Tot_Washed).copy(order="F")
    pc = np.zeros((ny, nx), order="F")
    tot = np.zeros((ny, nx), order="F")
    ml = C.c_double(mass_lost)
    oc, tm = C.c_double(), C.c_double()
    nm = None if nan_mask is None else np.asfortranarray(np.asarray(nan_mask, dtype=np.uint8))
    nmp = nm.ctypes.data_as(C.POINTER(C.c_uint8)) if nm is not None else C.POINTER(C.c_uint8)()
    rc = lib().oracle_washoff(ny, nx, _p(C3), _p(C4), *[_p(q) for q in qs], _p(B), time_step_min,
                              cell_area, nmp, flag_wq_model, C.byref(ml), _p(tw), Bmin, Bmax,
                              min_Bt, _p(pc), C.byref(oc), C.byref(tm), _p(tot))
    if rc < 0:
        raise ValueError("washoff: tmin_wq gives more than 1e7 sub-steps")
    return {"B_t": B, "P_conc": pc, "Out_Conc": oc.value, "tmin_wq": tm.value, "tot_W_out": tot,
            "mass_lost": ml.value, "Tot_Washed": tw}


D8_PAGES = ("qout_left", "qout_right", "qout_up", "qout_down", "outlet_flow",
            "qout_ne", "qout_se", "qout_sw", "qout_nw")


def _pages(a, ny, nx, k=9):
    """(ny, nx, k) page stack in Fortran order (zeros for None)."""
    if a is None:
        return np.zeros((ny, nx, k), order="F")
    return np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx, k)))


def bates_8d(z, d_tot, d_p, roughness, cell_area, time_step_min, resolution, outlet_type,
             slope_outlet, row_outlet, col_outlet, outflow_prev=None, flag_critical=0):
    """Bates_Inertial_8D on the oracle: dict of outflow (ny, nx, 9) [mm/h], d_t,
    I_tot_end_cell and Hf (ny, nx, 9)."""
    z = _f(z)
    ny, nx = z.shape
    oidx = outlet_linear(row_outlet, col_outlet, ny)
    rough = np.asfortranarray(np.broadcast_to(np.asarray(roughness, dtype=np.float64), (ny, nx)))
    qp = _pages(outflow_prev, ny, nx)
    out = {"outflow": np.zeros((ny, nx, 9), order="F"), "d_t": np.zeros((ny, nx), order="F"),
           "I_tot_end_cell": np.zeros((ny, nx), order="F"), "Hf": np.zeros((ny, nx, 9), order="F")}
    lib().oracle_bates_8d(ny, nx, _p(z), _p(_f(d_tot)), _p(_f(d_p)), _p(rough), cell_area,
                          time_step_min, resolution, outlet_type, slope_outlet,
                          oidx.ctypes.data_as(_i64p), oidx.size, _p(qp), flag_critical,
                          _p(out["outflow"]), _p(out["d_t"]), _p(out["I_tot_end_cell"]),
                          _p(out["Hf"]))
    return out


def ca_8d(z, d_tot, roughness, cell_area, time_step_min, h_0, resolution, I_tot_end_cell_prev,
          outlet_type, slope_outlet, row_outlet, col_outlet, flag_critical=0, d_tolerance=1.0):
    """CA_Routing_8D on the oracle: dict of outflow (ny, nx, 9) [mm/h], d_t and
    I_tot_end_cell."""
    z = _f(z)
    ny, nx = z.shape
    oidx = outlet_linear(row_outlet, col_outlet, ny)
    sh = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))
    h0 = sh(h_0) if h_0 is not None else None
    out = {"outflow": np.zeros((ny, nx, 9), order="F"), "d_t": np.zeros((ny, nx), order="F"),
           "I_tot_end_cell": np.zeros((ny, nx), order="F")}
    lib().oracle_ca_8d(ny, nx, _p(z), _p(_f(d_tot)), _p(sh(roughness)), cell_area, time_step_min,
                       _p(h0), resolution, _p(sh(I_tot_end_cell_prev)), outlet_type, slope_outlet,
                       oidx.ctypes.data_as(_i64p), oidx.size, flag_critical, d_tolerance,
                       _p(out["outflow"]), _p(out["d_t"]), _p(out["I_tot_end_cell"]))
    return out


def gather(outflow, time_step_min=0.0, d_t=None):
    """The caller's inflow gather (n_pages 5 or 9); returns (qin, d_t or None)."""
    q = np.asfortranarray(np.asarray(outflow, dtype=np.float64))
    ny, nx, k = q.shape
    qin = np.zeros((ny, nx), order="F")
    dt = None if d_t is None else _f(d_t).copy(order="F")
    lib().oracle_gather(ny, nx, _p(q), k, time_step_min, _p(qin), _p(dt))
    return qin, dt


def washoff_8d(C3, C4, q9, B_t, time_step_min, cell_area, nan_mask=None, flag_wq_model=0,
               mass_lost=0.0, Tot_Washed=None, Bmin=0.0, Bmax=0.0, min_Bt=0.0):
    """build_up_wash_off_8D on the oracle (q9: (ny, nx, 9) pages [mm/h]); returns a
    dict as washoff does."""
    B = _f(B_t).copy(order="F")
    ny, nx = B.shape
    sh = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))
    q = _pages(q9, ny, nx)
    tw = sh(0.0 if Tot_Washed is None else Tot_Washed).copy(order="F")
    pc, tot = np.zeros((ny, nx), order="F"), np.zeros((ny, nx), order="F")
    ml, oc, tm = C.c_double(mass_lost), C.c_double(), C.c_double()
    nm = None if nan_mask is None else np.asfortranarray(np.asarray(nan_mask, dtype=np.uint8))
    nmp = nm.ctypes.data_as(C.POINTER(C.c_uint8)) if nm is not None else C.POINTER(C.c_uint8)()
    rc = lib().oracle_washoff_8d(ny, nx, _p(sh(C3)), _p(sh(C4)), _p(q), _p(B), time_step_min,
                                 cell_area, nmp, flag_wq_model, C.byref(ml), _p(tw), Bmin, Bmax,
                                 min_Bt, _p(pc), C.byref(oc), C.byref(tm), _p(tot))
    if rc < 0:
        raise ValueError("washoff_8d: tmin_wq gives more than 1e7 sub-steps")
    return {"B_t": B, "P_conc": pc, "Out_Conc": oc.value, "tmin_wq": tm.value, "tot_W_out": tot,
            "mass_lost": ml.value, "Tot_Washed": tw, "split": rc}


def subgrid_build(dem, cellsize, rough, coarse_res, ny, nx, nr_levels=20):
    """Subgrid_Properties_Lookup on the oracle: a SubgridTables of the coarse
    ny x nx grid (rough: fine- or coarse-shaped Manning n)."""
    from hydropol2d_amd.subgrid import SubgridTables
    dem = _f(dem)
    rough = _f(rough)
    fine = 1 if rough.shape == dem.shape else 0
    T = SubgridTables(ny, nx, nr_levels, coarse_res ** 2)
    t = T.c_struct()
    rc = lib().oracle_subgrid_build(_p(dem), dem.shape[0], dem.shape[1], cellsize, _p(rough), fine,
                                    coarse_res, ny, nx, nr_levels, C.byref(t))
    if rc != 0:
        raise ValueError(f"subgrid_build: error {rc}")
    return T


def subgrid_queries(T, zs):
    """(V from zs, zs back from that V, wet area from zs) of every cell."""
    zs = _f(zs)
    V, zb, A = (np.zeros((T.ny, T.nx), order="F") for _ in range(3))
    t = T.c_struct()
    lib().oracle_subgrid_queries(C.byref(t), T.ny, T.nx, _p(zs), _p(V), _p(zb), _p(A))
    return V, zb, A


def subgrid_face_state(T, prefix, zu):
    """(HG, nrep, phi) of a face set ('u', 'v', 'north', ...) at levels zu
    (shaped as the set's faces)."""
    zu = _f(zu)
    t = T.c_struct()
    F = getattr(t, prefix) if prefix in ("u", "v") else t.side[abi.SUBGRID_SIDES.index(prefix)]
    HG, nr, phi = (np.zeros(zu.shape, order="F") for _ in range(3))
    lib().oracle_subgrid_face_state(C.byref(F), zu.size, T.nr_levels, _p(zu), _p(HG), _p(nr),
                                    _p(phi))
    return HG, nr, phi


def subgrid_volume(T, drep_old, delta_mm, cell_area):
    sh = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (T.ny, T.nx)))
    outs = [np.zeros((T.ny, T.nx), order="F") for _ in range(5)]
    t = T.c_struct()
    lib().oracle_subgrid_volume(C.byref(t), T.ny, T.nx, _p(sh(drep_old)), _p(sh(delta_mm)),
                                cell_area, *[_p(o) for o in outs])
    return tuple(outs)


def li_subgrid(T, d_tot, time_step_min, resolution, outlet_type=2, slope_outlet=0.0,
               row_outlet=(), col_outlet=(), outflow_prev=None, idx_nan=None):
    """Local_Inertial_Model_D4_Subgrid on the oracle: dict of outflow (3 pages),
    d_t, I_tot_end_cell, Hf (3 pages), C_a, eta_t, V_t, n_substeps."""
    from hydropol2d_amd.subgrid import outlet_sides
    ny, nx = T.ny, T.nx
    d = np.asfortranarray(np.broadcast_to(np.asarray(d_tot, dtype=np.float64), (ny, nx)))
    oidx = outlet_linear(row_outlet, col_outlet, ny)
    sides = outlet_sides(row_outlet, col_outlet, ny, nx)
    prev, npg = None, 0
    if outflow_prev is not None:
        prev = _f(outflow_prev)
        if prev.ndim == 2:
            prev = prev.reshape(ny, nx, 1, order="F")
        npg = prev.shape[2]
    m = None if idx_nan is None else np.asfortranarray(np.asarray(idx_nan, dtype=np.uint8))
    out = {"outflow": np.zeros((ny, nx, 3), order="F"), "Hf": np.zeros((ny, nx, 3), order="F")}
    for k in ("d_t", "I_tot_end_cell", "C_a", "eta_t", "V_t"):
        out[k] = np.zeros((ny, nx), order="F")
    ns = C.c_int32()
    t = T.c_struct()
    rc = lib().oracle_li_subgrid(ny, nx, C.byref(t), _p(d), time_step_min, resolution, outlet_type,
                                 slope_outlet, oidx.ctypes.data_as(_i64p),
                                 sides.ctypes.data_as(C.POINTER(C.c_int32)), oidx.size, _p(prev),
                                 npg, m.ctypes.data_as(C.POINTER(C.c_uint8)) if m is not None
                                 else C.POINTER(C.c_uint8)(), _p(out["outflow"]), _p(out["d_t"]),
                                 _p(out["I_tot_end_cell"]), _p(out["Hf"]), _p(out["C_a"]),
                                 _p(out["eta_t"]), _p(out["V_t"]), C.byref(ns))
    if rc != 0:
        raise ValueError("li_subgrid: sub-step loop did not end")
    out["n_substeps"] = ns.value
    return out


def unique_outlets(row_outlet, col_outlet, ny):
    """Linear outlet indices without repeats, in first-seen order."""
    idx = outlet_linear(row_outlet, col_outlet, ny)
    _, first = np.unique(idx, return_index=True)
    return np.ascontiguousarray(idx[np.sort(first)])


def full_momentum(z, d_tot, roughness, cell_area, time_step_min, resolution, outlet_type=2,
                  slope_outlet=0.0, row_outlet=(), col_outlet=(), outflow_prev=None,
                  idx_nan=None, d_tolerance=1.0, roughness_squared=None, outlet_index=None,
                  flag_HR=0, face_masks=None):
    """Full_Momentum_Model_D4 on the oracle: dict of outflow (5 pages), d_t,
    I_tot_end_cell, Hf (3 pages), C_a.  face_masks: None or (right, left, up,
    down) cell masks (None entries allowed)."""
    z = _f(z)
    ny, nx = z.shape
    sh = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))
    u8 = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.uint8), (ny, nx)))
    oidx = unique_outlets(row_outlet, col_outlet, ny)
    prev, npg = None, 0
    if outflow_prev is not None:
        prev = _f(outflow_prev)
        if prev.ndim == 3 and prev.shape[2] >= 2:
            npg = prev.shape[2]
    n2 = None if roughness_squared is None else sh(roughness_squared)
    keep = []
    up = C.POINTER(C.c_uint8)

    def mp(a):
        if a is None:
            return up()
        m = u8(a)
        keep.append(m)
        return m.ctypes.data_as(up)
    fm = None
    if face_masks is not None:
        fm = (up * 4)(*[mp(m) for m in face_masks])
    out = {"outflow": np.zeros((ny, nx, 5), order="F"), "Hf": np.zeros((ny, nx, 3), order="F")}
    for k in ("d_t", "I_tot_end_cell", "C_a"):
        out[k] = np.zeros((ny, nx), order="F")
    lib().oracle_full_momentum(ny, nx, _p(z), _p(sh(d_tot)), _p(sh(roughness)), _p(n2),
                               _p(sh(cell_area)), time_step_min, resolution, mp(outlet_index),
                               outlet_type, slope_outlet, oidx.ctypes.data_as(_i64p), oidx.size,
                               d_tolerance, _p(prev), npg, mp(idx_nan), int(flag_HR), fm,
                               _p(out["outflow"]), _p(out["d_t"]), _p(out["I_tot_end_cell"]),
                               _p(out["Hf"]), _p(out["C_a"]))
    del keep
    return out


def _mask(m, shape):
    if m is None:
        return None, C.POINTER(C.c_uint8)()
    a = np.asfortranarray(np.broadcast_to(np.asarray(m, dtype=bool), shape).astype(np.uint8))
    return a, a.ctypes.data_as(C.POINTER(C.c_uint8))


def boussinesq_2d(dt, dx, dy, h, z0, Sy, R, K, courant, h_soil, catchment_mask=None,
                  dirichlet_mask=None, h_dirichlet=None):
    """Boussinesq_2D_explicit on the oracle; returns a dict (h_t, u_x, u_y, q_exf,
    error, n_sub).  Scalars broadcast to the grid of h."""
    h = _f(h)
    ny, nx = h.shape
    g = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))
    z0, Sy, R, K, hs = g(z0), g(Sy), g(R), g(K), g(h_soil)
    hd = g(0.0 if h_dirichlet is None else h_dirichlet)
    cm, cmp = _mask(catchment_mask, (ny, nx))
    dm, dmp = _mask(dirichlet_mask, (ny, nx))
    out = {k: np.zeros((ny, nx), order="F") for k in ("h_t", "u_x", "u_y", "q_exf")}
    err, nsub = C.c_double(), C.c_int32()
    rc = lib().oracle_boussinesq_2d(ny, nx, dt, dx, dy, _p(h), _p(z0), _p(Sy), _p(R), _p(K),
                                    courant, _p(hs), cmp, dmp, _p(hd), _p(out["h_t"]),
                                    _p(out["u_x"]), _p(out["u_y"]), _p(out["q_exf"]),
                                    C.byref(err), C.byref(nsub))
    if rc != 0:
        raise ValueError(f"boussinesq_2d: oracle returned {rc}")
    out["error"] = err.value
    out["n_sub"] = nsub.value
    return out


def gw_recharge(infiltration_rate, S0, dt, min_S, max_S, theta_s, theta_r, n_vg, alpha_vg,
                ksat_mmh, l_vg=0.5, cumulative=None):
    """simulate_groundwater_recharge on the oracle; returns (R_mb [m/s], S_new [m], cumulative [mm])."""
    S0 = _f(S0)
    shape = S0.shape
    g = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape))
    args = [g(x) for x in (infiltration_rate, S0, min_S, max_S, theta_s, theta_r, n_vg,
                           alpha_vg, ksat_mmh)]
    Rr, Sn = np.zeros(shape, order="F"), np.zeros(shape, order="F")
    cum = g(0.0 if cumulative is None else cumulative).copy(order="F")
    lib().oracle_gw_recharge(S0.size, _p(args[0]), _p(args[1]), dt, _p(args[2]), _p(args[3]),
                             _p(args[4]), _p(args[5]), _p(args[6]), _p(args[7]), l_vg,
                             _p(args[8]), _p(Rr), _p(Sn), _p(cum))
    return Rr, Sn, cum


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else C.POINTER(C.c_float)()


def conv3x3(x, w, b=None, relu=False):
    """3x3 zero-padded convolution, NHWC x [N,H,W,Cin], w [Cout,3,3,Cin] -> [N,H,W,Cout]."""
    x, w = _f32(x), _f32(w)
    N, H, W, Cin = x.shape
    Cout = w.shape[0]
    b = None if b is None else _f32(b)
    y = np.zeros((N, H, W, Cout), dtype=np.float32)
    lib().oracle_conv3x3_fwd(N, H, W, Cin, Cout, _fp(x), _fp(w), _fp(b), int(relu), _fp(y))
    return y


def conv3x3_grad(x, w, gy):
    """Gradients (dx, dw, db) of conv3x3 for the output gradient gy (before activation)."""
    x, w, gy = _f32(x), _f32(w), _f32(gy)
    N, H, W, Cin = x.shape
    Cout = w.shape[0]
    gx = np.zeros_like(x)
    gw = np.zeros_like(w)
    gb = np.zeros(Cout, dtype=np.float32)
    lib().oracle_conv3x3_bwd(N, H, W, Cin, Cout, _fp(x), _fp(w), _fp(gy), _fp(gx), _fp(gw),
                             _fp(gb))
    return gx, gw, gb


def infiltration(time_step_min, resolution, d_t, I_t, elevation, h_t, soil_depth, theta_sat,
                 theta_r, n_vg, alpha_vg, ksat_mmh, idx_imp=None, C_a=None, l_vg=0.5, Ltop=0.05,
                 dh_max=1.0, Ksurf_weight=0.5, cum_inf=None):
    """Infiltration_Module (flag_infiltration = 1, no subgrid) on the oracle; returns a dict."""
    d = _f(d_t).copy(order="F")
    ny, nx = d.shape
    g = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))
    I = g(I_t).copy(order="F")
    args = [g(a) for a in (elevation, h_t, soil_depth, theta_sat, theta_r, n_vg, alpha_vg, ksat_mmh)]
    im, imp = _mask(idx_imp, (ny, nx))
    Ip, f = np.zeros((ny, nx), order="F"), np.zeros((ny, nx), order="F")
    cum = None if cum_inf is None else g(cum_inf).copy(order="F")
    vol, dS = C.c_double(), C.c_double()
    Ca = resolution * resolution if C_a is None else C_a
    lib().oracle_infiltration(ny, nx, time_step_min, resolution, Ca, *[_p(a) for a in args], imp,
                              l_vg, Ltop, dh_max, Ksurf_weight, _p(d), _p(I), _p(Ip), _p(f),
                              _p(cum), C.byref(vol), C.byref(dS))
    return {"d_t": d, "I_t": I, "I_p": Ip, "f": f, "cum_inf": cum, "inf_volume": vol.value,
            "dS_inf": dS.value}


class GroundwaterModule:
    """Groundwater_Module (no layered soil) on the oracle, keeping GW_States
    between calls like the reference's workspace does."""

    def __init__(self, elevation, soil_depth, theta_sat, theta_r, n_vg, alpha_vg, ksat_mmh,
                 ksat_gw_mmh, Sy, resolution, domain=None, flag_baseflow=1, flag_async=1,
                 target_dt_min=1440.0, min_dt_min=1.0, max_head_change_m=0.25, courant=0.25,
                 l_vg=0.5):
        z = _f(elevation)
        self.shape = z.shape
        g = lambda a: np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), z.shape))
        self._arr = {k: g(v) for k, v in dict(elevation=z, soil_depth=soil_depth, theta_sat=theta_sat,
                                              theta_r=theta_r, n_vg=n_vg, alpha_vg=alpha_vg,
                                              ksat_mmh=ksat_mmh, ksat_gw_mmh=ksat_gw_mmh,
                                              Sy=Sy).items()}
        self._dom, dp = _mask(domain, z.shape)
        self.p = GwmParams(z.shape[0], z.shape[1], 0.0, resolution, 0.0, 0.0, flag_baseflow,
                           flag_async, target_dt_min, min_dt_min, max_head_change_m, courant, l_vg,
                           *[_p(self._arr[k]) for k in ("elevation", "soil_depth", "theta_sat",
                                                         "theta_r", "n_vg", "alpha_vg", "ksat_mmh",
                                                         "ksat_gw_mmh", "Sy")], dp)
        nanm = np.isnan(z)
        self.pending = {k: np.where(nanm, np.nan, 0.0).copy(order="F")
                        for k in ("net", "recharge", "capillary")}
        self.s = GwmState(_p(self.pending["net"]), _p(self.pending["recharge"]),
                          _p(self.pending["capillary"]), 0.0, 0.0, 0.0, 0.0, 0.0, 0)

    def step(self, time_step_min, t_min, routing_time_min, h_t, I_t, d_t, cumulative_recharge,
             max_gw_depth):
        """One call; h_t, I_t, d_t, cumulative_recharge, max_gw_depth (Fortran float64
        arrays) are updated in place.  Returns (q_exf, error, ran, dt_gw)."""
        self.p.time_step_min = time_step_min
        self.p.t_min = t_min
        self.p.routing_time_min = routing_time_min
        q = np.zeros(self.shape, order="F")
        err, ran, dtg = C.c_double(), C.c_int32(), C.c_double()
        rc = lib().oracle_groundwater_module(C.byref(self.p), C.byref(self.s), _p(h_t), _p(I_t),
                                             _p(d_t), _p(cumulative_recharge), _p(max_gw_depth),
                                             _p(q), C.byref(err), C.byref(ran), C.byref(dtg))
        if rc != 0:
            raise ValueError(f"groundwater module: oracle returned {rc}")
        return q, err.value, bool(ran.value), dtg.value


def idw(x_obs, y_obs, f_obs, x_grid, y_grid, p=2.0, rad=np.inf):
    """idw_function on the grid meshgrid(x_grid, y_grid) (ny x nx)."""
    xo, yo, fo = (np.ascontiguousarray(np.ravel(a), dtype=np.float64) for a in (x_obs, y_obs, f_obs))
    xg = np.ascontiguousarray(np.ravel(x_grid), dtype=np.float64)
    yg = np.ascontiguousarray(np.ravel(y_grid), dtype=np.float64)
    out = np.zeros((yg.size, xg.size), order="F")
    lib().oracle_idw(xo.size, _p(xo), _p(yo), _p(fo), xg.size, _p(xg), yg.size, _p(yg), p, rad,
                     _p(out))
    return out
