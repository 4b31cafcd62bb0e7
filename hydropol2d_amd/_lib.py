"""Loader of the in-tree HIP library hydropol2d_amd/libhp2d.so (ctypes).

There is no CPU fallback: if the library is missing or no HIP device is
visible, every compute entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhp2d.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p

_lib = None


class HP2DError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"hp2d status {status}: {message}")
        self.status = status


# name -> (restype, argtypes); the list is the whole exported ABI of include/hp2d.h
SIGNATURES = {
    "hp2d_version": (C.c_char_p, []),
    "hp2d_device_count": (C.c_int32, []),
    "hp2d_global_last_error": (C.c_char_p, []),
    "hp2d_create": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "hp2d_destroy": (None, [_vp]),
    "hp2d_last_error": (C.c_char_p, [_vp]),
    "hp2d_inertial_solver": (C.c_int32, [_vp, C.c_int32, _dp, C.c_double, _dp, _dp, _dp,
                                         C.c_int32, _dp]),
    "hp2d_local_inertial_d4": (C.c_int32, [_vp, C.c_int32, _dp, _dp, _dp, _dp, C.c_double,
                                           C.c_double, C.c_double, C.c_int32, C.c_double, _i32p,
                                           _i32p, C.c_int32, _dp, C.c_int32] + [_dp] * 9),
    "hp2d_local_inertial_d4_f32": (C.c_int32, [_vp, C.c_int32, _fp, _fp, _fp, _fp, C.c_float,
                                               C.c_float, C.c_float, C.c_int32, C.c_float, _i32p,
                                               _i32p, C.c_int32, _fp, C.c_int32] + [_fp] * 9),
    "hp2d_ca_routing": (C.c_int32, [_vp, _dp, _dp, _dp, C.c_double, C.c_double, _dp,
                                    C.c_double, _dp, C.c_int32, C.c_double, _i32p, _i32p,
                                    C.c_int32, C.c_int32] + [_dp] * 7),
    "hp2d_build_up_wash_off": (C.c_int32, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                           C.c_double, C.c_double, _u8p, C.c_int32, _dp,
                                           C.c_double, C.c_double, C.c_double] + [_dp] * 7),
    "hp2d_bates_inertial_8d": (C.c_int32, [_vp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                           C.c_double, C.c_int32, C.c_double, _i32p, _i32p,
                                           C.c_int32, _dp, C.c_int32, _dp, _dp, _dp, _dp]),
    "hp2d_ca_routing_8d": (C.c_int32, [_vp, _dp, _dp, _dp, C.c_double, C.c_double, _dp,
                                       C.c_double, _dp, C.c_int32, C.c_double, _i32p, _i32p,
                                       C.c_int32, C.c_int32, C.c_double, _dp, _dp, _dp]),
    "hp2d_inflow_gather": (C.c_int32, [_vp, _dp, C.c_int32, C.c_double, _dp, _dp]),
    "hp2d_last_kernel_time": (C.c_int32, [_vp, _dp]),
    "hp2d_subgrid_tables_build": (C.c_int32, [_vp, _dp, C.c_int32, C.c_int32, C.c_double, _dp,
                                              C.c_int32, C.c_double, C.c_int32, _vp]),
    "hp2d_subgrid_set_tables": (C.c_int32, [_vp, _vp]),
    "hp2d_full_momentum_d4": (C.c_int32, [_vp, _dp, _dp, _dp, _dp, C.c_double, _dp, C.c_double,
                                          C.c_double, _u8p, C.c_int32, C.c_double, _i32p, _i32p,
                                          C.c_int32, C.c_double, _dp, C.c_int32, _u8p, C.c_int32,
                                          _u8p, _u8p, _u8p, _u8p] + [_dp] * 5),
    "hp2d_local_inertial_d4_subgrid": (C.c_int32, [_vp, _dp, C.c_double, C.c_double, C.c_int32,
                                                   C.c_double, _i32p, _i32p, C.c_int32, _dp,
                                                   C.c_int32, _u8p] + [_dp] * 7 + [_i32p]),
    "hp2d_subgrid_apply_volume_change": (C.c_int32, [_vp, _dp, _dp, C.c_double] + [_dp] * 5),
    "hp2d_build_up_wash_off_8d": (C.c_int32, [_vp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                              _u8p, C.c_int32, _dp, C.c_double, C.c_double,
                                              C.c_double] + [_dp] * 7),
    "hp2d_boussinesq_2d": (C.c_int32, [_vp, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
                                       _dp, _dp, C.c_double, _dp, _u8p, _u8p, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _dp, _i32p]),
    "hp2d_sim_record_time_to_peak": (C.c_int32, [_vp]),
    "hp2d_sim_get_time_to_peak": (C.c_int32, [_vp, _dp]),
    "hp2d_ensemble_create": (C.c_int32, [C.c_int32, C.POINTER(abi.RunConfig),
                                         C.POINTER(abi.RunInputs), C.c_int32, C.POINTER(_vp)]),
    "hp2d_ensemble_destroy": (None, [_vp]),
    "hp2d_ensemble_last_error": (C.c_char_p, [_vp]),
    "hp2d_ensemble_run": (C.c_int32, [_vp, C.c_int64, C.c_int32, _i32p]),
    "hp2d_ensemble_member": (C.c_int32, [_vp, C.c_int32, C.POINTER(_vp)]),
    "hp2d_ensemble_record_time_to_peak": (C.c_int32, [_vp]),
    "hp2d_ensemble_elapsed": (C.c_int32, [_vp, _dp]),
    "hp2d_infiltration_module": (C.c_int32, [_vp, C.c_double, C.c_double, C.c_double] + [_dp] * 8
                                 + [_u8p, C.c_double, C.c_double, C.c_double, C.c_double]
                                 + [_dp] * 9),
    "hp2d_groundwater_module": (C.c_int32, [_vp, _vp, _vp] + [_dp] * 7 + [_i32p]),
    "hp2d_fill_sinks": (C.c_int32, [C.c_int32, C.c_int32, _dp, _u8p, C.c_double, _dp]),
    "hp2d_flow_accumulation": (C.c_int32, [C.c_int32, C.c_int32, _dp, _dp]),
    "hp2d_nn_conv3x3_fwd": (C.c_int32, [C.c_int32] * 5 + [_vp, _vp, _vp, C.c_int32, _vp, _vp, _vp,
                                         _vp]),
    "hp2d_nn_conv3x3_dgrad": (C.c_int32, [C.c_int32] * 5 + [_vp] * 5),
    "hp2d_sim_get_max_risk": (C.c_int32, [_vp, _dp]),
    "hp2d_sim_get_map_records": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, _dp,
                                             C.POINTER(C.c_int32)]),
    "hp2d_geotiff_info": (C.c_int32, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      _dp]),
    "hp2d_geotiff_read": (C.c_int32, [C.c_char_p, _dp]),
    "hp2d_mat_write": (C.c_int32, [C.c_char_p, C.c_int32, _vp]),
    "hp2d_idw_interpolate": (C.c_int32, [C.c_int32, _dp, _dp, _dp, C.c_int32, _dp, C.c_int32,
                                         _dp, C.c_double, C.c_double, _dp]),
    "hp2d_nn_conv3x3_dgrad_noalloc": (C.c_int32, [C.c_int32] * 5 + [_vp] * 4),
    "hp2d_nn_conv3x3_wgrad": (C.c_int32, [C.c_int32] * 5 + [_vp] * 5),
    "hp2d_nn_relu_mask": (C.c_int32, [_vp, _vp, C.c_int64, _vp]),
    "hp2d_nn_add": (C.c_int32, [_vp, _vp, _vp, C.c_int64, _vp]),
    "hp2d_nn_avgpool2": (C.c_int32, [_vp] + [C.c_int32] * 4 + [_vp, C.c_int32, _vp]),
    "hp2d_nn_upsample2": (C.c_int32, [_vp] + [C.c_int32] * 4 + [_vp, C.c_int32, _vp]),
    "hp2d_nn_head": (C.c_int32, [_vp, C.c_int64, C.c_int32, _vp, _vp, C.c_int32, _vp, _vp, _vp,
                                 _vp, _vp, _vp]),
    "hp2d_nn_mse": (C.c_int32, [_vp, _vp, _vp, C.c_int64, C.c_int32, C.c_float, _vp, _vp, _vp]),
    "hp2d_nn_adam": (C.c_int32, [_vp, _vp, _vp, _vp, C.c_int64, C.c_float, C.c_float, C.c_float,
                                 C.c_float, C.c_int32, _vp, _vp]),
    "hp2d_nn_to_bf16": (C.c_int32, [_vp, C.c_int64, _vp, _vp]),
    "hp2d_sim_create": (C.c_int32, [C.POINTER(abi.RunConfig), C.POINTER(abi.RunInputs),
                                    C.POINTER(_vp)]),
    "hp2d_sim_destroy": (None, [_vp]),
    "hp2d_sim_last_error": (C.c_char_p, [_vp]),
    "hp2d_sim_run": (C.c_int32, [_vp, C.c_int64, _i32p]),
    "hp2d_sim_run_fixed": (C.c_int32, [_vp, C.c_int64, C.c_int32]),
    "hp2d_sim_report": (C.c_int32, [_vp, C.POINTER(abi.RunReport)]),
    "hp2d_sim_get_state": (C.c_int32, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "hp2d_sim_get_wq_maps": (C.c_int32, [_vp, _dp, _dp]),
    "hp2d_sim_get_hydrograph": (C.c_int32, [_vp, _dp, _dp, _dp, C.c_int32, _i32p]),
    "hp2d_sim_get_dt_history": (C.c_int32, [_vp, _dp, C.c_int32, _i32p]),
    "hp2d_debug_inject_failure": (C.c_int32, [_vp, C.c_int64, C.c_int32]),
    "hp2d_sim_kernel_time": (C.c_int32, [_vp, _dp, _dp, C.POINTER(C.c_int64)]),
    "hp2d_sim_autotune": (C.c_int32, [_vp, _i32p, _i32p, _dp]),
    "hp2d_sim_launch_shape": (C.c_int32, [_vp, _i32p, _i32p, _i32p]),
    "hp2d_sim_set_launch_shape": (C.c_int32, [_vp, C.c_int32, C.c_int32, C.c_int32]),
}


def load(path: str = LIB_PATH):
    """Load libhp2d.so and bind every symbol of include/hp2d.h (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HP2DError(-1, f"{path} is not built: run `python -m hydropol2d_amd.build`")
    # PyTorch ships its own HIP runtime under the same soname as the system's
    # (libamdhip64.so.7): whichever loads first serves the whole process, and
    # PyTorch finds no device on another build of it.  Loading PyTorch first
    # makes libhp2d bind to PyTorch's runtime, so both can share the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)  # AttributeError = missing export: fail loudly
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def device_count() -> int:
    return int(load().hp2d_device_count())


def require_device():
    if device_count() < 1:
        raise HP2DError(abi.HP2D_E_NO_DEVICE, "no HIP device visible; the hp2d library has no "
                                              "CPU fallback")


def check(status, handle=None, sim=None):
    if status != abi.HP2D_OK:
        L = load()
        if sim is not None:
            msg = L.hp2d_sim_last_error(sim)
        elif handle is not None:
            msg = L.hp2d_last_error(handle)
        else:
            msg = L.hp2d_global_last_error()
        raise HP2DError(status, (msg or b"").decode(errors="replace"))
