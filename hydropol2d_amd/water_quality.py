"""MATLAB-signature mirror of the reference's pollutant routine on the HIP library.

`build_up_wash_off` keeps the 20-argument signature and the 7 return values of
HydroPol2D_Functions/build_up_wash_off.m:1, as called at
HydroPol2D_Main_While.m:685 and by the validation driver
Validation/WaterQuality/BuildUpWashOff_MassBalance/run_water_quality_washoff_validation.m:262-276.
Arrays are column-major (Fortran order) as MATLAB holds them.
`initial_buildup` is the build-up initialisation of
HydroPol2D_preprocessing.m:2170-2174.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, abi
from .local_inertial import _Handle

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def initial_buildup(C_1, C_2, ADD, cell_area, idx_not_nan=None):
    """B_t [kg] = C_1*(1 - e^(-C_2*ADD)) [kg/ha] * cell_area/1e4, zero outside the domain."""
    B0 = np.asarray(C_1, dtype=np.float64) * (1.0 - np.power(np.e, -np.asarray(C_2, dtype=np.float64) * ADD))
    B = B0 * cell_area / 1e4
    if idx_not_nan is not None:
        B = B * np.asarray(idx_not_nan, dtype=np.float64)
    return B


def build_up_wash_off(C_3, C_4, qout_left_t, qout_right_t, qout_up_t, qout_down_t, outlet_flow,
                      B_t, time_step, nx_max, ny_max, cell_area, outlet_index, idx_nan_5,
                      flag_wq_model, mass_lost, Tot_Washed, Bmin, Bmax, min_Bt):
    """build_up_wash_off.m:1 on the gfx950 kernels.

    Returns (B_t, P_conc, Out_Conc, tmin_wq, tot_W_out, mass_lost, Tot_Washed).
    Scalars broadcast to the (ny_max, nx_max) grid; outlet_index is accepted
    and ignored, as in the reference; idx_nan_5 may be (ny, nx, 5), (ny, nx)
    or empty.
    """
    ny, nx = int(ny_max), int(nx_max)

    def grid(a):
        a = np.asarray(a, dtype=np.float64)
        if a.size == ny * nx:
            return np.asfortranarray(a.reshape(ny, nx, order="F"))
        return np.asfortranarray(np.broadcast_to(a, (ny, nx)))

    C3, C4 = grid(C_3), grid(C_4)
    qs = [grid(q) for q in (qout_left_t, qout_right_t, qout_up_t, qout_down_t, outlet_flow)]
    B = grid(B_t)
    tw = None if Tot_Washed is None else grid(Tot_Washed)
    mask = None
    if idx_nan_5 is not None and np.size(idx_nan_5) > 0:
        m = np.asarray(idx_nan_5).astype(bool)
        if m.ndim == 3:
            m = m[:, :, 0]
        mask = np.asfortranarray(m.reshape(ny, nx, order="F").astype(np.uint8))
    outs = [np.zeros((ny, nx), order="F") for _ in range(4)]  # B_t, P_conc, tot_W_out, Tot_Washed
    oc, tm, ml = C.c_double(), C.c_double(), C.c_double(float(mass_lost))
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    status = L.hp2d_build_up_wash_off(
        h, p(C3), p(C4), *[p(q) for q in qs], p(B), float(time_step), float(cell_area),
        mask.ctypes.data_as(_u8p) if mask is not None else _u8p(), int(flag_wq_model), p(tw),
        float(Bmin), float(Bmax), float(min_Bt), p(outs[0]), p(outs[1]), C.byref(oc),
        C.byref(tm), p(outs[2]), C.byref(ml), p(outs[3]))
    _lib.check(status, handle=h)
    return outs[0], outs[1], oc.value, tm.value, outs[2], ml.value, outs[3]


def build_up_wash_off_8D(C_3, C_4, qout_left_t, qout_right_t, qout_up_t, qout_down_t, outlet_flow,
                         qout_ne_t, qout_se_t, qout_sw_t, qout_nw_t, B_t, time_step, nx_max, ny_max,
                         cell_area, outlet_index, idx_nan_5, flag_wq_model, mass_lost, Tot_Washed,
                         Bmin, Bmax, min_Bt):
    """build_up_wash_off_8D.m:1 on the gfx950 kernels.

    Returns (B_t, P_conc, Out_Conc, tmin_wq, tot_W_out, mass_lost, Tot_Washed).
    Scalars broadcast to the (ny_max, nx_max) grid; outlet_index is accepted
    and ignored, as in the reference; idx_nan_5 may be (ny, nx, k), (ny, nx)
    or empty.
    """
    ny, nx = int(ny_max), int(nx_max)

    def grid(a):
        a = np.asarray(a, dtype=np.float64)
        if a.size == ny * nx:
            return np.asfortranarray(a.reshape(ny, nx, order="F"))
        return np.asfortranarray(np.broadcast_to(a, (ny, nx)))

    q = np.zeros((ny, nx, 9), order="F")
    for k, page in enumerate((qout_left_t, qout_right_t, qout_up_t, qout_down_t, outlet_flow,
                              qout_ne_t, qout_se_t, qout_sw_t, qout_nw_t)):
        q[:, :, k] = grid(page)
    C3, C4, B = grid(C_3), grid(C_4), grid(B_t)
    tw = None if Tot_Washed is None else grid(Tot_Washed)
    mask = None
    if idx_nan_5 is not None and np.size(idx_nan_5) > 0:
        m = np.asarray(idx_nan_5).astype(bool)
        if m.ndim == 3:
            m = m[:, :, 0]
        mask = np.asfortranarray(m.reshape(ny, nx, order="F").astype(np.uint8))
    outs = [np.zeros((ny, nx), order="F") for _ in range(4)]  # B_t, P_conc, tot_W_out, Tot_Washed
    oc, tm, ml = C.c_double(), C.c_double(), C.c_double(float(mass_lost))
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    status = L.hp2d_build_up_wash_off_8d(
        h, p(C3), p(C4), p(q), p(B), float(time_step), float(cell_area),
        mask.ctypes.data_as(_u8p) if mask is not None else _u8p(), int(flag_wq_model), p(tw),
        float(Bmin), float(Bmax), float(min_Bt), p(outs[0]), p(outs[1]), C.byref(oc),
        C.byref(tm), p(outs[2]), C.byref(ml), p(outs[3]))
    _lib.check(status, handle=h)
    return outs[0], outs[1], oc.value, tm.value, outs[2], ml.value, outs[3]
