"""Mirror of the reference's mainline infiltration step on the HIP library.

`Infiltration_Module` restates the inputs and outputs of the script
HydroPol2D_Functions/Infiltration_Module.m:101-583 (flag_infiltration = 1, no
subgrid, no layered soil), which Hydrological_Model.m runs on the workspace:
it reads depths.d_t, Soil_Properties.{I_t, Soil_Depth, theta_sat, theta_r,
n_vg, alpha_vg, ksat, l_vg, Ltop, dh_max, Ksurf_weight}, BC_States.h_t,
elevation, LULC_Properties.idx_imp, time_step and C_a, and writes depths.d_t,
Soil_Properties.{I_t, I_p}, Hydro_States.f, inf_volume, errors(4) and
cumulative_infiltration.  Arrays are column-major (Fortran order).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, abi
from .local_inertial import _Handle

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def Infiltration_Module(d_t, I_t, h_t, elevation, Soil_Depth, theta_sat, theta_r, n_vg, alpha_vg,
                        ksat, idx_imp, time_step, Resolution, C_a=None, l_vg=0.5, Ltop=0.05,
                        dh_max=1.0, Ksurf_weight=0.5, cumulative_infiltration=None):
    """Returns a dict with d_t, I_t, I_p, f [mm/h], inf_volume [m3], errors4 [m3]
    and cumulative_infiltration (when given)."""
    d = np.asfortranarray(d_t, dtype=np.float64)
    ny, nx = d.shape

    def grid(a):
        return np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))

    ins = [grid(a) for a in (elevation, h_t, Soil_Depth, theta_sat, theta_r, n_vg, alpha_vg, ksat)]
    I = grid(I_t)
    imp = None
    if idx_imp is not None and np.size(idx_imp) > 0:
        imp = np.asfortranarray(np.broadcast_to(np.asarray(idx_imp, dtype=bool),
                                                (ny, nx)).astype(np.uint8))
    out = {k: np.zeros((ny, nx), order="F") for k in ("d_t", "I_t", "I_p", "f")}
    cum = None if cumulative_infiltration is None else grid(cumulative_infiltration).copy(order="F")
    vol, dS = C.c_double(), C.c_double()
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    Ca = float(Resolution) ** 2 if C_a is None else float(C_a)
    st = L.hp2d_infiltration_module(h, float(time_step), float(Resolution), Ca, *[p(a) for a in ins],
                                    imp.ctypes.data_as(_u8p) if imp is not None else _u8p(),
                                    float(l_vg), float(Ltop), float(dh_max), float(Ksurf_weight),
                                    p(d), p(I), p(out["d_t"]), p(out["I_t"]), p(out["I_p"]),
                                    p(out["f"]), p(cum), C.byref(vol), C.byref(dS))
    _lib.check(st, handle=h)
    out.update(inf_volume=vol.value, errors4=dS.value, cumulative_infiltration=cum)
    return out
