"""MATLAB-signature mirror of the reference's Boussinesq groundwater solver.

`Boussinesq_2D_explicit` keeps the 18-argument signature and the 6 return
values of HydroPol2D_Functions/Boussinesq_2D_explicit.m:1, as called at
Groundwater_Module.m:229-247 and by
Validation/Groundwater/Hillslope_Analytical/run_hillslope_analytical_validation.m:81,159.
Arrays are column-major (Fortran order); scalars broadcast to the grid.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, abi
from .local_inertial import _Handle

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def Boussinesq_2D_explicit(dt, dx, dy, h, z0, Sy, R, K, river_mask, K_river, h_river, z_river,
                           Courant, h_soil, catchment_mask, dirichlet_mask, h_dirichlet,
                           perimeter):
    """Boussinesq_2D_explicit.m:1 on the gfx950 kernels (FP64).

    Returns (h_t, u_x, u_y, q_exf, q_river, error).  The river arguments and
    the perimeter are accepted and ignored, as in the reference.  The number
    of sub-steps taken is available as `Boussinesq_2D_explicit.last_substeps`.
    """
    h = np.asfortranarray(h, dtype=np.float64)
    ny, nx = h.shape

    def grid(a):
        return np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))

    def mask(m):
        if m is None or np.size(m) == 0:
            return None
        return np.asfortranarray(np.broadcast_to(np.asarray(m, dtype=bool), (ny, nx)).astype(np.uint8))

    z0, Sy, R, K, hs = grid(z0), grid(Sy), grid(R), grid(K), grid(h_soil)
    cm = mask(catchment_mask)
    dm = mask(dirichlet_mask)
    if dm is None or h_dirichlet is None or np.size(h_dirichlet) == 0:
        dm, hd = None, None      # :64-67: an empty mask or value disables Dirichlet cells
    else:
        hd = grid(h_dirichlet)
    outs = [np.zeros((ny, nx), order="F") for _ in range(5)]  # h_t u_x u_y q_exf q_river
    err, nsub = C.c_double(), C.c_int32()
    L = _lib.load()
    hdl = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    pm = (lambda a: a.ctypes.data_as(_u8p) if a is not None else _u8p())
    status = L.hp2d_boussinesq_2d(hdl, float(dt), float(dx), float(dy), p(h), p(z0), p(Sy), p(R),
                                  p(K), float(Courant), p(hs), pm(cm), pm(dm), p(hd),
                                  *[p(o) for o in outs], C.byref(err), C.byref(nsub))
    _lib.check(status, handle=hdl)
    Boussinesq_2D_explicit.last_substeps = nsub.value
    return outs[0], outs[1], outs[2], outs[3], outs[4], err.value


Boussinesq_2D_explicit.last_substeps = 0
