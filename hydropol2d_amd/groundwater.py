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
                           perimeter, precision="fp64"):
    """Boussinesq_2D_explicit.m:1 on the gfx950 kernels (FP64 arrays).

    Returns (h_t, u_x, u_y, q_exf, q_river, error).  The river arguments and
    the perimeter are accepted and ignored, as in the reference.  The number
    of sub-steps taken is available as `Boussinesq_2D_explicit.last_substeps`.
    precision="fp32" takes the sub-steps as FP32 runs take them (two passes
    that recompute the face fluxes, reciprocal multiplies); the default keeps
    the reference's operation order.
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
    hdl = _Handle.get(ny, nx, abi.HP2D_FP32 if precision == "fp32" else abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    pm = (lambda a: a.ctypes.data_as(_u8p) if a is not None else _u8p())
    status = L.hp2d_boussinesq_2d(hdl, float(dt), float(dx), float(dy), p(h), p(z0), p(Sy), p(R),
                                  p(K), float(Courant), p(hs), pm(cm), pm(dm), p(hd),
                                  *[p(o) for o in outs], C.byref(err), C.byref(nsub))
    _lib.check(status, handle=hdl)
    Boussinesq_2D_explicit.last_substeps = nsub.value
    return outs[0], outs[1], outs[2], outs[3], outs[4], err.value


Boussinesq_2D_explicit.last_substeps = 0


class _GwParams(C.Structure):
    """hp2d_gw_params of include/hp2d.h."""
    _fields_ = [("ny", C.c_int32), ("nx", C.c_int32), ("time_step_min", C.c_double),
                ("resolution", C.c_double), ("t_min", C.c_double), ("routing_time_min", C.c_double),
                ("flag_baseflow", C.c_int32), ("flag_async", C.c_int32),
                ("target_dt_min", C.c_double), ("min_dt_min", C.c_double),
                ("max_head_change_m", C.c_double), ("courant", C.c_double), ("l_vg", C.c_double)] + \
               [(k, _dp) for k in ("elevation", "soil_depth", "theta_sat", "theta_r", "n_vg",
                                   "alpha_vg", "ksat_mmh", "ksat_gw_mmh", "Sy")] + [("domain", _u8p)]


class _GwState(C.Structure):
    """hp2d_gw_state of include/hp2d.h (GW_States)."""
    _fields_ = [("pending_net", _dp), ("pending_recharge", _dp), ("pending_capillary", _dp),
                ("elapsed_s", C.c_double), ("total_recharge_m3", C.c_double),
                ("total_capillary_m3", C.c_double), ("total_exfiltration_m3", C.c_double),
                ("last_dt_s", C.c_double), ("n_updates", C.c_int32)]


class Groundwater_Module:
    """Groundwater_Module.m:37-404 (flag_groundwater_modeling = 1, no layered
    soil) on the gfx950 kernels.

    The reference runs it as a script on Main_While's workspace; here the
    GW_States struct (initialize_groundwater_scheduler.m:38-66: pending maps,
    elapsed time, totals, update count) lives in the object and the workspace
    arrays are passed to `step`.  Soil and aquifer maps are fixed at
    construction: elevation [m], soil depth [m], theta_sat / theta_r, n_vg,
    alpha_vg [1/m], ksat [mm/h] of the soil and of the aquifer, specific
    yield; NaN elevation marks cells outside the domain.
    """

    def __init__(self, elevation, soil_depth, theta_sat, theta_r, n_vg, alpha_vg, ksat_mmh,
                 ksat_gw_mmh, Sy, resolution, domain=None, flag_baseflow=1, flag_async=1,
                 target_dt_min=1440.0, min_dt_min=1.0, max_head_change_m=0.25, courant=0.25,
                 l_vg=0.5):
        z = np.asfortranarray(np.asarray(elevation, dtype=np.float64))
        self.shape = z.shape

        def grid(a):
            return np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), z.shape))

        names = ("elevation", "soil_depth", "theta_sat", "theta_r", "n_vg", "alpha_vg",
                 "ksat_mmh", "ksat_gw_mmh", "Sy")
        vals = (z, soil_depth, theta_sat, theta_r, n_vg, alpha_vg, ksat_mmh, ksat_gw_mmh, Sy)
        self._arr = {k: grid(v) for k, v in zip(names, vals)}
        self._dom = None
        if domain is not None and np.size(domain):
            self._dom = np.asfortranarray(np.broadcast_to(np.asarray(domain, dtype=bool),
                                                          z.shape).astype(np.uint8))
        self._p = _GwParams(z.shape[0], z.shape[1], 0.0, float(resolution), 0.0, 0.0,
                            int(flag_baseflow), int(flag_async), float(target_dt_min),
                            float(min_dt_min), float(max_head_change_m), float(courant), float(l_vg),
                            *[self._arr[k].ctypes.data_as(_dp) for k in names],
                            self._dom.ctypes.data_as(_u8p) if self._dom is not None else _u8p())
        nan = np.isnan(z)
        self.pending = {k: np.where(nan, np.nan, 0.0).copy(order="F")
                        for k in ("net", "recharge", "capillary")}
        self._s = _GwState(*[self.pending[k].ctypes.data_as(_dp)
                             for k in ("net", "recharge", "capillary")], 0.0, 0.0, 0.0, 0.0, 0.0, 0)

    @property
    def state(self):
        """The scalar GW_States fields."""
        s = self._s
        return dict(elapsed_s=s.elapsed_s, total_recharge_m3=s.total_recharge_m3,
                    total_capillary_m3=s.total_capillary_m3,
                    total_exfiltration_m3=s.total_exfiltration_m3, last_dt_s=s.last_dt_s,
                    n_updates=s.n_updates)

    def step(self, time_step_min, t_min, routing_time_min, h_t, I_t, d_t,
             cumulative_recharge=None, max_gw_depth=None):
        """One pass of the module at Main_While time t_min [min] after a step of
        time_step_min.  h_t [m], I_t [mm], d_t [mm], cumulative_recharge [mm] and
        max_gw_depth [m] (Fortran-ordered float64 ny x nx arrays) are updated in
        place.  Returns (q_exf [m/s], error [m3] = errors(5), ran)."""
        arrs = [h_t, I_t, d_t, cumulative_recharge, max_gw_depth]
        for a in arrs:
            if a is not None and (a.dtype != np.float64 or a.shape != self.shape
                                  or not a.flags.f_contiguous):
                raise ValueError("groundwater module: state arrays must be Fortran float64 "
                                 f"{self.shape}")
        self._p.time_step_min = float(time_step_min)
        self._p.t_min = float(t_min)
        self._p.routing_time_min = float(routing_time_min)
        q = np.zeros(self.shape, order="F")
        err, ran = C.c_double(), C.c_int32()
        L = _lib.load()
        hdl = _Handle.get(self.shape[0], self.shape[1], abi.HP2D_FP64)
        p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
        status = L.hp2d_groundwater_module(hdl, C.byref(self._p), C.byref(self._s), *[p(a) for a in arrs],
                                           p(q), C.byref(err), C.byref(ran))
        _lib.check(status, handle=hdl)
        return q, err.value, bool(ran.value)
