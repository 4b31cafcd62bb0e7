"""MATLAB-signature mirrors of the reference's routing functions on the HIP library.

`Local_Inertial_Model_D4` keeps the 45-argument signature and the 14 return
values of HydroPol2D_Functions/Local_Inertial_Model_D4.m:1-5, and
`Inertial_Solver` the one of Inertial_Solver.m:1, so a caller written against
the reference (e.g. the validation driver
run_local_inertial_ca_hydrodynamics_validation.m:406-411) can switch to this
module unchanged.  Arrays are column-major (Fortran order) exactly as MATLAB
holds them; outputs come back as Fortran-ordered numpy arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, abi

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)


class _Handle:
    """A library handle cached per (ny, nx, precision), freed at interpreter exit."""

    _cache: dict = {}

    @classmethod
    def get(cls, ny, nx, precision):
        key = (ny, nx, precision)
        h = cls._cache.get(key)
        if h is None:
            L = _lib.load()
            _lib.require_device()
            ptr = C.c_void_p()
            _lib.check(L.hp2d_create(ny, nx, precision, C.byref(ptr)))
            h = cls._cache[key] = ptr
        return h

    @classmethod
    def release_all(cls):
        L = _lib.load()
        for h in cls._cache.values():
            L.hp2d_destroy(h)
        cls._cache.clear()


def _empty_or(x):
    return x is None or (hasattr(x, "__len__") and len(x) == 0)


def Inertial_Solver(flag_numerical_scheme, q_p, dt, Hf, S, n_sq, dx=None, idx_nan=None):
    """q = Inertial_Solver(flag_numerical_scheme,q_p,dt,Hf,S,n_sq,dx,idx_nan)."""
    q_p = np.asfortranarray(q_p, dtype=np.float64)
    Hf = np.asfortranarray(Hf, dtype=np.float64)
    S = np.asfortranarray(S, dtype=np.float64)
    n_sq = np.asfortranarray(n_sq, dtype=np.float64)
    ny, nx = n_sq.shape[:2]
    n_pages = q_p.size // (ny * nx)
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    out = np.zeros_like(q_p, order="F")
    L = _lib.load()
    _lib.check(L.hp2d_inertial_solver(h, int(flag_numerical_scheme), q_p.ctypes.data_as(_dp),
                                      float(dt), Hf.ctypes.data_as(_dp), S.ctypes.data_as(_dp),
                                      n_sq.ctypes.data_as(_dp), n_pages, out.ctypes.data_as(_dp)),
               handle=h)
    return out


def Local_Inertial_Model_D4(flag_numerical_scheme, reservoir_x, reservoir_y, k1, h1, k2, k3, h2,
                            k4, yds1, xds1, yds2, xds2, flag_reservoir, z, d_tot, d_p,
                            roughness_cell, roughness_squared, cell_area, time_step, Resolution,
                            outlet_index, outlet_type, slope_outlet, row_outlet, col_outlet,
                            d_tolerance, outflow, idx_nan, flag_critical, flag_subgrid, nc, nf,
                            River_Width, River_Depth, Qc_prev, Qf_prev, Qci_prev, Qfi_prev,
                            C_a_prev, Subgrid_Properties, flag_overbanks, flag_inflow,
                            SubgridTables):
    """Local_Inertial_Model_D4.m:1-5 on the gfx950 kernel.

    Returns (qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t,
    I_tot_end_cell, outflow, Hf, Qc, Qf, Qci, Qfi, C_a).  The dtype of z
    selects the precision (float64 = MATLAB double, float32 = flag_single).
    Subgrid, reservoir and river-channel branches raise HP2DError
    (status HP2D_E_UNSUPPORTED), as the build covers the coarse-grid path.
    """
    if flag_reservoir == 1 or flag_subgrid == 1:
        raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED,
                             "reservoir / subgrid branches are outside this build")
    if not _empty_or(River_Width) and np.any(np.asarray(River_Width) > 0):
        outlet_mask = np.asarray(outlet_index, dtype=bool)
        rw = np.array(River_Width, dtype=float, copy=True)
        rw[outlet_mask] = 0   # Local_Inertial_Model_D4.m:18
        if np.any(rw > 0):
            raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED, "river channels are outside this build")
    single = np.asarray(z).dtype == np.float32
    T = np.float32 if single else np.float64
    z = np.asfortranarray(z, dtype=T)
    ny, nx = z.shape
    d_tot = np.asfortranarray(d_tot, dtype=T)
    rc = np.asfortranarray(np.broadcast_to(np.asarray(roughness_cell, dtype=T), (ny, nx)))
    rs = np.asfortranarray(np.broadcast_to(np.asarray(roughness_squared, dtype=T), (ny, nx)))
    of = np.asfortranarray(np.asarray(outflow, dtype=T).reshape(ny, nx, -1, order="F"))
    if of.shape[2] < 3:
        of = np.asfortranarray(np.concatenate([of, np.zeros((ny, nx, 3 - of.shape[2]), T)], axis=2))
    rows = np.ascontiguousarray(np.asarray(row_outlet, dtype=np.int32).ravel())
    cols = np.ascontiguousarray(np.asarray(col_outlet, dtype=np.int32).ravel())
    n_out = int(min(rows.size, cols.size))
    outs = [np.zeros((ny, nx), dtype=T, order="F") for _ in range(7)]
    outflow_new = np.zeros((ny, nx, 3), dtype=T, order="F")
    Hf = np.zeros((ny, nx, 3), dtype=T, order="F")
    L = _lib.load()
    prec = abi.HP2D_FP32 if single else abi.HP2D_FP64
    h = _Handle.get(ny, nx, prec)
    P = _fp if single else _dp
    fn = L.hp2d_local_inertial_d4_f32 if single else L.hp2d_local_inertial_d4
    cast = (lambda a: a.ctypes.data_as(P))
    status = fn(h, int(flag_numerical_scheme), cast(z), cast(d_tot), cast(rc), cast(rs),
                float(cell_area), float(time_step), float(Resolution), int(outlet_type),
                float(slope_outlet), rows.ctypes.data_as(_i32p), cols.ctypes.data_as(_i32p),
                n_out, cast(of), int(flag_critical), *[cast(o) for o in outs], cast(outflow_new),
                cast(Hf))
    _lib.check(status, handle=h)
    qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t, I_tot_end_cell = outs
    C_a = T(Resolution) ** 2
    return (qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t, I_tot_end_cell,
            outflow_new, Hf, 0, 0, 0, 0, C_a)


def CA_Routing(reservoir_x, reservoir_y, k1, h1, k2, k3, h2, k4, yds1, xds1, yds2, xds2,
               flag_reservoir, elevation_cell, d_tot, roughness_cell, cell_area, time_step,
               h_0_cell, Resolution, I_tot_end_cell, outlet_check, outlet_type, slope_outlet,
               row_outlet, col_outlet, idx_nan, flag_critical):
    """CA_Routing.m:1 on the gfx950 kernel (flag_reservoir = 0).

    Returns (qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t,
    I_tot_end_cell).  idx_nan is taken from NaN elevations (the reference's
    preprocessing sets both together); outlet_check is taken from the outlet
    rows/cols.
    """
    if flag_reservoir == 1:
        raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED, "reservoir branch is outside this build")
    z = np.asfortranarray(elevation_cell, dtype=np.float64)
    ny, nx = z.shape
    d_tot = np.asfortranarray(d_tot, dtype=np.float64)
    rc = np.asfortranarray(np.broadcast_to(np.asarray(roughness_cell, dtype=np.float64), (ny, nx)))
    h0 = None
    if not _empty_or(h_0_cell):
        h0 = np.asfortranarray(np.broadcast_to(np.asarray(h_0_cell, dtype=np.float64), (ny, nx)))
    Ip = np.asfortranarray(np.broadcast_to(np.asarray(I_tot_end_cell, dtype=np.float64), (ny, nx)))
    rows = np.ascontiguousarray(np.asarray(row_outlet, dtype=np.int32).ravel())
    cols = np.ascontiguousarray(np.asarray(col_outlet, dtype=np.int32).ravel())
    outs = [np.zeros((ny, nx), order="F") for _ in range(7)]
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    status = L.hp2d_ca_routing(h, p(z), p(d_tot), p(rc), float(cell_area), float(time_step), p(h0),
                               float(Resolution), p(Ip), int(outlet_type), float(slope_outlet),
                               rows.ctypes.data_as(_i32p), cols.ctypes.data_as(_i32p),
                               int(min(rows.size, cols.size)), int(flag_critical),
                               *[p(o) for o in outs])
    _lib.check(status, handle=h)
    return tuple(outs)


def _grid(a, ny, nx):
    return np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))


def _pages9(a, ny, nx):
    if _empty_or(a):
        return np.zeros((ny, nx, 9), order="F")
    return np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx, 9)))


def _outlets(row_outlet, col_outlet):
    rows = np.ascontiguousarray(np.asarray(row_outlet, dtype=np.int32).ravel())
    cols = np.ascontiguousarray(np.asarray(col_outlet, dtype=np.int32).ravel())
    return rows, cols, int(min(rows.size, cols.size))


def Bates_Inertial_8D(reservoir_x, reservoir_y, k1, h1, k2, k3, h2, k4, yds1, xds1, yds2, xds2,
                      flag_reservoir, z, d_tot, d_p, roughness_cell, cell_area, time_step,
                      Resolution, outlet_index, outlet_type, slope_outlet, row_outlet, col_outlet,
                      d_tolerance, outflow, idx_nan, flag_critical):
    """Bates_Inertial_8D.m:2 on the gfx950 kernel (flag_reservoir = 0).

    Returns (qout_left, qout_right, qout_up, qout_down, outlet_flow, qout_ne,
    qout_se, qout_sw, qout_nw, d_t, I_tot_end_cell, outflow, Hf): outflow and
    Hf are (ny, nx, 9) page stacks; outflow (the argument) is the previous
    call's page stack [mm/h] or empty/0 for none.  idx_nan is taken from NaN
    elevations; outlet_index from the outlet rows/cols.
    """
    if flag_reservoir == 1:
        raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED, "reservoir branch is outside this build")
    z = np.asfortranarray(z, dtype=np.float64)
    ny, nx = z.shape
    rows, cols, n_out = _outlets(row_outlet, col_outlet)
    q = np.zeros((ny, nx, 9), order="F")
    Hf = np.zeros((ny, nx, 9), order="F")
    d_t = np.zeros((ny, nx), order="F")
    I_tot = np.zeros((ny, nx), order="F")
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp))
    status = L.hp2d_bates_inertial_8d(
        h, p(z), p(_grid(d_tot, ny, nx)), p(_grid(d_p, ny, nx)), p(_grid(roughness_cell, ny, nx)),
        float(cell_area), float(time_step), float(Resolution), int(outlet_type),
        float(slope_outlet), rows.ctypes.data_as(_i32p), cols.ctypes.data_as(_i32p), n_out,
        p(_pages9(outflow, ny, nx)), int(flag_critical), p(q), p(d_t), p(I_tot), p(Hf))
    _lib.check(status, handle=h)
    return tuple(q[:, :, k].copy(order="F") for k in range(9)) + (d_t, I_tot, q, Hf)


def CA_Routing_8D(reservoir_x, reservoir_y, k1, h1, k2, k3, h2, k4, yds1, xds1, yds2, xds2,
                  flag_reservoir, elevation_cell, d_tot, roughness_cell, cell_area, time_step,
                  h_0_cell, Resolution, I_tot_end_cell, outlet_index, outlet_type, slope_outlet,
                  row_outlet, col_outlet, idx_nan, flag_critical, d_tolerance):
    """CA_Routing_8D.m:2 on the gfx950 kernel (flag_reservoir = 0).

    Returns (qout_left, qout_right, qout_up, qout_down, outlet_flow, qout_ne,
    qout_se, qout_sw, qout_nw, d_t, I_tot_end_cell).  idx_nan is taken from
    NaN elevations; outlet_index from the outlet rows/cols.
    """
    if flag_reservoir == 1:
        raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED, "reservoir branch is outside this build")
    z = np.asfortranarray(elevation_cell, dtype=np.float64)
    ny, nx = z.shape
    rows, cols, n_out = _outlets(row_outlet, col_outlet)
    h0 = None if _empty_or(h_0_cell) else _grid(h_0_cell, ny, nx)
    q = np.zeros((ny, nx, 9), order="F")
    d_t = np.zeros((ny, nx), order="F")
    I_new = np.zeros((ny, nx), order="F")
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    status = L.hp2d_ca_routing_8d(
        h, p(z), p(_grid(d_tot, ny, nx)), p(_grid(roughness_cell, ny, nx)), float(cell_area),
        float(time_step), p(h0), float(Resolution), p(_grid(I_tot_end_cell, ny, nx)),
        int(outlet_type), float(slope_outlet), rows.ctypes.data_as(_i32p),
        cols.ctypes.data_as(_i32p), n_out, int(flag_critical), float(d_tolerance), p(q), p(d_t),
        p(I_new))
    _lib.check(status, handle=h)
    return tuple(q[:, :, k].copy(order="F") for k in range(9)) + (d_t, I_new)


def outflows_to_inflows(outflow, time_step=0.0, d_t=None):
    """The inflow gather of HydroPol2D_Main_While.m:631-660 on the GPU.

    outflow: (ny, nx, 5) pages (left, right, up, down, outlet) or (ny, nx, 9)
    with the diagonal pages; returns qin_t [mm/h] and, when d_t is given,
    d_t + qin_t * time_step / 60 (the cellular-automata depth update).
    """
    q = np.asfortranarray(np.asarray(outflow, dtype=np.float64))
    ny, nx, k = q.shape
    qin = np.zeros((ny, nx), order="F")
    dt_new = None if d_t is None else _grid(d_t, ny, nx).copy(order="F")
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    status = L.hp2d_inflow_gather(h, q.ctypes.data_as(_dp), int(k), float(time_step),
                                  qin.ctypes.data_as(_dp),
                                  dt_new.ctypes.data_as(_dp) if dt_new is not None else _dp())
    _lib.check(status, handle=h)
    return qin, dt_new


def _mask_u8(a, ny, nx):
    if a is None or np.size(a) == 0:
        return None
    m = np.asarray(a)
    if m.size == 1:
        m = np.broadcast_to(m, (ny, nx))
    return np.asfortranarray(m.reshape(ny, nx, order="F").astype(bool).astype(np.uint8))


def Full_Momentum_Model_D4(flag_numerical_scheme, reservoir_x, reservoir_y, k1, h1, k2, k3, h2, k4,
                           yds1, xds1, yds2, xds2, flag_reservoir, z, d_tot, d_p, roughness_cell,
                           roughness_squared, cell_area, time_step, Resolution, outlet_index,
                           outlet_type, slope_outlet, row_outlet, col_outlet, d_tolerance, outflow,
                           idx_nan, flag_critical, flag_subgrid, nc, nf, River_Width, River_Depth,
                           Qc_prev, Qf_prev, Qci_prev, Qfi_prev, C_a_prev, Subgrid_Properties,
                           flag_overbanks, flag_inflow, SubgridTables, Outlet_Properties=None):
    """Full_Momentum_Model_D4.m:1 on the gfx950 kernels (ordinary grid,
    flag_reservoir = 0).

    Returns (qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t,
    I_tot_end_cell, outflow, Hf, Qc, Qf, Qci, Qfi, C_a).  Outlet_Properties is
    None or a dict / object with flag_HR_full_momentum and the face_right,
    face_left, face_up, face_down masks.
    """
    if flag_reservoir == 1:
        raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED, "reservoir branch is outside this build")
    if flag_subgrid == 1 and flag_overbanks != 1 and SubgridTables is not None:
        raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED,
                             "the lookup-subgrid storage branch of the full-momentum model is "
                             "outside this build")
    z = np.asfortranarray(z, dtype=np.float64)
    ny, nx = z.shape
    g = (lambda a: None if a is None or np.size(a) == 0 else _grid(a, ny, nx))
    ca = np.asarray(cell_area, dtype=np.float64)
    ca_grid = None if ca.size == 1 else _grid(ca, ny, nx)
    prev, n_prev = None, 0
    if outflow is not None and np.size(outflow) > 0:
        prev = np.asfortranarray(np.asarray(outflow, dtype=np.float64))
        if prev.ndim == 3 and prev.shape[2] >= 2:   # MATLAB: a single page is a 2-D matrix
            n_prev = prev.shape[2]
        else:
            prev = None

    def prop(name):
        if Outlet_Properties is None:
            return None
        if isinstance(Outlet_Properties, dict):
            return Outlet_Properties.get(name)
        return getattr(Outlet_Properties, name, None)
    faces = [_mask_u8(prop(k), ny, nx) for k in ("face_right", "face_left", "face_up", "face_down")]
    flag_hr = int(prop("flag_HR_full_momentum") or 0)
    rows, cols, n_out = _outlets(row_outlet, col_outlet)
    oi, nan = _mask_u8(outlet_index, ny, nx), _mask_u8(idx_nan, ny, nx)
    out = np.zeros((ny, nx, 5), order="F")
    Hf = np.zeros((ny, nx, 3), order="F")
    d_t, I_tot, C_a = (np.zeros((ny, nx), order="F") for _ in range(3))
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    u = (lambda a: a.ctypes.data_as(C.POINTER(C.c_uint8)) if a is not None
         else C.POINTER(C.c_uint8)())
    st = L.hp2d_full_momentum_d4(
        h, p(z), p(_grid(d_tot, ny, nx)), p(_grid(roughness_cell, ny, nx)), p(g(roughness_squared)),
        float(ca.ravel()[0]), p(ca_grid), float(time_step), float(Resolution), u(oi),
        int(outlet_type), float(np.asarray(slope_outlet, dtype=np.float64).ravel()[0]),
        rows.ctypes.data_as(_i32p), cols.ctypes.data_as(_i32p), n_out, float(d_tolerance),
        p(prev), n_prev, u(nan), flag_hr, *[u(f) for f in faces], p(out), p(d_t), p(I_tot), p(Hf),
        p(C_a))
    _lib.check(st, handle=h)
    ms = out[:, :, :2]
    qout_left = np.asfortranarray(-np.concatenate([np.zeros((ny, 1)), ms[:, :-1, 0]], axis=1))
    qout_down = np.asfortranarray(-np.concatenate([ms[1:, :, 1], np.zeros((1, nx))], axis=0))
    return (qout_left, ms[:, :, 0].copy(order="F"), ms[:, :, 1].copy(order="F"), qout_down,
            out[:, :, 2].copy(order="F"), d_t, I_tot, out, Hf, 0, 0, 0, 0, C_a)
