"""Spatial forcing of the step loop, with the reference's interfaces.

* `Rainfall_Interpolator` / `idw_function` (Rainfall_Interpolator.m:40-52,
  idw_function.m:18-63): inverse-distance weighting of rain gauges onto the
  grid, on the GPU (hp2d_idw_interpolate).
* `spatial_gauge_rain_maps`: the per-interval rate maps that
  update_spatial_BC.m:469-490 interpolates from the gauges' series (the
  gauges with a NaN value in an interval are left out of that interval's map,
  as localGetSpatialGaugeRainCached does, :853-901), ready for
  RunSpec.rain_maps_mmh.
* `inflow_cells_to_gauge_map`: Wshed_Properties.inflow_cells (a stack of 0/1
  masks) as the one-byte gauge index map of the device loop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_dp = C.POINTER(C.c_double)


def _c(a):
    return np.ascontiguousarray(np.ravel(a), dtype=np.float64)


def idw_function(X0, F0, x_grid, y_grid, p=2.0, rad=np.inf):
    """Fint on meshgrid(x_grid, y_grid) from observations X0 = [x, y] (n x 2)
    with values F0 (idw_function.m, L = 2)."""
    X0 = np.asarray(X0, dtype=np.float64).reshape(-1, 2)
    xo, yo, fo = _c(X0[:, 0]), _c(X0[:, 1]), _c(F0)
    xg, yg = _c(x_grid), _c(y_grid)
    out = np.zeros((yg.size, xg.size), order="F")
    L = _lib.load()
    _lib.check(L.hp2d_idw_interpolate(xo.size, xo.ctypes.data_as(_dp), yo.ctypes.data_as(_dp),
                                      fo.ctypes.data_as(_dp), xg.size, xg.ctypes.data_as(_dp),
                                      yg.size, yg.ctypes.data_as(_dp), float(p), float(rad),
                                      out.ctypes.data_as(_dp)))
    return out


def Rainfall_Interpolator(x_coordinate, y_coordinate, var_obs, x_grid, y_grid):
    """[Var_Interpolated, X, Y] = Rainfall_Interpolator(x_coordinate, y_coordinate,
    var_obs, x_grid, y_grid)."""
    X, Y = np.meshgrid(np.ravel(x_grid), np.ravel(y_grid))
    V = idw_function(np.column_stack([np.ravel(x_coordinate), np.ravel(y_coordinate)]),
                     var_obs, x_grid, y_grid)
    return V, X, Y


def spatial_gauge_rain_maps(gauge_rates, x_coordinate, y_coordinate, x_grid, y_grid,
                            nan_mask=None):
    """Rate maps [mm/h] (ny x nx x n_times) of the gauges' series
    gauge_rates (n_times x n_gauges): each interval interpolated from its
    gauges with a finite value, zeros when none has one, NaN on nan_mask."""
    R = np.atleast_2d(np.asarray(gauge_rates, dtype=np.float64))
    xc, yc = np.ravel(x_coordinate), np.ravel(y_coordinate)
    ny, nx = np.size(y_grid), np.size(x_grid)
    maps = np.zeros((ny, nx, R.shape[0]), order="F")
    for k in range(R.shape[0]):
        ok = np.isfinite(R[k])
        if ok.any():
            maps[:, :, k] = Rainfall_Interpolator(xc[ok], yc[ok], R[k, ok], x_grid, y_grid)[0]
        if nan_mask is not None:
            maps[:, :, k][np.asarray(nan_mask, dtype=bool)] = np.nan
    return maps


def inflow_cells_to_gauge_map(inflow_cells):
    """Wshed_Properties.inflow_cells (ny x nx x n_gauges of 0/1) as the gauge
    index map (0 = none, g = gauge g); a cell may belong to one gauge only."""
    c = np.asarray(inflow_cells)
    if c.ndim == 2:
        c = c[:, :, None]
    if np.any(np.sum(c != 0, axis=2) > 1):
        raise ValueError("inflow_cells: a cell belongs to more than one gauge")
    idx = np.zeros(c.shape[:2], dtype=np.uint8)
    for g in range(c.shape[2]):
        idx[c[:, :, g] != 0] = g + 1
    return np.asfortranarray(idx)
