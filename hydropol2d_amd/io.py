"""On-disk formats either side of the step loop.

* `GRIDobj` reads a GeoTIFF raster as HydroPol2D_preprocessing.m:192-194 does
  (TopoToolbox GRIDobj: Z ny x nx with the first image row on top, the cell
  size, the georeference, nodata as NaN), and `crop` trims its NaN border
  (HydroPol2D_preprocessing.m:202-204); the decoding is libhp2d's
  (hp2d_geotiff_read).
* `write_mat` writes MATLAB Level 5 MAT-files (hp2d_mat_write) with nested
  structs from dotted names; `save_map_hydro` and `save_map_stacks` write the
  Temporary_Files/save_map_hydro_N.mat chunks of 12 map records that
  save_output_maps.m:575-580 writes (the Maps struct), and
  `write_time_series_outputs` the current_time_series_outputs.mat snapshot
  (TempTS, save_output_maps.m:620-830) that recover_hydropol2d_from_temp.m and
  post_processing.m read.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from .abi import MAP_FIELDS

_dp = C.POINTER(C.c_double)

MAT_DOUBLE, MAT_SINGLE, MAT_INT32, MAT_UINT8, MAT_CHAR = range(5)


class MatVar(C.Structure):
    _fields_ = [("name", C.c_char_p), ("mclass", C.c_int32), ("ndims", C.c_int32),
                ("dims", C.POINTER(C.c_int64)), ("data", C.c_void_p)]


@dataclass
class GRIDobj:
    """A raster read from a GeoTIFF (the fields HydroPol2D reads of TopoToolbox's
    GRIDobj): Z (ny x nx), cellsize, size, the upper-left corner, EPSG code."""

    Z: np.ndarray
    cellsize: float
    x_ul: float
    y_ul: float
    dy: float
    epsg: int
    nodata: float

    def __init__(self, path=None, Z=None, cellsize=1.0, x_ul=0.0, y_ul=0.0, dy=None, epsg=0,
                 nodata=np.nan):
        if path is not None:
            L = _lib.load()
            ny, nx = C.c_int32(), C.c_int32()
            geo = np.zeros(6)
            p = os.fsencode(path)
            _lib.check(L.hp2d_geotiff_info(p, C.byref(ny), C.byref(nx), geo.ctypes.data_as(_dp)))
            Z = np.zeros((ny.value, nx.value), order="F")
            _lib.check(L.hp2d_geotiff_read(p, Z.ctypes.data_as(_dp)))
            x_ul, y_ul, cellsize, dy, nodata, epsg = geo
        self.Z = np.asfortranarray(Z, dtype=np.float64)
        self.cellsize = float(cellsize)
        self.x_ul, self.y_ul = float(x_ul), float(y_ul)
        self.dy = float(cellsize if dy is None else dy)
        self.epsg = int(epsg)
        self.nodata = float(nodata)

    @property
    def size(self):
        return self.Z.shape

    def x(self):
        """Cell-centre eastings of the columns."""
        return self.x_ul + self.cellsize * (np.arange(self.Z.shape[1]) + 0.5)

    def y(self):
        """Cell-centre northings of the rows (the first row on top)."""
        return self.y_ul - self.dy * (np.arange(self.Z.shape[0]) + 0.5)

    def crop(self):
        """The raster without its outer rows and columns that hold NaN only."""
        ok = ~np.isnan(self.Z)
        rows, cols = np.flatnonzero(ok.any(axis=1)), np.flatnonzero(ok.any(axis=0))
        if rows.size == 0:
            return GRIDobj(Z=self.Z[:0, :0], cellsize=self.cellsize, x_ul=self.x_ul,
                           y_ul=self.y_ul, dy=self.dy, epsg=self.epsg, nodata=self.nodata)
        r0, r1, c0, c1 = rows[0], rows[-1] + 1, cols[0], cols[-1] + 1
        return GRIDobj(Z=self.Z[r0:r1, c0:c1], cellsize=self.cellsize,
                       x_ul=self.x_ul + c0 * self.cellsize, y_ul=self.y_ul - r0 * self.dy,
                       dy=self.dy, epsg=self.epsg, nodata=self.nodata)


def write_mat(path, variables):
    """Write variables {dotted name: array or str} to a Level 5 MAT-file;
    names sharing a prefix become fields of nested structs."""
    keep = []
    vs = (MatVar * len(variables))()
    for i, (name, val) in enumerate(variables.items()):
        if isinstance(val, str):
            a = np.frombuffer(val.encode("ascii"), dtype=np.uint8).reshape(1, -1)
            cls = MAT_CHAR
        else:
            a = np.asarray(val)
            if a.dtype == np.float32:
                cls = MAT_SINGLE
            elif a.dtype == np.int32:
                cls = MAT_INT32
            elif a.dtype == np.uint8 or a.dtype == bool:
                a, cls = a.astype(np.uint8), MAT_UINT8
            else:
                a, cls = a.astype(np.float64), MAT_DOUBLE
            if a.ndim == 0:
                a = a.reshape(1, 1)
            elif a.ndim == 1:
                a = a.reshape(-1, 1)
        a = np.asfortranarray(a)
        dims = np.array(a.shape, dtype=np.int64)
        keep += [a, dims]
        vs[i].name = name.encode("ascii")
        vs[i].mclass = cls
        vs[i].ndims = dims.size
        vs[i].dims = dims.ctypes.data_as(C.POINTER(C.c_int64))
        vs[i].data = a.ctypes.data_as(C.c_void_p)
    L = _lib.load()
    _lib.check(L.hp2d_mat_write(os.fsencode(path), len(variables), vs))
    del keep


# where each map-stack field lives in the Maps struct (save_output_maps.m:148-258)
MAPS_PATH = {"d": "Maps.Hydro.d", "velocity": "Maps.Hydro.velocity",
             "hazard_dv": "Maps.Hydro.hazard_dv", "risk": "Maps.Hydro.risk",
             "I_t": "Maps.Hydro.I_t", "GWdepth": "Maps.Hydro.GWdepth_save",
             "Pol_Conc_Map": "Maps.WQ_States.Pol_Conc_Map",
             "Pol_Mass_Map": "Maps.WQ_States.Pol_Mass_Map",
             "Pol_Load_Map": "Maps.WQ_States.Pol_Load_Map"}


def save_map_hydro(path, records, first, count, single=True):
    """One save_map_hydro_N.mat: records first .. first + count - 1 of every
    recorded field (a dict as Simulation.map_records returns) in the Maps struct."""
    out = {}
    for name in MAP_FIELDS:
        if name in records:
            a = records[name][:, :, first:first + count]
            out[MAPS_PATH[name]] = a.astype(np.float32) if single else a
    write_mat(path, out)


def save_map_stacks(folder, records, chunk=12, single=True):
    """The save_map_hydro_N.mat files of a run's records: records 1..12 in
    file 1, 13..24 in file 2, ... (save_output_maps.m:574-580: a file is written
    each time the 12-slot stack fills; a partial last chunk is written too, as
    the run's end leaves it in the Maps structure).  Returns the paths."""
    os.makedirs(folder, exist_ok=True)
    n = len(records["t_min"])
    paths = []
    for store, first in enumerate(range(0, n, chunk), start=1):
        p = os.path.join(folder, f"save_map_hydro_{store}.mat")
        save_map_hydro(p, records, first, min(chunk, n - first), single=single)
        paths.append(p)
    return paths


def write_time_series_outputs(folder, sim):
    """current_time_series_outputs.mat (TempTS) of a Simulation: the outlet
    hydrograph and outlet depth at the hydrograph records, the runoff and the
    run's metadata (save_output_maps.m:620-830)."""
    rep = sim.report()
    hy = sim.hydrograph()
    n = max(1, int(hy["n_hydro"]))
    out = {
        "TempTS.meta.k": float(rep["steps"] + 1),
        "TempTS.meta.t_current_min": float(rep["t_min"]),
        "TempTS.meta.actual_record_hydrograph": float(n),
        "TempTS.meta.record_time_hydrographs_min": float(sim.spec.record_time_hydrographs_min),
        "TempTS.meta.description": "Temporary HydroPol2D time-series snapshot. "
                                   "Raster map stacks are not stored here.",
        "TempTS.time.time_hydrograph_min": hy["hydro_t"][:n],
        "TempTS.outlet.hydrograph_m3_s": hy["hydro_q"][:n],
        "TempTS.outlet.depth_outlet_mm": hy["hydro_depth"][:n],
        "TempTS.diagnostics.outlet_runoff_volume_mm": float(rep["outlet_runoff_volume_mm"]),
    }
    if sim.spec.water_quality is not None:
        out["TempTS.waterquality.mass_outlet"] = float(rep["wq_mass_outlet_mg"])
        out["TempTS.waterquality.volume_outlet"] = float(rep["wq_vol_outlet"])
    os.makedirs(folder, exist_ok=True)
    path = os.path.join(folder, "current_time_series_outputs.mat")
    write_mat(path, out)
    return path
