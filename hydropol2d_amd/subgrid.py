"""The lookup-table subgrid local-inertial model on the gfx950 kernels, with
the MATLAB signatures of the reference:

* `Subgrid_Properties_Lookup` (Subgrid_Properties_Lookup.m:1): the SFINCS-style
  cell, velocity-point and boundary tables of a coarse grid, built from a fine
  DEM on the GPU (hp2d_subgrid_tables_build);
* `Local_Inertial_Model_D4_Subgrid` (Local_Inertial_Model_D4_Subgrid.m:1): one
  routing step with adaptive internal sub-steps (hp2d_local_inertial_d4_subgrid);
* `hp2d_subgrid_apply_volume_change` (hp2d_subgrid_apply_volume_change.m:1,
  SFINCS-exact branch): a depth change (rainfall) applied through the cell
  volume tables.

`SubgridTables` holds the struct's fields as Fortran-ordered numpy arrays with
MATLAB's names and shapes (u_* are ny x (nx-1) [x nr_levels], v_* are
(ny-1) x nx, north_*/south_*/west_*/east_* are ny x nx).  The tables live on
the GPU beside the handle of their grid; a table object made or changed on the
host is uploaded when a step first uses it.
"""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np

from . import _lib, abi
from .local_inertial import _Handle

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)

CELL_FIELDS = ("z_zmin", "z_zmax", "z_volmax")
_tokens = itertools.count(1)
_resident = {}   # handle address -> token of the tables resident on it
CELL_LEVEL_FIELDS = ("z_level", "z_volume", "z_area")
FACE_SETS = ("u", "v") + abi.SUBGRID_SIDES


def _face_shape(prefix, ny, nx):
    if prefix == "u":
        return ny, max(nx - 1, 1)
    if prefix == "v":
        return max(ny - 1, 1), nx
    return ny, nx


class SubgridTables:
    """SubgridTables of an ny x nx coarse grid (MATLAB field names)."""

    def __init__(self, ny, nx, nr_levels, cell_area, fill=np.nan):
        self.ny, self.nx, self.nr_levels = int(ny), int(nx), int(nr_levels)
        self.cell_area = float(cell_area)
        self.sfincs_exact = True
        f = {}
        for k in CELL_FIELDS:
            f[k] = np.full((ny, nx), fill, order="F")
        for k in CELL_LEVEL_FIELDS:
            f[k] = np.full((ny, nx, nr_levels), fill, order="F")
        for s in FACE_SETS:
            sh = _face_shape(s, ny, nx)
            for k in abi.SUBGRID_FACE_FIELDS:
                lev = k in ("havg", "nrep", "pwet")
                f[f"{s}_{k}"] = np.full(sh + ((nr_levels,) if lev else ()), fill, order="F")
        self.fields = f
        self._token = next(_tokens)

    def __getattr__(self, name):
        fields = self.__dict__.get("fields")
        if fields is not None and name in fields:
            return fields[name]
        raise AttributeError(name)

    @property
    def invert_el(self):
        return self.fields["z_zmin"]

    def touch(self):
        """Mark the host arrays as changed (the next step uploads them)."""
        self._token = next(_tokens)

    def c_struct(self):
        """The hp2d_subgrid_tables view of these arrays (they must stay alive)."""
        t = abi.SubgridTablesC()
        t.nr_levels = self.nr_levels
        t.cell_area = self.cell_area
        f = self.fields
        for k in CELL_FIELDS + CELL_LEVEL_FIELDS:
            a = f[k]
            assert a.flags.f_contiguous and a.dtype == np.float64
            setattr(t, k, a.ctypes.data_as(_dp))
        for s in FACE_SETS:
            faces = abi.SubgridFaces(*[f[f"{s}_{k}"].ctypes.data_as(_dp)
                                       for k in abi.SUBGRID_FACE_FIELDS])
            if s in ("u", "v"):
                setattr(t, s, faces)
            else:
                t.side[abi.SUBGRID_SIDES.index(s)] = faces
        return t


def _Z(r):
    return np.asfortranarray(np.asarray(getattr(r, "Z", r), dtype=np.float64))


def Subgrid_Properties_Lookup(DEM_raster, Roughness_raster, Reference_raster, coarse_res,
                              nr_levels=20, nlevels=None, **_ignored):
    """[SubgridTables, invert_el] = Subgrid_Properties_Lookup(DEM_raster,
    Roughness_raster, Reference_raster, coarse_res, 'nr_levels', k).

    The rasters are GRIDobj-like (.Z, .cellsize) or arrays (then DEM_raster
    needs a `cellsize` keyword through a GRIDobj).  Roughness_raster.Z has the
    fine or the coarse (Reference_raster) shape.  'dz', 'maxDepth', 'tol_rel'
    and 'verbose' do not change the tables and are accepted and ignored.
    """
    dem = _Z(DEM_raster)
    cellsize = float(DEM_raster.cellsize)
    rough = _Z(Roughness_raster)
    ny, nx = _Z(Reference_raster).shape
    if rough.shape == dem.shape:
        fine = 1
    elif rough.shape == (ny, nx):
        fine = 0
    else:
        raise _lib.HP2DError(abi.HP2D_E_ARG, "Roughness_raster.Z must match the fine DEM or the "
                                             "coarse Reference_raster")
    nl = int(round(nlevels if nlevels is not None else nr_levels))
    T = SubgridTables(ny, nx, nl, float(coarse_res) ** 2)
    keep = T.c_struct()
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    st = L.hp2d_subgrid_tables_build(h, dem.ctypes.data_as(_dp), dem.shape[0], dem.shape[1],
                                     cellsize, rough.ctypes.data_as(_dp), fine, float(coarse_res),
                                     nl, C.byref(keep))
    _lib.check(st, handle=h)
    _resident[h.value] = T._token
    return T, T.invert_el.copy(order="F")


def _ensure_uploaded(T, h):
    """Make T the tables resident on handle h (uploads when another table set,
    or an older state of T, is there)."""
    if _resident.get(h.value) != T._token:
        keep = T.c_struct()
        _lib.check(_lib.load().hp2d_subgrid_set_tables(h, C.byref(keep)), handle=h)
        _resident[h.value] = T._token


def _grid(a, ny, nx):
    return np.asfortranarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (ny, nx)))


def Local_Inertial_Model_D4_Subgrid(flag_numerical_scheme, reservoir_x, reservoir_y, k1, h1, k2,
                                    k3, h2, k4, yds1, xds1, yds2, xds2, flag_reservoir, z, d_tot,
                                    d_p, roughness_cell, roughness_squared, cell_area, time_step,
                                    Resolution, outlet_index, outlet_type, slope_outlet,
                                    row_outlet, col_outlet, d_tolerance, outflow, idx_nan,
                                    flag_critical, nc, nf, River_Width, River_Depth, Qc_prev,
                                    Qf_prev, Qci_prev, Qfi_prev, C_a_prev, Subgrid_Properties,
                                    flag_inflow, SubgridTables):
    """Local_Inertial_Model_D4_Subgrid.m:1 on the gfx950 kernels
    (flag_reservoir = 0).

    Returns (qout_left, qout_right, qout_up, qout_down, outlet_flow, d_t,
    I_tot_end_cell, outflow, Hf, Qc, Qf, Qci, Qfi, C_a, eta_t, V_t).  The
    arguments the reference leaves unused (:14-16) are accepted and ignored.
    """
    if flag_reservoir == 1:
        raise _lib.HP2DError(abi.HP2D_E_UNSUPPORTED, "reservoir branch is outside this build")
    T = SubgridTables
    ny, nx = T.ny, T.nx
    d = _grid(d_tot, ny, nx)
    prev = None
    n_prev = 0
    if outflow is not None and np.size(outflow) > 0:
        prev = np.asfortranarray(np.asarray(outflow, dtype=np.float64))
        if prev.ndim == 2:
            prev = prev.reshape(ny, nx, 1, order="F")
        n_prev = prev.shape[2]
    mask = None
    if idx_nan is not None and np.size(idx_nan) > 0:
        mask = np.asfortranarray(np.asarray(idx_nan, dtype=bool).reshape(ny, nx, order="F")
                                 .astype(np.uint8))
    rows = np.ascontiguousarray(np.asarray(row_outlet, dtype=np.int32).ravel())
    cols = np.ascontiguousarray(np.asarray(col_outlet, dtype=np.int32).ravel())
    out = np.zeros((ny, nx, 3), order="F")
    Hf = np.zeros((ny, nx, 3), order="F")
    d_t, I_tot, C_a, eta_t, V_t = (np.zeros((ny, nx), order="F") for _ in range(5))
    nsub = C.c_int32()
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    _ensure_uploaded(T, h)
    p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
    st = L.hp2d_local_inertial_d4_subgrid(
        h, p(d), float(time_step), float(Resolution), int(outlet_type),
        float(np.asarray(slope_outlet, dtype=np.float64).ravel()[0]),
        rows.ctypes.data_as(_i32p), cols.ctypes.data_as(_i32p), int(min(rows.size, cols.size)),
        p(prev), n_prev, mask.ctypes.data_as(_u8p) if mask is not None else _u8p(),
        p(out), p(d_t), p(I_tot), p(Hf), p(C_a), p(eta_t), p(V_t), C.byref(nsub))
    _lib.check(st, handle=h)
    ms = out[:, :, :2]
    qout_left = -np.concatenate([np.zeros((ny, 1)), ms[:, :-1, 0]], axis=1)
    qout_right = ms[:, :, 0].copy(order="F")
    qout_up = ms[:, :, 1].copy(order="F")
    qout_down = -np.concatenate([ms[1:, :, 1], np.zeros((1, nx))], axis=0)
    return (np.asfortranarray(qout_left), qout_right, qout_up, np.asfortranarray(qout_down),
            out[:, :, 2].copy(order="F"), d_t, I_tot, out, Hf, 0, 0, 0, 0, C_a, eta_t, V_t)


def hp2d_subgrid_apply_volume_change(drep_old, delta_depth_mm, SubgridTables, cell_area):
    """[drep_new, V_old, V_new, A_new, residual_m3] =
    hp2d_subgrid_apply_volume_change(drep_old, delta_depth_mm, SubgridTables,
    cell_area) (SFINCS-exact branch); drep in m, delta_depth_mm scalar or grid."""
    T = SubgridTables
    ny, nx = T.ny, T.nx
    d0 = _grid(drep_old, ny, nx)
    delta = _grid(delta_depth_mm, ny, nx)
    area = float(np.asarray(cell_area).ravel()[0])
    outs = [np.zeros((ny, nx), order="F") for _ in range(5)]
    L = _lib.load()
    h = _Handle.get(ny, nx, abi.HP2D_FP64)
    _ensure_uploaded(T, h)
    st = L.hp2d_subgrid_apply_volume_change(h, d0.ctypes.data_as(_dp), delta.ctypes.data_as(_dp),
                                            area, *[o.ctypes.data_as(_dp) for o in outs])
    _lib.check(st, handle=h)
    return tuple(outs)


def outlet_sides(row_outlet, col_outlet, ny, nx):
    """outlet_boundary_side (Local_Inertial_Model_D4_Subgrid.m:286-308): 0 north,
    1 south, 2 west, 3 east; an interior outlet takes the nearest side."""
    sides = []
    for r, c in zip(np.asarray(row_outlet).ravel(), np.asarray(col_outlet).ravel()):
        if r <= 1:
            sides.append(0)
        elif r >= ny:
            sides.append(1)
        elif c <= 1:
            sides.append(2)
        elif c >= nx:
            sides.append(3)
        else:
            dist = [r - 1, ny - r, c - 1, nx - c]
            sides.append(int(np.argmin(dist)))
    return np.asarray(sides, dtype=np.int32)
