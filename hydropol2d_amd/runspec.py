"""Host-side description of one HydroPol2D run (flags, running_control, rasters).

`RunSpec` is the Python mirror of the workspace structs that
HydroPol2D_Main_While.m reads (flags, running_control, Elevation_Properties,
LULC_Properties, Soil_Properties, Rainfall_Parameters, WQ/GW states).  It
converts itself into the plain-C `hp2d_run_config` / `hp2d_run_inputs` pair of
include/hp2d.h.  No compute happens here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi


def _f64(a):
    return None if a is None else np.asfortranarray(np.asarray(a, dtype=np.float64))


def _u8(a):
    return None if a is None else np.asfortranarray(np.asarray(a, dtype=np.uint8))


def _ptr(a, ctype):
    if a is None:
        return C.POINTER(ctype)()
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class GreenAmpt:
    """Class-indexed Green-Ampt soil parameters (HydroPol2D_preprocessing.m:2249-2253)."""

    ksat_mmh: np.ndarray
    psi_mm: np.ndarray
    dtheta: np.ndarray
    theta_sat_minus_r: np.ndarray
    soil_class: Optional[np.ndarray] = None   # ny x nx uint8, None = class 0
    lu_m: Optional[np.ndarray] = None         # per class, NaN = derived from ksat
    I0_mm: Optional[np.ndarray] = None        # ny x nx initial infiltrated depth
    I_min_mm: float = 1e-3                    # min_soil_moisture (preprocessing.m:2386)


@dataclass
class VanGenuchten:
    """The mainline Infiltration_Module (Infiltration_Module.m:182-473,
    flag_infiltration = 1 there): van Genuchten / Mualem surface-interface
    capacity with a Darcy driving head, limited by the supply and by the
    unsaturated zone between the ground and the water table.

    Per soil class: ksat_mmh [mm/h], theta_sat, theta_r, n_vg (> 1), alpha_vg
    [1/m].  h0: water-table head [m] (ny x nx); soil_depth: soil above the
    aquifer base [m] (grid or scalar).  With a `Groundwater` module in the run,
    the head follows its updates and its soil parameters must be these.
    impervious: idx_imp (ny x nx, 1 = no infiltration).  Ltop, dh_max,
    Ksurf_weight, l_vg: Soil_Properties' scalars (the reference's defaults)."""

    ksat_mmh: np.ndarray
    theta_sat: np.ndarray
    theta_r: np.ndarray
    n_vg: np.ndarray
    alpha_vg: np.ndarray
    h0: np.ndarray
    soil_depth: object
    soil_class: Optional[np.ndarray] = None
    impervious: Optional[np.ndarray] = None
    I0_mm: Optional[np.ndarray] = None
    I_min_mm: float = 0.0
    l_vg: float = 0.5
    Ltop: float = 0.05
    dh_max: float = 1.0
    Ksurf_weight: float = 0.5


@dataclass
class WaterQuality:
    """Class-indexed build-up / wash-off parameters (preprocessing.m:2159-2174)."""

    C1: np.ndarray
    C2: np.ndarray
    C3: np.ndarray
    C4: np.ndarray
    lulc_class: Optional[np.ndarray] = None
    ADD_days: float = 7.0
    flag_wq_model: int = 0
    Bmin: float = 0.0
    Bmax: float = 1e30
    min_Bt: float = 0.0


@dataclass
class Groundwater:
    """Groundwater_Module (vadose bucket on I_t, Boussinesq baseflow) run every
    `every_steps` accepted surface steps over the interval since its last call.

    h0: initial water-table head [m]; soil_depth: soil above the aquifer base
    [m] (grid or scalar); ksat_mmh: aquifer conductivity [mm/h]; Sy: specific
    yield; theta_sat / theta_r / n_vg / alpha_vg [1/m]: van Genuchten
    parameters per soil class (the Green-Ampt classes carry the soil Ks)."""

    h0: np.ndarray
    soil_depth: object
    theta_sat: np.ndarray
    theta_r: np.ndarray
    n_vg: np.ndarray
    alpha_vg: np.ndarray
    ksat_mmh: float = 36.0
    Sy: float = 0.2
    every_steps: int = 10
    courant: float = 0.25
    l_vg: float = 0.5
    exact: bool = False  # FP32 runs: the FP64 runs' four-pass sub-steps


@dataclass
class Inflow:
    """Inflow hydrographs (update_spatial_BC.m:362-398): end-stamped times
    [min] shared by the gauges, rates [mm/h over a gauge's cells] as an
    n_times x n_gauges array, and the gauge cells, either as an index map
    (ny x nx, 1..n_gauges, 0 = none) or as the reference's ny x nx x n_gauges
    stack of 0/1 masks (Wshed_Properties.inflow_cells; gauges may not share a
    cell)."""

    time_min: np.ndarray
    rate_mmh: np.ndarray
    cells: np.ndarray

    def gauge_map(self):
        c = np.asarray(self.cells)
        if c.ndim == 3:
            from .forcing import inflow_cells_to_gauge_map
            return inflow_cells_to_gauge_map(c)
        return np.asfortranarray(c.astype(np.uint8))


@dataclass
class RunSpec:
    elevation: np.ndarray
    roughness: np.ndarray
    resolution: float
    outlet_mask: Optional[np.ndarray] = None
    d0: Optional[np.ndarray] = None
    rain_time_min: np.ndarray = field(default_factory=lambda: np.array([1e9]))
    rain_intensity_mmh: np.ndarray = field(default_factory=lambda: np.array([0.0]))
    rain_pattern: Optional[np.ndarray] = None
    precision: str = "fp64"
    routing: str = "inertial"
    flag_numerical_scheme: int = 1
    flag_critical: int = 0
    outlet_type: int = 2
    slope_outlet: float = 0.0
    routing_time_min: float = 60.0
    min_time_step_s: float = 0.01
    max_time_step_s: float = 10.0
    time_step_increments_s: float = 0.001
    time_step_change_s: float = 1e-4
    alfa_min: float = 0.7
    alfa_max: float = 0.7
    flag_timestep: int = 1
    flag_adaptive_timestepping: int = 1
    depth_tolerance_mm: float = 1.0
    record_time_hydrographs_min: float = 1.0
    report_every_percent: float = 5.0
    factor_reduction: float = 2.0
    infiltration: Optional[object] = None   # GreenAmpt or VanGenuchten
    water_quality: Optional[WaterQuality] = None
    groundwater: Optional[Groundwater] = None
    inflow: Optional[Inflow] = None
    # rainfall as rate maps [mm/h] (ny x nx x K) ending at rain_map_time_min [min];
    # replaces the lumped series and rain_pattern when set
    rain_maps_mmh: Optional[np.ndarray] = None
    rain_map_time_min: Optional[np.ndarray] = None
    # Human_Instability_Module mode 1 (a person of weight [kg], body widths [m],
    # friction coefficient and drag coefficient; water density, gravity)
    human_instability: Optional[dict] = None
    # map stacks every record_time_maps_min of model time (0: off)
    record_time_maps_min: float = 0.0
    # routing = "ca": LULC_Properties.h_0 [mm] per cell (CA_Routing.m:110-113; None = 0)
    ca_h0_mm: Optional[np.ndarray] = None

    @property
    def shape(self):
        return np.asarray(self.elevation).shape

    def n_hydro_records(self) -> int:
        return int(np.floor(self.routing_time_min / self.record_time_hydrographs_min + 1e-12)) + 1

    def to_c(self):
        """Return (RunConfig, RunInputs, keepalive) for the C ABI."""
        ny, nx = self.shape
        keep = []

        def f(a):
            arr = _f64(a)
            if arr is not None:
                keep.append(arr)
            return arr

        def u(a):
            arr = _u8(a)
            if arr is not None:
                keep.append(arr)
            return arr

        cfg = abi.RunConfig()
        cfg.ny, cfg.nx = ny, nx
        cfg.precision = abi.HP2D_FP64 if self.precision == "fp64" else abi.HP2D_FP32
        cfg.routing = abi.HP2D_ROUTING_INERTIAL if self.routing == "inertial" else abi.HP2D_ROUTING_CA
        cfg.flag_numerical_scheme = self.flag_numerical_scheme
        cfg.flag_critical = self.flag_critical
        cfg.outlet_type = self.outlet_type
        cfg.slope_outlet = self.slope_outlet
        cfg.resolution = self.resolution
        cfg.routing_time_min = self.routing_time_min
        cfg.min_time_step_s = self.min_time_step_s
        cfg.max_time_step_s = self.max_time_step_s
        cfg.time_step_increments_s = self.time_step_increments_s
        cfg.time_step_change_s = self.time_step_change_s
        cfg.alfa_min = self.alfa_min
        cfg.alfa_max = self.alfa_max
        cfg.flag_timestep = self.flag_timestep
        cfg.flag_adaptive_timestepping = self.flag_adaptive_timestepping
        cfg.depth_tolerance_mm = self.depth_tolerance_mm
        cfg.record_time_hydrographs_min = self.record_time_hydrographs_min
        cfg.report_every_percent = self.report_every_percent
        cfg.factor_reduction = self.factor_reduction
        vgi = self.infiltration if isinstance(self.infiltration, VanGenuchten) else None
        cfg.flag_infiltration = 0 if self.infiltration is None else (2 if vgi is not None else 1)
        cfg.flag_waterquality = 1 if self.water_quality is not None else 0
        cfg.flag_wq_model = self.water_quality.flag_wq_model if self.water_quality else 0
        cfg.flag_groundwater = 1 if self.groundwater is not None else 0
        cfg.gw_every_steps = self.groundwater.every_steps if self.groundwater else 0
        cfg.gw_courant = self.groundwater.courant if self.groundwater else 0.25
        cfg.gw_exact = int(bool(self.groundwater.exact)) if self.groundwater else 0

        ins = abi.RunInputs()
        ins.elevation = _ptr(f(self.elevation), C.c_double)
        ins.roughness = _ptr(f(self.roughness), C.c_double)
        ins.d0 = _ptr(f(self.d0), C.c_double)
        ins.outlet_mask = _ptr(u(self.outlet_mask), C.c_uint8)
        rt = f(np.atleast_1d(self.rain_time_min))
        ri = f(np.atleast_1d(self.rain_intensity_mmh))
        ins.rain_time_min = _ptr(rt, C.c_double)
        ins.rain_intensity_mmh = _ptr(ri, C.c_double)
        ins.n_rain = int(min(rt.size, ri.size))
        ins.rain_pattern = _ptr(f(self.rain_pattern), C.c_double)
        ins.ca_h0_mm = _ptr(f(self.ca_h0_mm), C.c_double)
        ga = self.infiltration if vgi is None else None
        if ga is not None:
            cfg.n_soil_classes = int(np.atleast_1d(ga.ksat_mmh).size)
            ins.soil_class = _ptr(u(ga.soil_class), C.c_uint8)
            ins.ksat_mmh = _ptr(f(np.atleast_1d(ga.ksat_mmh)), C.c_double)
            ins.psi_mm = _ptr(f(np.atleast_1d(ga.psi_mm)), C.c_double)
            ins.dtheta = _ptr(f(np.atleast_1d(ga.dtheta)), C.c_double)
            ins.theta_sat_minus_r = _ptr(f(np.atleast_1d(ga.theta_sat_minus_r)), C.c_double)
            ins.lu_m = _ptr(f(None if ga.lu_m is None else np.atleast_1d(ga.lu_m)), C.c_double)
            ins.I0_mm = _ptr(f(ga.I0_mm), C.c_double)
            ins.I_min_mm = ga.I_min_mm
        wq = self.water_quality
        if wq is not None:
            cfg.n_lulc_classes = int(np.atleast_1d(wq.C1).size)
            ins.lulc_class = _ptr(u(wq.lulc_class), C.c_uint8)
            ins.wq_C1 = _ptr(f(np.atleast_1d(wq.C1)), C.c_double)
            ins.wq_C2 = _ptr(f(np.atleast_1d(wq.C2)), C.c_double)
            ins.wq_C3 = _ptr(f(np.atleast_1d(wq.C3)), C.c_double)
            ins.wq_C4 = _ptr(f(np.atleast_1d(wq.C4)), C.c_double)
            ins.wq_ADD_days = wq.ADD_days
            ins.wq_Bmin = wq.Bmin
            ins.wq_Bmax = wq.Bmax
            ins.wq_min_Bt = wq.min_Bt
        hi = self.human_instability
        if hi is not None:
            cfg.flag_human_instability = 1
            cfg.hi_weight_kg = hi.get("weight_person", 70.0)
            cfg.hi_width1_m = hi.get("width1_person", 0.35)
            cfg.hi_width2_m = hi.get("width2_person", 0.25)
            cfg.hi_mu = hi.get("mu", 0.5)
            cfg.hi_Cd = hi.get("Cd", 1.1)
            cfg.hi_ro_water = hi.get("ro_water", 1000.0)
            cfg.hi_gravity = hi.get("gravity", 9.81)
        cfg.record_time_maps_min = self.record_time_maps_min
        fl = self.inflow
        if fl is not None:
            rate = np.asarray(fl.rate_mmh, dtype=np.float64)
            if rate.ndim == 1:
                rate = rate[:, None]
            cfg.flag_inflow = 1
            cfg.n_inflow_gauges = int(rate.shape[1])
            ins.inflow_time_min = _ptr(f(np.atleast_1d(fl.time_min)), C.c_double)
            ins.inflow_rate_mmh = _ptr(f(rate), C.c_double)
            ins.n_inflow_times = int(rate.shape[0])
            ins.inflow_gauge = _ptr(u(fl.gauge_map()), C.c_uint8)
        if self.rain_maps_mmh is not None:
            maps = np.asarray(self.rain_maps_mmh, dtype=np.float64)
            if maps.ndim == 2:
                maps = maps[:, :, None]
            cfg.n_rain_maps = int(maps.shape[2])
            ins.rain_maps_mmh = _ptr(f(maps), C.c_double)
            ins.rain_map_time_min = _ptr(f(np.atleast_1d(self.rain_map_time_min)), C.c_double)
        if vgi is not None:
            ny, nx = np.shape(self.elevation)
            cfg.n_soil_classes = int(np.atleast_1d(vgi.ksat_mmh).size)
            cfg.vg_Ltop, cfg.vg_dh_max = vgi.Ltop, vgi.dh_max
            cfg.vg_Ksurf_weight = vgi.Ksurf_weight
            ins.soil_class = _ptr(u(vgi.soil_class), C.c_uint8)
            ins.ksat_mmh = _ptr(f(np.atleast_1d(vgi.ksat_mmh)), C.c_double)
            ins.I0_mm = _ptr(f(vgi.I0_mm), C.c_double)
            ins.I_min_mm = vgi.I_min_mm
            ins.impervious = _ptr(u(vgi.impervious), C.c_uint8)
            ins.gw_h0 = _ptr(f(vgi.h0), C.c_double)
            sd = np.broadcast_to(np.asarray(vgi.soil_depth, dtype=np.float64), (ny, nx))
            ins.gw_soil_depth = _ptr(f(sd), C.c_double)
            ins.gw_theta_sat = _ptr(f(np.atleast_1d(vgi.theta_sat)), C.c_double)
            ins.gw_theta_r = _ptr(f(np.atleast_1d(vgi.theta_r)), C.c_double)
            ins.gw_n_vg = _ptr(f(np.atleast_1d(vgi.n_vg)), C.c_double)
            ins.gw_alpha_vg = _ptr(f(np.atleast_1d(vgi.alpha_vg)), C.c_double)
            ins.gw_l_vg = vgi.l_vg
        gw = self.groundwater
        if gw is not None and vgi is not None:
            # one set of soil parameters per run: the module's must be the closure's
            same = (np.array_equal(np.asarray(gw.h0, float), np.asarray(vgi.h0, float))
                    and np.array_equal(np.broadcast_to(np.asarray(gw.soil_depth, float), (ny, nx)),
                                       np.broadcast_to(np.asarray(vgi.soil_depth, float), (ny, nx)))
                    and all(np.array_equal(np.atleast_1d(getattr(gw, k)).astype(float),
                                           np.atleast_1d(getattr(vgi, k)).astype(float))
                            for k in ("theta_sat", "theta_r", "n_vg", "alpha_vg"))
                    and gw.l_vg == vgi.l_vg)
            if not same:
                raise ValueError("Groundwater and VanGenuchten must carry the same head, soil depth "
                                 "and van Genuchten parameters")
        if gw is not None:
            ny, nx = np.shape(self.elevation)
            ins.gw_h0 = _ptr(f(gw.h0), C.c_double)
            sd = np.broadcast_to(np.asarray(gw.soil_depth, dtype=np.float64), (ny, nx))
            ins.gw_soil_depth = _ptr(f(sd), C.c_double)
            ins.gw_ksat_mmh = gw.ksat_mmh
            ins.gw_Sy = gw.Sy
            ins.gw_theta_sat = _ptr(f(np.atleast_1d(gw.theta_sat)), C.c_double)
            ins.gw_theta_r = _ptr(f(np.atleast_1d(gw.theta_r)), C.c_double)
            ins.gw_n_vg = _ptr(f(np.atleast_1d(gw.n_vg)), C.c_double)
            ins.gw_alpha_vg = _ptr(f(np.atleast_1d(gw.alpha_vg)), C.c_double)
            ins.gw_l_vg = gw.l_vg
        return cfg, ins, keep
