"""ctypes mirrors of the plain-C structs declared in include/hp2d.h.

These are layout definitions only (no compute).  Field order and types must
match include/hp2d.h exactly; tests/test_abi.py checks the sizes against the
compiled library.
"""
import ctypes as C

HP2D_OK = 0
HP2D_E_ARG = 1
HP2D_E_HIP = 2
HP2D_E_NOMEM = 3
HP2D_E_STATE = 4
HP2D_E_UNSUPPORTED = 5
HP2D_E_INSTABILITY = 6
HP2D_E_NO_DEVICE = 7

HP2D_FP64 = 0
HP2D_FP32 = 1
HP2D_ROUTING_INERTIAL = 1
HP2D_ROUTING_CA = 2

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class RunConfig(C.Structure):
    """hp2d_run_config: flags and running_control of one run."""

    _fields_ = [
        ("ny", C.c_int32),
        ("nx", C.c_int32),
        ("precision", C.c_int32),
        ("routing", C.c_int32),
        ("flag_numerical_scheme", C.c_int32),
        ("flag_critical", C.c_int32),
        ("outlet_type", C.c_int32),
        ("slope_outlet", C.c_double),
        ("resolution", C.c_double),
        ("routing_time_min", C.c_double),
        ("min_time_step_s", C.c_double),
        ("max_time_step_s", C.c_double),
        ("time_step_increments_s", C.c_double),
        ("time_step_change_s", C.c_double),
        ("alfa_min", C.c_double),
        ("alfa_max", C.c_double),
        ("flag_timestep", C.c_int32),
        ("flag_adaptive_timestepping", C.c_int32),
        ("depth_tolerance_mm", C.c_double),
        ("record_time_hydrographs_min", C.c_double),
        ("report_every_percent", C.c_double),
        ("factor_reduction", C.c_double),
        ("flag_infiltration", C.c_int32),
        ("flag_waterquality", C.c_int32),
        ("flag_wq_model", C.c_int32),
        ("flag_groundwater", C.c_int32),
        ("gw_every_steps", C.c_int32),
        ("gw_courant", C.c_double),
        ("n_soil_classes", C.c_int32),
        ("n_lulc_classes", C.c_int32),
        ("flag_inflow", C.c_int32),
        ("n_inflow_gauges", C.c_int32),
        ("n_rain_maps", C.c_int32),
        ("flag_human_instability", C.c_int32),
        ("hi_weight_kg", C.c_double), ("hi_width1_m", C.c_double), ("hi_width2_m", C.c_double),
        ("hi_mu", C.c_double), ("hi_Cd", C.c_double), ("hi_ro_water", C.c_double),
        ("hi_gravity", C.c_double),
        ("record_time_maps_min", C.c_double),
        ("gw_exact", C.c_int32),
        ("vg_Ltop", C.c_double), ("vg_dh_max", C.c_double), ("vg_Ksurf_weight", C.c_double),
    ]


MAP_FIELDS = ("d", "velocity", "hazard_dv", "risk", "I_t", "GWdepth", "Pol_Conc_Map",
              "Pol_Mass_Map", "Pol_Load_Map")


class RunInputs(C.Structure):
    """hp2d_run_inputs: host arrays of one run (pointers may be NULL)."""

    _fields_ = [
        ("elevation", _dp),
        ("roughness", _dp),
        ("d0", _dp),
        ("outlet_mask", _u8p),
        ("rain_time_min", _dp),
        ("rain_intensity_mmh", _dp),
        ("n_rain", C.c_int32),
        ("rain_pattern", _dp),
        ("soil_class", _u8p),
        ("ksat_mmh", _dp),
        ("psi_mm", _dp),
        ("dtheta", _dp),
        ("theta_sat_minus_r", _dp),
        ("lu_m", _dp),
        ("I0_mm", _dp),
        ("I_min_mm", C.c_double),
        ("lulc_class", _u8p),
        ("wq_C1", _dp),
        ("wq_C2", _dp),
        ("wq_C3", _dp),
        ("wq_C4", _dp),
        ("wq_ADD_days", C.c_double),
        ("wq_Bmin", C.c_double),
        ("wq_Bmax", C.c_double),
        ("wq_min_Bt", C.c_double),
        ("gw_h0", _dp),
        ("gw_soil_depth", _dp),
        ("gw_ksat_mmh", C.c_double),
        ("gw_Sy", C.c_double),
        ("gw_theta_sat", _dp),
        ("gw_theta_r", _dp),
        ("gw_n_vg", _dp),
        ("gw_alpha_vg", _dp),
        ("gw_l_vg", C.c_double),
        ("inflow_time_min", _dp),
        ("inflow_rate_mmh", _dp),
        ("n_inflow_times", C.c_int32),
        ("inflow_gauge", _u8p),
        ("rain_map_time_min", _dp),
        ("rain_maps_mmh", _dp),
        ("impervious", _u8p),
        ("ca_h0_mm", _dp),
    ]


class RunReport(C.Structure):
    """hp2d_run_report: FP64 ledgers and counters of a run."""

    _fields_ = [
        ("steps", C.c_int64),
        ("rejected_steps", C.c_int64),
        ("t_min", C.c_double),
        ("time_step_s", C.c_double),
        ("max_velocity", C.c_double),
        ("volume_error_m3", C.c_double),
        ("n_reports", C.c_int32),
        ("outflow_volume_m3", C.c_double),
        ("inflow_volume_m3", C.c_double),
        ("outlet_runoff_volume_mm", C.c_double),
        ("cum_rain_m3", C.c_double),
        ("cum_infiltration_m3", C.c_double),
        ("cum_outflow_m3", C.c_double),
        ("storage_m3", C.c_double),
        ("catch_index", C.c_int32),
        ("wq_mass_kg", C.c_double),
        ("wq_mass_lost_kg", C.c_double),
        ("wq_out_conc_mgl", C.c_double),
        ("wq_tmin_s", C.c_double),
        ("wq_mass_outlet_mg", C.c_double),
        ("wq_vol_outlet", C.c_double),
        ("wq_substep_steps", C.c_int64),
        ("gw_storage_m3", C.c_double),
        ("gw_total_recharge_m3", C.c_double),
        ("gw_total_exfiltration_m3", C.c_double),
        ("gw_error_m3", C.c_double),
        ("gw_updates", C.c_int32),
        ("cum_inflow_m3", C.c_double),
        ("n_map_records", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


SUBGRID_FACE_FIELDS = ("zmin", "zmax", "havg", "nrep", "pwet", "navg", "ffit")
SUBGRID_SIDES = ("north", "south", "west", "east")


class SubgridFaces(C.Structure):
    """hp2d_subgrid_faces: one set of velocity-point tables."""

    _fields_ = [(k, _dp) for k in SUBGRID_FACE_FIELDS]


class SubgridTablesC(C.Structure):
    """hp2d_subgrid_tables: the SFINCS-style SubgridTables of a coarse grid."""

    _fields_ = ([("nr_levels", C.c_int32), ("cell_area", C.c_double)]
                + [(k, _dp) for k in ("z_zmin", "z_zmax", "z_volmax", "z_level", "z_volume",
                                      "z_area")]
                + [("u", SubgridFaces), ("v", SubgridFaces), ("side", SubgridFaces * 4)])
