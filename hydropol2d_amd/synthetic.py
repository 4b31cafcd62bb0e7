"""Seeded synthetic inputs of the benchmark configurations.

All generation happens in FP64 on the host; the arrays are column-major
(ny x nx, MATLAB layout).  Each builder returns a RunSpec.
"""
from __future__ import annotations

import numpy as np

from .runspec import GreenAmpt, Groundwater, RunSpec, WaterQuality


def tilted_plane_c1(ny=256, nx=256, dx=5.0, seed=1, precision="fp64", routing_time_min=60.0):
    """Tilted plane: z = 0.01 x + 0.002 y + N(0, 0.02^2), n = 0.02, 50 mm/h for 30 min,
    Green-Ampt Ks = 5 mm/h, psi = 110 mm, dtheta = 0.3, critical-depth outlet at the
    lowest cell, alpha = 0.7, 1 h."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[None, :] * dx            # along columns
    y = np.arange(ny)[:, None] * dx            # along rows
    z = 0.01 * x + 0.002 * y + rng.normal(0.0, 0.02, (ny, nx))
    z = np.asfortranarray(z)
    outlet = np.zeros((ny, nx), dtype=np.uint8, order="F")
    # HydroPol2D_preprocessing.m:1708-1711: outlet = cell(s) with the minimum elevation
    outlet[np.unravel_index(np.argmin(z), z.shape)] = 1
    ga = GreenAmpt(ksat_mmh=np.array([5.0]), psi_mm=np.array([110.0]), dtheta=np.array([0.3]),
                   theta_sat_minus_r=np.array([0.4]), I_min_mm=1e-3)
    return RunSpec(
        elevation=z, roughness=np.full((ny, nx), 0.02, order="F"), resolution=dx,
        outlet_mask=outlet, rain_time_min=np.array([30.0, 60.0]),
        rain_intensity_mmh=np.array([50.0, 0.0]), precision=precision, outlet_type=2,
        routing_time_min=routing_time_min, min_time_step_s=0.01, max_time_step_s=10.0,
        time_step_increments_s=0.001, time_step_change_s=1e-4, alfa_min=0.7, alfa_max=0.7,
        flag_timestep=1, flag_adaptive_timestepping=1, depth_tolerance_mm=1.0,
        record_time_hydrographs_min=1.0, infiltration=ga)


def fill_sinks(z, seed_mask=None, epsilon=1e-4):
    """Depression-filled copy of z (Priority-Flood+epsilon in libhp2d, host code)."""
    import ctypes as C
    from . import _lib
    L = _lib.load()
    z = np.asfortranarray(z, dtype=np.float64)
    out = np.empty_like(z, order="F")
    m = None
    if seed_mask is not None:
        m = np.asfortranarray(np.asarray(seed_mask, dtype=np.uint8))
    st = L.hp2d_fill_sinks(z.shape[0], z.shape[1], z.ctypes.data_as(C.POINTER(C.c_double)),
                           m.ctypes.data_as(C.POINTER(C.c_uint8)) if m is not None
                           else C.POINTER(C.c_uint8)(), float(epsilon),
                           out.ctypes.data_as(C.POINTER(C.c_double)))
    _lib.check(st)
    return out


def flow_accumulation(z):
    """D8 upslope contributing cells of every cell (libhp2d, host code)."""
    import ctypes as C
    from . import _lib
    L = _lib.load()
    z = np.asfortranarray(z, dtype=np.float64)
    out = np.empty_like(z, order="F")
    _lib.check(L.hp2d_flow_accumulation(z.shape[0], z.shape[1],
                                        z.ctypes.data_as(C.POINTER(C.c_double)),
                                        out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def d8_slope(z, dx):
    """Steepest downhill D8 slope of every cell [m/m] (0 at pits and flats,
    NaN at NaN cells); diagonal neighbours at dx * sqrt(2)."""
    z = np.asarray(z, dtype=np.float64)
    ny, nx = z.shape
    zp = np.pad(z, 1, constant_values=np.nan)
    s = np.zeros_like(z)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            if di == 0 and dj == 0:
                continue
            nb = zp[1 + di:1 + di + ny, 1 + dj:1 + dj + nx]
            drop = (z - nb) / (dx * (np.sqrt(2.0) if di and dj else 1.0))
            s = np.fmax(s, np.where(np.isnan(drop), 0.0, drop))
    s[np.isnan(z)] = np.nan
    return s


def spectral_dem(ny, nx, dx, hurst=0.8, relief=200.0, seed=2):
    """Fractal (fBm) surface by spectral synthesis, scaled to `relief` metres."""
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(ny, d=dx)[:, None]
    kx = np.fft.rfftfreq(nx, d=dx)[None, :]
    k = np.sqrt(kx ** 2 + ky ** 2)
    k[0, 0] = np.inf
    amp = k ** (-(hurst + 1.0))
    phase = rng.uniform(0, 2 * np.pi, amp.shape)
    spec = amp * (np.cos(phase) + 1j * np.sin(phase))
    z = np.fft.irfft2(spec, s=(ny, nx))
    z = (z - z.min()) / (z.max() - z.min()) * relief
    return z


def voronoi_classes(ny, nx, n_classes, n_seeds, seed):
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.uniform(0, ny, n_seeds), rng.uniform(0, nx, n_seeds)], axis=1)
    cls = rng.integers(0, n_classes, n_seeds)
    out = np.empty((ny, nx), dtype=np.uint8)
    # nearest seed on a coarse lattice, upsampled (fast, deterministic)
    step = max(1, min(ny, nx) // 512)
    yy = np.arange(0, ny, step)
    xx = np.arange(0, nx, step)
    best = np.full((yy.size, xx.size), np.inf)
    lab = np.zeros((yy.size, xx.size), dtype=np.uint8)
    for (py, px), c in zip(pts, cls):
        d = (yy[:, None] - py) ** 2 + (xx[None, :] - px) ** 2
        m = d < best
        best[m] = d[m]
        lab[m] = c
    out[:] = np.repeat(np.repeat(lab, step, axis=0), step, axis=1)[:ny, :nx]
    return out


def gaussian_random_field(ny, nx, dx, mean, cv, corr_m, seed):
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(ny, d=dx)[:, None]
    kx = np.fft.rfftfreq(nx, d=dx)[None, :]
    k2 = kx ** 2 + ky ** 2
    filt = np.exp(-2 * (np.pi ** 2) * k2 * corr_m ** 2)
    noise = np.fft.rfft2(rng.normal(size=(ny, nx)))
    f = np.fft.irfft2(noise * np.sqrt(filt), s=(ny, nx))
    f = (f - f.mean()) / (f.std() + 1e-300)
    field = mean * (1.0 + cv * f)
    return np.maximum(field, 0.0)


def c2_water_quality(lulc):
    """One pollutant with exponential build-up (7 dry days) and a mass-based
    power-law wash-off, parameters per LULC class; cells under 0.01 g/m2 carry
    no wash-off (min_Bt; with 0 a washed-out cell would give tmin_wq = 0)."""
    k = int(np.max(lulc)) + 1
    return WaterQuality(C1=np.linspace(20.0, 120.0, k), C2=np.full(k, 0.3),
                        C3=np.linspace(5.0, 30.0, k), C4=np.full(k, 1.1),
                        lulc_class=np.asfortranarray(lulc.astype(np.uint8)), ADD_days=7.0,
                        min_Bt=0.01)


def c2_groundwater(z, n_soil, soil_depth=2.0, wt_depth=1.0, every_steps=10):
    """Aquifer under a 2 m soil (Sy = 0.2, K = 1e-5 m/s = 36 mm/h), water table 1 m
    below the ground, van Genuchten soils from sand to clay over the soil classes."""
    return Groundwater(h0=np.asfortranarray(z - wt_depth), soil_depth=soil_depth,
                       theta_sat=np.linspace(0.40, 0.47, n_soil),
                       theta_r=np.linspace(0.05, 0.09, n_soil),
                       n_vg=np.linspace(2.6, 1.3, n_soil), alpha_vg=np.linspace(14.0, 1.0, n_soil),
                       ksat_mmh=36.0, Sy=0.2, every_steps=every_steps)


def full_physics_c2(ny=4096, nx=4096, dx=2.0, precision="fp32", routing_time_min=180.0,
                    fill=True, water_quality=True, groundwater=True):
    """Sink-filled fractal DEM (Hurst 0.8, 200 m relief), 8 LULC classes (n in [0.015, 0.15]), 5 soil
    classes (Ks in [1, 50] mm/h), GRF rainfall pattern (mean 40 mm/h, CV 0.5, 5 km
    correlation) on a 2 h triangular hyetograph, perimeter outlets at the lowest edge,
    one pollutant with build-up / wash-off per LULC class and the groundwater module
    (Sy = 0.2, K = 1e-5 m/s) every 10 surface steps."""
    z = spectral_dem(ny, nx, dx, 0.8, 200.0, seed=2)
    # tilt towards the lowest column so that the catchment drains to one edge
    z = z + np.linspace(50.0, 0.0, nx)[None, :]
    outlet = np.zeros((ny, nx), dtype=np.uint8, order="F")
    outlet[:, nx - 1] = 1
    if fill:  # sinks filled: every depression drains to the outlet column
        z = fill_sinks(z, seed_mask=outlet)
    lulc = voronoi_classes(ny, nx, 8, 64, seed=3)
    n_cls = np.linspace(0.015, 0.15, 8)
    rough = n_cls[lulc]
    soil = voronoi_classes(ny, nx, 5, 48, seed=4)
    pattern = gaussian_random_field(ny, nx, dx, 1.0, 0.5, 5000.0, seed=5)
    # 2 h triangular hyetograph peaking at 40/mean... mean intensity 40 mm/h
    t_end = np.arange(10.0, 130.0, 10.0)        # end-stamped 10-min blocks over 2 h
    mid = t_end - 5.0
    tri = np.where(mid <= 60.0, mid / 60.0, (120.0 - mid) / 60.0) * 80.0   # mean 40 mm/h
    t_end = np.append(t_end, 1e9)
    tri = np.append(tri, 0.0)
    ga = GreenAmpt(ksat_mmh=np.linspace(1.0, 50.0, 5), psi_mm=np.linspace(200.0, 50.0, 5),
                   dtheta=np.full(5, 0.3), theta_sat_minus_r=np.full(5, 0.4),
                   soil_class=np.asfortranarray(soil), I_min_mm=1e-3)
    return RunSpec(
        elevation=np.asfortranarray(z), roughness=np.asfortranarray(rough), resolution=dx,
        outlet_mask=outlet, rain_time_min=t_end, rain_intensity_mmh=tri,
        rain_pattern=np.asfortranarray(pattern), precision=precision, outlet_type=2,
        routing_time_min=routing_time_min, min_time_step_s=0.01, max_time_step_s=10.0,
        time_step_increments_s=0.001, time_step_change_s=1e-4, alfa_min=0.7, alfa_max=0.7,
        record_time_hydrographs_min=1.0, infiltration=ga,
        water_quality=c2_water_quality(lulc) if water_quality else None,
        groundwater=c2_groundwater(z, 5) if groundwater else None)


def dam_break_c3(ny=2048, nx=2048, dx=1.0, precision="fp32", routing_time_min=30.0):
    """V valley (5 % side slope, 0.5 % longitudinal slope along the rows), 10 m reservoir
    upstream of row 256 behind a removed 50 m breach, n = 0.033, dry bed downstream."""
    i = np.arange(ny)[:, None] * dx
    j = np.arange(nx)[None, :] * dx
    centre = (nx - 1) * dx / 2
    z = 0.05 * np.abs(j - centre) + 0.005 * (ny * dx - i)
    wall = 256
    d0 = np.zeros((ny, nx))
    pool = z[:wall, :]
    level = z[wall, int(nx // 2)] + 10.0
    d0[:wall, :] = np.maximum(level - pool, 0.0) * 1000.0
    # dam across row `wall` except a 50 m breach around the valley axis
    half = int(round(25.0 / dx))
    jc = nx // 2
    z[wall, :] = level + 1.0
    z[wall, jc - half:jc + half] = z[wall + 1, jc - half:jc + half]
    d0[wall, :] = 0.0
    outlet = np.zeros((ny, nx), dtype=np.uint8, order="F")
    outlet[ny - 1, :] = 1
    return RunSpec(
        elevation=np.asfortranarray(z), roughness=np.full((ny, nx), 0.033, order="F"),
        resolution=dx, outlet_mask=outlet, d0=np.asfortranarray(d0), precision=precision,
        outlet_type=2, routing_time_min=routing_time_min, min_time_step_s=0.001,
        max_time_step_s=10.0, time_step_increments_s=0.001, time_step_change_s=1e-4,
        alfa_min=0.7, alfa_max=0.7, record_time_hydrographs_min=1.0)


def ensemble_c4(n_members=512, seed=1000, min_log2=7, max_log2=10, dx=5.0, precision="fp32",
                routing_time_min=120.0, vary_soil_moisture=False, filled=False, rain_mmh=50.0,
                rain_duration_min=30.0):
    """Ensemble members around the C1 plane: grid side 2^U(min_log2, max_log2)
    (log-uniform), a seeded fractal DEM (Hurst 0.8, relief 5 % of the domain
    length) on a 1 % tilt towards the last column (critical-depth outlets
    there), and Manning n, Green-Ampt Ks and the rainfall intensity perturbed
    by e^U(ln 0.5, ln 1.5) around n = 0.02, Ks = 5 mm/h and rain_mmh (50 mm/h)
    for rain_duration_min (30 min)."""
    rng = np.random.default_rng(seed)
    specs = []
    for m in range(n_members):
        side = int(round(2.0 ** rng.uniform(min_log2, max_log2)))
        f_n, f_ks, f_rain = np.exp(rng.uniform(np.log(0.5), np.log(1.5), 3))
        # initial soil moisture deficit theta_s - theta_i (fixed 0.3 unless varied)
        dtheta = rng.uniform(0.15, 0.35) if vary_soil_moisture else 0.3
        ny = nx = side
        z = spectral_dem(ny, nx, dx, 0.8, 0.05 * side * dx, seed=seed + m)
        z = z + 0.01 * (nx - 1 - np.arange(nx))[None, :] * dx
        outlet = np.zeros((ny, nx), dtype=np.uint8, order="F")
        outlet[:, nx - 1] = 1
        if filled:  # every depression drains to the outlet column
            z = fill_sinks(z, seed_mask=outlet)
        ga = GreenAmpt(ksat_mmh=np.array([5.0 * f_ks]), psi_mm=np.array([110.0]),
                       dtheta=np.array([dtheta]), theta_sat_minus_r=np.array([0.4]), I_min_mm=1e-3)
        specs.append(RunSpec(
            elevation=np.asfortranarray(z), roughness=np.full((ny, nx), 0.02 * f_n, order="F"),
            resolution=dx, outlet_mask=outlet, rain_time_min=np.array([rain_duration_min, 1e9]),
            rain_intensity_mmh=np.array([rain_mmh * f_rain, 0.0]), precision=precision,
            outlet_type=2, routing_time_min=routing_time_min, min_time_step_s=0.01,
            max_time_step_s=10.0, time_step_increments_s=0.001, time_step_change_s=1e-4,
            alfa_min=0.7, alfa_max=0.7, flag_timestep=1, flag_adaptive_timestepping=1,
            depth_tolerance_mm=1.0, record_time_hydrographs_min=1.0, infiltration=ga))
    return specs
