"""Device-loop runs with pollutant wash-off and the groundwater module, shared
by the oracle pins and the GPU parity tests: small tilted planes (the 256 x 256
plane of the single-domain benchmark cut down) with a wet start, three LULC
classes of build-up / wash-off parameters and the groundwater module called
every few steps."""
import numpy as np

from hydropol2d_amd import synthetic
from hydropol2d_amd.runspec import WaterQuality

C1 = np.array([30.0, 60.0, 90.0])  # kg/ha
C2 = np.full(3, 0.4)                # 1/day
ADD = 5.0                           # dry days


def lulc_map(ny, nx, seed):
    rng = np.random.default_rng(seed + 7)
    return np.asfortranarray((rng.random((ny, nx)) * 3).astype(np.uint8))


def plane(ny=32, nx=40, C3=3e4, C4=1.1, min_Bt=0.01, wq=True, gw=True, T=5.0, every=10, seed=1,
          d0=5.0, nan_corner=False, outlets=True, precision="fp64", model=0, Bmin=0.0, Bmax=1e30,
          h0_offset=-1.0):
    spec = synthetic.tilted_plane_c1(ny, nx, routing_time_min=T, seed=seed, precision=precision)
    spec.d0 = np.full((ny, nx), d0, order="F")
    if nan_corner:
        z = spec.elevation.copy(order="F")
        z[: ny // 4, nx - nx // 4:] = np.nan
        spec.elevation = z
        spec.outlet_mask = np.where(np.isnan(z), 0, spec.outlet_mask).astype(np.uint8)
        if not spec.outlet_mask.any():
            spec.outlet_mask[ny - 1, 0] = 1
    if not outlets:
        spec.outlet_mask = np.zeros((ny, nx), np.uint8, order="F")
    if wq:
        spec.water_quality = WaterQuality(C1=C1, C2=C2, C3=np.array([C3, 2 * C3, 3 * C3]),
                                          C4=np.full(3, C4), lulc_class=lulc_map(ny, nx, seed),
                                          ADD_days=ADD, min_Bt=min_Bt, flag_wq_model=model,
                                          Bmin=Bmin, Bmax=Bmax)
    if gw:
        z = np.nan_to_num(spec.elevation, nan=0.0)
        spec.groundwater = synthetic.c2_groundwater(z, 1, wt_depth=-h0_offset, every_steps=every)
    return spec


def initial_mass(spec):
    """sum of C1 (1 - e^(-C2 ADD)) A / 1e4 over the valid cells (build-up closed form)."""
    wq = spec.water_quality
    cls = wq.lulc_class
    A = spec.resolution ** 2
    b = wq.C1[cls] * (1.0 - np.exp(-wq.C2[cls] * wq.ADD_days)) * A / 1e4
    return float(np.sum(np.where(np.isnan(spec.elevation), 0.0, b)))
