"""GPU parity of the fused local-inertial step against the FP64 oracle.

Tolerances (stated per test): FP64 mode <= 1e-9 relative (the kernel uses
the same operation order; remaining differences are libm pow vs the device
pow, ~1 ulp); FP32 mode <= 1e-3 m depth against the FP64 oracle.
"""
import numpy as np
import pytest

import li_cases
from oracle import oracle

pytestmark = pytest.mark.gpu

RTOL64 = 1e-9


def _lim():
    from hydropol2d_amd import local_inertial
    return local_inertial


def _random_case(seed, ny, nx, dtype=np.float64, with_nan=True, outlet_type=2, critical=0):
    rng = np.random.default_rng(seed)
    z = rng.uniform(0.0, 2.0, (ny, nx)) + np.linspace(0, 1, nx)[None, :]
    d = rng.uniform(0, 300.0, (ny, nx)) * (rng.uniform(size=(ny, nx)) > 0.3)
    if with_nan:
        z[: ny // 5, : nx // 6] = np.nan
        d[np.isnan(z)] = np.nan
    n = rng.uniform(0.01, 0.1, (ny, nx))
    of = np.zeros((ny, nx, 3))
    of[:, :, 0] = rng.normal(0, 500, (ny, nx))
    of[:, :, 1] = rng.normal(0, 500, (ny, nx))
    rows = np.array([ny, ny, max(ny // 2, 1), 1], dtype=np.int32)
    cols = np.array([nx, max(nx // 2, 1), nx, 1], dtype=np.int32)
    return dict(z=z.astype(dtype), d=d.astype(dtype), n=n.astype(dtype), of=of.astype(dtype),
                rows=rows, cols=cols, outlet_type=outlet_type, critical=critical)


def _call_gpu(c, dt_s, dx, slope=0.003):
    L = _lim()
    ny, nx = c["z"].shape
    mask = np.zeros((ny, nx), dtype=bool)
    mask[c["rows"] - 1, c["cols"] - 1] = True
    return L.Local_Inertial_Model_D4(1, [], [], [], [], [], [], [], [], [], [], [], [], 0, c["z"],
                                     c["d"], c["d"], c["n"], c["n"] ** 2, dx * dx, dt_s / 60, dx,
                                     mask, c["outlet_type"], slope, c["rows"], c["cols"], 1e-6,
                                     c["of"], np.zeros((ny, nx), bool), c["critical"], 0, [], [],
                                     np.zeros((ny, nx)), np.zeros((ny, nx)), 0, 0, 0, 0, dx * dx,
                                     [], 0, 0, [])


def _call_oracle(c, dt_s, dx, slope=0.003):
    f = lambda a: np.asarray(a, dtype=np.float64)
    n = f(c["n"])
    return oracle.li_d4(f(c["z"]), f(c["d"]), n, n ** 2, dx * dx, dt_s / 60, dx, c["outlet_type"],
                        slope, c["rows"], c["cols"], f(c["of"]), c["critical"])


NAMES = ["qout_left", "qout_right", "qout_up", "qout_down", "outlet_flow", "d_t",
         "I_tot_end_cell", "outflow", "Hf"]


@pytest.mark.parametrize("seed,ny,nx,otype,crit", [
    (0, 37, 53, 2, 0), (1, 130, 71, 1, 1), (2, 64, 200, 2, 1), (3, 5, 9, 1, 0),
    (4, 1, 17, 2, 0), (5, 123, 1, 2, 0), (6, 250, 130, 2, 0)])
def test_dropin_fp64_matches_oracle(seed, ny, nx, otype, crit):
    c = _random_case(seed, ny, nx, outlet_type=otype, critical=crit)
    print(f"case seed={seed} shape={ny}x{nx} outlet={otype} critical={crit}", flush=True)
    g = _call_gpu(c, 0.7, 3.0)
    o = _call_oracle(c, 0.7, 3.0)
    for k, name in enumerate(NAMES):
        gv, ov = np.asarray(g[k]), o[name]
        np.testing.assert_array_equal(np.isnan(gv), np.isnan(ov), err_msg=name)
        np.testing.assert_allclose(gv, ov, rtol=RTOL64, atol=1e-9 * (np.nanmax(np.abs(ov)) + 1e-300),
                                   equal_nan=True, err_msg=name)


def test_dropin_fp32_within_1mm():
    c64 = _random_case(11, 96, 140)
    c32 = {k: (v.astype(np.float32) if isinstance(v, np.ndarray) and v.dtype == np.float64 else v)
           for k, v in c64.items()}
    g = _call_gpu(c32, 0.5, 2.0)
    o = _call_oracle(c64, 0.5, 2.0)
    err_m = np.nanmax(np.abs(np.asarray(g[5], dtype=np.float64) - o["d_t"])) / 1000
    assert err_m <= 1e-3, err_m


def _gpu_li_step():
    L = _lim()

    def step(z, d_tot, n, dt_s, dx, state, outlet, outlet_slope):
        ny, nx = z.shape
        if state is None:
            state = np.zeros((ny, nx, 3), order="F")
        mask, rows, cols = outlet
        out = L.Local_Inertial_Model_D4(1, [], [], [], [], [], [], [], [], [], [], [], [], 0, z,
                                        d_tot, d_tot, n, n ** 2, dx ** 2, dt_s / 60, dx, mask, 1,
                                        outlet_slope, rows, cols, 1e-6, state,
                                        np.zeros((ny, nx), bool), 0, 0, [], [], np.zeros((ny, nx)),
                                        np.zeros((ny, nx)), 0, 0, 0, 0, dx ** 2, [], 0, 0, [])
        return out[5], out[4], out[7]

    return step


@pytest.mark.parametrize("case", ["plane", "vtilted", "stage", "ritter"])
def test_validation_cases_gpu_equal_oracle(case):
    fn = {"plane": li_cases.plane_case, "vtilted": li_cases.vtilted_case,
          "stage": li_cases.stage_case, "ritter": li_cases.ritter_case}[case]
    limit = {"plane": 1200, "vtilted": 900, "stage": 900, "ritter": None}[case]
    kw = {} if limit is None else {"n_steps_limit": limit}
    g = fn(_gpu_li_step(), **kw)
    o = fn(li_cases.oracle_li_step(), **kw)
    np.testing.assert_allclose(g["h"], o["h"], rtol=RTOL64, atol=1e-12)


def test_validation_thresholds_on_gpu():
    step = _gpu_li_step()
    r = li_cases.plane_case(step)
    assert r["mass_abs_pct"] < 0.1 and r["nse"] > 0.70 and r["rel_l2"] < 0.45
    r = li_cases.vtilted_case(step)
    assert r["mass_abs_pct"] < 0.1 and r["symmetry_max"] < 1e-8 and r["channel_final"] > 0
    r = li_cases.stage_case(step)
    assert r["max_stage_error"] < 1e-12 and r["mass_abs_pct"] < 0.1 and r["response_amp"] > 1e-4


# ------------------------------------------------------ CA routing -----------

CA_NAMES = ["qout_left", "qout_right", "qout_up", "qout_down", "outlet_flow", "d_t",
            "I_tot_end_cell"]


@pytest.mark.parametrize("seed,ny,nx,otype,crit,zero_prev", [
    (0, 33, 47, 1, 0, True), (1, 80, 61, 2, 1, False), (2, 7, 130, 2, 0, False),
    (3, 1, 9, 1, 1, True)])
def test_ca_dropin_matches_oracle(seed, ny, nx, otype, crit, zero_prev):
    c = _random_case(seed, ny, nx, outlet_type=otype, critical=crit)
    rng = np.random.default_rng(seed + 100)
    Ip = np.zeros((ny, nx)) if zero_prev else rng.uniform(0, 5.0, (ny, nx))
    h0 = rng.uniform(0, 2.0, (ny, nx))
    print(f"ca case seed={seed} shape={ny}x{nx} outlet={otype} critical={crit}", flush=True)
    L = _lim()
    mask = np.zeros((ny, nx), dtype=bool)
    mask[c["rows"] - 1, c["cols"] - 1] = True
    g = L.CA_Routing([], [], [], [], [], [], [], [], [], [], [], [], 0, c["z"], c["d"], c["n"], 9.0,
                     0.4 / 60, h0, 3.0, Ip, mask, otype, 0.004, c["rows"], c["cols"],
                     np.isnan(c["z"]), crit)
    o = oracle.ca_routing(c["z"], c["d"], c["n"], 9.0, 0.4 / 60, h0, 3.0, Ip, otype, 0.004,
                          c["rows"], c["cols"], crit)
    for k, name in enumerate(CA_NAMES):
        gv, ov = np.asarray(g[k]), o[name]
        np.testing.assert_array_equal(np.isnan(gv), np.isnan(ov), err_msg=name)
        np.testing.assert_allclose(gv, ov, rtol=RTOL64,
                                   atol=1e-9 * (np.nanmax(np.abs(ov)) + 1e-300), equal_nan=True,
                                   err_msg=name)


def _gpu_ca_step():
    L = _lim()

    def step(z, d_tot, n, dt_s, dx, state, outlet, outlet_slope):
        ny, nx = z.shape
        if state is None:
            state = np.zeros((ny, nx), order="F")
        mask, rows, cols = outlet
        out = L.CA_Routing([], [], [], [], [], [], [], [], [], [], [], [], 0, z, d_tot, n, dx ** 2,
                           dt_s / 60, np.zeros((ny, nx)), dx, state, mask, 1, outlet_slope, rows,
                           cols, np.zeros((ny, nx), bool), 0)
        qin = li_cases.ca_gather(out[0], out[1], out[2], out[3])
        return out[5] + qin * (dt_s / 60) / 60, out[4], out[6]

    return step


@pytest.mark.parametrize("case", ["plane", "vtilted", "stage"])
def test_ca_validation_cases_gpu_equal_oracle(case):
    fn = {"plane": li_cases.plane_case, "vtilted": li_cases.vtilted_case,
          "stage": li_cases.stage_case}[case]
    g = fn(_gpu_ca_step(), n_steps_limit=600)
    o = fn(li_cases.oracle_ca_step(), n_steps_limit=600)
    np.testing.assert_allclose(g["h"], o["h"], rtol=RTOL64, atol=1e-12)
