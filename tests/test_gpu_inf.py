"""GPU parity of the mainline infiltration step (Infiltration_Module.m) with the
FP64 oracle.  Tolerance 1e-12 relative (the device pow against libm pow,
about 1 ulp, is the only difference in the arithmetic)."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def _random(seed, ny, nx):
    rng = np.random.default_rng(seed)
    elev = rng.uniform(50, 60, (ny, nx))
    depth = rng.uniform(0.5, 3.0, (ny, nx))
    zwt = rng.uniform(-0.2, 1.0, (ny, nx)) * depth          # some cells without an unsaturated zone
    h_t = elev - zwt
    ts = rng.uniform(0.35, 0.5, (ny, nx))
    tr = rng.uniform(0.02, 0.1, (ny, nx))
    uz = np.maximum(zwt, 0) * (ts - tr) * 1000
    I = rng.uniform(0, 1.0, (ny, nx)) * uz
    d = rng.uniform(0, 30, (ny, nx)) * (rng.uniform(size=(ny, nx)) > 0.3)
    d[0, : nx // 3] = np.nan                                 # outside the domain
    return dict(d_t=d, I_t=I, elevation=elev, h_t=h_t, soil_depth=depth, theta_sat=ts, theta_r=tr,
                n_vg=rng.uniform(1.2, 2.5, (ny, nx)), alpha_vg=rng.uniform(0.5, 8, (ny, nx)),
                ksat_mmh=rng.uniform(0.5, 60, (ny, nx)),
                idx_imp=rng.uniform(size=(ny, nx)) < 0.1, cum_inf=np.ones((ny, nx)))


@pytest.mark.parametrize("seed,ny,nx,ts", [(0, 23, 31, 0.5), (1, 128, 96, 2.0), (2, 1, 7, 0.1),
                                           (3, 300, 257, 1.0)])
def test_infiltration_matches_oracle(seed, ny, nx, ts):
    from hydropol2d_amd.infiltration import Infiltration_Module
    c = _random(seed, ny, nx)
    ref = oracle.infiltration(time_step_min=ts, resolution=4.0, **c)
    got = Infiltration_Module(c["d_t"], c["I_t"], c["h_t"], c["elevation"], c["soil_depth"],
                              c["theta_sat"], c["theta_r"], c["n_vg"], c["alpha_vg"], c["ksat_mmh"],
                              c["idx_imp"], ts, 4.0, cumulative_infiltration=c["cum_inf"])
    for k in ("d_t", "I_t", "I_p", "f"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, atol=1e-12, err_msg=k)
    np.testing.assert_allclose(got["cumulative_infiltration"], ref["cum_inf"], rtol=1e-12)
    np.testing.assert_allclose(got["inf_volume"], ref["inf_volume"], rtol=1e-12)
    scale = np.nansum(c["d_t"] + c["I_t"]) * 16 / 1000
    assert abs(got["errors4"] - ref["dS_inf"]) <= 1e-12 * scale
