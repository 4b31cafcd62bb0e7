"""GPU tests of the batched ensemble: every member of one batched run equals
the same member run alone, bit for bit (same kernels, same tile shapes, same
reduction order), FP64 members equal the oracle, and the time-to-peak map
equals the one derived on the host from per-step states."""
import numpy as np
import pytest

from hydropol2d_amd import synthetic
from oracle import oracle

pytestmark = pytest.mark.gpu


def _specs(precision, sizes=(37, 64, 100, 128, 45, 200), minutes=(8.0, 5.0, 10.0, 3.0, 12.0, 6.0)):
    out = []
    for k, (n, tmin) in enumerate(zip(sizes, minutes)):
        sp = synthetic.tilted_plane_c1(n, n + 3 * k, seed=50 + k, precision=precision,
                                       routing_time_min=tmin)
        out.append(sp)
    return out


def _solo(spec, shape):
    """The member run alone with the member's launch shape (the shape sets the
    order of the domain sums of the report)."""
    from hydropol2d_amd.simulation import Simulation
    with Simulation(spec) as s:
        s.set_launch_shape(shape["tj"], shape["waves_per_block"], shape["rows"])
        s.run()
        return s.state(infiltration=True), s.report(), s.hydrograph(), s.dt_history()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("graph_steps", [0, 20])
def test_members_equal_solo_runs(precision, graph_steps):
    from hydropol2d_amd.simulation import Ensemble
    specs = _specs(precision)
    with Ensemble(specs) as ens:
        n_done = ens.run(graph_steps=graph_steps)
        assert n_done == len(specs)
        for i, sp in enumerate(specs):
            m = ens.member(i)
            st, rep, hy, dth = m.state(infiltration=True), m.report(), m.hydrograph(), m.dt_history()
            assert m.launch_shape()["rows"] == 1
            st0, rep0, hy0, dth0 = _solo(sp, m.launch_shape())
            for k in ("d_t", "outflow", "dmax", "I_t"):
                np.testing.assert_array_equal(st[k], st0[k], err_msg=f"member {i} {k}")
            assert rep == rep0, i
            np.testing.assert_array_equal(hy["hydro_q"], hy0["hydro_q"])
            np.testing.assert_array_equal(dth, dth0)


def test_fp64_members_match_oracle():
    from hydropol2d_amd.simulation import Ensemble
    specs = _specs("fp64", sizes=(40, 72), minutes=(6.0, 4.0))
    with Ensemble(specs) as ens:
        ens.run()
        for i, sp in enumerate(specs):
            o = oracle.run(sp)
            st = ens.member(i).state()
            ok = ~np.isnan(o["d_t"])
            np.testing.assert_allclose(st["d_t"][ok], o["d_t"][ok], rtol=1e-9, atol=1e-9)
            assert ens.member(i).report()["steps"] == o["report"]["steps"]


def test_partial_runs_and_compaction():
    # members finish at different times; stepping in small chunks (which
    # rebuilds the tile table as members drop out) gives the same results
    from hydropol2d_amd.simulation import Ensemble
    specs = _specs("fp32")
    with Ensemble(specs) as a, Ensemble(specs) as b:
        a.run()
        done = 0
        while done < len(specs):
            done = b.run(max_steps=37, graph_steps=10)
        for i in range(len(specs)):
            np.testing.assert_array_equal(a.member(i).state()["d_t"], b.member(i).state()["d_t"])


def test_time_to_peak_matches_host_derivation():
    from hydropol2d_amd.simulation import Simulation
    spec = synthetic.tilted_plane_c1(40, 52, seed=3, precision="fp64", routing_time_min=3.0)
    # the recorded time is the physical end of the step: its start plus the dt
    # it used (the loop's clock then advances by the next dt, as the
    # reference's does)
    with Simulation(spec) as s:
        s.record_time_to_peak()
        best = None
        tbest = np.zeros((40, 52))
        rep = s.report()
        while True:
            t_end = rep["t_min"] + rep["time_step_s"] / 60.0
            done = s.run(1)
            rep = s.report()
            d = s.state()["d_t"]
            if best is None:
                best, tbest[:] = d.copy(), t_end
            else:
                up = d > best
                best = np.maximum(best, d)
                tbest[up] = t_end
            if done:
                break
        tp = s.time_to_peak()
    np.testing.assert_allclose(tp, tbest, rtol=0, atol=1e-9)


def test_c4_members_run():
    # a small slice of the C4 ensemble (log-uniform sizes, perturbed n, Ks, rain)
    from hydropol2d_amd.simulation import Ensemble
    specs = synthetic.ensemble_c4(12, max_log2=8, routing_time_min=10.0)
    with Ensemble(specs) as ens:
        ens.record_time_to_peak()
        assert ens.run() == len(specs)
        for i in range(len(specs)):
            m = ens.member(i)
            dm, tp = m.state()["dmax"], m.time_to_peak()
            assert np.isfinite(dm).all() and (dm >= 0).all()
            assert (tp >= 0).all() and (tp <= 10.0 + 1.0).all()


def test_c4_full_batch_members_equal_solo_runs():
    """The whole C4 ensemble: 512 members with sides log-uniform in 128..1024
    (1.3e8 cells), perturbed n, Ks and rainfall, 2 h each, FP32, in one
    batched run.  Every member ends with finite, non-negative depths and peak
    depths, a time to peak inside the run, and its run clock at 2 h; eight
    members across the size range, run alone with the member's launch shape,
    equal their batched selves bit for bit."""
    from hydropol2d_amd.simulation import Ensemble
    specs = synthetic.ensemble_c4(512)
    sides = np.array([sp.shape[0] for sp in specs])
    print(f"C4 batch: 512 members, sides {sides.min()}..{sides.max()}, "
          f"{int((sides.astype(np.int64) ** 2).sum())} cells", flush=True)
    order = np.argsort(sides, kind="stable")
    sample = [int(order[k]) for k in np.linspace(0, len(specs) - 1, 8).astype(int)]
    with Ensemble(specs) as ens:
        ens.record_time_to_peak()
        assert ens.run() == len(specs)
        got = {}
        for i in range(len(specs)):
            m = ens.member(i)
            st, rep = m.state(infiltration=True), m.report()
            assert np.isfinite(st["d_t"]).all() and (st["d_t"] >= 0).all(), i
            assert np.isfinite(st["dmax"]).all() and (st["dmax"] >= st["d_t"]).all(), i
            tp = m.time_to_peak()
            # the last step may end up to one max_dt (10 s) past the 2 h
            assert (tp >= 0).all() and (tp <= 120.0 + 10.0 / 60.0 + 1e-9).all(), i
            assert rep["t_min"] == pytest.approx(120.0, abs=10.0 / 60.0), i
            if i in sample:
                got[i] = (st, rep, m.launch_shape())
    for i in sample:
        st, rep, shape = got[i]
        st0, rep0, _, _ = _solo(specs[i], shape)
        for k in ("d_t", "outflow", "dmax", "I_t"):
            np.testing.assert_array_equal(st[k], st0[k], err_msg=f"member {i} {k}")
        assert rep == rep0, i


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("graph_steps", [0, 20])
def test_wash_off_and_groundwater_members_equal_solo_runs(precision, graph_steps):
    """Members with the wash-off (Calibration_Water_Quality.m's batched
    members), with the groundwater module (each called every N of its own
    accepted steps), with both, and plain ones, in one batched run: each equals
    the same member run alone, bit for bit, including the pollutant and head
    fields and the ledgers."""
    import loop_cases
    from hydropol2d_amd.simulation import Ensemble, Simulation
    specs = [loop_cases.plane(24, 30, C3=3e4, gw=False, precision=precision, T=3.0),
             loop_cases.plane(32, 40, wq=False, every=3, precision=precision, T=4.0),
             loop_cases.plane(40, 28, C3=3e5, min_Bt=5.0, every=5, precision=precision, seed=2,
                              T=3.5),
             loop_cases.plane(20, 20, wq=False, gw=False, precision=precision, seed=3, T=2.0)]
    with Ensemble(specs) as ens:
        assert ens.run(graph_steps=graph_steps) == len(specs)
        for i, sp in enumerate(specs):
            wq, gw = sp.water_quality is not None, sp.groundwater is not None
            m = ens.member(i)
            st = m.state(infiltration=True, water_quality=wq, groundwater=gw)
            rep, shape = m.report(), m.launch_shape()
            with Simulation(sp) as s:
                s.set_launch_shape(shape["tj"], shape["waves_per_block"], shape["rows"])
                s.run()
                st0 = s.state(infiltration=True, water_quality=wq, groundwater=gw)
                rep0 = s.report()
            for k in st0:
                np.testing.assert_array_equal(st[k], st0[k], err_msg=f"member {i} {k}")
            assert rep == rep0, i
            if gw:
                assert rep["gw_updates"] > 0
            if wq:
                assert rep["wq_mass_kg"] > 0
