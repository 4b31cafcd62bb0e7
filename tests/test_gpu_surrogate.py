"""The surrogate against the solver that trains and judges it (GPU).

A small study that runs in seconds: 64 x 64 C4-style members (sink-filled
fractal DEM, perturbed n, Ks, rainfall and soil moisture, 2 h each) solved as
one batched ensemble; the network is trained on 288 of them and predicts the
peak depth of 32 held-out members.  It must beat both maps a network-free
forecaster has: the all-dry map and the training mean map.  The judge
re-runs a sample of the members on the solver and falls back to the solver
for every member when the sample's RMSE exceeds the tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def study():
    import torch
    from hydropol2d_amd import surrogate as S
    from hydropol2d_amd import synthetic
    specs = synthetic.ensemble_c4(320, seed=777, min_log2=6, max_log2=6, routing_time_min=120.0,
                                  vary_soil_moisture=True, filled=True)
    pk, tp, _ = S.solve_members(specs)
    print("64 x 64 surrogate study: 288 training members, 32 held out", flush=True)
    net = S.UNetSurrogate(64, 64, batch=32, seed=1)
    x = S.stack_inputs(net, specs)
    y = S.stack_targets(net, pk, tp)
    hist = S.train(net, x[:288], y[:288], epochs=20, batch=32, lr=1e-3, lr_final=5e-5)
    pred = net.predict(x[288:])
    torch.cuda.synchronize()
    peak = (pred[..., 0] / S.DEPTH_SCALE).clamp(min=0).cpu().numpy()
    return S, net, specs, pk, peak, hist


def test_surrogate_beats_the_network_free_maps(study):
    S, net, specs, pk, peak, hist = study
    ref = pk[288:]
    rmse = np.sqrt(np.mean((peak - ref) ** 2))
    zero = np.sqrt(np.mean(ref ** 2))
    mean_map = np.sqrt(np.mean((pk[:288].mean(axis=0)[None] - ref) ** 2))
    print(f"held-out peak-depth RMSE {rmse * 100:.2f} cm, mean map {mean_map * 100:.2f} cm, "
          f"dry map {zero * 100:.2f} cm", flush=True)
    assert np.isfinite(hist).all() and hist[-1] < 0.7 * hist[0]
    # measured 1.73 cm against 2.05 cm (mean map) and 2.31 cm (dry map); the
    # bounds leave room for the run-to-run spread of the FP32 atomics in the
    # weight gradients
    assert rmse < 0.92 * mean_map and rmse < 0.85 * zero


def test_judge_keeps_or_replaces_the_predictions(study):
    S, net, specs, pk, peak, hist = study
    held = specs[288:]
    ref = pk[288:]
    # a tolerance the surrogate meets: the predictions are kept and the judged
    # members carry the solver's maps
    ok = S.judged_predictions(net, held, judge_fraction=0.25, tol_m=1.0)
    assert not ok["fallback"] and len(ok["judged"]) == 8
    np.testing.assert_array_equal(ok["peak_m"][ok["judged"]], ref[ok["judged"]])
    others = np.setdiff1d(np.arange(32), ok["judged"])
    np.testing.assert_allclose(ok["peak_m"][others], peak[others], rtol=0, atol=1e-6)
    assert ok["judge_rmse_m"] == pytest.approx(
        np.sqrt(np.mean((peak[ok["judged"]] - ref[ok["judged"]]) ** 2)), rel=1e-5)
    # a tolerance it cannot meet: every member falls back to the solver
    bad = S.judged_predictions(net, held, judge_fraction=0.25, tol_m=1e-6)
    assert bad["fallback"]
    np.testing.assert_array_equal(bad["peak_m"], ref)
