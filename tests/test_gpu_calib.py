"""Truncated-backprop calibration (pywrf_amd.calib) on the device: a
water-vapour source bias is recovered from a truth run through the
hand-written transport adjoint (120 x 120 x 41 and 240 x 240 x 41, 12 RK3
steps in windows of 6, 20 epochs of Adam with cosine decay: the misfit falls
by at least 5x and the relative error of the fitted bias from 1 to below
0.35), and a theta-forcing bias through an RK stage's acoustic loop."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config", ["config4", "config2"])
def test_calibration_recovers_the_bias(config):
    from pywrf_amd import calib
    out = calib.calibrate(config, nsteps=12, window=6, epochs=20, lr=2e-7, log=print)
    h = out["history"]
    print({k: v for k, v in out.items() if k != "history"})
    assert h[0]["bias_error"] == 1.0
    assert h[-1]["loss"] < 0.2 * h[1]["loss"]
    assert h[-1]["bias_error"] < 0.35


@pytest.mark.parametrize("config", ["config4", "config2"])
def test_theta_forcing_bias_recovered_through_the_acoustic_loop(config):
    """A smooth theta-forcing bias (10 % of max |t_tend|) is recovered from
    the theta at the end of the third RK stage's acoustic loop (four
    substeps, 120 x 120 x 41 and 240 x 240 x 41) through the checkpointed
    loop adjoint: 60 Adam steps take the misfit down by at least 100x and
    the relative error of the bias below 0.1."""
    from pywrf_amd import calib
    out = calib.fit_forcing_bias(config, stage=3, iters=60, log=print)
    h = out["history"]
    print({k: v for k, v in out.items() if k != "history"})
    assert h[0]["bias_error"] == 1.0
    assert h[-1]["loss"] < 0.01 * h[0]["loss"]
    assert h[-1]["bias_error"] < 0.1
