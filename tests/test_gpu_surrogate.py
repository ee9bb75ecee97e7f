"""The WSM6 column surrogate on the device: training, inline use, fallback.

* Trained by the library's own kernels on column pairs that the plain WSM6
  kernel writes inside the step (two seeded 40 x 40 x 41 runs), the network's
  error on held-out pairs (a third run) falls well below the untrained
  network's: what the training gains.
* Inline, with every column forced to fail the checks (a water tolerance
  no column can meet), the surrogate step equals the plain step bit for bit
  and every column is counted as a fallback: the fallback is the plain scheme.
* With no check able to fail, no column falls back and the result differs
  from the plain step; with the default tolerances the fallback count lies
  between the two and the step stays close to the plain one.
"""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    from pywrf_amd import surrogate as sg
    data = sg.collect_pairs(runs=2, steps=30, seed0=7)
    valid = sg.collect_pairs(runs=1, steps=30, seed0=50)
    sur = sg.Surrogate.fit(data, epochs=6, batch=2048, lr=2e-3, log=print)
    return sg, sur, data, valid


def test_training_gains_over_the_untrained_network(trained):
    sg, sur, data, valid = trained
    assert len(data) == 2 * 30 * 27 * 27
    untrained = sg.Surrogate.untrained(data)
    l0 = untrained.loss(valid, 2048)
    l1 = sur.loss(valid, 2048)
    print("held-out loss: untrained %.4g  trained %.4g" % (l0, l1))
    print("history", sur.history)
    assert sur.history[-1]["train_loss"] < sur.history[0]["train_loss"]
    assert l1 < 0.25 * l0


def _step(config, ops, nsteps=1):
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config, device="cuda")
    for _ in range(nsteps):
        rk3_step(s, ops, RunConfig())
    torch.cuda.synchronize()
    return s


FIELDS = ("t", "p", "al", "u", "w")


def test_all_columns_failing_gives_the_plain_scheme(trained):
    import pywrf_amd
    sg, sur, _, _ = trained
    plain = _step("config1c", pywrf_amd.ops())
    ops = sg.surrogate_ops(sur, wtol=-2.0)        # w_new - w_old > -2 w_old holds for every column
    s = _step("config1c", ops)
    assert int(ops.nfallback) == 27 * 27
    for n in FIELDS:
        assert torch.equal(getattr(s, n), getattr(plain, n)), n
    assert torch.equal(s.moist, plain.moist)
    assert torch.equal(s.rainnc, plain.rainnc)


def test_fallback_counts_and_inline_accuracy(trained):
    import pywrf_amd
    sg, sur, _, _ = trained
    plain = _step("config1c", pywrf_amd.ops(), 5)
    none = sg.surrogate_ops(sur, qtol=1e30, wtol=1e30, envelope=1e30, dtmax=1e30)
    s_none = _step("config1c", none, 5)
    assert int(none.nfallback) == 0
    assert not torch.equal(s_none.moist, plain.moist)
    ops = sg.surrogate_ops(sur)
    s = _step("config1c", ops, 5)
    nf = int(ops.nfallback)
    print("fallback columns at step 5: %d of %d" % (nf, 27 * 27))
    assert 0 <= nf <= 27 * 27
    rms = lambda x: float(torch.sqrt(torch.mean(x.double() ** 2)))
    eq = rms(s.moist[1] - plain.moist[1]) / rms(plain.moist[1])
    et = rms(s.t - plain.t)
    print("after 5 steps: qv rms rel %.2e, theta rms %.2e K" % (eq, et))
    assert bool(torch.isfinite(s.moist).all()) and bool(torch.isfinite(s.t).all())
    assert eq < 1e-2 and et < 0.5
