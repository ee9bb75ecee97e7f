"""The HIP-graph replay of a step (pywrf_amd.graph.StepGraph, used by the
benchmark) gives bit-identical states to the eager step, replay after replay,
with eager work and synchronisations interleaved between the replays -- with
constant lateral boundaries (config1c) and with time-varying ones (bdy24,
where the boundary time dtbc the graph forms on the device changes every
step, and with a 2-step boundary interval also jumps back when a new
boundary record starts)."""
import pytest
import torch

from parity import state, run, fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config,interval", [("config1c", 360), ("bdy24", 360), ("bdy24", 2)])
def test_graph_replays_match_eager(config, interval):
    import pywrf_amd
    from pywrf_amd.graph import StepGraph
    from pywrf_amd.step import RunConfig
    cfg = RunConfig(bdy_interval=interval)
    ops = pywrf_amd.ops()
    ref = state(config, device="cuda")
    run(ref, ops, 1, cfg)
    s = state(config, device="cuda")
    run(s, ops, 1, cfg)                     # eager first step
    g = StepGraph(s, cfg, ops)
    for n in range(3):
        g.replay()
        torch.cuda.synchronize()
        run(ref, ops, 1, cfg)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()     # unrelated eager work between replays
        a, b = fields(s), fields(ref)
        for name in b:
            assert torch.equal(a[name], b[name]), (n, name)
        assert s.step == ref.step
    del junk


@pytest.mark.parametrize("config", ["config1c", "bdy24"])
def test_two_step_graph_matches_eager(config):
    """A graph of two steps (the copies back happen once per replay)."""
    import pywrf_amd
    from pywrf_amd.graph import StepGraph
    from pywrf_amd.step import RunConfig
    cfg = RunConfig()
    ops = pywrf_amd.ops()
    ref = state(config, device="cuda")
    run(ref, ops, 1)
    s = state(config, device="cuda")
    run(s, ops, 1)
    g = StepGraph(s, cfg, ops, steps=2)
    for n in range(2):
        g.replay()
        torch.cuda.synchronize()
        run(ref, ops, 2)
        torch.cuda.synchronize()
        a, b = fields(s), fields(ref)
        for name in b:
            assert torch.equal(a[name], b[name]), (n, name)
    assert s.step == ref.step


@pytest.mark.parametrize("interval", [1, 2, 3, 360])
def test_device_clock_matches_the_host_counters(interval):
    """pywrf_clock_advance against the driver's own host arithmetic
    (solver.py:457-458, 713: itimestep += 1; it_bdy = int(itimestep /
    bdy_interval) when itimestep % bdy_interval == 1; dtbc = dt * (itimestep -
    it_bdy * bdy_interval), taken by the kernels as float32) for 12 steps from
    a non-zero start."""
    import numpy as np
    from pywrf_amd.step import DeviceClock, RunConfig

    class Host:
        step, it_bdy = 5, 1

    cfg = RunConfig(dt=54.0, bdy_interval=interval)
    clock = DeviceClock(Host)
    step, it_bdy = Host.step, Host.it_bdy
    for _ in range(12):
        step += 1
        if step % interval == 1:
            it_bdy = int(step / interval)
        dtbc = np.float32(cfg.dt * (step - it_bdy * interval))
        clock.advance(cfg)
        assert int(clock.step) == step and int(clock.it_bdy) == it_bdy
        assert np.float32(clock.dtbc.item()) == dtbc, (step, float(clock.dtbc), float(dtbc))
