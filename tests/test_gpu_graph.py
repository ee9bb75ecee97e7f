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
