"""The HIP-graph replay of a step (pywrf_amd.graph.StepGraph, used by the
benchmark) gives bit-identical states to the eager step, replay after replay,
with eager work and synchronisations interleaved between the replays."""
import pytest
import torch

from parity import state, run, fields

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager():
    import pywrf_amd
    from pywrf_amd.graph import StepGraph
    from pywrf_amd.step import RunConfig
    cfg = RunConfig()
    ops = pywrf_amd.ops()
    ref = state("config1c", device="cuda")
    run(ref, ops, 1)
    s = state("config1c", device="cuda")
    run(s, ops, 1)                          # eager first step
    g = StepGraph(s, cfg, ops)
    for n in range(3):
        g.replay()
        torch.cuda.synchronize()
        run(ref, ops, 1)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()     # unrelated eager work between replays
        a, b = fields(s), fields(ref)
        for name in b:
            assert torch.equal(a[name], b[name]), (n, name)
        assert s.step == ref.step
    del junk


def test_two_step_graph_matches_eager():
    """A graph of two steps (the copies back happen once per replay)."""
    import pywrf_amd
    from pywrf_amd.graph import StepGraph
    from pywrf_amd.step import RunConfig
    cfg = RunConfig()
    ops = pywrf_amd.ops()
    ref = state("config1c", device="cuda")
    run(ref, ops, 1)
    s = state("config1c", device="cuda")
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
