"""The device NaN flag (step.NanFlag): the reference's per-stage NaN exits
(solver.py:1648-1654, 1768-1774) without a synchronisation in the step.
A NaN is planted in the input temperature from the host (no kernel is made
to fault); the first RK stage's pressure then holds NaNs, and the record
names that check and a NaN point, eagerly and from a HIP-graph replay.  A
clean step leaves the flag clear."""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _state():
    from pywrf_amd.gen import synthetic_state
    return synthetic_state("config1c", device="cuda")


def test_clean_step_leaves_the_flag_clear():
    import pywrf_amd
    from pywrf_amd.step import rk3_step, RunConfig, NanFlag
    s = _state()
    f = NanFlag()
    f.clear()
    rk3_step(s, pywrf_amd.ops(), RunConfig(), nan_flag=f)
    torch.cuda.synchronize()
    assert f.record() == (0, 0, 0)
    f.check(tuple(s.p.shape))


@pytest.mark.parametrize("graph", [False, True])
def test_planted_nan_is_reported_by_the_first_stage(graph):
    import pywrf_amd
    from pywrf_amd.graph import StepGraph
    from pywrf_amd.step import rk3_step, RunConfig, NanFlag
    s = _state()
    f = NanFlag()
    cfg = RunConfig()
    if graph:
        g = StepGraph(s, cfg, pywrf_amd.ops(), steps=1, nan_flag=f)
    s.t[10, 20, 20] = math.nan
    f.clear()
    if graph:
        g.replay()
    else:
        rk3_step(s, pywrf_amd.ops(), cfg, nan_flag=f)
    torch.cuda.synchronize()
    code, first, count = f.record()
    print("record", code, first, count)
    assert code == 1 and count > 0
    assert bool(torch.isnan(s.p.reshape(-1)[first]))
    with pytest.raises(FloatingPointError, match="rk_step 1"):
        f.check(tuple(s.p.shape))
