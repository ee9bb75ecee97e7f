"""The batched entry points (pywrf_zero_group, pywrf_set_physical_bc_group)
give exactly what the individual calls they replace give."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_zero_group_zeroes_exactly_its_tensors():
    import pywrf_amd.wrf_dynamics as dyn
    torch.manual_seed(0)
    shapes = [(41, 40, 40), (40, 40), (7, 41, 40, 40), (7,), (1,), (3, 5, 11)]
    ts = [torch.randn(sh, device="cuda") + 3.0 for sh in shapes]
    keep = torch.randn(41, 40, 40, device="cuda")
    keep0 = keep.clone()
    dyn.zero_group(ts)
    torch.cuda.synchronize()
    for t in ts:
        assert int(torch.count_nonzero(t)) == 0
    assert torch.equal(keep, keep0)
    dyn.zero_group([])                       # no-op


@pytest.mark.parametrize("n", [40, 120])
def test_bc_group_equals_individual_calls(n):
    import pywrf_amd.wrf_dynamics as dyn
    from pywrf_amd.step import Grid
    g = Grid(41, n, n)
    torch.manual_seed(1)
    items = [(torch.randn(41, n, n, device="cuda"), v, 0) for v in "uvwtpef"]
    items += [(torch.randn(n, n, device="cuda"), v, 1) for v in "uvt"]
    ref = [t.clone() for t, _, _ in items]
    for r, (_, v, two) in zip(ref, items):
        if two:
            dyn.set_physical_bc2d(r, v, *g.ix2d)
        else:
            dyn.set_physical_bc3d(r, v, *g.ix24)
    dyn.set_physical_bc_group(items, g.ix24, g.ix2d)
    torch.cuda.synchronize()
    for r, (t, v, two) in zip(ref, items):
        assert torch.equal(r, t), (v, two)
