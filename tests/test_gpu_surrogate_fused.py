"""The surrogate network in one launch (pywrf_wsm6_surrogate_outputs,
fused = 1: inputs and hidden layers kept in LDS) against the packed inputs
and three GEMM launches (fused = 0) on the same state: the outputs must be
the same bits.  Each 32 x 32 output tile is accumulated from zero over k in
the same order by the same MFMA in both, with the same bias, ReLU and BF16
roundings, so any difference is a fault of the fused kernel (a wrong row,
feature, bias or a missed tile).  States: one RK3 step of the seeded
240 x 240 x 41 (51529 columns), 120 x 120 x 41 (11449) and 40 x 40 x 41 with
clouds (841 columns, a last tile of 73 columns)."""
import ctypes
import os

import pytest
import torch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("config", ["config2", "config4", "config1c"])
def test_fused_network_equals_three_gemms(config):
    import pywrf_amd
    from pywrf_amd import _lib
    from pywrf_amd import surrogate as sg
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    sur = sg.Surrogate.load(os.path.join(ROOT, "pywrf_amd", "data", "wsm6_surrogate.pt"))
    net = sur.struct()
    s = synthetic_state(config, device="cuda")
    rk3_step(s, pywrf_amd.ops(), RunConfig())
    g = _lib.grid(s.th_phy.shape, s.grid.phys18())
    gc = g.contents
    ncol = ((gc.ite - 1) - gc.its) * ((gc.jte - 1) - gc.jts)
    ldy = (ncol + 127) // 128 * 128
    m = s.moist
    fields = [s.th_phy, s.rho, s.pi_phy, s.p_phy, m[1], m[2], m[3], m[4], m[5], m[6]]
    ptrs = [ctypes.c_void_p(f.data_ptr()) for f in fields]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    work = torch.empty(ldy * 2816, dtype=torch.uint8, device="cuda")
    for fused in (1, 0):
        y = torch.full((288, ldy), float("nan"), device="cuda")
        rc = _lib.LIB.pywrf_wsm6_surrogate_outputs(g, ctypes.addressof(net), *ptrs, ctypes.c_void_p(y.data_ptr()),
                                                   ldy, fused, ctypes.c_void_p(work.data_ptr()),
                                                   ctypes.c_longlong(work.numel()), stream)
        assert rc == 0, rc
        out[fused] = y
    torch.cuda.synchronize()
    a, b = out[1][:, :ncol], out[0][:, :ncol]
    print(config, "columns", ncol, "max |fused - gemm|", float((a - b).abs().max()))
    assert bool(torch.isfinite(a).all())
    assert torch.equal(a, b)
    # columns past the tile are not written
    assert bool(torch.isnan(out[1][:, ncol:]).all())
