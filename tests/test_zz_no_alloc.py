"""The allocation-free surrogate entry point (pywrf_wsm6_surrogate_noalloc)
against the workspace entry point (pywrf_wsm6_surrogate), through the C ABI.

Both run the same kernels on the same inputs; the allocation-free one takes
its buffers from the library's static device array.  Two calls in a row --
the 240 x 240 x 41 state (51529 columns), then the smaller 120 x 120 x 41 one
(11449 columns) -- must give bit-identical fields, precipitation and
fallback counts.  A tile with more columns than the array holds is refused
with PYWRF_E_RANGE before anything is launched.
"""
import ctypes
import os

import pytest
import torch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _inputs(config):
    """A state just before microphysics: one RK3 step of the seeded state."""
    import pywrf_amd
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config, device="cuda")
    rk3_step(s, pywrf_amd.ops(), RunConfig())
    torch.cuda.synchronize()
    return s


def _args(s):
    """The physics fields of a state, copied, in pywrf_wsm6_surrogate's order."""
    m = s.moist
    fields = [s.th_phy, s.rho, s.pi_phy, s.p_phy, s.dz8w, m[1], m[2], m[3], m[4], m[5], m[6], s.rainnc, s.rainncv,
              s.snownc, s.snowncv, s.sr, s.graupelnc, s.graupelncv]
    return [f.clone() for f in fields]


def _call(name, g, net, a, extra, nfall):
    from pywrf_amd import _lib
    ptrs = [ctypes.c_void_p(t.data_ptr()) for t in a]
    rc = getattr(_lib.LIB, name)(
        g, ctypes.addressof(net), *ptrs[:5], ctypes.c_double(54.0), *ptrs[5:], *extra,
        ctypes.c_void_p(nfall.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc


def test_noalloc_equals_workspace_entry_point():
    from pywrf_amd import _lib
    from pywrf_amd import surrogate as sg
    sur = sg.Surrogate.load(os.path.join(ROOT, "pywrf_amd", "data", "wsm6_surrogate.pt"))
    net = sur.struct()
    maxc, nbytes = ctypes.c_int(), ctypes.c_longlong()
    assert _lib.LIB.pywrf_wsm6_surrogate_noalloc_limits(ctypes.addressof(maxc), ctypes.byref(nbytes)) == 0
    print("noalloc limits: %d columns, %d bytes" % (maxc.value, nbytes.value))
    for config, ncol in (("config2", 227 * 227), ("config4", 107 * 107)):
        s = _inputs(config)
        g = _lib.grid(s.th_phy.shape, s.grid.phys18())
        a_ws, a_na = _args(s), _args(s)
        nf_ws = torch.zeros(1, dtype=torch.int32, device="cuda")
        nf_na = torch.full((1,), 12345, dtype=torch.int32, device="cuda")   # the call must reset it
        wsize = ctypes.c_longlong()
        assert _lib.LIB.pywrf_wsm6_surrogate_workspace(g, 65536, ctypes.byref(wsize)) == 0
        ws = torch.empty(wsize.value, dtype=torch.uint8, device="cuda")
        rc = _call("pywrf_wsm6_surrogate", g, net, a_ws, [ctypes.c_void_p(ws.data_ptr()), ctypes.c_longlong(wsize.value)],
                   nf_ws)
        assert rc == 0, rc
        rc = _call("pywrf_wsm6_surrogate_noalloc", g, net, a_na, [], nf_na)
        assert rc == 0, rc
        torch.cuda.synchronize()
        print(config, "fallback columns", int(nf_ws), int(nf_na), "of", ncol)
        assert int(nf_na) == int(nf_ws)
        assert 0 <= int(nf_ws) <= ncol
        for x, y in zip(a_ws, a_na):
            assert torch.equal(x, y)


def test_noalloc_refuses_too_many_columns():
    from pywrf_amd import _lib
    from pywrf_amd.step import Grid
    from pywrf_amd import surrogate as sg
    sur = sg.Surrogate.load(os.path.join(ROOT, "pywrf_amd", "data", "wsm6_surrogate.pt"))
    gr = Grid(41, 300, 300)                            # 287^2 = 82369 tile columns > 65536
    g = _lib.grid((41, 300, 300), gr.phys18())
    nf = torch.zeros(1, dtype=torch.int32, device="cuda")
    dummy = [torch.zeros(1, device="cuda")] * 18        # never read: the call is refused first
    rc = _call("pywrf_wsm6_surrogate_noalloc", g, sur.struct(), dummy, [], nf)
    assert rc == _lib.PYWRF_E_RANGE
