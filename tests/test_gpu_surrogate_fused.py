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


def _emulate_apply(sur, net, fields, Y, gc, km, dt):
    """The checks and the update of the surrogate step restated in torch FP32
    (elementwise, in the kernel's operation order: per level the tendency
    out * std + mean, the updated values and the checks; the column water
    summed over the six species in order, over the 3-level groups' levels in
    order, then over the 16 groups in order).  Returns (bad [ncol] bool,
    th_new [km, ncol], q_new [6, km, ncol], w_old, w_new)."""
    th, rho, pii, dz = fields["th"], fields["rho"], fields["pii"], fields["dz"]
    qs = fields["q"]
    ncol = Y.shape[1]
    f32 = lambda x: torch.tensor(x, dtype=torch.float32, device="cuda")
    scale, dtmax, qtol, wtol = f32(net.envelope), f32(net.dtmax), f32(net.qtol), f32(net.wtol)
    dtf = f32(dt)
    lo, hi, amax = sur.in_lo, sur.in_hi, sur.out_absmax
    ostd, omean = sur.out_std, sur.out_mean

    def in_range(feat, x):
        w = (hi[feat] - lo[feat]) * (scale - 1.0)
        return (x >= lo[feat] - w) & (x <= hi[feat] + w)

    def tend(v, k):
        o = v * 41 + k
        return Y[o] * ostd[o] + omean[o]

    bad = torch.zeros(ncol, dtype=torch.bool, device="cuda")
    th_new = torch.empty(km, ncol, device="cuda")
    q_new = torch.empty(6, km, ncol, device="cuda")
    gwo = [torch.zeros(ncol, device="cuda") for _ in range(16)]
    gwn = [torch.zeros(ncol, device="cuda") for _ in range(16)]
    for k in range(km):
        g = k // 3
        mass = rho[k] * dz[k]
        t0 = th[k] * pii[k]
        tt = tend(0, k)
        tn = t0 + dtf * tt
        bad |= ~torch.isfinite(tn) | ~((tn - t0).abs() <= dtmax)
        bad |= ~in_range(k, t0) | ~in_range(2 * 41 + k, rho[k]) | ~(tt.abs() <= scale * amax[k])
        th_new[k] = tn / pii[k]
        so = torch.zeros(ncol, device="cuda")
        sn = torch.zeros(ncol, device="cuda")
        for v in range(1, 7):
            old = qs[v - 1][k]
            tv = tend(v, k)
            x = old + dtf * tv
            bad |= ~torch.isfinite(x) | (x < -qtol)
            bad |= ~in_range((v + 2) * 41 + k, old) | ~(tv.abs() <= scale * amax[v * 41 + k])
            q_new[v - 1, k] = torch.where(x > 0, x, torch.zeros_like(x))
            so = so + old
            sn = sn + q_new[v - 1, k]
        gwo[g] = gwo[g] + mass * so
        gwn[g] = gwn[g] + mass * sn
    w_old = torch.zeros(ncol, device="cuda")
    w_new = torch.zeros(ncol, device="cuda")
    for g in range(16):
        w_old = w_old + gwo[g]
        w_new = w_new + gwn[g]
    bad |= (w_new - w_old) > wtol * w_old
    return bad, th_new, q_new, w_old, w_new


@pytest.mark.parametrize("config", ["config2", "config4"])
def test_fused_checks_and_update_equal_their_restatement(config):
    """The surrogate step (pywrf_wsm6_surrogate: network, checks and update
    in one launch, then the plain scheme on the failing columns) against the
    network's outputs (pywrf_wsm6_surrogate_outputs) put through a torch
    restatement of the checks and the update: the same columns fail (the
    fallback list, as a set), and on every passing column theta, the six
    mixing ratios, rainncv and rain are the same bits."""
    import pywrf_amd
    from pywrf_amd import _lib
    from pywrf_amd import surrogate as sg
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    sur = sg.Surrogate.load(os.path.join(ROOT, "pywrf_amd", "data", "wsm6_surrogate.pt"))
    net = sur.struct()
    s = synthetic_state(config, device="cuda")
    rk3_step(s, pywrf_amd.ops(), RunConfig())
    torch.cuda.synchronize()
    g = _lib.grid(s.th_phy.shape, s.grid.phys18())
    gc = g.contents
    ni, nj = (gc.ite - 1) - gc.its, (gc.jte - 1) - gc.jts
    ncol, km = ni * nj, (gc.kte - 1) - gc.kts
    m = s.moist
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the network's outputs
    ldy = (ncol + 127) // 128 * 128
    Y = torch.empty((288, ldy), device="cuda")
    ins = [s.th_phy, s.rho, s.pi_phy, s.p_phy, m[1], m[2], m[3], m[4], m[5], m[6]]
    rc = _lib.LIB.pywrf_wsm6_surrogate_outputs(g, ctypes.addressof(net), *[ctypes.c_void_p(t.data_ptr()) for t in ins],
                                               ctypes.c_void_p(Y.data_ptr()), ldy, 1, None, 0, stream)
    assert rc == 0, rc
    # the tile's columns as [level, column]
    jj = torch.arange(nj, device="cuda").repeat_interleave(ni) + gc.jts
    ii = torch.arange(ni, device="cuda").repeat(nj) + gc.its
    col = lambda a: a[:km, jj, ii]
    fields = {"th": col(s.th_phy), "rho": col(s.rho), "pii": col(s.pi_phy), "dz": col(s.dz8w),
              "q": [col(m[v]) for v in range(1, 7)]}
    rain0 = s.rainnc[jj, ii].clone()
    bad, th_new, q_new, w_old, w_new = _emulate_apply(sur, net, fields, Y[:, :ncol], gc, km, 54.0)
    # the surrogate step
    outs = [s.th_phy, s.rho, s.pi_phy, s.p_phy, s.dz8w, m[1], m[2], m[3], m[4], m[5], m[6], s.rainnc, s.rainncv,
            s.snownc, s.snowncv, s.sr, s.graupelnc, s.graupelncv]
    wsize = ctypes.c_longlong()
    assert _lib.LIB.pywrf_wsm6_surrogate_workspace(g, 65536, ctypes.byref(wsize)) == 0
    ws = torch.zeros(wsize.value // 4 + 1, dtype=torch.int32, device="cuda")
    nf = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = _lib.LIB.pywrf_wsm6_surrogate(g, ctypes.addressof(net), *[ctypes.c_void_p(t.data_ptr()) for t in outs[:5]],
                                       ctypes.c_double(54.0), *[ctypes.c_void_p(t.data_ptr()) for t in outs[5:]],
                                       ctypes.c_void_p(ws.data_ptr()), ctypes.c_longlong(wsize.value),
                                       ctypes.c_void_p(nf.data_ptr()), stream)
    assert rc == 0, rc
    torch.cuda.synchronize()
    n = int(nf)
    fell = torch.zeros(ncol, dtype=torch.bool, device="cuda")
    fell[ws[:n].long()] = True
    print(config, "columns", ncol, "fallback", n, "restated", int(bad.sum()))
    assert torch.equal(fell, bad)
    ok = ~bad
    assert int(ok.sum()) > 0
    assert torch.equal(col(s.th_phy)[:, ok], th_new[:, ok])
    for v in range(6):
        assert torch.equal(col(m[v + 1])[:, ok], q_new[v][:, ok]), v
    pr = torch.where(w_old - w_new > 0, w_old - w_new, torch.zeros_like(w_old))
    assert torch.equal(s.rainncv[jj, ii][ok], pr[ok])
    assert torch.equal(s.rainnc[jj, ii][ok], (rain0 + pr)[ok])
