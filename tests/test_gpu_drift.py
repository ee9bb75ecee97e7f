"""Accuracy of the device step at the benchmark's shape, against FP64.

Ten RK3 steps of the 240 x 240 x 41 seeded state (config2) are run three
ways: on the device through the HIP kernels (FP32), by the oracle on the host
in FP32 (the reference's own arithmetic) and by the oracle in FP64 (the
accuracy yardstick; the reference's FP64 switch is its default dtype).  Per
prognostic field the drift rms(x - x64) / rms(x64) of the device run must be
within 15 % of the FP32 oracle's (plus 1e-7): the device arithmetic loses
no more accuracy than the reference's does, up to the spread between two
FP32 roundings of the same arithmetic (measured: the device drift is 0.97 to
1.06 times the oracle's over the ten fields).
"""
import os
import time

import pytest
import torch

pytestmark = pytest.mark.gpu
NSTEPS = 10
FIELDS = ("u", "v", "w", "t", "ph", "p", "al", "mu")


def _fields(s):
    out = {n: getattr(s, n).detach().double().cpu() for n in FIELDS}
    out["qv"] = s.moist[1].detach().double().cpu()
    out["condensate"] = s.moist[2:7].sum(0).detach().double().cpu()
    return out


@pytest.mark.timeout(1200)
def test_drift_240_no_larger_than_reference_fp32():
    import oracle
    import pywrf_amd
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.state import clone_state
    from pywrf_amd.step import rk3_step, RunConfig
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    cfg = RunConfig()
    s32 = synthetic_state("config2")
    dev = clone_state(s32, device="cuda")
    o = pywrf_amd.ops()
    for _ in range(NSTEPS):
        rk3_step(dev, o, cfg)
    torch.cuda.synchronize()
    hip = _fields(dev)
    assert all(bool(torch.isfinite(v).all()) for v in hip.values())
    ref = {}
    for dt in (torch.float64, torch.float32):
        s = synthetic_state("config2", dtype=dt)
        t0 = time.perf_counter()
        for n in range(NSTEPS):
            rk3_step(s, oracle.ops(), cfg)
            print("oracle %s step %d" % (dt, n + 1), flush=True)
        print("oracle %s: %d steps in %.1f s" % (dt, NSTEPS, time.perf_counter() - t0), flush=True)
        ref[dt] = _fields(s)
    rms = lambda x: float(torch.sqrt(torch.mean(x * x)))
    r64 = ref[torch.float64]
    bad = {}
    for n in r64:
        den = rms(r64[n]) or 1.0
        e_hip = rms(hip[n] - r64[n]) / den
        e_ref = rms(ref[torch.float32][n] - r64[n]) / den
        print("  %-10s device %.3e   reference fp32 %.3e" % (n, e_hip, e_ref))
        if not e_hip <= 1.15 * e_ref + 1e-7:
            bad[n] = (e_hip, e_ref)
    assert not bad, bad
