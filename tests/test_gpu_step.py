"""Whole-step parity of the HIP path (every call on the device) with the
reference's FP32 trajectory (golden fixtures written by the reference under
tests/golden/make_golden.py), judged against an FP64 yardstick.

Two faithful FP32 implementations differ by rounding (transcendentals, the
ulp of full pressure inside the perturbation pressure p); fields of small
magnitude (v, w at the first steps) amplify that into normwise differences
above 1e-4 even between the reference and its CPU restatement.  So each
field must satisfy

    err(hip32, ref32) <= max(1e-4, 3 * err(ref32, oracle64))

where oracle64 is the oracle run in FP64 on the same seeded state -- the
reference's own FP32 distance to FP64 sets the scale.  Mass-like fields
(t, ph, mu, moist) are held to 1e-4 regardless.
"""
import os

import pytest
import torch

from parity import state, run, fields, golden_fields, errors

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STRICT = ("t", "ph", "mu", "moist")


def _hip(config, nsteps):
    import pywrf_amd
    s = state(config, device="cuda")
    run(s, pywrf_amd.ops(), nsteps)
    torch.cuda.synchronize()
    return fields(s)


def _oracle64(config, nsteps):
    import oracle
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    s = state(config, dtype=torch.float64)
    run(s, oracle.ops(), nsteps)
    return fields(s)


@pytest.mark.parametrize("config,nsteps", [("config1", 1), ("config1c", 1), ("config1", 10)])
def test_step_matches_reference(config, nsteps):
    ref32 = golden_fields(os.path.join(GOLD, "state_%s_s%d.npz" % (config, nsteps)))
    hip = _hip(config, nsteps)
    o64 = _oracle64(config, nsteps)
    e_hip = errors(hip, ref32)
    e_ref = errors({n: ref32[n] for n in ref32}, {n: o64[n] for n in ref32})
    e_hip64 = errors({n: hip[n] for n in ref32}, {n: o64[n] for n in ref32})
    print("%s x%d" % (config, nsteps))
    for n in ref32:
        print("  %-10s hip-ref32 %.2e   ref32-fp64 %.2e   hip-fp64 %.2e" % (n, e_hip[n], e_ref[n], e_hip64[n]))
    for n in ref32:
        tol = 1e-4 if n in STRICT else max(1e-4, 3 * e_ref[n])
        assert e_hip[n] <= tol, (n, e_hip[n], tol)
