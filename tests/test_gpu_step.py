"""Whole-step parity of the HIP path (every call on the device) with the
reference's FP32 trajectory (golden fixtures written by the reference under
tests/golden/make_golden.py), judged against an FP64 yardstick.

Metrics, per field: normwise max|hip - ref32| / max|ref32| (bound
STEP_BOUND in parity.py: 1e-4 for the full fields u, t, ph, mu, moisture,
precipitation, 3e-4 for the perturbation fields v, w, p, al, whose FP32
rounding noise is set by the large quantities they are differences of -- the
reference's own FP32 run is 1.3e-4 to 5.3e-4 from its FP64 run on them), and
pointwise |hip - ref32| <= B (|ref32| + 0.01 max|ref32|) at every point
(POINT_BOUND: 1e-4 for u, t, ph, moisture, 1e-3 for precipitation, 5e-3 for
mu, 3e-2 for v, w, p, al, each at least twice the largest value measured
and at or above the reference's own FP32-against-FP64 distance).
The distances of both FP32 runs to the FP64 run of the oracle are printed
beside it as a diagnostic.
"""
import os

import pytest
import torch

from parity import (state, run, fields, golden_fields, errors, pointwise_errors, STEP_BOUND, POINT_BOUND,
                    POINT_FLOOR)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


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


@pytest.mark.parametrize("config,nsteps", [("config1", 1), ("config1c", 1), ("config1", 10), ("bdy24", 1),
                                            ("bdy24", 3)])
def test_step_matches_reference(config, nsteps):
    ref32 = golden_fields(os.path.join(GOLD, "state_%s_s%d.npz" % (config, nsteps)))
    hip = _hip(config, nsteps)
    o64 = _oracle64(config, nsteps)
    e_hip = errors(hip, ref32)
    e_ref = errors({n: ref32[n] for n in ref32}, {n: o64[n] for n in ref32})
    e_hip64 = errors({n: hip[n] for n in ref32}, {n: o64[n] for n in ref32})
    print("%s x%d" % (config, nsteps))
    p_hip = pointwise_errors(hip, ref32, POINT_FLOOR)
    p_ref = pointwise_errors({n: ref32[n] for n in ref32}, {n: o64[n] for n in ref32}, POINT_FLOOR)
    for n in ref32:
        print("  %-10s hip-ref32 %.2e   ref32-fp64 %.2e   hip-fp64 %.2e   pointwise hip-ref32 %.2e (ref32-fp64 %.2e)"
              % (n, e_hip[n], e_ref[n], e_hip64[n], p_hip[n], p_ref[n]))
    for n in ref32:
        assert e_hip[n] <= STEP_BOUND[n], (n, e_hip[n], STEP_BOUND[n])
        assert p_hip[n] <= POINT_BOUND[n], ("pointwise", n, p_hip[n], POINT_BOUND[n])
