"""Call-by-call parity of the HIP kernels with the oracle.

The oracle drives RK3 steps; at every call of a hot-path function the HIP
entry point runs on a device copy of the same arguments and every returned
tensor and every (possibly updated) argument is compared with the oracle's,
normwise (max|hip - oracle| / max|oracle|).  The oracle's results then carry
the step on, so each call is judged on identical inputs.

Bounds:
* every function not listed in INEXACT: bit-exact, every output of every
  call (error exactly 0) -- the kernels keep the reference's operation order,
  rounding and summation order;
* the functions in INEXACT (transcendentals from the device library, or a
  sum the kernel forms in another order): normwise error <= 1e-4;
* the lateral-boundary functions are also judged on the points they change
  (the strips), against the strip's own values: bit-exact.

Workloads: 40 x 40 x 41 (config1, config1c: one step), 24 x 24 x 41 with
time-varying lateral boundaries (bdy24: three steps, so the boundary
tendencies, the relaxation forcing and the boundary time dtbc are live in
every call), 120 x 120 x 41 (config4: one step) and the benchmark's
240 x 240 x 41 state (config2: one step).
"""
import pytest
import torch

from parity import state, run

pytestmark = pytest.mark.gpu
TOL = 1e-4
BOUNDARY = ("relax_bdytend", "relax_bdy_scalar", "spec_bdytend", "spec_bdy_scalar", "spec_bdy_final",
            "spec_bdyupdate", "spec_bdyupdate_ph", "zero_grad_bdy", "flow_dep_bdy")
# the functions whose results may differ from the oracle's in the last bits, and why
INEXACT = {
    "advance_w": "sinf of the damping profile (device library)",
    "calc_p_rho_phi": "powf of the equation of state (device library)",
    "calculate_km_kh": "expf of the stability limiter (device library)",
    "microphysics_driver": "expf / logf / fractional powers of WSM6 (device library)",
    "moist_physics_prep_em": "powf of the Exner function (device library)",
    "phy_prep": "expf / logf of the hydrostatic pressure (device library)",
    "calc_cq": "moisture sum over the species in another order",
    "rhs_ph": "the vertical advection terms summed in another order",
}
CALLS_PER_STEP = 401


def _dual(config, nsteps):
    import oracle
    import pywrf_amd
    from dual import DualOps
    torch.manual_seed(0)
    torch.set_num_threads(16)
    ops = DualOps(oracle.ops(), pywrf_amd.ops(), b_device="cuda", local=BOUNDARY)
    s = state(config)
    run(s, ops, nsteps)
    torch.cuda.synchronize()
    return ops.records


def _worst(records, local=False):
    worst = {}
    for r in records:
        for which, e in r["errs"]:
            if which.startswith("local_") != local:
                continue
            key = r["name"]
            if e > worst.get(key, (-1,))[0]:
                worst[key] = (e, r["call"], which)
    return worst


def _check(config, records, nsteps):
    worst = _worst(records)
    print(config, "calls:", len(records))
    for name in sorted(worst, key=lambda n: -worst[n][0]):
        print("  %-32s %.2e (call %d, %s)" % ((name,) + worst[name]))
    bad = {n: w for n, w in worst.items() if not w[0] <= TOL}
    assert not bad, bad
    inexact = {n: w for n, w in worst.items() if n not in INEXACT and w[0] != 0.0}
    assert not inexact, inexact
    strips = _worst(records, local=True)
    for name in sorted(strips):
        print("  strip %-26s %.2e (call %d, %s)" % ((name,) + strips[name]))
    bad = {n: w for n, w in strips.items() if w[0] != 0.0}
    assert not bad, bad
    assert len(records) == CALLS_PER_STEP * nsteps


@pytest.mark.parametrize("config", ["config1", "config1c"])
def test_every_call_matches_oracle(config):
    _check(config, _dual(config, 1), 1)


def test_every_call_matches_oracle_live_boundaries():
    """Three steps of the 24 x 24 x 41 state whose boundary tendencies are
    not zero: relax_bdytend / relax_bdy_scalar see a real forcing, and
    spec_bdytend / spec_bdy_scalar / spec_bdy_final copy non-zero tendency
    strips, in steps 2 and 3 as well as 1."""
    records = _dual("bdy24", 3)
    _check("bdy24", records, 3)
    names = [r["name"] for r in records]
    for fn, per_step in (("relax_bdytend", 5), ("relax_bdy_scalar", 1), ("spec_bdytend", 15),
                         ("spec_bdy_scalar", 1), ("spec_bdy_final", 5)):
        assert names.count(fn) == 3 * per_step, fn


def test_every_call_matches_oracle_120():
    """A second grid size (120 x 120 x 41, config4), so that index-set logic is
    exercised away from the 40-point case."""
    _check("config4", _dual("config4", 1), 1)


@pytest.mark.timeout(900)
def test_every_call_matches_oracle_240():
    """The benchmark's own state (240 x 240 x 41, config2), one step."""
    _check("config2", _dual("config2", 1), 1)
