"""CPU oracle for the PyWRF ARW RK3 step -- TEST INFRASTRUCTURE ONLY.

A restatement, in PyTorch on the CPU, of the reference functions on the hot
path (pywrf/wrf_dynamics.py, pywrf/wrf_physics.py, pywrf/wrf_boundary.py) with
the reference's positional signatures, index boxes and expression order.  It
runs in FP32 or FP64 (the dtype of its inputs).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it, and only
as the checker; the product path (pywrf_amd) never imports it.

Pinning: every function is checked against golden vectors written by driving
the real reference (tests/golden/make_golden.py) and, where mathematics fixes
the answer, against closed forms (tests/test_oracle_invariants.py).
"""
from types import SimpleNamespace

from . import wrf_dynamics, wrf_physics, wrf_boundary, constants


def ops():
    """Namespace with every hot-path function under its reference name."""
    ns = {}
    for mod in (wrf_dynamics, wrf_physics, wrf_boundary):
        for k, v in vars(mod).items():
            if callable(v) and not k.startswith('_'):
                ns[k] = v
    return SimpleNamespace(**ns)
