"""The HIP step from files against the reference's own time loop at the
benchmark's size.

The seeded 240 x 240 x 41 state (config2) is written as WRF NetCDF files,
loaded by pywrf_amd.wrfio into a device state and advanced one RK3 step by
the HIP path; tests/golden/state_driver240_s1.npz holds the reference
driver's state after its first step from the same datasets
(make_driver_golden.py): at 4096 seeded points per field the values must
agree within the per-field step bounds (max|d| / max|ref| over the points,
STEP_BOUND), and the fields' sum|x| within the same bound.
"""
import os

import numpy as np
import pytest
import torch

from parity import STEP_BOUND
from test_oracle_driver import _loaded_state

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_hip_step_from_files_matches_the_reference_driver(tmp_path):
    import pywrf_amd
    from pywrf_amd import wrfio
    from pywrf_amd.step import RunConfig
    s, paths = _loaded_state(tmp_path, device="cuda")
    wrfio.run(s, pywrf_amd.ops(), RunConfig(), 1, wrfbdy=paths["wrfbdy"])
    torch.cuda.synchronize()
    ref = np.load(os.path.join(GOLD, "state_driver240_s1.npz"))
    fields = {n: getattr(s, n) for n in ("u", "v", "w", "t", "ph", "p", "al", "mu")}
    fields["moist"] = s.moist[1:7]
    for n, x in fields.items():
        flat = x.detach().double().cpu().reshape(-1).numpy()
        idx, val = ref[n + ".idx"], ref[n + ".val"].astype(np.float64)
        err = np.abs(flat[idx] - val).max() / max(np.abs(val).max(), 1e-30)
        st = ref[n + ".stats"]
        serr = abs(np.abs(flat).sum() - st[1]) / max(st[1], 1e-30)
        print("%-6s sample %.2e  sum|x| %.2e" % (n, err, serr))
        bound = STEP_BOUND.get(n, 1e-4)
        assert err <= bound, n
        assert serr <= bound, n
