"""Write the reference driver's own first step at 240 x 240 x 41 as fixtures.

Run in the build container only (needs /root/reference):

    python tests/golden/make_driver_golden.py

The reference's time loop (WrfSolver.solve, solver.py:40-1910) is run as it
is, reading its input, boundary and output datasets through a stand-in for
xarray.open_dataset that serves arrays written from the seeded config2 state
with the slices the driver itself copies (solver.py:213-281 for the input,
461-703 for the boundary records, 62 for w); the stand-ins are
tests/golden/driver_files.py.  Every hot-path function the driver calls is
wrapped to record its name and the statistics of its results ([sum, sum|x|,
max|x|, non-finite count] of each returned tensor), and the loop is stopped
after the first step's last call (set_w_surface, solver.py:1906), where the
driver's local state is read from its frame.

Written:
* calls_driver240.npz -- the names of the calls of the first step, in order,
  and the statistics of each call's returned tensors;
* state_driver240_s1.npz -- after the step, per prognostic field (u v w t ph
  p al mu and the six moisture species), [sum, sum|x|, max|x|] and the values
  at 4096 seeded sample points.

The fixtures are data; no reference source is copied.
"""
import json
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from ref_shim import load_reference, REF_ROOT, _GRID_GLOBALS   # noqa: E402
from dual import stats_of                                        # noqa: E402
import driver_files                                              # noqa: E402
import oracle                                                    # noqa: E402
from pywrf_amd.gen import synthetic_state                       # noqa: E402

N = 240
FIELDS = ("u", "v", "w", "t", "ph", "p", "al", "mu")


class StopAfterStep(Exception):
    pass


def main():
    t0 = time.time()
    src = synthetic_state("config2")
    files = driver_files.datasets(src)
    import types
    xr = types.ModuleType("xarray")
    xr.open_dataset = lambda path: files[driver_files.kind_of(path)]
    sys.modules["xarray"] = xr
    load_reference(N, N, dt=54)
    import pywrf.solver as solver
    torch.set_default_device("cpu")
    # the driver's own copies of the grid globals and the device
    ref_vals = {k: v for k, v in vars(sys.modules["pywrf.wrf_dynamics"]).items() if k in _GRID_GLOBALS}
    for k, v in ref_vals.items():
        if k in solver.__dict__:
            solver.__dict__[k] = v
    solver.__dict__["device"] = torch.device("cpu")
    names = set(vars(oracle.ops())) - {"Box", "flux3", "flux5", "zeros_like3"}
    rows = []
    final = {}

    def wrap(name, f):
        def call(*args):
            r = f(*args)
            outs = r if isinstance(r, tuple) else (r,)
            rows.append((name, [stats_of(o) for o in outs if isinstance(o, torch.Tensor)]))
            if name == "set_w_surface":
                frame = sys._getframe(1)
                loc = frame.f_locals
                for n in FIELDS:
                    final[n] = (r if n == "w" else loc[n]).detach().clone()
                final["moist"] = loc["moist"][1:7].detach().clone()
                raise StopAfterStep()
            return r
        return call

    for n in names:
        if n in solver.__dict__ and callable(solver.__dict__[n]):
            solver.__dict__[n] = wrap(n, solver.__dict__[n])
    torch.set_num_threads(8)
    try:
        solver.WrfSolver().solve()
    except StopAfterStep:
        pass
    print("driver step: %d calls in %.1f s" % (len(rows), time.time() - t0), flush=True)
    call_names = np.array([n for n, _ in rows])
    stats = {"names": call_names}
    for i, (_, st) in enumerate(rows):
        stats["c%d" % i] = np.array(st, dtype=np.float64).reshape(len(st), 4)
    np.savez_compressed(os.path.join(HERE, "calls_driver240.npz"), **stats)
    rng = np.random.default_rng(2024)
    state = {}
    for n, a in final.items():
        x = a.double().numpy()
        flat = x.reshape(-1)
        idx = rng.integers(0, flat.size, 4096)
        state[n + ".stats"] = np.array([flat.sum(), np.abs(flat).sum(), np.abs(flat).max()])
        state[n + ".idx"] = idx
        state[n + ".val"] = flat[idx].astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "state_driver240_s1.npz"), **state)
    meta_path = os.path.join(HERE, "meta.json")
    meta = json.load(open(meta_path))
    meta["driver240"] = {"calls": int(len(rows)), "torch": torch.__version__, "seconds": round(time.time() - t0, 1),
                         "source": "WrfSolver.solve (solver.py:40-1910) on stand-in datasets of config2"}
    json.dump(meta, open(meta_path, "w"), indent=1)


if __name__ == "__main__":
    main()
