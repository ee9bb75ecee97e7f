"""This build's driver and file loader against the reference's own time loop
(CPU).

tests/golden/calls_driver240.npz holds the reference's first RK3 step at
240 x 240 x 41 as its own driver (WrfSolver.solve, solver.py:40-1910) ran it
on datasets written from the seeded config2 state (make_driver_golden.py):
the name of every hot-path call in order and the statistics of what each
returned.  Here the same datasets are written as NetCDF files, loaded by
pywrf_amd.wrfio (the input fields, alb / al / mub, boundary record 0) into a
fresh state, and pywrf_amd.step.rk3_step drives the oracle through one step:
the call sequence must be the reference driver's, name for name, and every
call's results must agree in sum and sum|x| to 1e-4 of sum|x| (measured:
5.5e-6 at worst; ulp-level pow differences in the equation of state are
what remains).
"""
import os

import numpy as np
import pytest
import torch

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _loaded_state(tmp_path, device="cpu"):
    import driver_files
    from pywrf_amd import wrfio
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.state import allocate_state
    src = synthetic_state("config2")
    inp = driver_files.input_arrays(src)
    w = inp.pop("_W")
    paths = {k: str(tmp_path / k) for k in ("wrfinput", "wrfout", "wrfbdy")}
    driver_files.write_netcdf(paths["wrfinput"], inp)
    driver_files.write_netcdf(paths["wrfout"], {"W": w})
    driver_files.write_netcdf(paths["wrfbdy"], driver_files.bdy_arrays(src))
    g = src.grid
    s = allocate_state(g.nz, g.ny, g.nx, device=device)
    wrfio.load_input(s, paths["wrfinput"], paths["wrfout"])
    return s, paths


def _stats(t):
    x = t.detach().double()
    fin = torch.isfinite(x)
    xf = torch.where(fin, x, torch.zeros_like(x))
    return [float(xf.sum()), float(xf.abs().sum()), float(xf.abs().max()) if x.numel() else 0.0,
            float((~fin).sum())]


def test_driver_and_loader_match_the_reference_driver(tmp_path):
    import oracle
    from pywrf_amd import wrfio
    from pywrf_amd.step import rk3_step, RunConfig
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    ref = np.load(os.path.join(GOLD, "calls_driver240.npz"))
    ref_names = [str(n) for n in ref["names"]]
    s, paths = _loaded_state(tmp_path)
    rows = []
    O = oracle.ops()

    class Rec:
        def __getattr__(self, name):
            f = getattr(O, name)

            def call(*args):
                r = f(*args)
                outs = r if isinstance(r, tuple) else (r,)
                rows.append((name, [_stats(o) for o in outs if isinstance(o, torch.Tensor)]))
                return r
            return call

    cfg = RunConfig()
    wrfio.run(s, Rec(), cfg, 1, wrfbdy=paths["wrfbdy"])
    names = [n for n, _ in rows]
    assert len(names) == len(ref_names) == 401
    first = next((i for i, (a, b) in enumerate(zip(names, ref_names)) if a != b), None)
    assert first is None, (first, names[first - 2:first + 3], ref_names[first - 2:first + 3])
    worst = (0.0, None)
    for i, (name, st) in enumerate(rows):
        rs = ref["c%d" % i]
        assert len(st) == len(rs), (i, name)
        for a, b in zip(st, rs):
            assert a[3] == b[3], (i, name, "non-finite count")
            scale = max(b[1], 1e-30)
            err = max(abs(a[0] - b[0]), abs(a[1] - b[1])) / scale
            if err > worst[0]:
                worst = (err, (i, name))
    print("worst call: %.2e %s" % worst)
    assert worst[0] <= 1e-4, worst
