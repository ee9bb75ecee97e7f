"""Helpers shared by the parity tests: seeded states, trajectories through any
implementation of the hot-path functions, and the normwise error metric.

Error metric: max|x - y| / max|y| over finite entries, per field (rel_err in
tests/golden/dual.py); non-finite entries must coincide.
"""
import numpy as np
import torch

from dual import rel_err  # noqa: F401  (re-exported)

PROG = ("u", "v", "w", "t", "ph", "p", "al", "mu")


def state(config, device="cpu", dtype=torch.float32):
    from pywrf_amd.gen import synthetic_state
    return synthetic_state(config, device=device, dtype=dtype)


def run(s, ops, nsteps, cfg=None):
    from pywrf_amd.step import rk3_step, RunConfig
    cfg = cfg or RunConfig()
    for _ in range(nsteps):
        rk3_step(s, ops, cfg)
    return s


def fields(s):
    """The prognostic fields compared after a step, as CPU float64 tensors."""
    out = {n: getattr(s, n).detach().double().cpu() for n in PROG}
    out["moist"] = s.moist[1:7].detach().double().cpu()
    for n in ("rainnc", "snownc", "graupelnc"):
        out[n] = getattr(s, n).detach().double().cpu()
    return out


def golden_fields(path):
    g = np.load(path)
    return {n: torch.from_numpy(g[n]).double() for n in g.files}


def errors(a, b):
    return {n: rel_err(a[n], b[n]) for n in b}
