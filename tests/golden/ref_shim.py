"""Load the reference PyWRF kernels on the CPU at an arbitrary grid size.

Used ONLY by make_golden.py in the build container, where the reference is
mounted at /root/reference; nothing that runs on the GPU box imports this.

The reference modules import xarray and netCDF4 at top level but use them only
for file I/O, so empty stand-in modules are registered for the import;
config_params.py:24 makes cuda:0 the default device, so the default device is
reset to the CPU after import; the grid is frozen into module globals by the
star import (config_params.py:7-13), so those globals are patched to the
requested allocation size.
"""
import sys
import types
from types import SimpleNamespace

import torch

REF_ROOT = "/root/reference"
_GRID_GLOBALS = ("nzall", "nyall", "nxall", "nyfull", "nxfull", "nzfull", "ide", "jde", "ite", "jte",
                 "ipe", "jpe", "ime", "jme", "dt", "time_step", "p_top")


def load_reference(ny, nx, nz=41, dt=54, p_top=5000.0):
    for name in ("xarray", "netCDF4"):
        sys.modules.setdefault(name, types.ModuleType(name))
    if REF_ROOT not in sys.path:
        sys.path.insert(0, REF_ROOT)
    import pywrf.config_params as cp
    torch.set_default_device("cpu")
    import pywrf.wrf_dynamics as rd
    import pywrf.wrf_physics as rp
    import pywrf.wrf_boundary as rb
    vals = dict(nzall=nz, nyall=ny, nxall=nx, nyfull=ny, nxfull=nx, nzfull=nz, ide=nx - 5, jde=ny - 5,
                ite=nx - 5, jte=ny - 5, ipe=nx - 5, jpe=ny - 5, ime=nx, jme=ny, dt=dt, time_step=dt,
                p_top=p_top)
    for mod in (cp, rd, rp, rb):
        for k in _GRID_GLOBALS:
            if k in mod.__dict__:
                mod.__dict__[k] = vals[k]
        if "device" in mod.__dict__:
            mod.__dict__["device"] = torch.device("cpu")
    ns = {}
    for mod in (rd, rp, rb):
        for k, v in vars(mod).items():
            if callable(v) and not k.startswith("_") and getattr(v, "__module__", "").startswith("pywrf"):
                ns[k] = v
    return SimpleNamespace(**ns)
