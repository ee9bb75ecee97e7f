"""MI355X-native PyWRF ARW RK3 step: HIP kernels behind the reference's
wrf_dynamics / wrf_physics / wrf_boundary function signatures.

    import pywrf_amd
    ops = pywrf_amd.ops()          # every hot-path function, HIP-backed
    pywrf_amd.step.rk3_step(state, ops, pywrf_amd.step.RunConfig())

Importing the kernel modules loads libpywrf_amd.so (built by
`python -m pywrf_amd.build`); a missing library raises ImportError.
"""
from types import SimpleNamespace


def ops():
    """Namespace of the HIP-backed hot-path functions under the reference's names."""
    from . import wrf_dynamics, wrf_physics, wrf_boundary
    ns = {}
    for mod in (wrf_dynamics, wrf_physics, wrf_boundary):
        for name in dir(mod):
            f = getattr(mod, name)
            if callable(f) and not name.startswith("_") and getattr(f, "__module__", None) == mod.__name__:
                ns[name] = f
    return SimpleNamespace(**ns)
