"""The C ABI (CPU): the library loads, exports every entry point that
include/pywrf_amd.h declares, its constants agree with the oracle's, the host
mirror covers every hot-path function of the reference, and argument and
index-set validation rejects bad calls before anything touches a device."""
import ctypes
import subprocess

import pytest


def test_library_exports_every_declared_symbol():
    from pywrf_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = set(_lib.DECLS)
    assert declared, "no declarations parsed"
    assert declared <= exported, sorted(declared - exported)
    assert len(declared) >= 62


def test_constants_agree_with_oracle():
    from pywrf_amd import _lib
    import oracle.constants as oc
    v = ctypes.c_double()
    names = [n for n in vars(oc) if not n.startswith("_") and isinstance(getattr(oc, n), float)]
    checked = 0
    for n in names:
        rc = _lib.LIB.pywrf_constant(n.encode(), ctypes.byref(v))
        if rc == 0:
            assert v.value == getattr(oc, n), n
            checked += 1
    assert checked >= 70
    assert _lib.LIB.pywrf_constant(b"no_such_constant", ctypes.byref(v)) == _lib.PYWRF_E_ARG


def test_host_mirror_covers_the_reference_functions():
    import oracle
    import pywrf_amd
    helpers = {"Box", "flux3", "flux5", "zeros_like3", "nislfv_rain_plm", "nislfv_rain_plm6"}
    need = set(vars(oracle.ops())) - helpers
    have = set(vars(pywrf_amd.ops()))
    assert need <= have, sorted(need - have)


def test_signatures_match_oracle():
    """Same positional parameter names as the reference (the oracle keeps them)."""
    import inspect
    import oracle
    import pywrf_amd
    o, h = oracle.ops(), pywrf_amd.ops()
    for name in vars(h):
        if hasattr(o, name):
            assert list(inspect.signature(getattr(h, name)).parameters) == \
                list(inspect.signature(getattr(o, name)).parameters), name


def _grid(n=40, nz=41):
    from pywrf_amd import _lib
    from pywrf_amd.step import Grid
    return _lib.grid((nz, n, n), Grid(nz, n, n).ix24)


def test_bad_flag_is_rejected_without_launch():
    from pywrf_amd import _lib
    g = _grid()
    assert _lib.LIB.pywrf_set_physical_bc3d(g, None, b"z", None) == _lib.PYWRF_E_ARG
    assert _lib.LIB.pywrf_set_physical_bc2d(_lib.grid((40, 40), (5, 35, 5, 35, 0, 40, 0, 40, 5, 35, 5, 35, 5, 35, 5,
                                                                 35)), None, b"q", None) == _lib.PYWRF_E_ARG
    with pytest.raises(NameError):
        _lib.check("pywrf_set_physical_bc3d", _lib.PYWRF_E_ARG)


def test_index_sets_outside_allocation_are_rejected():
    from pywrf_amd import _lib
    from pywrf_amd.step import Grid
    gr = Grid(41, 40, 40)
    bad = list(gr.ix18)
    bad[1] = 39                       # ide too close to the edge for a 5-point halo
    g = _lib.grid((41, 40, 40), bad)
    assert _lib.LIB.pywrf_advect_u(g, None, None, None, None, None, None, None, None, 1.0, 1.0, None, None) \
        == _lib.PYWRF_E_RANGE
    bad[1] = 40                       # ide at the array edge: the i+1 stencil would leave it
    g = _lib.grid((41, 40, 40), bad)
    assert _lib.LIB.pywrf_calc_ww_cp(g, None, None, None, None, None, None, 1.0, 1.0, None, None, None, None,
                                     None) == _lib.PYWRF_E_RANGE
    with pytest.raises(IndexError):
        _lib.check("pywrf_advect_u", _lib.PYWRF_E_RANGE)


def test_grid_packing():
    from pywrf_amd import _lib
    from pywrf_amd.step import Grid
    gr = Grid(41, 240, 240)
    g = _lib.grid((41, 240, 240), gr.ix24).contents
    assert (g.nz, g.ny, g.nx, g.ids, g.ide, g.kte, g.its, g.ite) == (41, 240, 240, 5, 235, 41, 5, 235)
    g2 = _lib.grid((240, 240), gr.ix2d).contents
    assert (g2.nz, g2.kds, g2.kde, g2.jte) == (1, 0, 1, 235)


def test_physics_hooks_leave_the_state_alone():
    """pbl_driver and sixth_order_diffusion are hooks with no arithmetic in the
    reference (wrf_dynamics.py:413-415, 3088-3097: None returned, tendency
    untouched); the mirror and the oracle keep that."""
    import torch
    import oracle
    import pywrf_amd
    for o in (pywrf_amd.ops(), oracle.ops()):
        assert o.pbl_driver() is None
        tend = torch.arange(24.0).reshape(2, 3, 4)
        before = tend.clone()
        idx = (0, 4, 0, 3, 0, 2) * 3
        assert o.sixth_order_diffusion("m", tend.clone(), tend, None, 54.0, None, None, None, 2, 0.12, *idx) is None
        assert torch.equal(tend, before)
