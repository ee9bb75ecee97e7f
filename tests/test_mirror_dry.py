"""The host mirror's argument handling, without a device (CPU).

The mirror (pywrf_amd.wrf_dynamics / wrf_physics / wrf_boundary) is driven
through whole RK3 steps with the library's entry points replaced by a stub
that records each call and launches nothing, and with host tensors accepted
as if they were device tensors.  This checks, for every call a step makes:
the operand checks accept the step's own arguments (so they can only reject
mistakes), the lateral-boundary arrays of the reference driver's layout are
passed in place (no copy), and mismatched operands -- a 2-D array where a
3-D one is read, a field of another grid, a boundary array of the wrong
width or layout -- are rejected with IndexError before any launch.
"""
import ctypes

import pytest
import torch


@pytest.fixture
def dry(monkeypatch):
    from pywrf_amd import _lib, wrf_dynamics, wrf_physics, wrf_boundary
    calls = []

    class Stub:
        def __getattr__(self, name):
            def f(*args):
                calls.append((name, args))
                return 0
            return f

    monkeypatch.setattr(_lib, "LIB", Stub())
    monkeypatch.setattr(_lib, "_require_device", lambda t: None)
    for mod in (_lib, wrf_dynamics, wrf_physics, wrf_boundary):
        monkeypatch.setattr(mod, "stream", lambda: ctypes.c_void_p(0), raising=False)
    return calls


@pytest.mark.parametrize("config", ["config1c", "bdy24"])
def test_every_call_of_a_step_passes_the_operand_checks(dry, config):
    import pywrf_amd
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config)
    rk3_step(s, pywrf_amd.ops(), RunConfig())
    names = {n for n, _ in dry}
    assert len(dry) > 150
    for fn in ("pywrf_relax_bdytend", "pywrf_spec_bdytend_group", "pywrf_boundary_epilogue",
               "pywrf_microphysics_driver",
               "pywrf_advance_w", "pywrf_rk_scalar_tend_pair"):
        assert any(n.startswith(fn) for n in names), fn
    # the driver's boundary arrays are read in place: their pointers reach the library unchanged
    mine = {a.data_ptr() for arrs in s.bdy.values() for a in arrs}
    passed = set()
    for n, args in dry:
        for a in args:
            if isinstance(a, ctypes.c_void_p) and type(a).__name__ == "BoundaryPtr":
                passed.add(a.value)
    assert passed and passed <= mine


def test_mismatched_operands_are_rejected(dry):
    import pywrf_amd
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import SPEC_BDY_WIDTH, SPEC_ZONE, RELAX_ZONE
    o = pywrf_amd.ops()
    s = synthetic_state("bdy24")
    g = s.grid
    I18, I24 = g.ix18, g.ix24
    # a 2-D array where the kernel reads a 3-D one
    with pytest.raises(IndexError):
        o.calc_alt(s.alt, s.mu, s.alb, *I18)
    # a field of another horizontal size
    with pytest.raises(IndexError):
        o.calc_alt(s.alt, s.al[:, :-1, :].contiguous(), s.alb, *I18)
    # fewer levels than the grid
    with pytest.raises(IndexError):
        o.calc_php(s.php, s.ph[:20].contiguous(), s.phb, *I18)
    # a level array that is too short
    with pytest.raises(IndexError):
        o.calc_ww_cp(s.u, s.v, s.mu, s.mub, s.muu, s.muv, s.c1h[:10].contiguous(), s.c2h, s.ww, 1e-4, 1e-4,
                     s.msftx, s.msfty, s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.dnw, *I18)
    b = s.bdy["u"]
    args = lambda arrs, width=SPEC_BDY_WIDTH: (s.ru, s.u_save, *arrs, 'u', width, SPEC_ZONE, RELAX_ZONE, 54.0,
                                                s.fcx, s.gcx, *I24)
    n = len(dry)
    o.relax_bdytend(*args(b))
    assert len(dry) == n + 1
    # boundary strips of another width, or x and y arrays exchanged
    with pytest.raises(IndexError):
        o.relax_bdytend(*args(b, width=4))
    with pytest.raises(IndexError):
        o.relax_bdytend(*args((b[0][:, :, :4],) + b[1:]))
    with pytest.raises(IndexError):
        o.relax_bdytend(*args((b[2], b[3], b[0], b[1]) + b[4:]))
    assert len(dry) == n + 1


def test_boundary_arrays_in_another_layout_are_copied_once(dry):
    """A contiguous (kz, ny, 5) array (not the driver's layout) is accepted
    through a copy in the kernels' layout, made once while it is unchanged."""
    from pywrf_amd import _lib
    t = torch.arange(41 * 24 * 5, dtype=torch.float32).reshape(41, 24, 5)
    c1 = _lib.boundary_layout(t, "x", 41, 24)
    c2 = _lib.boundary_layout(t, "x", 41, 24)
    assert c1 is c2 and torch.equal(c1, t) and c1.stride() == (24, 1, 41 * 24)
    t.add_(1)                      # a new version: a new copy
    c3 = _lib.boundary_layout(t, "x", 41, 24)
    assert c3 is not c1 and torch.equal(c3, t)
    y = torch.zeros(5, 41, 24).permute(1, 0, 2)
    assert _lib.boundary_layout(y, "y", 41, 24) is y
