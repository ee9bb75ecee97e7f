"""Host-side logic of the gradient API (pywrf_amd.adjoint), without a GPU:
the acoustic loop's input layout, the recorder's call-sequence check and the
positions through which the recorded calls hand over the tendencies and the
saved fields (checked against the reference's own signatures, read from the
host mirror, which repeats them)."""
import inspect

import pytest
import torch


def test_acoustic_loop_rejects_a_wrong_number_of_inputs():
    from pywrf_amd import adjoint
    x = [torch.zeros(1) for _ in range(len(adjoint.ACOUSTIC_STATE) + 1)]
    with pytest.raises(ValueError):
        adjoint.AcousticLoop.apply([], *x)


def test_loop_recorder_refuses_an_unexpected_call_sequence():
    from pywrf_amd import adjoint

    class Ops:
        def __getattr__(self, name):
            return lambda *a: None

    rec = adjoint._LoopRecorder(Ops(), 1)
    rec.advance_uv(torch.zeros(1))
    rec.advance_w(torch.zeros(1))
    rec.calc_mu_uv_1()
    names = [n for n, _ in rec.calls]
    assert names == ["advance_uv", "advance_w"]
    assert names != list(adjoint.SUBSTEP_CALLS)


def _pos(fn, name):
    return list(inspect.signature(fn).parameters).index(name)


def test_recorded_positions_match_the_reference_signatures():
    """The argument positions the loop's Functions and loop_forcing /
    loop_saved use, against the parameter names of the mirror functions
    (which keep the reference's names and order)."""
    from pywrf_amd import adjoint, wrf_boundary, wrf_dynamics as d
    W = adjoint.AdvanceW
    assert (W.W, W.RWT, W.WW, W.WSAVE, W.U, W.V, W.MUAVE, W.MUTS, W.T2AVE, W.T2, W.T1, W.PH, W.PH1, W.PHT) == tuple(
        _pos(d.advance_w, n) for n in ("w_ori", "rw_tend", "ww", "w_save", "u", "v", "muave", "muts", "t_2ave_ori",
                                       "t_2", "t_1", "ph_ori", "ph_1", "ph_tend"))
    mt = ("u", "v", "ww_ori", "mu_ori", "muave_ori", "muts_ori", "mudf_ori", "t_ori", "t_ave_ori", "ft", "mu_tend")
    assert adjoint.AdvanceMuT.DIFF == tuple(_pos(d.advance_mu_t, n) for n in mt)
    assert adjoint.AdvanceMuT.EXTRA == tuple(_pos(d.advance_mu_t, n) for n in ("u_1", "v_1", "ww_1", "t_1"))
    assert adjoint.AdvanceUV.DIFF == tuple(
        _pos(d.advance_uv, n) for n in ("u_ori", "ru_tend", "v_ori", "rv_tend", "p", "ph", "al", "mu", "mudf"))
    assert adjoint.CalcPRho.DIFF == tuple(
        _pos(d.calc_p_rho, n) for n in ("al_ori", "p_ori", "ph", "t_2", "pm1", "mu", "mut"))
    assert _pos(d.calc_p_rho, "t_1") == 5
    assert adjoint.SumFlux.DIFF == tuple(_pos(d.sumflux, n) for n in ("ru", "rv", "ww", "ru_m", "rv_m", "ww_m"))
    assert [_pos(d.sumflux, n) for n in ("u_lin", "v_lin", "ww_lin")] == [3, 4, 5]
    assert [_pos(wrf_boundary.spec_bdyupdate_ph, n) for n in ("ph_save", "field", "field_tend", "mu_tend",
                                                               "muts")] == [0, 1, 2, 3, 4]
