"""Gradients through parts of the step: hand-written adjoint kernels behind
torch.autograd.Function.

The reference claims differentiability through torch autograd (README.md:37-40)
but runs its step under no_grad (solver.py:456, 710, 1748).  Here gradients
come from the library's own adjoint kernels (csrc/adjoint.hip), each the
transpose of a linear part of the step:

* AdvectScalar -- advect_scalar (wrf_dynamics.py:1841-2028) as a function of
  the field, the mass fluxes held fixed.  Backward: pywrf_advect_scalar_adjoint.
* AdvectU -- advect_u (wrf_dynamics.py:1112-1307) as a function of u, the
  mass fluxes held fixed.  Backward: pywrf_advect_u_adjoint (scatter form).
* AdvectV, AdvectW -- advect_v (1332-1532) and advect_w (1535-1838) likewise.
* AdvanceW -- the vertical acoustic step advance_w (wrf_dynamics.py:3641-3796)
  as a function of w, ph, t_2, t_2ave, u, v, muts, muave and ww.  Backward: pywrf_advance_w_adjoint
  (the damping, the transposed column solve, the explicit predictor and the
  geopotential right-hand side, in reverse).
* AdvanceUV -- the horizontal acoustic step advance_uv
  (wrf_dynamics.py:3391-3525) as a function of u, v, ru_tend, rv_tend, p, ph,
  al, mu and mudf.  Backward: pywrf_advance_uv_adjoint (each u or v column
  adds its sensitivities into the two columns its differences read).
* AdvanceMuT -- the column-mass and theta acoustic step advance_mu_t
  (wrf_dynamics.py:3528-3638) as a function of u, v, ww, mu, muave, muts,
  mudf, t, t_ave, ft and mu_tend.  Backward: pywrf_advance_mu_t_adjoint (the
  eta-velocity recurrence as suffix sums down each column).
* SumFlux, SpecBdyUpdatePh, ZeroGradBdy, CalcPRho, OpenBoundaryFill -- the
  calls that close an acoustic substep: sumflux (wrf_dynamics.py:3799-3839),
  spec_bdyupdate_ph (wrf_boundary.py:287-358), zero_grad_bdy (361-405),
  calc_p_rho (wrf_dynamics.py:3269-3313) and a run of open-boundary fills
  (164-278), each on copies of its in-place operands.  With the three above
  and the zone updates (fixed increments) they chain into the gradient of the
  whole acoustic loop.
* SmallStepPrep, CalcMuUV1, SmallStepFinish -- small_step_prep past the
  first stage (wrf_dynamics.py:3152-3266), calc_mu_uv_1 (3844-3860) and
  small_step_finish (3863-3978); with the loop they make acoustic_stage, an
  RK stage's whole acoustic part (record_acoustic_stage records it).
* TridiagSolve -- the column tridiagonal solve of advance_w
  (wrf_dynamics.py:3758-3765) as a function of its coefficients a, alpha,
  gamma and right-hand side.  Backward: pywrf_tridiag_solve_adjoint
  (the transposed Thomas sweeps).
* TracerStep -- one RK3 step of the scalar transport of a tracer
  (rk_scalar_tend's advection, wrf_dynamics.py:3982-4051, and
  rk_update_scalar, 4055-4119, for the three stages of solver.py:1548-1636)
  with the step's mass fluxes and column masses recorded from a model run
  (record_transport).  The map is linear in the tracer, so its backward needs
  no stored tracer states: per stage in reverse, pywrf_rk_update_scalar_adjoint
  then pywrf_advect_scalar_adjoint.
* SurrogateNet -- the WSM6 column surrogate (pywrf_amd.surrogate) as a
  function of its standardised inputs: the backward products on the matrix
  cores (the same GEMM kernel as training, the ReLU derivative in its
  epilogue).

    from pywrf_amd import adjoint
    rec = adjoint.record_transport(state, ops, cfg, nsteps=6)
    q = q0 + bias                                  # bias.requires_grad_()
    for step in rec:
        q = adjoint.TracerStep.apply(q, step)
    loss = ((q - target) ** 2).mean(); loss.backward()
"""
import torch

from ._lib import call, flag, grid, ptr, stream
from . import surrogate as _sg


def _idx18(g):
    return (g.ids, g.ide, g.jds, g.jde, g.kds, g.kde, g.ims, g.ime, g.jms, g.jme, g.kms, g.kme,
            g.its, g.ite, g.jts, g.jte, g.kts, g.kte)


def _f32(t):
    if not (t.is_cuda and t.dtype == torch.float32):
        raise TypeError("expected a float32 CUDA tensor")
    return t.contiguous()


# ----------------------------------------------------------------------- advection
def advect_scalar_tendency(field, ru, rv, rom, msftx, fzm, fzp, rdx, rdy, rdzw, idx):
    """The tendency advect_scalar adds to a zeroed array (HIP)."""
    g = grid(field.shape, idx)
    tend = torch.empty_like(field)
    call("pywrf_advect_scalar_fresh", g, ptr(field), ptr(tend), ptr(ru), ptr(rv), ptr(rom), ptr(msftx), ptr(fzm),
         ptr(fzp), float(rdx), float(rdy), ptr(rdzw), stream())
    return tend


def advect_scalar_adjoint(lam_tend, ru, rv, rom, msftx, fzm, fzp, rdx, rdy, rdzw, idx, out=None):
    """A^T lam_tend for the advection map A of advect_scalar_tendency; added to
    `out` when given, else a fresh array."""
    g = grid(lam_tend.shape, idx)
    acc = 1 if out is not None else 0
    if out is None:
        out = torch.empty_like(lam_tend)
    call("pywrf_advect_scalar_adjoint", g, ptr(lam_tend), ptr(ru), ptr(rv), ptr(rom), ptr(msftx), ptr(fzm), ptr(fzp),
         float(rdx), float(rdy), ptr(rdzw), ptr(out), acc, stream())
    return out


class AdvectScalar(torch.autograd.Function):
    """tendency = advect_scalar(field) with fixed fluxes; gradient w.r.t. field."""

    @staticmethod
    def forward(ctx, field, ru, rv, rom, msftx, fzm, fzp, rdx, rdy, rdzw, idx):
        field = _f32(field)
        ctx.args = (ru, rv, rom, msftx, fzm, fzp, rdx, rdy, rdzw, tuple(idx))
        return advect_scalar_tendency(field, ru, rv, rom, msftx, fzm, fzp, rdx, rdy, rdzw, idx)

    @staticmethod
    def backward(ctx, lam):
        g = advect_scalar_adjoint(_f32(lam), *ctx.args)
        return (g,) + (None,) * 10


class AdvectU(torch.autograd.Function):
    """The tendency advect_u (wrf_dynamics.py:1112-1307) adds to a zeroed
    array, as a function of u (the mass fluxes held fixed).  args: advect_u's
    positional arguments in the reference's order (its tendency is not
    read).  Backward: pywrf_advect_u_adjoint."""

    @staticmethod
    def forward(ctx, u, args):
        from . import wrf_dynamics
        args = list(args)
        args[0] = _f32(u)
        args[2] = torch.empty_like(args[0])
        ctx.args = args
        return wrf_dynamics.advect_u_fresh(*args)

    @staticmethod
    def backward(ctx, lam):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        out = torch.empty_like(a[0])
        # (u, u_old, tendency, ru, rv, rom, c1, c2, mut, time_step, msfux, msfuy, msfvx, msfvy, msftx, msfty,
        #  fzm, fzp, rdx, rdy, rdzw, idx)
        call("pywrf_advect_u_adjoint", g, ptr(_f32(lam)), ptr(a[0]), ptr(a[3]), ptr(a[4]), ptr(a[5]), ptr(a[10]),
             ptr(a[16]), ptr(a[17]), float(a[18]), float(a[19]), ptr(a[20]), ptr(out), stream())
        return out, None


class AdvectV(torch.autograd.Function):
    """The tendency advect_v (wrf_dynamics.py:1332-1532) adds to a zeroed
    array, as a function of v (the mass fluxes held fixed).  args: advect_v's
    positional arguments in the reference's order.  Backward:
    pywrf_advect_v_adjoint."""

    @staticmethod
    def forward(ctx, v, args):
        from . import wrf_dynamics
        args = list(args)
        args[0] = _f32(v)
        args[2] = torch.empty_like(args[0])
        ctx.args = args
        return wrf_dynamics.advect_v_fresh(*args)

    @staticmethod
    def backward(ctx, lam):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        out = torch.empty_like(a[0])
        call("pywrf_advect_v_adjoint", g, ptr(_f32(lam)), ptr(a[0]), ptr(a[3]), ptr(a[4]), ptr(a[5]), ptr(a[12]),
             ptr(a[13]), ptr(a[16]), ptr(a[17]), float(a[18]), float(a[19]), ptr(a[20]), ptr(out), stream())
        return out, None


class AdvectW(torch.autograd.Function):
    """The tendency advect_w (wrf_dynamics.py:1535-1838) adds to a zeroed
    array, as a function of w (the mass fluxes held fixed).  args: advect_w's
    positional arguments in the reference's order.  Backward:
    pywrf_advect_w_adjoint."""

    @staticmethod
    def forward(ctx, w, args):
        from . import wrf_dynamics
        args = list(args)
        args[0] = _f32(w)
        args[2] = torch.empty_like(args[0])
        ctx.args = args
        return wrf_dynamics.advect_w_fresh(*args)

    @staticmethod
    def backward(ctx, lam):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        out = torch.empty_like(a[0])
        call("pywrf_advect_w_adjoint", g, ptr(_f32(lam)), ptr(a[3]), ptr(a[4]), ptr(a[5]), ptr(a[14]), ptr(a[16]),
             ptr(a[17]), float(a[18]), float(a[19]), ptr(a[20]), ptr(out), stream())
        return out, None


# ----------------------------------------------------------------------- tridiagonal solve
def tridiag_solve(a, alpha, gamma, r, idx):
    """x with the column solve of advance_w applied to r (HIP; r unchanged)."""
    x = r.clone()
    g = grid(r.shape, idx)
    call("pywrf_tridiag_solve", g, ptr(a), ptr(alpha), ptr(gamma), ptr(x), stream())
    return x


class TridiagSolve(torch.autograd.Function):
    """x = solve(a, alpha, gamma; r) over advance_w's columns (levels
    1..kte-1; other points pass r through); gradients w.r.t. all four."""

    @staticmethod
    def forward(ctx, a, alpha, gamma, r, idx):
        a, alpha, gamma, r = (_f32(t) for t in (a, alpha, gamma, r))
        x = tridiag_solve(a, alpha, gamma, r, idx)
        ctx.save_for_backward(a, alpha, gamma, r, x)
        ctx.idx = tuple(idx)
        return x

    @staticmethod
    def backward(ctx, lam):
        a, alpha, gamma, r, x = ctx.saved_tensors
        g = grid(r.shape, ctx.idx)
        lr = _f32(lam).clone()
        ga, galpha, ggamma = (torch.zeros_like(r) for _ in range(3))
        work = torch.empty_like(r)
        call("pywrf_tridiag_solve_adjoint", g, ptr(a), ptr(alpha), ptr(gamma), ptr(r), ptr(x), ptr(lr), ptr(ga),
             ptr(galpha), ptr(ggamma), ptr(work), stream())
        return ga, galpha, ggamma, lr, None


# ----------------------------------------------------------------------- acoustic step: advance_w
class AdvanceW(torch.autograd.Function):
    """(t_2ave, w, ph) = advance_w(...) (wrf_dynamics.py:3641-3796) with the
    gradient with respect to w, ph, t_2, t_2ave, u, v, muts, muave, ww, the
    forcing rw_tend and ph_tend and w_save (every other argument is a
    coefficient held fixed).  args: advance_w's positional arguments in the
    reference's order; rw_tend, ph_tend, w_save None: those of args."""

    # positions in advance_w's signature
    W, RWT, WW, WSAVE, U, V, MUAVE, MUTS, T2AVE, T2, T1, PH, PH1, PHT = 0, 1, 2, 3, 4, 5, 8, 9, 18, 19, 20, 21, 22, 24
    DIFF = (W, PH, T2, T2AVE, U, V, MUTS, MUAVE, WW)        # the inputs before args
    EXTRA = (RWT, PHT, WSAVE, T1, PH1)                      # the optional ones after it

    @staticmethod
    def forward(ctx, w, ph, t_2, t_2ave, u, v, muts, muave, ww, args, rw_tend=None, ph_tend=None, w_save=None,
                t_1=None, ph_1=None):
        from . import wrf_dynamics
        args = list(args)
        rw_tend = args[AdvanceW.RWT] if rw_tend is None else rw_tend
        ph_tend = args[AdvanceW.PHT] if ph_tend is None else ph_tend
        w_save = args[AdvanceW.WSAVE] if w_save is None else w_save
        t_1 = args[AdvanceW.T1] if t_1 is None else t_1
        ph_1 = args[AdvanceW.PH1] if ph_1 is None else ph_1
        for pos, x in zip(AdvanceW.DIFF + AdvanceW.EXTRA,
                          (w, ph, t_2, t_2ave, u, v, muts, muave, ww, rw_tend, ph_tend, w_save, t_1, ph_1)):
            args[pos] = _f32(x)
        ctx.args = args
        t2a, wo, pho = wrf_dynamics.advance_w(*args)
        ctx.outs = (t2a, wo)
        return t2a, wo, pho

    @staticmethod
    def backward(ctx, lT, lW, lP):
        from .wrf_dynamics import _cf
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        z = lambda t: torch.zeros_like(a[0]) if t is None else _f32(t)
        lT, lW, lP = z(lT), z(lW), z(lP)
        cf, _keep = _cf(a[43], a[44], a[45])
        out = [torch.empty_like(a[0]) for _ in range(4)]
        uv = [torch.empty_like(a[4]), torch.empty_like(a[5])]
        planes = [torch.empty_like(a[9]), torch.empty_like(a[8])]
        lww = torch.empty_like(a[2])
        forcing = [torch.empty_like(a[0]) for _ in range(4)]       # rw_tend, ph_tend, w_save, t_1
        lph1 = torch.empty_like(a[0])
        work = [torch.empty_like(a[0]) for _ in range(2)]
        t2a_out, w_out = ctx.outs
        # advance_w(w, rw_tend, ww, w_save, u, v, mu1, mut, muave, muts, c1h, c2h, c1f, c2f, c3h, c4h, c3f, c4f,
        #           t_2ave, t_2, t_1, ph, ph_1, phb, ph_tend, ht, c2a, cqw, alt, alb, a, alpha, gamma,
        #           rdx, rdy, dts, t0, epssm, dnw, fnm, fnp, rdnw, rdn, cf1, cf2, cf3, msftx, msfty, top_lid, idx)
        call("pywrf_advance_w_adjoint", g, ptr(lT), ptr(lW), ptr(lP), ptr(a[7]), ptr(a[9]), ptr(a[20]),
             ptr(a[22]), ptr(a[23]), ptr(a[26]), ptr(a[27]), ptr(a[28]), ptr(a[30]), ptr(a[31]), ptr(a[32]),
             float(a[35]), float(a[36]), float(a[37]), ptr(a[41]), ptr(a[42]), ptr(a[47]), ptr(t2a_out),
             ptr(w_out), ptr(a[25]), ptr(a[46]), cf, float(a[33]), float(a[34]), ptr(a[39]), ptr(a[40]),
             *[ptr(t) for t in out], *[ptr(t) for t in uv], *[ptr(t) for t in planes], ptr(lww),
             *[ptr(t) for t in forcing], ptr(a[2]), ptr(a[3]), ptr(lph1), *[ptr(t) for t in work], stream())
        lw, lph, lt2, lt2a = out
        lu, lv = uv
        lmuts, lmuave = planes
        res = (lw, lph, lt2, lt2a, lu, lv, lmuts, lmuave, lww, None, *forcing, lph1)
        need = ctx.needs_input_grad          # the optional forcing inputs: only those given as tensors
        return tuple(r if (n < 10 or need[n]) else None for n, r in enumerate(res[:len(need)]))


class AdvanceUV(torch.autograd.Function):
    """(u, v) = advance_uv(...) (wrf_dynamics.py:3391-3525) with the gradient
    with respect to the inputs at DIFF (u, ru_tend, v, rv_tend, p, ph, al, mu,
    mudf); every other argument is a coefficient held fixed.  args:
    advance_uv's positional arguments in the reference's order."""

    # positions in advance_uv's signature
    DIFF = (0, 1, 2, 3, 4, 6, 9, 10, 15)

    @staticmethod
    def forward(ctx, u, ru_tend, v, rv_tend, p, ph, al, mu, mudf, args):
        from . import wrf_dynamics
        args = list(args)
        for pos, t in zip(AdvanceUV.DIFF, (u, ru_tend, v, rv_tend, p, ph, al, mu, mudf)):
            args[pos] = _f32(t)
        ctx.args = args
        return wrf_dynamics.advance_uv(*args)

    @staticmethod
    def backward(ctx, lu, lv):
        from .wrf_dynamics import _cf
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        lu = torch.zeros_like(a[0]) if lu is None else _f32(lu)
        lv = torch.zeros_like(a[2]) if lv is None else _f32(lv)
        cf, _keep = _cf(a[32], a[33], a[34])
        # outputs in the order u, v, p, ph, al, mu, mudf, ru_tend, rv_tend
        out = [torch.empty_like(a[n]) for n in (0, 2, 4, 6, 9, 10, 15, 1, 3)]
        call("pywrf_advance_uv_adjoint", g, ptr(lu), ptr(lv), ptr(a[5]), ptr(a[7]), ptr(a[8]), ptr(a[11]),
             ptr(a[12]), ptr(a[13]), ptr(a[14]), ptr(a[24]), ptr(a[25]), ptr(a[26]), ptr(a[27]), ptr(a[28]),
             float(a[29]), float(a[30]), float(a[31]), cf, ptr(a[35]), ptr(a[36]), float(a[37]), ptr(a[38]),
             int(a[39]), *[ptr(t) for t in out], stream())
        gu, gv, gp, gph, gal, gmu, gmudf, gru, grv = out
        return gu, gru, gv, grv, gp, gph, gal, gmu, gmudf, None


class AdvanceMuT(torch.autograd.Function):
    """The seven fresh results of advance_mu_t (wrf_dynamics.py:3528-3638):
    (ww, t, t_ave, muave, muts, mudf, mu), with the gradient with respect to
    the inputs at DIFF (u, v, ww, mu, muave, muts, mudf, t, t_ave, ft,
    mu_tend) and, given after args, the large-step u_1, v_1, ww_1 and t_1;
    mut, muu, muv and the map factors are held fixed.  The reference runs the
    mass updates under no_grad (its autograd would leave them out); this is
    the full derivative.  args: advance_mu_t's positional arguments in the
    reference's order; u_1 .. t_1 None: those of args."""

    # positions in advance_mu_t's signature
    DIFF = (2, 4, 0, 6, 8, 9, 12, 24, 26, 27, 28)
    EXTRA = (3, 5, 1, 25)                   # u_1, v_1, ww_1, t_1
    OUT = (0, 2, 3, 7, 8, 9, 10)            # positions of the fresh results in its 11-tuple

    @staticmethod
    def forward(ctx, u, v, ww, mu, muave, muts, mudf, t, t_ave, ft, mu_tend, args, u_1=None, v_1=None, ww_1=None,
                t_1=None):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(AdvanceMuT.DIFF, (u, v, ww, mu, muave, muts, mudf, t, t_ave, ft, mu_tend)):
            args[pos] = _f32(x)
        for pos, x in zip(AdvanceMuT.EXTRA, (u_1, v_1, ww_1, t_1)):
            if x is not None:
                args[pos] = _f32(x)
        ctx.args = args
        r = wrf_dynamics.advance_mu_t(*args)
        ctx.wwo = r[0]
        return tuple(r[n] for n in AdvanceMuT.OUT)

    @staticmethod
    def backward(ctx, lww, lt, ltave, lmuave, lmuts, lmudf, lmu):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        like = (a[0], a[24], a[26], a[8], a[9], a[12], a[6])
        lam = [torch.zeros_like(x) if l is None else _f32(l)
               for l, x in zip((lww, lt, ltave, lmuave, lmuts, lmudf, lmu), like)]
        # (ww_ori, ww_1, u, u_1, v, v_1, mu_ori, mut, muave_ori, muts_ori, muu, muv, mudf_ori, c1h, c2h, c1f, c2f,
        #  c3h, c4h, c3f, c4f, uam, vam, wwam, t_ori, t_1, t_ave_ori, ft, mu_tend, rdx, rdy, dts, epssm,
        #  dnw, fnm, fnp, rdnw, msfux, msfuy, msfvx, msfvx_inv, msfvy, msftx, msfty, step, idx)
        out = [torch.empty_like(a[n]) for n in (0, 2, 4, 6, 8, 9, 12, 24, 26, 27, 28)]
        big = [torch.empty_like(a[n]) for n in (3, 5, 1, 25)]
        call("pywrf_advance_mu_t_adjoint", g, *[ptr(x) for x in lam], ptr(a[25]), ptr(a[42]), ptr(a[43]),
             ptr(a[33]), ptr(a[34]), ptr(a[35]), ptr(a[36]), float(a[29]), float(a[30]), float(a[31]),
             float(a[32]), *[ptr(x) for x in out], ptr(ctx.wwo), ptr(a[2]), ptr(a[4]), ptr(a[10]), ptr(a[11]),
             ptr(a[38]), ptr(a[40]), *[ptr(x) for x in big], stream())
        gww, gu, gv, gmu, gmuave, gmuts, gmudf, gt, gtave, gft, gmut = out
        res = (gu, gv, gww, gmu, gmuave, gmuts, gmudf, gt, gtave, gft, gmut, None, *big)
        need = ctx.needs_input_grad          # the optional inputs: only those given as tensors
        return tuple(r if (n < 12 or need[n]) else None for n, r in enumerate(res[:len(need)]))


class CalcPRho(torch.autograd.Function):
    """(al, p, pm1) = calc_p_rho(...) (wrf_dynamics.py:3269-3313) with the
    gradient with respect to al, p, ph, t_2, pm1, mu and mut (alt, t_1 and c2a
    held fixed).  The incoming pm1 is left as it is (the call's in-place update
    goes to a copy).  args: calc_p_rho's positional arguments in the
    reference's order."""

    # positions in calc_p_rho's signature
    DIFF = (0, 1, 2, 4, 7, 8, 9)

    @staticmethod
    def forward(ctx, al, p, ph, t_2, pm1, mu, mut, args, t_1=None):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(CalcPRho.DIFF, (al, p, ph, t_2, pm1, mu, mut)):
            args[pos] = _f32(x)
        if t_1 is not None:
            args[5] = _f32(t_1)
        args[7] = args[7].clone()
        ctx.args = args
        return wrf_dynamics.calc_p_rho(*args)

    @staticmethod
    def backward(ctx, lal, lp, lpm1):
        a = ctx.args
        g = grid(a[1].shape, tuple(a[-18:]))
        lal, lp, lpm1 = [torch.zeros_like(a[1]) if l is None else _f32(l) for l in (lal, lp, lpm1)]
        out = [torch.empty_like(a[1]) for _ in range(5)] + [torch.empty_like(a[8]), torch.empty_like(a[9]),
                                                             torch.empty_like(a[1])]
        # (al_ori, p_ori, ph, alt, t_2, t_1, c2a, pm1, mu, mut, c1h, c2h, c1f, c2f, c3h, c4h, c3f, c4f, znu, t0,
        #  rdnw, dnw, smdiv, non_hydrostatic, step, idx)
        call("pywrf_calc_p_rho_adjoint", g, ptr(lal), ptr(lp), ptr(lpm1), ptr(a[2]), ptr(a[3]), ptr(a[4]),
             ptr(a[5]), ptr(a[6]), ptr(a[8]), ptr(a[9]), float(a[19]), ptr(a[20]), float(a[22]), int(a[24]),
             *[ptr(t) for t in out], stream())
        gal, gp, gpm1, gph, gt2, gmu, gmut, gt1 = out
        need = ctx.needs_input_grad
        return (gal, gp, gph, gt2, gpm1, gmu, gmut, None, gt1 if need[-1] and len(need) > 8 else None)[:len(need)]


class SumFlux(torch.autograd.Function):
    """(ru_m, rv_m, ww_m) = sumflux(...) (wrf_dynamics.py:3799-3839) with the
    gradient with respect to ru, rv, ww and the three accumulators (the call's
    in-place update goes to copies).  args: sumflux's positional arguments in
    the reference's order."""

    # positions in sumflux's signature
    DIFF = (0, 1, 2, 16, 17, 18)

    @staticmethod
    def forward(ctx, ru, rv, ww, ru_m, rv_m, ww_m, args, u_lin=None, v_lin=None, ww_lin=None):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(SumFlux.DIFF, (ru, rv, ww, ru_m, rv_m, ww_m)):
            args[pos] = _f32(x)
        for pos, x in ((3, u_lin), (4, v_lin), (5, ww_lin)):
            if x is not None:
                args[pos] = _f32(x)
        for pos in (16, 17, 18):
            args[pos] = args[pos].clone()
        ctx.args = args
        return wrf_dynamics.sumflux(*args)

    @staticmethod
    def backward(ctx, l1, l2, l3):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        lam = [torch.zeros_like(a[n]) if l is None else _f32(l) for l, n in zip((l1, l2, l3), (16, 17, 18))]
        out = [torch.empty_like(a[n]) for n in (0, 1, 2, 16, 17, 18)]
        lin = [torch.empty_like(a[n]) for n in (3, 4, 5)]
        call("pywrf_sumflux_adjoint", g, *[ptr(x) for x in lam], int(a[25]), int(a[26]),
             *[ptr(x) for x in out], ptr(a[6]), ptr(a[7]), ptr(a[21]), ptr(a[23]), *[ptr(x) for x in lin],
             stream())
        need = ctx.needs_input_grad
        return tuple(r if (n < 7 or need[n]) else None for n, r in enumerate((*out, None, *lin)[:len(need)]))


class SpecBdyUpdatePh(torch.autograd.Function):
    """spec_bdyupdate_ph (wrf_boundary.py:287-358) with the gradient with
    respect to the field, muts, field_tend and mu_tend (ph_save held fixed);
    the update goes to a copy of the field.  args: its positional arguments
    in the reference's order; field_tend, mu_tend None: those of args."""

    @staticmethod
    def forward(ctx, field, muts, args, field_tend=None, mu_tend=None, ph_save=None):
        from . import wrf_boundary
        args = list(args)
        ctx.f_in = _f32(field)
        args[1], args[4] = ctx.f_in.clone(), _f32(muts)
        if ph_save is not None:
            args[0] = _f32(ph_save)
        if field_tend is not None:
            args[2] = _f32(field_tend)
        if mu_tend is not None:
            args[3] = _f32(mu_tend)
        ctx.args = args
        return wrf_boundary.spec_bdyupdate_ph(*args)

    @staticmethod
    def backward(ctx, lf):
        a = ctx.args
        g = grid(a[1].shape, tuple(a[-24:]))
        lam = torch.zeros_like(a[1]) if lf is None else _f32(lf).clone()
        lmuts, lft, lmt = torch.zeros_like(a[4]), torch.zeros_like(a[2]), torch.zeros_like(a[3])
        lps = torch.zeros_like(a[0])
        call("pywrf_spec_bdyupdate_ph_adjoint", g, ptr(a[0]), ptr(ctx.f_in), ptr(a[2]), ptr(a[3]), ptr(a[4]),
             float(a[7]), flag(a[8]), ptr(lam), ptr(lmuts), ptr(lft), ptr(lmt), ptr(lps), stream())
        need = ctx.needs_input_grad          # the optional inputs: only those given as tensors
        return tuple(r if (n < 3 or need[n]) else None
                     for n, r in enumerate((lam, lmuts, None, lft, lmt, lps)[:len(need)]))


class ZeroGradBdy(torch.autograd.Function):
    """zero_grad_bdy (wrf_boundary.py:361-405) on a copy of the field, with
    its gradient.  args: its positional arguments in the reference's order."""

    @staticmethod
    def forward(ctx, field, args):
        from . import wrf_boundary
        args = list(args)
        args[0] = _f32(field).clone()
        ctx.args = args
        return wrf_boundary.zero_grad_bdy(*args)

    @staticmethod
    def backward(ctx, lf):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-24:]))
        lam = torch.zeros_like(a[0]) if lf is None else _f32(lf)
        out = torch.empty_like(a[0])
        call("pywrf_zero_grad_bdy_adjoint", g, ptr(lam), flag(a[1]), ptr(out), stream())
        return out, None


class OpenBoundaryFill(torch.autograd.Function):
    """A run of set_physical_bc3d / set_physical_bc2d calls
    (wrf_dynamics.py:164-278) on copies of distinct fields, with the gradient
    with respect to each.  kinds: [(variable, is_2d)] per field; idx24 /
    idx2d: the index tuples the individual calls would take."""

    @staticmethod
    def forward(ctx, kinds, idx24, idx2d, *fields):
        from . import wrf_dynamics
        outs = [_f32(f).clone() for f in fields]
        wrf_dynamics.set_physical_bc_group([(o, v, two) for o, (v, two) in zip(outs, kinds)], idx24, idx2d)
        ctx.kinds, ctx.idx = kinds, (idx24, idx2d)
        ctx.shapes = [o.shape for o in outs]
        ctx.like = outs[0]
        return tuple(outs)

    @staticmethod
    def backward(ctx, *lams):
        import ctypes
        kinds, (idx24, idx2d) = ctx.kinds, ctx.idx
        lam = [torch.zeros(sh, dtype=torch.float32, device=ctx.like.device) if l is None else _f32(l).clone()
               for l, sh in zip(lams, ctx.shapes)]
        sh3 = next((sh for sh, (_, two) in zip(ctx.shapes, kinds) if not two), (idx24[5],) + tuple(ctx.shapes[0][-2:]))
        sh2 = next((sh for sh, (_, two) in zip(ctx.shapes, kinds) if two), tuple(ctx.shapes[0][-2:]))
        n = len(lam)
        ptrs = (ctypes.c_void_p * n)(*[ptr(t).value for t in lam])
        flags = b"".join(v.encode() for v, _ in kinds)
        two = (ctypes.c_int * n)(*[1 if t else 0 for _, t in kinds])
        call("pywrf_set_physical_bc_group_adjoint", grid(sh3, idx24), grid(sh2, idx2d), n, ptrs, flags, two,
             stream())
        return (None, None, None, *lam)


class SpecBdyUpdateGroup(torch.autograd.Function):
    """A run of spec_bdyupdate calls (wrf_boundary.py:228-284) on copies of
    distinct fields: each adds its tendency times dt on the outer ring, so the
    Jacobian with respect to each field is the identity and that with respect
    to each tendency dt on the ring (the same update applied to zeros).  call:
    the grouped call's arguments (items, dt, idx24, idx2d) as the step makes
    it; args: the n fields that replace the items' fields, then n tendencies
    (None: the item's own, held fixed).

        u2, v2 = SpecBdyUpdateGroup.apply(call, u, v, None, None)
    """

    @staticmethod
    def forward(ctx, call, *args):
        from . import wrf_boundary
        items, dt, ix24, ix2d = call
        n = len(items)
        fields, tends = args[:n], args[n:2 * n]
        tends = [it[1] if t is None else _f32(t) for t, it in zip(tends, items)]
        outs = [_f32(f).clone() for f in fields]
        wrf_boundary.spec_bdyupdate_group([(o, t, it[2], it[3]) for o, t, it in zip(outs, tends, items)], dt, ix24,
                                          ix2d)
        ctx.call, ctx.live = call, [t is not None for t in args[n:2 * n]]
        ctx.tends = tends
        return tuple(outs)

    @staticmethod
    def backward(ctx, *lams):
        from . import wrf_boundary
        items, dt, ix24, ix2d = ctx.call
        lt = [torch.zeros_like(t) if (on and l is not None) else None
              for on, l, t in zip(ctx.live, lams, ctx.tends)]
        run = [(z, _f32(l), it[2], it[3]) for z, l, it in zip(lt, lams, items) if z is not None]
        if run:
            wrf_boundary.spec_bdyupdate_group(run, dt, ix24, ix2d)
        return (None, *lams, *lt)


class SmallStepPrep(torch.autograd.Function):
    """small_step_prep (wrf_dynamics.py:3152-3266) past the first RK stage as a
    function of u_1, u_2, v_1, v_2, w_1, w_2, t_1, t_2, ph_1, ph_2, mu_1, mu_2
    and ww (pb, p, alt -- so c2a -- and the column masses mub, muu, muv, mut
    held fixed), on copies of its in-place operands.  Returns (u_save,
    v_save, w_save, t_save, ph_save, ww_save, u, v, w, t, ph, mu_2, muts,
    muus, muvs, mu_save).  args: its positional arguments in the reference's
    order."""

    POS = (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 34)
    INPLACE = (11, 12, 14, 16, 18, 19, 28, 29, 30, 31, 32, 33, 35, 36)
    OUT = (5, 6, 7, 8, 9, 16, 10, 11, 12, 13, 14, 18, 21, 19, 20, 23)   # positions in its 24 results

    @staticmethod
    def forward(ctx, u_1, u_2, v_1, v_2, w_1, w_2, t_1, t_2, ph_1, ph_2, mu_1, mu_2, ww, args):
        from . import wrf_dynamics
        args = list(args)
        if int(args[49]) == 1:
            raise ValueError("SmallStepPrep: the first RK stage (copies) is not differentiated")
        for pos, x in zip(SmallStepPrep.POS, (u_1, u_2, v_1, v_2, w_1, w_2, t_1, t_2, ph_1, ph_2, mu_1, mu_2, ww)):
            args[pos] = _f32(x)
        for pos in SmallStepPrep.INPLACE:
            args[pos] = args[pos].clone()
        ctx.args = args
        r = wrf_dynamics.small_step_prep(*args)
        ctx.res = (r[19], r[20], r[21])          # muus, muvs, muts
        return tuple(r[n] for n in SmallStepPrep.OUT)

    @staticmethod
    def backward(ctx, *lams):
        a = ctx.args
        g = grid(a[1].shape, tuple(a[-18:]))
        like = [a[28], a[29], a[30], a[31], a[32], a[35], a[1], a[3], a[5], a[7], a[9], a[12], a[18], a[14], a[16],
                a[33]]
        lam = [torch.zeros_like(x) if l is None else _f32(l) for l, x in zip(lams, like)]
        out3 = [torch.empty_like(a[n]) for n in (1, 3, 5, 7, 9, 0, 2, 4, 6, 8, 34)]
        out2 = [torch.empty_like(a[11]), torch.empty_like(a[12])]
        work = torch.empty((3,) + tuple(a[12].shape), device=a[12].device, dtype=torch.float32)
        muus, muvs, muts = ctx.res
        call("pywrf_small_step_prep_adjoint", g, int(a[49]), *[ptr(x) for x in lam], ptr(a[0]), ptr(a[2]),
             ptr(a[4]), ptr(a[6]), ptr(muus), ptr(muvs), ptr(muts), ptr(a[13]), ptr(a[15]), ptr(a[17]), ptr(a[41]),
             ptr(a[43]), ptr(a[46]), *[ptr(x) for x in out3], *[ptr(x) for x in out2], ptr(work), stream())
        gu2, gv2, gw2, gt2, gph2, gu1, gv1, gw1, gt1, gph1, gww = out3
        gmu1, gmu2 = out2
        return gu1, gu2, gv1, gv2, gw1, gw2, gt1, gt2, gph1, gph2, gmu1, gmu2, gww, None


class CalcMuUV1(torch.autograd.Function):
    """(muu, muv) = calc_mu_uv_1(mu, muu, muv, ...) (wrf_dynamics.py:3844-3860)
    on copies of muu, muv, with the gradient with respect to the three."""

    @staticmethod
    def forward(ctx, mu, muu, muv, args):
        from . import wrf_dynamics
        args = list(args)
        args[0], args[1], args[2] = _f32(mu), _f32(muu).clone(), _f32(muv).clone()
        ctx.args = args
        return wrf_dynamics.calc_mu_uv_1(*args)

    @staticmethod
    def backward(ctx, lu, lv):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        lu = torch.zeros_like(a[1]) if lu is None else _f32(lu)
        lv = torch.zeros_like(a[2]) if lv is None else _f32(lv)
        gmu, gu, gv = torch.empty_like(a[0]), torch.empty_like(a[1]), torch.empty_like(a[2])
        call("pywrf_calc_mu_uv_1_adjoint", g, ptr(lu), ptr(lv), ptr(gmu), ptr(gu), ptr(gv), stream())
        return gmu, gu, gv, None


class SmallStepFinish(torch.autograd.Function):
    """(u, v, w, ph, ww, t, mu_2) = small_step_finish(...)
    (wrf_dynamics.py:3863-3978) as a function of the acoustic u, v, w, t, ph,
    ww, mu_2, the saved u_save .. ph_save, ww1, mu_save and the column masses
    muus, muvs, muts (mut, muu, muv, the map factors and h_diabatic held
    fixed), on copies of its in-place operands.  args: its positional
    arguments in the reference's order."""

    POS = (0, 2, 4, 6, 8, 10, 12, 28, 29, 30, 31, 32, 11, 33, 17, 19, 15)

    @staticmethod
    def forward(ctx, u, v, w, t, ph, ww, mu, u_save, v_save, w_save, t_save, ph_save, ww1, mu_save, muus, muvs,
                muts, args):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(SmallStepFinish.POS, (u, v, w, t, ph, ww, mu, u_save, v_save, w_save, t_save, ph_save,
                                                ww1, mu_save, muus, muvs, muts)):
            args[pos] = _f32(x)
        args[10], args[12] = args[10].clone(), args[12].clone()
        ctx.args = args
        r = wrf_dynamics.small_step_finish(*args)
        ctx.res = (r[0], r[1], r[2], r[5])       # u, v, w, t
        return r

    @staticmethod
    def backward(ctx, lu, lv, lw, lph, lww, lt, lmu):
        a = ctx.args
        g = grid(a[0].shape, tuple(a[-18:]))
        like = (a[0], a[2], a[4], a[8], a[10], a[6], a[12])
        lam = [torch.zeros_like(x) if l is None else _f32(l) for l, x in zip((lu, lv, lw, lph, lww, lt, lmu), like)]
        out3 = [torch.empty_like(a[n]) for n in (0, 2, 4, 8, 10, 6, 28, 29, 30, 32, 11, 31)]
        out2 = [torch.empty_like(a[n]) for n in (17, 19, 15, 12, 33)]
        uo, vo, wo, to = ctx.res
        call("pywrf_small_step_finish_adjoint", g, *[ptr(x) for x in lam], ptr(uo), ptr(vo), ptr(wo), ptr(to),
             ptr(a[16]), ptr(a[17]), ptr(a[18]), ptr(a[19]), ptr(a[14]), ptr(a[15]), ptr(a[35]), ptr(a[36]),
             ptr(a[39]), *[ptr(x) for x in out3], *[ptr(x) for x in out2], stream())
        gu, gv, gw, gph, gww, gt, gus, gvs, gws, gps, gww1, gts = out3
        gmuus, gmuvs, gmuts, gmu, gmusv = out2
        return gu, gv, gw, gt, gph, gww, gmu, gus, gvs, gws, gts, gps, gww1, gmusv, gmuus, gmuvs, gmuts, None


# ----------------------------------------------------------------------- the acoustic loop
# The 17 fields an RK stage's acoustic loop (solver.py:1349-1493) carries from
# one substep to the next.  Everything else its calls read is fixed for the
# stage: the large-step tendencies (which can also be given as inputs, FORCING
# below), the saved *_save / *_1 fields, the coefficients a, alpha, gamma,
# c2a, alt, php, cq*, mut and the map factors.
ACOUSTIC_STATE = ("u", "v", "p", "ph", "al", "mu", "mudf", "ww", "muave", "muts", "t", "t_ave", "w", "pm1",
                  "ru_m", "rv_m", "ww_m")
SUBSTEP_CALLS = ("advance_uv", "spec_bdyupdate_group", "advance_mu_t", "spec_bdyupdate_group", "advance_w",
                 "sumflux", "spec_bdyupdate_ph_zero_grad", "calc_p_rho", "set_physical_bc_group")


class _LoopRecorder:
    """The step's functions, keeping copies of the arguments of every call of
    the acoustic loop that starts at the n-th call of advance_uv, up to the
    loop's end (the next calc_mu_uv_1)."""

    def __init__(self, ops, n):
        self._ops, self._n, self._seen, self.calls, self._done = ops, n, 0, [], False

    @staticmethod
    def _copy(a):
        if isinstance(a, torch.Tensor):
            return a.clone()
        if isinstance(a, (list, tuple)):
            return type(a)(_LoopRecorder._copy(x) for x in a)
        return a

    def __getattr__(self, name):
        f = getattr(self._ops, name)

        def rec(*args):
            if name == "advance_uv":
                self._seen += 1
            if name == "calc_mu_uv_1" and self._seen >= self._n:
                self._done = True
            if self._seen >= self._n and not self._done:
                self.calls.append((name, self._copy(list(args))))
            return f(*args)
        return rec


def record_acoustic_loop(s, ops, cfg, stage=2):
    """Run one RK3 step on state s and keep the acoustic loop of RK stage
    `stage` (1, 2 or 3): returns (x, substeps) with x the 17 carried fields at
    the loop's start (ACOUSTIC_STATE order) and substeps one list of the nine
    calls' arguments per substep.  The step must take the grouped calls (the
    library's own ops do)."""
    from .step import rk3_step
    rec = _LoopRecorder(ops, {1: 1, 2: 2, 3: 2 + cfg.soundsteps // 2}[stage])
    rk3_step(s, rec, cfg)
    names = [n for n, _ in rec.calls]
    m = len(names) // len(SUBSTEP_CALLS)
    if m == 0 or names != list(SUBSTEP_CALLS) * m:
        raise RuntimeError("record_acoustic_loop: unexpected call sequence %s" % names[:12])
    subs = [[a for _, a in rec.calls[9 * q:9 * q + 9]] for q in range(m)]
    uv, _, mt, _, wa, _, _, cp, _ = subs[0]
    first = {"u": uv[0], "v": uv[2], "p": uv[4], "ph": uv[6], "al": uv[9], "mu": uv[10], "mudf": uv[15],
             "ww": mt[0], "muave": mt[8], "muts": mt[9], "t": mt[24], "t_ave": mt[26], "w": wa[0], "pm1": cp[7],
             "ru_m": mt[21], "rv_m": mt[22], "ww_m": mt[23]}
    return [first[n] for n in ACOUSTIC_STATE], subs


FORCING = ("ru_tend", "rv_tend", "rw_tend", "t_tend", "ph_tend", "mu_tend")


SAVED = ("u_save", "v_save", "w_save", "t_save", "ph_save", "ww1")


def loop_forcing(substeps):
    """The six large-step tendencies (FORCING order) a recorded loop reads."""
    uv, _, mt, _, wa = substeps[0][:5]
    return [uv[1], uv[3], wa[1], mt[27], wa[24], mt[28]]


def loop_saved(substeps):
    """The six fields small_step_prep saved for the stage (SAVED order) that a
    recorded loop reads: u_save, v_save (advance_mu_t's u_1, v_1, sumflux's
    u_lin, v_lin), w_save (advance_w's damping), t_save (the t_1 of
    advance_mu_t, advance_w and calc_p_rho), ph_save (advance_w's ph_1,
    spec_bdyupdate_ph's ph_save) and ww1 (advance_mu_t's ww_1, sumflux's
    ww_lin)."""
    _, _, mt, _, wa = substeps[0][:5]
    return [mt[3], mt[5], wa[3], mt[25], wa[22], mt[1]]


def acoustic_substep(x, calls, forcing=None, saved=None):
    """One acoustic substep (solver.py:1349-1493) as autograd Functions:
    advance_uv, the u/v zone update, advance_mu_t, the t/mu/muts zone update,
    advance_w, sumflux, spec_bdyupdate_ph, zero_grad_bdy, calc_p_rho and the
    six open-boundary fills.  x: the 17 carried fields (ACOUSTIC_STATE
    order); calls: the nine recorded calls' arguments; forcing: None, or the
    six large-step tendencies (FORCING order) to use, and differentiate, in
    place of the recorded ones wherever the substep reads them; saved: the
    same for the six saved fields (SAVED order).  Returns the 17 fields
    after the substep."""
    st = dict(zip(ACOUSTIC_STATE, x))
    uv, z1, mt, z2, wa, sf, pz, cp, bc = calls
    fc = dict(zip(FORCING, forcing)) if forcing is not None else {
        "ru_tend": uv[1], "rv_tend": uv[3], "rw_tend": None, "t_tend": mt[27], "ph_tend": None, "mu_tend": mt[28]}
    given = forcing is not None
    sv = dict(zip(SAVED, saved)) if saved is not None else dict.fromkeys(SAVED)
    u1, v1 = AdvanceUV.apply(st["u"], fc["ru_tend"], st["v"], fc["rv_tend"], st["p"], st["ph"], st["al"], st["mu"],
                             st["mudf"], uv)
    u2, v2 = SpecBdyUpdateGroup.apply(z1, u1, v1, *((fc["ru_tend"], fc["rv_tend"]) if given else (None, None)))
    ww, t, t_ave, muave, muts, mudf, mu = AdvanceMuT.apply(
        u2, v2, st["ww"], st["mu"], st["muave"], st["muts"], st["mudf"], st["t"], st["t_ave"], fc["t_tend"],
        fc["mu_tend"], mt, sv["u_save"], sv["v_save"], sv["ww1"], sv["t_save"])
    t2, mu2, muts2 = SpecBdyUpdateGroup.apply(z2, t, mu, muts, *((fc["t_tend"], fc["mu_tend"], fc["mu_tend"])
                                                                 if given else (None, None, None)))
    t2a, w, ph = AdvanceW.apply(st["w"], st["ph"], t2, t_ave, u2, v2, muts2, muave, ww, wa, fc["rw_tend"],
                                fc["ph_tend"], sv["w_save"], sv["t_save"], sv["ph_save"])
    ru_m, rv_m, ww_m = SumFlux.apply(u2, v2, ww, st["ru_m"], st["rv_m"], st["ww_m"], sf, sv["u_save"],
                                     sv["v_save"], sv["ww1"])
    ph = SpecBdyUpdatePh.apply(ph, muts2, pz[0], fc["ph_tend"], fc["mu_tend"] if given else None, sv["ph_save"])
    w = ZeroGradBdy.apply(w, pz[1])
    al, p, pm1 = CalcPRho.apply(st["al"], st["p"], ph, t2, st["pm1"], mu2, muts2, cp, sv["t_save"])
    items, ix24, ix2d = bc
    ph, al, p, muts2, mu2, mudf = OpenBoundaryFill.apply([(v, two) for _, v, two in items], ix24, ix2d, ph, al, p,
                                                         muts2, mu2, mudf)
    new = {"u": u2, "v": v2, "p": p, "ph": ph, "al": al, "mu": mu2, "mudf": mudf, "ww": ww, "muave": muave,
           "muts": muts2, "t": t2, "t_ave": t2a, "w": w, "pm1": pm1, "ru_m": ru_m, "rv_m": rv_m, "ww_m": ww_m}
    return [new[n] for n in ACOUSTIC_STATE]


STAGE_IN = ("u_2", "v_2", "w_2", "t_2", "ph_2", "mu_2", "u_1", "v_1", "w_1", "t_1", "ph_1", "mu_1", "ww", "mudf",
            "muave", "t_ave", "ru_m", "rv_m", "ww_m")
STAGE_OUT = ("u", "v", "w", "ph", "ww", "t", "mu", "p", "al", "pm1", "mudf", "muave", "muts", "t_ave", "ru_m",
             "rv_m", "ww_m")


class _StageRecorder(_LoopRecorder):
    """The step's functions, keeping copies of the arguments of every call
    from the n-th small_step_prep through the next small_step_finish."""

    def __getattr__(self, name):
        f = getattr(self._ops, name)

        def rec(*args):
            if name == "small_step_prep":
                self._seen += 1
            on = self._seen == self._n and not self._done
            if on:
                self.calls.append((name, self._copy(list(args))))
            if on and name == "small_step_finish":
                self._done = True
            return f(*args)
        return rec


def record_acoustic_stage(s, ops, cfg, stage=2):
    """Run one RK3 step on state s and keep the acoustic part of RK stage
    `stage` (2 or 3): small_step_prep, the pressure and density of calc_p_rho,
    the open-boundary fills, the acoustic loop, calc_mu_uv_1 and
    small_step_finish (solver.py:1258-1543).  Returns (x, calls): x the
    stage's inputs in STAGE_IN order, calls the recorded calls' arguments."""
    from .step import rk3_step
    if stage not in (2, 3):
        raise ValueError("record_acoustic_stage: stage 2 or 3 (the first stage's preparation copies)")
    rec = _StageRecorder(ops, stage)
    rk3_step(s, rec, cfg)
    names = [n for n, _ in rec.calls]
    m = (len(names) - 6) // len(SUBSTEP_CALLS)
    head = ["small_step_prep", "calc_p_rho", "calc_coef_w", "set_physical_bc_group"]
    if m < 1 or names != head + list(SUBSTEP_CALLS) * m + ["calc_mu_uv_1", "small_step_finish"]:
        raise RuntimeError("record_acoustic_stage: unexpected call sequence %s" % names[:8])
    calls = [a for _, a in rec.calls]
    pr, mt = calls[0], calls[4 + 2]
    x = [pr[1], pr[3], pr[5], pr[7], pr[9], pr[12], pr[0], pr[2], pr[4], pr[6], pr[8], pr[11], pr[34], pr[19],
         mt[8], mt[26], mt[21], mt[22], mt[23]]
    return x, calls


def acoustic_stage(x, calls):
    """The acoustic part of an RK stage past the first, as autograd Functions
    (small_step_prep, calc_p_rho at step 0, the open-boundary fills, the
    acoustic substeps, calc_mu_uv_1, small_step_finish); x: STAGE_IN, calls:
    from record_acoustic_stage.  Returns the STAGE_OUT fields.  pb, p, al
    and alt enter only through fixed coefficients (c2a, and al and p outside
    calc_p_rho's box), the large-step tendencies and masses are fixed."""
    v = dict(zip(STAGE_IN, x))
    pr, cp, _, bc = calls[:4]
    subs = [calls[4 + 9 * q:4 + 9 * q + 9] for q in range((len(calls) - 6) // 9)]
    mc, fin = calls[-2], calls[-1]
    (us, vs, ws, ts, ps, ww1, u, vv, w, t, ph, mu, muts, muus, muvs, musv) = SmallStepPrep.apply(
        v["u_1"], v["u_2"], v["v_1"], v["v_2"], v["w_1"], v["w_2"], v["t_1"], v["t_2"], v["ph_1"], v["ph_2"],
        v["mu_1"], v["mu_2"], v["ww"], pr)
    al, p, pm1 = CalcPRho.apply(cp[0], cp[1], ph, t, cp[7], mu, muts, cp, ts)
    items, ix24, ix2d = bc
    _, _, ph, al, p, _, ts, _, mu, mudf = OpenBoundaryFill.apply(
        [(k, two) for _, k, two in items], ix24, ix2d, items[0][0], items[1][0], ph, al, p, v["t_1"], ts,
        v["mu_1"], mu, v["mudf"])
    st = [u, vv, p, ph, al, mu, mudf, v["ww"], v["muave"], muts, t, v["t_ave"], w, pm1, v["ru_m"], v["rv_m"],
          v["ww_m"]]
    saved = [us, vs, ws, ts, ps, ww1]
    for c in subs:
        st = acoustic_substep(st, c, None, saved)
    st = dict(zip(ACOUSTIC_STATE, st))
    muus, muvs = CalcMuUV1.apply(st["muts"], muus, muvs, mc)
    uo, vo, wo, pho, wwo, to, muo = SmallStepFinish.apply(
        st["u"], st["v"], st["w"], st["t"], st["ph"], st["ww"], st["mu"], us, vs, ws, ts, ps, ww1, musv, muus, muvs,
        st["muts"], fin)
    out = {"u": uo, "v": vo, "w": wo, "ph": pho, "ww": wwo, "t": to, "mu": muo}
    for n in ("p", "al", "pm1", "mudf", "muave", "muts", "t_ave", "ru_m", "rv_m", "ww_m"):
        out[n] = st[n]
    return [out[n] for n in STAGE_OUT]


def acoustic_substeps(x, substeps):
    """The substeps of a recorded loop in turn (acoustic_substep)."""
    for calls in substeps:
        x = acoustic_substep(x, calls)
    return x


class AcousticLoop(torch.autograd.Function):
    """A stage's whole acoustic loop, x -> the 17 carried fields after its
    last substep, with one checkpoint: the forward keeps only the loop's
    inputs; the backward runs the substeps again from them (storing each
    substep's operands for that pass only) and returns the gradient through
    the hand-written adjoints.  substeps: from record_acoustic_loop; args:
    the 17 fields (ACOUSTIC_STATE order), then optionally the six large-step
    tendencies (FORCING order) and after them optionally the six saved
    fields (SAVED order), which then replace the recorded ones in every
    substep and get gradients too.

        x0, subs = adjoint.record_acoustic_loop(state, ops, cfg)
        x = [f.clone().requires_grad_() for f in x0]
        out = adjoint.AcousticLoop.apply(subs, *x)
    """

    @staticmethod
    def _run(substeps, inputs):
        n, m = len(ACOUSTIC_STATE), len(ACOUSTIC_STATE) + len(FORCING)
        x, forcing, saved = list(inputs[:n]), inputs[n:m] or None, inputs[m:] or None
        for calls in substeps:
            x = acoustic_substep(x, calls, forcing, saved)
        return x

    @staticmethod
    def forward(ctx, substeps, *args):
        n = len(ACOUSTIC_STATE)
        if len(args) not in (n, n + len(FORCING), n + len(FORCING) + len(SAVED)):
            raise ValueError("AcousticLoop: expected the 17 carried fields, optionally then the 6 tendencies "
                             "and the 6 saved fields")
        ctx.substeps = substeps
        ctx.save_for_backward(*args)
        with torch.no_grad():
            y = AcousticLoop._run(substeps, args)
        return tuple(y)

    @staticmethod
    def backward(ctx, *lams):
        x = [t.detach().requires_grad_() for t in ctx.saved_tensors]
        with torch.enable_grad():
            y = AcousticLoop._run(ctx.substeps, x)
            pairs = [(o, l) for o, l in zip(y, lams) if l is not None]
            grads = torch.autograd.grad([o for o, _ in pairs], x, [l for _, l in pairs], allow_unused=True)
        return (None, *grads)


# ----------------------------------------------------------------------- tracer transport
class _TransportRecorder:
    """The step's functions, recording for each RK stage the arguments the
    scalar transport of that stage uses (the mass fluxes ru_m, rv_m, ww_m, the
    column masses mu_1, mu and mub, the map factors, the stage's dt)."""

    def __init__(self, ops):
        self._ops = ops
        self.stages = []
        self._pending = None

    def __getattr__(self, name):
        return getattr(self._ops, name)

    def rk_scalar_tend_pair(self, moist_args, scalar_args):
        a = moist_args
        # (scs, sce, tenddec, rk_step, dt, ru, rv, ww, mut, mub, mu_old, c1h, c2h, alt, scalar_old, scalar,
        #  scalar_tends, advect_tend, h_tendency, z_tendency, RQVFTEN, base, moist_step, fnm, fnp,
        #  msfux, msfuy, msfvx, msfvx_inv, msfvy, msftx, msfty, rdx, rdy, rdn, rdnw, ...)
        self._pending = dict(ru=a[5].clone(), rv=a[6].clone(), ww=a[7].clone(), fnm=a[23], fnp=a[24],
                             msftx=a[30], rdx=float(a[32]), rdy=float(a[33]), rdnw=a[35], idx=tuple(a[-18:]))
        return self._ops.rk_scalar_tend_pair(moist_args, scalar_args)

    def rk_update_scalar_pair(self, moist_args, scalar_args):
        a = moist_args
        # (scs, sce, scalar_1, scalar_2, sc_tend, advect_tend, h_tendency, z_tendency, msftx, msfty, c1, c2,
        #  mu_old, mu_new, mu_base, rk_step, dt, spec_zone, tenddec, *idx18)
        st = dict(self._pending, msfty=a[9], mu_old=a[12].clone(), mu_new=a[13].clone(), mu_base=a[14],
                  rk_step=int(a[15]), dt=float(a[16]), spec_zone=int(a[17]))
        self.stages.append(st)
        return self._ops.rk_update_scalar_pair(moist_args, scalar_args)


def record_transport(s, ops, cfg, nsteps):
    """Run nsteps RK3 steps of state s with ops and return, per step, the list
    of its three transport stages (see _TransportRecorder)."""
    from .step import rk3_step
    rec = _TransportRecorder(ops)
    steps = []
    for _ in range(nsteps):
        rec.stages = []
        rk3_step(s, rec, cfg)
        if len(rec.stages) != cfg.rk_order:
            raise RuntimeError("expected %d transport stages, recorded %d" % (cfg.rk_order, len(rec.stages)))
        steps.append(rec.stages)
    return steps


def tracer_step(q, stages):
    """One RK3 step of the transport of tracer q (3-D) with recorded stages
    (HIP): per stage the advection tendency of the current tracer, then
    rk_update_scalar from the step's starting tracer."""
    g = grid(q.shape, stages[0]["idx"])
    q1 = q.clone()
    q2 = q.clone()
    sc = torch.zeros_like(q)
    for st in stages:
        t = advect_scalar_tendency(q2, st["ru"], st["rv"], st["ww"], st["msftx"], st["fnm"], st["fnp"], st["rdx"],
                                   st["rdy"], st["rdnw"], st["idx"])
        call("pywrf_rk_update_scalar", g, 0, 0, ptr(q1), ptr(q2), ptr(sc), ptr(t), ptr(st["msfty"]),
             ptr(st["mu_old"]), ptr(st["mu_new"]), ptr(st["mu_base"]), st["rk_step"], st["dt"], st["spec_zone"],
             stream())
    return q2


def tracer_step_adjoint(lam, stages):
    """The sensitivity of the step's starting tracer given that of its result."""
    g = grid(lam.shape, stages[0]["idx"])
    lam = lam.clone()
    lam_n = torch.zeros_like(lam)
    lam_adv = torch.empty_like(lam)
    for st in reversed(stages):
        prev = torch.empty_like(lam)
        call("pywrf_rk_update_scalar_adjoint", g, ptr(lam), ptr(st["msfty"]), ptr(st["mu_old"]), ptr(st["mu_new"]),
             ptr(st["mu_base"]), st["dt"], st["spec_zone"], ptr(lam_n), ptr(lam_adv), None, ptr(prev), stream())
        advect_scalar_adjoint(lam_adv, st["ru"], st["rv"], st["ww"], st["msftx"], st["fnm"], st["fnp"], st["rdx"],
                              st["rdy"], st["rdnw"], st["idx"], out=prev)
        lam = prev
    return lam_n + lam


class TracerStep(torch.autograd.Function):
    """q_next = one RK3 transport step of q (recorded stages)."""

    @staticmethod
    def forward(ctx, q, stages):
        ctx.stages = stages
        return tracer_step(_f32(q), stages)

    @staticmethod
    def backward(ctx, lam):
        return tracer_step_adjoint(_f32(lam), ctx.stages), None


# ----------------------------------------------------------------------- surrogate
class SurrogateNet(torch.autograd.Function):
    """Standardised tendencies Y [M x 287] of standardised column inputs
    X [M x 369] (FP32; BF16 on the matrix cores, FP32 accumulation), with the
    gradient w.r.t. X.  M must be a multiple of 32."""

    @staticmethod
    def forward(ctx, X, net):
        M = X.shape[0]
        ws = _sg.Workspace(M, device=X.device)
        Xb = torch.zeros(M, _sg.KIN, dtype=torch.bfloat16, device=X.device)
        Xb[:, :_sg.FIN] = X.to(torch.bfloat16)
        Y = _sg.forward(net, Xb, ws)
        ctx.net, ctx.ws = net, ws
        return Y[:, :_sg.FOUT].clone()

    @staticmethod
    def backward(ctx, dY):
        net, ws = ctx.net, ctx.ws
        M = ws.M
        dYb = torch.zeros(M, _sg.NO, dtype=torch.bfloat16, device=dY.device)
        dYb[:, :_sg.FOUT] = dY.to(torch.bfloat16)
        _sg.gemm(2, dYb, net.Wt[2], ws.dH2, H=ws.H2)        # (dY W3) [H2 > 0]
        _sg.gemm(2, ws.dH2, net.Wt[1], ws.dH1, H=ws.H1)     # (dH2 W2) [H1 > 0]
        dX = torch.empty(M, _sg.KIN, dtype=torch.float32, device=dY.device)
        _sg.gemm(0, ws.dH1, net.Wt[0], dX)                  # dH1 W1
        return dX[:, :_sg.FIN].clone(), None


class AcousticStage(torch.autograd.Function):
    """An RK stage's whole acoustic part (acoustic_stage) with one
    checkpoint: the forward keeps only the stage's 19 inputs (STAGE_IN), the
    backward runs the stage again from them and returns the gradient through
    the hand-written adjoints.  calls: from record_acoustic_stage.

        x0, calls = adjoint.record_acoustic_stage(state, ops, cfg, stage=2)
        out = adjoint.AcousticStage.apply(calls, *[f.clone().requires_grad_() for f in x0])
    """

    @staticmethod
    def forward(ctx, calls, *x):
        if len(x) != len(STAGE_IN):
            raise ValueError("AcousticStage: expected the %d stage inputs" % len(STAGE_IN))
        ctx.calls = calls
        ctx.save_for_backward(*x)
        with torch.no_grad():
            y = acoustic_stage(list(x), calls)
        return tuple(y)

    @staticmethod
    def backward(ctx, *lams):
        x = [t.detach().requires_grad_() for t in ctx.saved_tensors]
        with torch.enable_grad():
            y = acoustic_stage(x, ctx.calls)
            pairs = [(o, l) for o, l in zip(y, lams) if l is not None]
            grads = torch.autograd.grad([o for o, _ in pairs], x, [l for _, l in pairs], allow_unused=True)
        return (None, *grads)


def time_advection_adjoints(s, repeat=20):
    """advect_u, advect_v and advect_w (the tendency into a zeroed array) and
    their adjoints (AdvectU / AdvectV / AdvectW's backward) on the arguments
    of their first calls in one RK3 step of the state s, each timed over
    `repeat` calls between HIP events on the current stream.  bytes: the 3D
    arrays each adjoint entry reads (lam, the field whose sign picks the edge
    faces' upwind side -- u and v only --, ru, rv, rom) and writes (its
    gradient) once each, FP32."""
    from . import ops
    from .step import rk3_step, RunConfig
    base, first = ops(), {}

    class _First:
        def __getattr__(self, n):
            f = getattr(base, n)
            w = n[7:-6] if n.startswith("advect_") and n.endswith("_fresh") else None
            if w not in ("u", "v", "w"):
                return f

            def run(*args):
                first.setdefault(w, [x.clone() if isinstance(x, torch.Tensor) else x for x in args])
                return f(*args)
            return run

    rk3_step(s, _First(), RunConfig())
    fns, arrays = {"u": AdvectU, "v": AdvectV, "w": AdvectW}, {"u": 6, "v": 6, "w": 5}
    out = {}
    for w in ("u", "v", "w"):
        args = first[w]
        x = args[0].clone().requires_grad_()
        y = fns[w].apply(x, args)
        lam = torch.randn_like(y)

        def per_call(fn):
            fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(repeat):
                fn()
            e1.record()
            e1.synchronize()
            return 1e3 * e0.elapsed_time(e1) / repeat

        with torch.no_grad():
            fwd = per_call(lambda: fns[w].apply(x, args))
        adj = per_call(lambda: torch.autograd.grad(y, x, lam, retain_graph=True))
        nbytes = arrays[w] * x.numel() * 4
        out["advect_" + w] = {"forward_us": fwd, "adjoint_us": adj, "adjoint_bytes": nbytes,
                              "adjoint_GBs": nbytes / adj / 1e3}
    return out


def main(argv=None):
    """Time the gradient of an RK stage's acoustic loop (--what loop) or of its
    whole acoustic part (--what stage) on a seeded state: the forward, the
    checkpointed forward + backward (AcousticLoop / AcousticStage) and the
    same with every operand kept, with the device memory each holds; or
    (--what advect) the momentum advection adjoints per call.

        python -m pywrf_amd.adjoint --config config2 --stage 3 --repeat 5 [--what stage|advect]
    """
    import argparse
    import json
    import time
    from . import ops
    from .gen import synthetic_state
    from .step import RunConfig
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="config2")
    ap.add_argument("--stage", type=int, default=3)
    ap.add_argument("--repeat", type=int, default=5)
    ap.add_argument("--what", choices=("loop", "stage", "advect"), default="loop")
    a = ap.parse_args(argv)
    s = synthetic_state(a.config, device="cuda")
    if a.what == "advect":
        r = {"what": a.what, "config": a.config, "repeat": a.repeat}
        r.update(time_advection_adjoints(s, a.repeat))
        print(json.dumps(r))
        return
    if a.what == "loop":
        x0, subs = record_acoustic_loop(s, ops(), RunConfig(), stage=a.stage)
        chain, ckpt = (lambda x: acoustic_substeps(x, subs)), (lambda x: AcousticLoop.apply(subs, *x))
    else:
        x0, subs = record_acoustic_stage(s, ops(), RunConfig(), stage=a.stage)
        chain, ckpt = (lambda x: acoustic_stage(x, subs)), (lambda x: AcousticStage.apply(subs, *x))
    del s
    with torch.no_grad():
        lams = [torch.randn_like(o) for o in chain(list(x0))]
    fields = sum(f.numel() * f.element_size() for f in x0)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(a.repeat):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / a.repeat

    def fwd():
        with torch.no_grad():
            chain(list(x0))

    held = {}

    def grad(kept):
        def run():
            torch.cuda.synchronize()
            base = torch.cuda.memory_allocated()
            torch.cuda.reset_peak_memory_stats()
            x = [f.clone().requires_grad_() for f in x0]
            y = chain(x) if kept else ckpt(x)
            torch.cuda.synchronize()
            held["kept" if kept else "checkpoint"] = torch.cuda.memory_allocated() - base
            torch.autograd.backward(y, lams)
            torch.cuda.synchronize()
            held[("kept" if kept else "checkpoint") + "_peak"] = torch.cuda.max_memory_allocated() - base
        return run

    r = {"what": a.what, "config": a.config, "stage": a.stage, "fields_MB": fields / 1e6,
         "forward_ms": 1e3 * timed(fwd), "checkpoint_grad_ms": 1e3 * timed(grad(False)),
         "kept_grad_ms": 1e3 * timed(grad(True))}
    r.update({k + "_MB": v / 1e6 for k, v in held.items()})
    print(json.dumps(r))


if __name__ == "__main__":
    main()
