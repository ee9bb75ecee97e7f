"""Gradients through parts of the step: hand-written adjoint kernels behind
torch.autograd.Function.

The reference claims differentiability through torch autograd (README.md:37-40)
but runs its step under no_grad (solver.py:456, 710, 1748).  Here gradients
come from the library's own adjoint kernels (csrc/adjoint.hip), each the
transpose of a linear part of the step:

* AdvectScalar -- advect_scalar (wrf_dynamics.py:1841-2028) as a function of
  the field, the mass fluxes held fixed.  Backward: pywrf_advect_scalar_adjoint.
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
    gradient with respect to w, ph, t_2, t_2ave, u, v, muts, muave and ww (every
    other argument is a coefficient held fixed).  args: advance_w's
    positional arguments in the reference's order."""

    # positions in advance_w's signature
    W, WW, U, V, MUAVE, MUTS, T2AVE, T2, PH = 0, 2, 4, 5, 8, 9, 18, 19, 21
    DIFF = (W, PH, T2, T2AVE, U, V, MUTS, MUAVE, WW)

    @staticmethod
    def forward(ctx, w, ph, t_2, t_2ave, u, v, muts, muave, ww, args):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(AdvanceW.DIFF, (w, ph, t_2, t_2ave, u, v, muts, muave, ww)):
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
             *[ptr(t) for t in work], stream())
        lw, lph, lt2, lt2a = out
        lu, lv = uv
        lmuts, lmuave = planes
        return lw, lph, lt2, lt2a, lu, lv, lmuts, lmuave, lww, None


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
    mu_tend); t_1, u_1, v_1, ww_1, mut and the map factors are held fixed.
    The reference runs the mass updates under no_grad (its autograd would
    leave them out); this is the full derivative.  args: advance_mu_t's
    positional arguments in the reference's order."""

    # positions in advance_mu_t's signature
    DIFF = (2, 4, 0, 6, 8, 9, 12, 24, 26, 27, 28)
    OUT = (0, 2, 3, 7, 8, 9, 10)            # positions of the fresh results in its 11-tuple

    @staticmethod
    def forward(ctx, u, v, ww, mu, muave, muts, mudf, t, t_ave, ft, mu_tend, args):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(AdvanceMuT.DIFF, (u, v, ww, mu, muave, muts, mudf, t, t_ave, ft, mu_tend)):
            args[pos] = _f32(x)
        ctx.args = args
        r = wrf_dynamics.advance_mu_t(*args)
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
        call("pywrf_advance_mu_t_adjoint", g, *[ptr(x) for x in lam], ptr(a[25]), ptr(a[42]), ptr(a[43]),
             ptr(a[33]), ptr(a[34]), ptr(a[35]), ptr(a[36]), float(a[29]), float(a[30]), float(a[31]),
             float(a[32]), *[ptr(x) for x in out], stream())
        gww, gu, gv, gmu, gmuave, gmuts, gmudf, gt, gtave, gft, gmut = out
        return gu, gv, gww, gmu, gmuave, gmuts, gmudf, gt, gtave, gft, gmut, None


class CalcPRho(torch.autograd.Function):
    """(al, p, pm1) = calc_p_rho(...) (wrf_dynamics.py:3269-3313) with the
    gradient with respect to al, p, ph, t_2, pm1, mu and mut (alt, t_1 and c2a
    held fixed).  The incoming pm1 is left as it is (the call's in-place update
    goes to a copy).  args: calc_p_rho's positional arguments in the
    reference's order."""

    # positions in calc_p_rho's signature
    DIFF = (0, 1, 2, 4, 7, 8, 9)

    @staticmethod
    def forward(ctx, al, p, ph, t_2, pm1, mu, mut, args):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(CalcPRho.DIFF, (al, p, ph, t_2, pm1, mu, mut)):
            args[pos] = _f32(x)
        args[7] = args[7].clone()
        ctx.args = args
        return wrf_dynamics.calc_p_rho(*args)

    @staticmethod
    def backward(ctx, lal, lp, lpm1):
        a = ctx.args
        g = grid(a[1].shape, tuple(a[-18:]))
        lal, lp, lpm1 = [torch.zeros_like(a[1]) if l is None else _f32(l) for l in (lal, lp, lpm1)]
        out = [torch.empty_like(a[1]) for _ in range(5)] + [torch.empty_like(a[8]), torch.empty_like(a[9])]
        # (al_ori, p_ori, ph, alt, t_2, t_1, c2a, pm1, mu, mut, c1h, c2h, c1f, c2f, c3h, c4h, c3f, c4f, znu, t0,
        #  rdnw, dnw, smdiv, non_hydrostatic, step, idx)
        call("pywrf_calc_p_rho_adjoint", g, ptr(lal), ptr(lp), ptr(lpm1), ptr(a[2]), ptr(a[3]), ptr(a[4]),
             ptr(a[5]), ptr(a[6]), ptr(a[8]), ptr(a[9]), float(a[19]), ptr(a[20]), float(a[22]), int(a[24]),
             *[ptr(t) for t in out], stream())
        gal, gp, gpm1, gph, gt2, gmu, gmut = out
        return gal, gp, gph, gt2, gpm1, gmu, gmut, None


class SumFlux(torch.autograd.Function):
    """(ru_m, rv_m, ww_m) = sumflux(...) (wrf_dynamics.py:3799-3839) with the
    gradient with respect to ru, rv, ww and the three accumulators (the call's
    in-place update goes to copies).  args: sumflux's positional arguments in
    the reference's order."""

    # positions in sumflux's signature
    DIFF = (0, 1, 2, 16, 17, 18)

    @staticmethod
    def forward(ctx, ru, rv, ww, ru_m, rv_m, ww_m, args):
        from . import wrf_dynamics
        args = list(args)
        for pos, x in zip(SumFlux.DIFF, (ru, rv, ww, ru_m, rv_m, ww_m)):
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
        call("pywrf_sumflux_adjoint", g, *[ptr(x) for x in lam], int(a[25]), int(a[26]),
             *[ptr(x) for x in out], stream())
        return (*out, None)


class SpecBdyUpdatePh(torch.autograd.Function):
    """spec_bdyupdate_ph (wrf_boundary.py:287-358) with the gradient with
    respect to the field and muts (ph_save, field_tend and mu_tend held
    fixed); the update goes to a copy of the field.  args: its positional
    arguments in the reference's order."""

    @staticmethod
    def forward(ctx, field, muts, args):
        from . import wrf_boundary
        args = list(args)
        ctx.f_in = _f32(field)
        args[1], args[4] = ctx.f_in.clone(), _f32(muts)
        ctx.args = args
        return wrf_boundary.spec_bdyupdate_ph(*args)

    @staticmethod
    def backward(ctx, lf):
        a = ctx.args
        g = grid(a[1].shape, tuple(a[-24:]))
        lam = torch.zeros_like(a[1]) if lf is None else _f32(lf).clone()
        lmuts = torch.zeros_like(a[4])
        call("pywrf_spec_bdyupdate_ph_adjoint", g, ptr(a[0]), ptr(ctx.f_in), ptr(a[2]), ptr(a[3]), ptr(a[4]),
             float(a[7]), flag(a[8]), ptr(lam), ptr(lmuts), stream())
        return lam, lmuts, None


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
