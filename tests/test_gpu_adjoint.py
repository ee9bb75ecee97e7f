"""The hand-written adjoint kernels (pywrf_amd.adjoint) on the device.

Checkers, none of them the adjoint kernels' own formulas:
* the transpose property <A q, lam> = <q, A^T lam> on random vectors, with A
  the device's forward advection (FP64 sums);
* torch autograd through the oracle (the CPU restatement of advect_scalar and
  rk_update_scalar) in FP64, on the same mass fluxes -- the gradient the
  reference's own functions give under autograd;
* FP64 central differences of the oracle's transport, eps = 1e-3, along 20
  random directions, on the 120 x 120 x 41 workload over 6 RK3 steps;
* for the tridiagonal solve, the textbook factorisation: the sweeps solve
  L U x = r with L lower bidiagonal (1/alpha on the diagonal, a below) and U
  unit upper bidiagonal (gamma above), so x and its gradients are checked
  against torch.linalg.solve on that matrix, differentiated by autograd.
* for the surrogate, CPU autograd in FP32 of the same network on the BF16
  values the device uses.
Tolerances are stated per test.
"""
import os

import pytest
import torch

pytestmark = pytest.mark.gpu
N = None


def _recorded(config, nsteps):
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import RunConfig
    s = synthetic_state(config, device="cuda")
    steps = adjoint.record_transport(s, pywrf_amd.ops(), RunConfig(), nsteps)
    torch.cuda.synchronize()
    return s, steps


def _cpu64(st):
    return {k: (v.detach().double().cpu() if isinstance(v, torch.Tensor) else v) for k, v in st.items()}


def _oracle_advect(q, st):
    import oracle
    return oracle.ops().advect_scalar(q, q, torch.zeros_like(q), st["ru"], st["rv"], st["ww"], N, N, N, N, N, N, N, N,
                                      st["msftx"], N, st["fnm"], st["fnp"], st["rdx"], st["rdy"], st["rdnw"],
                                      *st["idx"])


def _oracle_tracer_step(q, stages):
    """One RK3 transport step of q by the oracle's functions (CPU)."""
    import oracle
    O = oracle.ops()
    s1 = q.unsqueeze(0).clone()
    s2 = q.unsqueeze(0).clone()
    zero = torch.zeros_like(s2)
    for st in stages:
        t = _oracle_advect(s2[0], st)
        O.rk_update_scalar(0, 0, s1, s2, zero, t.unsqueeze(0), N, N, N, st["msfty"], N, N, st["mu_old"],
                           st["mu_new"], st["mu_base"], st["rk_step"], st["dt"], st["spec_zone"], False, *st["idx"])
    return s2[0]


def _rel(x, y):
    return float((x - y).abs().max() / y.abs().max())


@pytest.mark.parametrize("config", ["config1c", "config4"])
def test_advection_adjoint_is_the_transpose(config):
    from pywrf_amd import adjoint
    s, steps = _recorded(config, 1)
    gen = torch.Generator(device="cuda").manual_seed(1)
    for st in steps[0]:
        q = torch.randn(s.t.shape, device="cuda", generator=gen)
        lam = torch.randn(s.t.shape, device="cuda", generator=gen)
        args = (st["ru"], st["rv"], st["ww"], st["msftx"], st["fnm"], st["fnp"], st["rdx"], st["rdy"], st["rdnw"],
                st["idx"])
        Aq = adjoint.advect_scalar_tendency(q, *args)
        ATl = adjoint.advect_scalar_adjoint(lam, *args)
        lhs = float((Aq.double() * lam.double()).sum())
        rhs = float((q.double() * ATl.double()).sum())
        scale = float((Aq.double().abs() * lam.double().abs()).sum())
        print("<Aq, l> %.9e  <q, A^T l> %.9e" % (lhs, rhs))
        assert abs(lhs - rhs) <= 1e-6 * scale


@pytest.mark.parametrize("config", ["config1c", "config4"])
def test_advection_adjoint_matches_oracle_autograd(config):
    from pywrf_amd import adjoint
    s, steps = _recorded(config, 1)
    st = steps[0][1]
    lam = torch.randn(s.t.shape, device="cuda", generator=torch.Generator(device="cuda").manual_seed(2))
    g = adjoint.advect_scalar_adjoint(lam, st["ru"], st["rv"], st["ww"], st["msftx"], st["fnm"], st["fnp"],
                                      st["rdx"], st["rdy"], st["rdnw"], st["idx"])
    st64 = _cpu64(st)
    q = s.moist[1].detach().double().cpu().requires_grad_()
    (_oracle_advect(q, st64) * lam.double().cpu()).sum().backward()
    err = _rel(g.double().cpu(), q.grad)
    print("A^T lam vs oracle autograd (FP64): %.2e" % err)
    assert err <= 1e-5


def test_tracer_step_forward_equals_oracle():
    """The device transport step is the oracle's arithmetic: bit for bit in FP32."""
    from pywrf_amd import adjoint
    s, steps = _recorded("config1c", 1)
    q = s.moist[1].clone()
    mine = adjoint.tracer_step(q, steps[0])
    st32 = [{k: (v.detach().cpu() if isinstance(v, torch.Tensor) else v) for k, v in st.items()} for st in steps[0]]
    ref = _oracle_tracer_step(q.cpu(), st32)
    assert torch.equal(mine.cpu(), ref)


def test_tracer_gradient_matches_fp64_differences_and_autograd():
    """Gradient of J(b) = mean((q_6(q_0 + b) - target)^2) at b = 0 on the
    120 x 120 x 41 workload over 6 RK3 steps; the target is the same
    transport of a perturbed start."""
    from pywrf_amd import adjoint
    nsteps = 6
    s, steps = _recorded("config4", nsteps)
    q0 = s.moist[1].clone()
    gen = torch.Generator(device="cuda").manual_seed(3)
    pert = 0.05 * float(q0.abs().max()) * torch.randn(q0.shape, device="cuda", generator=gen)
    with torch.no_grad():
        target = q0 + pert
        for st in steps:
            target = adjoint.tracer_step(target, st)
    b = torch.zeros_like(q0, requires_grad=True)
    q = q0 + b
    for st in steps:
        q = adjoint.TracerStep.apply(q, st)
    loss = ((q - target) ** 2).mean()
    loss.backward()
    grad = b.grad.double().cpu()

    st64 = [[_cpu64(st) for st in stp] for stp in steps]
    q064, tgt64 = q0.double().cpu(), target.double().cpu()

    def J(bias):
        x = q064 + bias
        for stp in st64:
            x = _oracle_tracer_step(x, stp)
        return ((x - tgt64) ** 2).mean()

    b64 = torch.zeros_like(q064, requires_grad=True)
    J(b64).backward()
    err = _rel(grad, b64.grad)
    print("gradient vs oracle autograd (FP64): %.2e" % err)
    assert err <= 1e-4
    g2 = torch.Generator().manual_seed(4)
    eps = 1e-3
    worst = 0.0
    with torch.no_grad():
        for _ in range(20):
            d = torch.randn(q064.shape, generator=g2, dtype=torch.float64)
            d *= float(q064.abs().max()) / float(d.abs().max())
            fd = float(J(eps * d) - J(-eps * d)) / (2 * eps)
            ad = float((grad * d).sum())
            worst = max(worst, abs(ad - fd) / max(abs(fd), 1e-300))
    print("20 directions: worst |<g,d> - FD| / |FD| = %.2e" % worst)
    assert worst <= 1e-4


def _bidiag_matrix(a, alpha, gamma, K):
    """M = L U for one batch of columns [..., K+1] (level 0 an identity row)."""
    n = K + 1
    shape = a.shape[:-1] + (n, n)
    L = torch.zeros(shape, dtype=a.dtype)
    U = torch.zeros(shape, dtype=a.dtype)
    idx = torch.arange(n)
    U[..., idx, idx] = 1.0
    L[..., 0, 0] = 1.0
    k = torch.arange(1, n)
    L[..., k, k] = 1.0 / alpha[..., 1:n]
    L[..., k, k - 1] = a[..., 1:n]
    k2 = torch.arange(1, K)
    U[..., k2, k2 + 1] = gamma[..., 1:K]
    return L @ U


def test_tridiag_solve_and_adjoint_against_the_factorisation():
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    s = synthetic_state("config1c", device="cuda")
    g = s.grid
    idx = g.ix18
    K = g.kte - 1
    gen = torch.Generator(device="cuda").manual_seed(5)
    # coefficients of the kind calc_coef_w forms: diagonally dominant
    a = 0.3 * torch.rand(s.w.shape, device="cuda", generator=gen) - 0.15
    gamma = 0.3 * torch.rand(s.w.shape, device="cuda", generator=gen) - 0.15
    alpha = 0.8 + 0.4 * torch.rand(s.w.shape, device="cuda", generator=gen)
    r = torch.randn(s.w.shape, device="cuda", generator=gen)
    ts = [x.clone().requires_grad_() for x in (a, alpha, gamma, r)]
    x = adjoint.TridiagSolve.apply(*ts, idx)
    lam = torch.randn(s.w.shape, device="cuda", generator=gen)
    (x * lam).sum().backward()
    j0, j1 = max(g.jts, g.jds + 1), min(g.jte, g.jde - 2)
    i0, i1 = max(g.its, g.ids + 1), min(g.ite, g.ide - 2)
    col = lambda t: t.detach().double().cpu()[:K + 1, j0:j1, i0:i1].permute(1, 2, 0)
    c64 = [col(t).requires_grad_() for t in (a, alpha, gamma, r)]
    M = _bidiag_matrix(c64[0], c64[1], c64[2], K)
    xr = torch.linalg.solve(M, c64[3].unsqueeze(-1)).squeeze(-1)
    ex = _rel(col(x), xr.detach())
    print("solve vs factorisation: %.2e" % ex)
    assert ex <= 1e-5
    (xr * col(lam)).sum().backward()
    for name, t, ref in zip(("a", "alpha", "gamma", "r"), ts, c64):
        e = _rel(col(t.grad), ref.grad)
        print("d/d%s: %.2e" % (name, e))
        assert e <= 1e-4, name
    # outside the solved columns r passes through, the coefficients get nothing
    m = torch.ones(s.w.shape, dtype=torch.bool, device="cuda")
    m[:K + 1, j0:j1, i0:i1] = False
    assert torch.equal(ts[3].grad[m], lam[m])
    assert float(ts[0].grad[m].abs().max()) == 0.0


def test_surrogate_input_gradient_matches_cpu_autograd():
    from pywrf_amd import adjoint, surrogate as sg
    net = sg.MLP(seed=3)
    M = 256
    gen = torch.Generator(device="cuda").manual_seed(6)
    X = torch.randn(M, sg.FIN, device="cuda", generator=gen).requires_grad_()
    Y = adjoint.SurrogateNet.apply(X, net)
    dY = torch.randn(M, sg.FOUT, device="cuda", generator=gen)
    (Y * dY).sum().backward()
    # CPU FP32 reference on the BF16 values the kernels see (inputs, weights,
    # hidden activations and output gradient rounded to BF16)
    bf = lambda t: t.detach().cpu().to(torch.bfloat16).float()
    W = [bf(w) for w in net.W]
    b = [x.detach().cpu() for x in net.b]
    Xc = bf(X).requires_grad_()
    xp = torch.nn.functional.pad(Xc, (0, sg.KIN - sg.FIN))
    h1 = torch.relu(xp @ W[0].t() + b[0])
    h1b = h1 + (bf(h1) - h1).detach()
    h2 = torch.relu(h1b @ W[1].t() + b[1])
    h2b = h2 + (bf(h2) - h2).detach()
    y = (h2b @ W[2].t() + b[2])[:, :sg.FOUT]
    assert _rel(Y.detach().cpu(), y.detach()) <= 2e-2
    (y * bf(dY)).sum().backward()
    err = _rel(X.grad.cpu(), Xc.grad)
    print("dX vs CPU autograd: %.2e" % err)
    assert err <= 3e-2


class _FirstCall:
    """The step's functions, keeping a copy of the arguments of the first call of one of them."""

    def __init__(self, ops, name):
        self._ops, self._name, self.args = ops, name, None

    def __getattr__(self, n):
        f = getattr(self._ops, n)
        if n != self._name:
            return f

        def rec(*args):
            if self.args is None:
                self.args = [a.clone() if isinstance(a, torch.Tensor) else a for a in args]
            return f(*args)
        return rec


@pytest.mark.parametrize("config", ["config1c", "bdy24"])
def test_advance_w_adjoint_matches_oracle_autograd(config):
    """Gradients of <lT, t_2ave> + <lW, w> + <lP, ph> after advance_w with
    respect to w, ph, t_2, t_2ave, u, v, muts, muave, ww, the forcing
    rw_tend and ph_tend and the saved w_save, t_1 and ph_1 (the last through
    the vertical advection and the damping weights), at the arguments of the
    step's first acoustic substep, against torch autograd through the oracle's advance_w
    in FP64 (the same coefficients).  Bound: 1e-4 normwise per gradient (the
    coefficients are FP32 values in both; the FP32 adjoint's rounding)."""
    import oracle
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config, device="cuda")
    rec = _FirstCall(pywrf_amd.ops(), "advance_w")
    rk3_step(s, rec, RunConfig())
    args = rec.args
    gen = torch.Generator(device="cuda").manual_seed(7)
    lams = [torch.randn(s.w.shape, device="cuda", generator=gen) for _ in range(3)]
    D = adjoint.AdvanceW.DIFF + adjoint.AdvanceW.EXTRA
    x = [args[n].clone().requires_grad_() for n in D]
    outs = adjoint.AdvanceW.apply(*x[:9], args, *x[9:])
    sum(float(0) + (o * l).sum() for o, l in zip(outs, lams)).backward()
    a64 = [a.detach().double().cpu() if isinstance(a, torch.Tensor) else a for a in args]
    x64 = [a64[n].clone().requires_grad_() for n in D]
    for n, t in zip(D, x64):
        a64[n] = t
    o64 = oracle.ops().advance_w(*a64)
    sum((o * l.double().cpu()).sum() for o, l in zip(o64, lams)).backward()
    names = ("w", "ph", "t_2", "t_2ave", "u", "v", "muts", "muave", "ww", "rw_tend", "ph_tend", "w_save", "t_1",
             "ph_1")
    assert len(names) == len(D)
    for name, mine, ref in zip(names, x, x64):
        assert float(ref.grad.abs().max()) > 0, name
        err = _rel(mine.grad.double().cpu(), ref.grad)
        print("d/d%s: %.2e" % (name, err))
        assert err <= 1e-4, name


@pytest.mark.parametrize("config", ["config1c", "bdy24"])
def test_advance_uv_adjoint_matches_oracle_autograd(config):
    """Gradients of <lU, u> + <lV, v> after advance_uv with respect to u,
    ru_tend, v, rv_tend, p, ph, al, mu and mudf, at the arguments of the
    step's first acoustic substep, against torch autograd through the
    oracle's advance_uv in FP64 (the same coefficients).  The map is affine
    in these inputs, so the check covers every term whatever their values.
    Bound: 1e-4 normwise per gradient (FP32 rounding and the atomic adds'
    order)."""
    import oracle
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config, device="cuda")
    rec = _FirstCall(pywrf_amd.ops(), "advance_uv")
    rk3_step(s, rec, RunConfig())
    args = rec.args
    gen = torch.Generator(device="cuda").manual_seed(11)
    lams = [torch.randn(args[n].shape, device="cuda", generator=gen) for n in (0, 2)]
    D = adjoint.AdvanceUV.DIFF
    x = [args[n].clone().requires_grad_() for n in D]
    outs = adjoint.AdvanceUV.apply(*x, args)
    ((outs[0] * lams[0]).sum() + (outs[1] * lams[1]).sum()).backward()
    a64 = [a.detach().double().cpu() if isinstance(a, torch.Tensor) else a for a in args]
    x64 = [a64[n].clone().requires_grad_() for n in D]
    for n, t in zip(D, x64):
        a64[n] = t
    o64 = oracle.ops().advance_uv(*a64)
    sum((o * l.double().cpu()).sum() for o, l in zip(o64, lams)).backward()
    names = ("u", "ru_tend", "v", "rv_tend", "p", "ph", "al", "mu", "mudf")
    for name, mine, ref in zip(names, x, x64):
        assert float(ref.grad.abs().max()) > 0, name
        err = _rel(mine.grad.double().cpu(), ref.grad)
        print("d/d%s: %.2e" % (name, err))
        assert err <= 1e-4, name


@pytest.mark.parametrize("config", ["config1c", "bdy24"])
def test_advance_mu_t_adjoint_matches_oracle_differences(config):
    """advance_mu_t's gradient with respect to each of u, v, ww, mu, muave,
    muts, mudf, t, t_ave, ft, mu_tend and the large-step u_1, v_1, ww_1, t_1,
    at the arguments of the step's first acoustic substep, against FP64
    differences of the oracle's advance_mu_t:
    for a seeded direction d in one input, sum_o <lam_o, f_o(x + d) - f_o(x)>
    (exact up to FP64 rounding: the map is affine in each of these inputs) against
    <grad, d>.  Autograd through the oracle cannot serve here: like the
    reference it updates the column masses under no_grad.  Bound: 1e-4
    per input, relative to the larger of |difference| and the 2-norm of the
    summands grad_i d_i (a sum over random signs can cancel far below the
    size of its terms, and the FP32 gradient is exact only to the latter)."""
    import oracle
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config, device="cuda")
    rec = _FirstCall(pywrf_amd.ops(), "advance_mu_t")
    rk3_step(s, rec, RunConfig())
    args = rec.args
    D, OUT = adjoint.AdvanceMuT.DIFF + adjoint.AdvanceMuT.EXTRA, adjoint.AdvanceMuT.OUT
    gen = torch.Generator().manual_seed(13)
    lams = [torch.randn(tuple(args[n].shape), generator=gen, dtype=torch.float64) for n in (0, 24, 26, 8, 9, 12, 6)]
    x = [args[n].clone().requires_grad_() for n in D]
    outs = adjoint.AdvanceMuT.apply(*x[:11], args, *x[11:])
    sum((o * l.float().cuda()).sum() for o, l in zip(outs, lams)).backward()
    a64 = [a.detach().double().cpu() if isinstance(a, torch.Tensor) else a for a in args]
    base = oracle.ops().advance_mu_t(*a64)
    names = ("u", "v", "ww", "mu", "muave", "muts", "mudf", "t", "t_ave", "ft", "mu_tend", "u_1", "v_1", "ww_1",
             "t_1")
    assert len(names) == len(D)
    for name, pos, mine in zip(names, D, x):
        d = torch.randn(tuple(a64[pos].shape), generator=gen, dtype=torch.float64)
        b = list(a64)
        b[pos] = a64[pos] + d
        moved = oracle.ops().advance_mu_t(*b)
        ref = sum(float(((moved[n] - base[n]) * l).sum()) for n, l in zip(OUT, lams))
        terms = mine.grad.double().cpu() * d
        got = float(terms.sum())
        scale = max(abs(ref), float(terms.norm()))      # a random-direction sum can cancel far below its terms
        err = abs(got - ref) / scale
        print("d/d%s: <grad, d> %.6e, difference %.6e, rel %.2e" % (name, got, ref, err))
        assert err <= 1e-4, name


def test_adjoints_match_the_reference_autograd_gradients():
    """Against the reference's own functions differentiated by torch autograd
    (tests/golden/adjoint_golden.npz, make_adjoint_golden.py): advance_w's
    gradient with respect to w, ph, t_2, t_2ave and advect_scalar's with
    respect to the field, at the arguments of the first call of each in a
    step of the 40 x 40 x 41 state with clouds, for the same seeded
    sensitivities.  The arguments here come from the device's own step (the
    reference's were formed by its functions on the host; they agree to the
    last bits), so the bound is normwise 1e-4."""
    import numpy as np
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "adjoint_golden.npz"))
    s = synthetic_state("config1c", device="cuda")
    rec = _FirstCall(_FirstCall(pywrf_amd.ops(), "advance_w"), "advect_scalar_fresh")
    rk3_step(s, rec, RunConfig())
    w_args = rec._ops.args
    g = torch.Generator().manual_seed(7)
    lT, lW, lP = (torch.randn(tuple(s.w.shape), generator=g).cuda() for _ in range(3))
    x = [w_args[n].clone().requires_grad_() for n in adjoint.AdvanceW.DIFF]
    outs = adjoint.AdvanceW.apply(*x, w_args)
    ((outs[0] * lT).sum() + (outs[1] * lW).sum() + (outs[2] * lP).sum()).backward()
    for name, t in zip(("w", "ph", "t_2", "t_2ave"), x[:4]):
        ref = torch.from_numpy(gold["advance_w.d" + name]).double()
        err = _rel(t.grad.double().cpu(), ref)
        print("advance_w d/d%s vs reference autograd: %.2e" % (name, err))
        assert err <= 1e-4, name
    a = rec.args
    (lam,) = [torch.randn(tuple(s.t.shape), generator=torch.Generator().manual_seed(2)).cuda()]
    gq = adjoint.advect_scalar_adjoint(lam, a[3], a[4], a[5], a[14], a[16], a[17], a[18], a[19], a[20],
                                       tuple(a[21:39]))
    err = _rel(gq.double().cpu(), torch.from_numpy(gold["advect_scalar.dfield"]).double())
    print("advect_scalar d/dfield vs reference autograd: %.2e" % err)
    assert err <= 1e-4


class _Substep:
    """The step's functions, keeping clones of the arguments of every call
    from the first advance_uv through the first advance_w."""

    def __init__(self, ops):
        self._ops, self.calls, self.on, self.done = ops, [], False, False

    def __getattr__(self, n):
        f = getattr(self._ops, n)

        def rec(*args):
            self.on = self.on or n == "advance_uv"
            if self.on and not self.done:
                clone = lambda a: a.clone() if isinstance(a, torch.Tensor) else (
                    [tuple(clone(x) for x in it) for it in a] if isinstance(a, list) else a)
                self.calls.append((n, [clone(a) for a in args]))
                self.done = n == "advance_w"
            return f(*args)
        return rec


@pytest.mark.parametrize("config", ["config1c", "bdy24"])
def test_acoustic_adjoints_chain_through_a_substep(config):
    """The three acoustic adjoints chained by autograd through the first
    acoustic substep's advance_uv, the zone update of u and v, advance_mu_t,
    the zone update of t, mu and muts, and advance_w (the zone updates add
    fixed tendencies on the outer ring: their Jacobian is the identity),
    against FP64 central differences of the same sequence through the
    oracle: for each of u, v, p, ph, al, mu, mudf, ww, muave, muts, t, t_ave
    and w a seeded direction d, <grad, d> against the difference of
    sum_o <lam_o, out_o>, relative to the larger of the difference and the
    2-norm of grad_i d_i; bound 1e-4."""
    import oracle
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config, device="cuda")
    rec = _Substep(pywrf_amd.ops())
    rk3_step(s, rec, RunConfig())
    names = [n for n, _ in rec.calls]
    assert names == ["advance_uv", "spec_bdyupdate_group", "advance_mu_t", "spec_bdyupdate_group", "advance_w"], names
    uv, z1, mt, z2, wa = (a for _, a in rec.calls)
    o = pywrf_amd.ops()

    def zone(items, dt, ix24, ix2d, fields):
        """the increments the zone update adds to `fields` (HIP, on clones)"""
        outs = [f.detach().clone() for f in fields]
        o.spec_bdyupdate_group([(f, it[1], it[2], it[3]) for f, it in zip(outs, items)], dt, ix24, ix2d)
        return [a - f.detach() for a, f in zip(outs, fields)]

    # inputs: advance_uv's u, v, p, ph, al, mu, mudf; advance_mu_t's ww, muave, muts, t, t_ave; advance_w's w
    src = {"u": (uv, 0), "v": (uv, 2), "p": (uv, 4), "ph": (uv, 6), "al": (uv, 9), "mu": (uv, 10),
           "mudf": (uv, 15), "ww": (mt, 0), "muave": (mt, 8), "muts": (mt, 9), "t": (mt, 24), "t_ave": (mt, 26),
           "w": (wa, 0)}
    x = {n: a[p].clone().requires_grad_() for n, (a, p) in src.items()}
    u1, v1 = adjoint.AdvanceUV.apply(x["u"], uv[1], x["v"], uv[3], x["p"], x["ph"], x["al"], x["mu"], x["mudf"], uv)
    du, dv = zone(z1[0], z1[1], z1[2], z1[3], [u1, v1])
    u2, v2 = u1 + du, v1 + dv
    ww_o, t_o, tave_o, muave_o, muts_o, mudf_o, mu_o = adjoint.AdvanceMuT.apply(
        u2, v2, x["ww"], x["mu"], x["muave"], x["muts"], x["mudf"], x["t"], x["t_ave"], mt[27], mt[28], mt)
    dt_, dmu, dmuts = zone(z2[0], z2[1], z2[2], z2[3], [t_o, mu_o, muts_o])
    t2, mu2, muts2 = t_o + dt_, mu_o + dmu, muts_o + dmuts
    t2a_w, w_o, ph_o = adjoint.AdvanceW.apply(x["w"], x["ph"], t2, tave_o, u2, v2, muts2, muave_o, ww_o, wa)
    outs = [u2, v2, ww_o, t2, mu2, muts2, mudf_o, muave_o, t2a_w, w_o, ph_o]
    gen = torch.Generator().manual_seed(17)
    lams = [torch.randn(tuple(t.shape), generator=gen, dtype=torch.float64) for t in outs]
    sum((t * l.float().cuda()).sum() for t, l in zip(outs, lams)).backward()

    # the same sequence through the oracle in FP64
    ro = oracle.ops()
    d64 = lambda args: [a.detach().double().cpu() if isinstance(a, torch.Tensor) else a for a in args]
    uv64, mt64, wa64 = d64(uv), d64(mt), d64(wa)
    z164 = [(it[0].double().cpu(), it[1].double().cpu(), it[2], it[3]) for it in z1[0]]
    z264 = [(it[0].double().cpu(), it[1].double().cpu(), it[2], it[3]) for it in z2[0]]

    def ozone(items, dt, ix24, ix2d, fields):
        return [ro.spec_bdyupdate(f.clone(), it[1], dt, it[2], 1, *(ix2d if it[3] else ix24))
                for f, it in zip(fields, items)]

    def chain(v):
        a = list(uv64)
        for n in ("u", "v", "p", "ph", "al", "mu", "mudf"):
            a[src[n][1]] = v[n]
        u1_, v1_ = ro.advance_uv(*a)
        u2_, v2_ = ozone(z164, z1[1], z1[2], z1[3], [u1_, v1_])
        m = list(mt64)
        m[2], m[4] = u2_, v2_
        # advance_mu_t's positions of its inputs (mu and mudf are advance_uv's too)
        for n, pos in (("ww", 0), ("mu", 6), ("muave", 8), ("muts", 9), ("mudf", 12), ("t", 24), ("t_ave", 26)):
            m[pos] = v[n]
        r = ro.advance_mu_t(*m)
        t2_, mu2_, muts2_ = ozone(z264, z2[1], z2[2], z2[3], [r[2], r[10], r[8]])
        w_ = list(wa64)
        w_[0], w_[2], w_[4], w_[5], w_[8], w_[9], w_[18], w_[19], w_[21] = (
            v["w"], r[0], u2_, v2_, r[7], muts2_, r[3], t2_, v["ph"])
        t2a_, wo_, pho_ = ro.advance_w(*w_)
        res = [u2_, v2_, r[0], t2_, mu2_, muts2_, r[9], r[7], t2a_, wo_, pho_]
        return sum(float((q * l).sum()) for q, l in zip(res, lams))

    base = {n: x[n].detach().double().cpu() for n in src}
    for n in src:
        d = torch.randn(tuple(base[n].shape), generator=gen, dtype=torch.float64)
        eps = 1e-4 * float(base[n].abs().max() + 1.0) / float(d.abs().max())
        vp, vm = dict(base), dict(base)
        vp[n], vm[n] = base[n] + eps * d, base[n] - eps * d
        ref = (chain(vp) - chain(vm)) / (2.0 * eps)
        terms = x[n].grad.double().cpu() * d
        got = float(terms.sum())
        err = abs(got - ref) / max(abs(ref), float(terms.norm()))
        print("d/d%s: <grad, d> %.6e  difference %.6e  rel %.2e" % (n, got, ref, err))
        assert err <= 1e-4, n


def _clone(a):
    if isinstance(a, torch.Tensor):
        return a.clone()
    if isinstance(a, list):
        return [_clone(x) for x in a]
    if isinstance(a, tuple):
        return tuple(_clone(x) for x in a)
    return a


def _double(a):
    if isinstance(a, torch.Tensor):
        return a.detach().double().cpu()
    if isinstance(a, list):
        return [_double(x) for x in a]
    if isinstance(a, tuple):
        return tuple(_double(x) for x in a)
    return a


@pytest.mark.parametrize("config", ["config1c", "bdy24", "config4"])
def test_acoustic_loop_adjoint_matches_fp64_differences(config):
    """The gradient of the second RK3 stage's whole acoustic loop (two
    substeps, each advance_uv, the u/v zone update, advance_mu_t, the
    t/mu/muts zone update, advance_w, sumflux, spec_bdyupdate_ph,
    zero_grad_bdy, calc_p_rho and the six open-boundary fills), taken by
    AcousticLoop (one checkpoint, the substeps run again in its backward)
    from the hand-written adjoints, with respect to all 17 fields the loop
    carries (u, v, p, ph, al, mu, mudf, ww, muave, muts, t, t_ave, w, pm1 and
    the three flux accumulators), the six large-step tendencies it reads
    (ru_tend, rv_tend, rw_tend, t_tend, ph_tend, mu_tend) and the six fields
    small_step_prep saved for the stage (u_save, v_save, w_save, t_save,
    ph_save, ww1), each wherever a call reads it, against FP64 central
    differences of the same sequence
    through the oracle: for each input a seeded direction d, <grad, d>
    against the difference of sum_o <lam_o, out_o> over the 17 results,
    relative to the larger of the difference and the 2-norm of grad_i d_i;
    bound 1e-4.  The two substeps take sumflux's first and last branches and
    calc_p_rho's damped one."""
    import oracle
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.adjoint import ACOUSTIC_STATE as STATE, FORCING, SAVED
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import RunConfig
    s = synthetic_state(config, device="cuda")
    x0, subs = adjoint.record_acoustic_loop(s, pywrf_amd.ops(), RunConfig(), stage=2)
    assert len(subs) == 2
    names = STATE + FORCING + SAVED
    x = {n: f.clone().requires_grad_()
         for n, f in zip(names, list(x0) + adjoint.loop_forcing(subs) + adjoint.loop_saved(subs))}
    out = adjoint.AcousticLoop.apply(subs, *[x[n] for n in names])
    st = dict(zip(STATE, out))
    gen = torch.Generator().manual_seed(23)
    lams = {n: torch.randn(tuple(st[n].shape), generator=gen, dtype=torch.float64) for n in STATE}
    sum((st[n] * lams[n].float().cuda()).sum() for n in STATE).backward()

    # the same sequence through the oracle in FP64
    ro = oracle.ops()
    subs64 = [_double(c) for c in subs]

    def ozone(call, fields, tends):
        items, dt, ix24, ix2d = call
        return [ro.spec_bdyupdate(f.clone(), td, dt, it[2], 1, *(ix2d if it[3] else ix24))
                for f, td, it in zip(fields, tends, items)]

    def loop(v):
        st = {n: v[n].clone() for n in STATE}
        fc = {n: v[n] for n in FORCING}
        sv = {n: v[n] for n in SAVED}
        for uv, z1, mt, z2, wa, sf, pz, cp, bc in subs64:
            a = list(uv)
            a[0], a[1], a[2], a[3], a[4], a[6], a[9], a[10], a[15] = (
                st["u"], fc["ru_tend"], st["v"], fc["rv_tend"], st["p"], st["ph"], st["al"], st["mu"], st["mudf"])
            u2, v2 = ozone(z1, ro.advance_uv(*[_clone(q) for q in a]), (fc["ru_tend"], fc["rv_tend"]))
            m = _clone(list(mt))
            m[0], m[2], m[4], m[6], m[8], m[9], m[12], m[24], m[26], m[27], m[28] = (
                st["ww"], u2, v2, st["mu"], st["muave"], st["muts"], st["mudf"], st["t"], st["t_ave"],
                fc["t_tend"], fc["mu_tend"])
            m[1], m[3], m[5], m[25] = sv["ww1"], sv["u_save"], sv["v_save"], sv["t_save"]
            m = [_clone(q) for q in m]
            r = ro.advance_mu_t(*m)
            t2, mu2, muts2 = ozone(z2, [r[2], r[10], r[8]], (fc["t_tend"], fc["mu_tend"], fc["mu_tend"]))
            w_ = _clone(list(wa))
            w_[0], w_[1], w_[2], w_[4], w_[5], w_[8], w_[9], w_[18], w_[19], w_[21], w_[24] = (
                st["w"].clone(), fc["rw_tend"], r[0], u2, v2, r[7], muts2, r[3].clone(), t2, st["ph"].clone(),
                fc["ph_tend"])
            w_[3], w_[20], w_[22] = sv["w_save"], sv["t_save"], sv["ph_save"]
            t2a, wo, pho = ro.advance_w(*w_)
            f = _clone(list(sf))
            f[0], f[1], f[2], f[16], f[17], f[18] = u2, v2, r[0], st["ru_m"].clone(), st["rv_m"].clone(), \
                st["ww_m"].clone()
            f[3], f[4], f[5] = sv["u_save"], sv["v_save"], sv["ww1"]
            ru_m, rv_m, ww_m = ro.sumflux(*f)
            pa, za = list(_clone(pz[0])), list(_clone(pz[1]))
            pa[0], pa[1], pa[2], pa[3], pa[4] = sv["ph_save"], pho.clone(), fc["ph_tend"], fc["mu_tend"], muts2
            ph_s = ro.spec_bdyupdate_ph(*pa)
            za[0] = wo.clone()
            w_z = ro.zero_grad_bdy(*za)
            c = _clone(list(cp))
            c[0], c[1], c[2], c[4], c[7], c[8], c[9] = st["al"], st["p"], ph_s, t2, st["pm1"].clone(), mu2, muts2
            c[5] = sv["t_save"]
            al, p, pm1 = ro.calc_p_rho(*c)
            items, ix24, ix2d = bc
            filled = []
            for q, (_, var, two) in zip((ph_s, al, p, muts2, mu2, r[9]), items):
                q = q.clone()
                filled.append(ro.set_physical_bc2d(q, var, *ix2d) if two else ro.set_physical_bc3d(q, var, *ix24))
            ph_f, al_f, p_f, muts_f, mu_f, mudf_f = filled
            st = {"u": u2, "v": v2, "p": p_f, "ph": ph_f, "al": al_f, "mu": mu_f, "mudf": mudf_f, "ww": r[0],
                  "muave": r[7], "muts": muts_f, "t": t2, "t_ave": t2a, "w": w_z, "pm1": pm1, "ru_m": ru_m,
                  "rv_m": rv_m, "ww_m": ww_m}
        return sum(float((st[n] * lams[n]).sum()) for n in STATE)

    base = {n: x[n].detach().double().cpu() for n in names}
    worst = 0.0
    for n in names:
        d = torch.randn(tuple(base[n].shape), generator=gen, dtype=torch.float64)
        eps = 1e-4 * float(base[n].abs().max() + 1.0) / float(d.abs().max())
        vp, vm = dict(base), dict(base)
        vp[n], vm[n] = base[n] + eps * d, base[n] - eps * d
        ref = (loop(vp) - loop(vm)) / (2.0 * eps)
        terms = x[n].grad.double().cpu() * d
        got = float(terms.sum())
        err = abs(got - ref) / max(abs(ref), float(terms.norm()), 1e-300)
        worst = max(worst, err)
        print("d/d%s: <grad, d> %.6e  difference %.6e  rel %.2e" % (n, got, ref, err), flush=True)
    for n in names:
        assert x[n].grad is not None, n
    assert worst <= 1e-4


def test_acoustic_loop_checkpoint_equals_the_stored_chain():
    """AcousticLoop (one checkpoint: the forward keeps the loop's input, the
    backward runs the substeps again) against the same substeps chained with
    every operand kept (acoustic_substep under autograd), on the third RK3
    stage's four substeps at 120 x 120 x 41: the gradients of all 17 inputs
    agree within 1e-5 relative, and after the forward the checkpointed form
    holds only its
    inputs and results (at most 2.1 times the 17 fields' bytes) while the
    kept chain holds more than twice that.  Both run the same kernels on the
    same operands; the adjoints' atomic adds sum in no fixed order, which
    over four substeps moves the column-summed gradients (mu, muts) by about
    1e-6 of their norm, so the bound is 1e-5."""
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import RunConfig
    s = synthetic_state("config4", device="cuda")
    x0, subs = adjoint.record_acoustic_loop(s, pywrf_amd.ops(), RunConfig(), stage=3)
    assert len(subs) == 4
    gen = torch.Generator().manual_seed(5)
    lams = [torch.randn(tuple(f.shape), generator=gen).cuda() for f in x0]
    fields = sum(f.numel() * f.element_size() for f in x0)

    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    x = [f.clone().requires_grad_() for f in x0]
    out = adjoint.AcousticLoop.apply(subs, *x)
    torch.cuda.synchronize()
    held_ck = torch.cuda.memory_allocated() - base
    torch.autograd.backward(out, lams)
    g_ck = [t.grad.clone() for t in x]
    del out, x

    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    x = [f.clone().requires_grad_() for f in x0]
    y = x
    for calls in subs:
        y = adjoint.acoustic_substep(y, calls)
    torch.cuda.synchronize()
    held_kept = torch.cuda.memory_allocated() - base
    torch.autograd.backward(y, lams)
    print("held after the forward: checkpoint %.1f MB, kept chain %.1f MB, 17 fields %.1f MB"
          % (held_ck / 1e6, held_kept / 1e6, fields / 1e6))
    for n, a, t in zip(adjoint.ACOUSTIC_STATE, g_ck, x):
        rel = float((a - t.grad).norm() / max(float(t.grad.norm()), 1e-30))
        print("d/d%s: checkpoint vs kept chain %.2e" % (n, rel))
        assert rel <= 1e-5, n
    assert held_ck <= 2.1 * fields
    assert held_kept > 2.0 * held_ck


def _oracle_substep(ro, st, sv, calls):
    """One acoustic substep through the oracle (FP64 operands), the recorded
    tendencies held, the carried fields st and the saved fields sv given."""
    uv, z1, mt, z2, wa, sf, pz, cp, bc = calls

    def ozone(call, fields):
        items, dt, ix24, ix2d = call
        return [ro.spec_bdyupdate(f.clone(), it[1], dt, it[2], 1, *(ix2d if it[3] else ix24))
                for f, it in zip(fields, items)]
    a = _clone(list(uv))
    a[0], a[2], a[4], a[6], a[9], a[10], a[15] = (st["u"], st["v"], st["p"], st["ph"], st["al"], st["mu"],
                                                  st["mudf"])
    u2, v2 = ozone(z1, ro.advance_uv(*[_clone(q) for q in a]))
    m = _clone(list(mt))
    m[0], m[2], m[4], m[6], m[8], m[9], m[12], m[24], m[26] = (
        st["ww"], u2, v2, st["mu"], st["muave"], st["muts"], st["mudf"], st["t"], st["t_ave"])
    m[1], m[3], m[5], m[25] = sv["ww1"], sv["u_save"], sv["v_save"], sv["t_save"]
    r = ro.advance_mu_t(*[_clone(q) for q in m])
    t2, mu2, muts2 = ozone(z2, [r[2], r[10], r[8]])
    w_ = _clone(list(wa))
    w_[0], w_[2], w_[3], w_[4], w_[5], w_[8], w_[9], w_[18], w_[19], w_[20], w_[21], w_[22] = (
        st["w"].clone(), r[0], sv["w_save"], u2, v2, r[7], muts2, r[3].clone(), t2, sv["t_save"], st["ph"].clone(),
        sv["ph_save"])
    t2a, wo, pho = ro.advance_w(*w_)
    f = _clone(list(sf))
    f[0], f[1], f[2], f[3], f[4], f[5] = u2, v2, r[0], sv["u_save"], sv["v_save"], sv["ww1"]
    f[16], f[17], f[18] = st["ru_m"].clone(), st["rv_m"].clone(), st["ww_m"].clone()
    ru_m, rv_m, ww_m = ro.sumflux(*f)
    pa, za = list(_clone(pz[0])), list(_clone(pz[1]))
    pa[0], pa[1], pa[4] = sv["ph_save"], pho.clone(), muts2
    ph_s = ro.spec_bdyupdate_ph(*pa)
    za[0] = wo.clone()
    w_z = ro.zero_grad_bdy(*za)
    c = _clone(list(cp))
    c[0], c[1], c[2], c[4], c[5], c[7], c[8], c[9] = (st["al"], st["p"], ph_s, t2, sv["t_save"], st["pm1"].clone(),
                                                      mu2, muts2)
    al, p, pm1 = ro.calc_p_rho(*c)
    items, ix24, ix2d = bc
    filled = [ro.set_physical_bc2d(q.clone(), var, *ix2d) if two else ro.set_physical_bc3d(q.clone(), var, *ix24)
              for q, (_, var, two) in zip((ph_s, al, p, muts2, mu2, r[9]), items)]
    ph_f, al_f, p_f, muts_f, mu_f, mudf_f = filled
    return {"u": u2, "v": v2, "p": p_f, "ph": ph_f, "al": al_f, "mu": mu_f, "mudf": mudf_f, "ww": r[0],
            "muave": r[7], "muts": muts_f, "t": t2, "t_ave": t2a, "w": w_z, "pm1": pm1, "ru_m": ru_m, "rv_m": rv_m,
            "ww_m": ww_m}


@pytest.mark.parametrize("config,stage", [("config1c", 2), ("bdy24", 2), ("config1c", 3), ("config4", 2)])
def test_acoustic_stage_adjoint_matches_fp64_differences(config, stage):
    """The gradient of the second (or third: four substeps, the diabatic term
    of the last stage) RK stage's whole acoustic part --
    small_step_prep, calc_p_rho at step 0, the open-boundary fills, the two
    acoustic substeps, calc_mu_uv_1 and small_step_finish -- from the
    hand-written adjoints chained by autograd (AcousticStage: one checkpoint,
    the stage run again in its backward), with respect to the stage's
    19 inputs (u_2, v_2, w_2, t_2, ph_2, mu_2, u_1 .. mu_1, ww, mudf, muave,
    t_ave and the flux accumulators), against FP64 central differences of the
    same sequence through the oracle, over the 17 results; bound 1e-4 as in
    the loop's test."""
    import oracle
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.adjoint import ACOUSTIC_STATE, STAGE_IN, STAGE_OUT
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import RunConfig
    s = synthetic_state(config, device="cuda")
    x0, calls = adjoint.record_acoustic_stage(s, pywrf_amd.ops(), RunConfig(), stage=stage)
    x = {n: f.clone().requires_grad_() for n, f in zip(STAGE_IN, x0)}
    out = dict(zip(STAGE_OUT, adjoint.AcousticStage.apply(calls, *[x[n] for n in STAGE_IN])))
    gen = torch.Generator().manual_seed(29)
    lams = {n: torch.randn(tuple(out[n].shape), generator=gen, dtype=torch.float64) for n in STAGE_OUT}
    sum((out[n] * lams[n].float().cuda()).sum() for n in STAGE_OUT).backward()

    ro = oracle.ops()
    c64 = [_double(c) for c in calls]
    pr, cp, _, bc = c64[:4]
    subs = [c64[4 + 9 * q:4 + 9 * q + 9] for q in range((len(c64) - 6) // 9)]
    mc, fin = c64[-2], c64[-1]

    def stage(v):
        a = _clone(list(pr))
        for n, pos in (("u_1", 0), ("u_2", 1), ("v_1", 2), ("v_2", 3), ("w_1", 4), ("w_2", 5), ("t_1", 6),
                       ("t_2", 7), ("ph_1", 8), ("ph_2", 9), ("mu_1", 11), ("mu_2", 12), ("ww", 34)):
            a[pos] = v[n].clone()
        r = ro.small_step_prep(*a)
        us, vs, ws, ts, ps, u, vv, w, t, ph, ww1, mu, muus, muvs, muts, musv = (
            r[5], r[6], r[7], r[8], r[9], r[10], r[11], r[12], r[13], r[14], r[16], r[18], r[19], r[20], r[21], r[23])
        c = _clone(list(cp))
        c[2], c[4], c[5], c[8], c[9] = ph, t, ts, mu, muts
        al, p, pm1 = ro.calc_p_rho(*c)
        items, ix24, ix2d = bc
        src = (items[0][0], items[1][0], ph, al, p, v["t_1"], ts, v["mu_1"], mu, v["mudf"])
        filled = [ro.set_physical_bc2d(q.clone(), var, *ix2d) if two else ro.set_physical_bc3d(q.clone(), var, *ix24)
                  for q, (_, var, two) in zip(src, items)]
        _, _, ph, al, p, _, ts, _, mu, mudf = filled
        st = {"u": u, "v": vv, "p": p, "ph": ph, "al": al, "mu": mu, "mudf": mudf, "ww": v["ww"].clone(),
              "muave": v["muave"], "muts": muts, "t": t, "t_ave": v["t_ave"], "w": w, "pm1": pm1,
              "ru_m": v["ru_m"], "rv_m": v["rv_m"], "ww_m": v["ww_m"]}
        sv = {"u_save": us, "v_save": vs, "w_save": ws, "t_save": ts, "ph_save": ps, "ww1": ww1}
        for cs in subs:
            st = _oracle_substep(ro, st, sv, cs)
        m = list(mc)
        muus, muvs = ro.calc_mu_uv_1(st["muts"], muus.clone(), muvs.clone(), *m[3:])
        f = _clone(list(fin))
        for n, pos in (("u", 0), ("v", 2), ("w", 4), ("t", 6), ("ph", 8), ("ww", 10), ("mu", 12), ("muts", 15)):
            f[pos] = st[n].clone()
        f[11], f[17], f[19], f[28], f[29], f[30], f[31], f[32], f[33] = ww1, muus, muvs, us, vs, ws, ts, ps, musv
        uo, vo, wo, pho, wwo, to, muo = ro.small_step_finish(*f)
        res = {"u": uo, "v": vo, "w": wo, "ph": pho, "ww": wwo, "t": to, "mu": muo}
        for n in ("p", "al", "pm1", "mudf", "muave", "muts", "t_ave", "ru_m", "rv_m", "ww_m"):
            res[n] = st[n]
        return sum(float((res[n] * lams[n]).sum()) for n in STAGE_OUT)

    base = {n: x[n].detach().double().cpu() for n in STAGE_IN}
    worst = 0.0
    for n in STAGE_IN:
        assert x[n].grad is not None, n
        d = torch.randn(tuple(base[n].shape), generator=gen, dtype=torch.float64)
        eps = 1e-4 * float(base[n].abs().max() + 1.0) / float(d.abs().max())
        vp, vm = dict(base), dict(base)
        vp[n], vm[n] = base[n] + eps * d, base[n] - eps * d
        ref = (stage(vp) - stage(vm)) / (2.0 * eps)
        terms = x[n].grad.double().cpu() * d
        got = float(terms.sum())
        err = abs(got - ref) / max(abs(ref), float(terms.norm()), 1e-300)
        worst = max(worst, err)
        print("d/d%s: <grad, d> %.6e  difference %.6e  rel %.2e" % (n, got, ref, err), flush=True)
    assert worst <= 1e-4


@pytest.mark.parametrize("which", ["u", "v", "w"])
@pytest.mark.parametrize("config", ["config1c", "bdy24", "config4"])
def test_advect_u_adjoint_matches_oracle_differences(config, which):
    """advect_u's (advect_v's, advect_w's) gradient with respect to u (v, w) (the mass fluxes
    of the step's first call held fixed), A^T lam by the scatter-form adjoint kernel,
    against FP64 central differences of the oracle's advect_u (whose
    in-place face updates autograd cannot follow) along three seeded
    directions, relative to the larger of the difference and the 2-norm of
    grad_i d_i; bound 1e-5.  The map is linear in u but for the edge
    x-faces' upwind choice, which the small steps do not flip."""
    import oracle
    import pywrf_amd
    from pywrf_amd import adjoint
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import rk3_step, RunConfig
    s = synthetic_state(config, device="cuda")
    name = "advect_%s_fresh" % which if hasattr(pywrf_amd.ops(), "advect_u_fresh") else "advect_" + which
    rec = _FirstCall(pywrf_amd.ops(), name)
    rk3_step(s, rec, RunConfig())
    args = rec.args
    if which == "w":
        # w is zero at the step's start in these states (advect_w's tendency
        # would be 0): a seeded field of 0.5 m/s stands in for it
        args = list(args)
        w0 = torch.randn(tuple(args[0].shape), generator=torch.Generator().manual_seed(5))
        args[0] = (0.5 * w0).to(args[0].device)
    x = args[0].clone().requires_grad_()
    T = {"u": adjoint.AdvectU, "v": adjoint.AdvectV, "w": adjoint.AdvectW}[which].apply(x, args)
    gen = torch.Generator().manual_seed(31)
    lam = torch.randn(tuple(T.shape), generator=gen, dtype=torch.float64)
    (T * lam.float().cuda()).sum().backward()
    a64 = [a.detach().double().cpu() if isinstance(a, torch.Tensor) else a for a in args]
    u0 = a64[0]

    def A(u):
        b = list(a64)
        b[0], b[1], b[2] = u, u, torch.zeros_like(a64[2])
        return getattr(oracle.ops(), "advect_" + which)(*b)
    ferr = _rel(T.detach().double().cpu(), A(u0.clone()))
    print("advect_%s forward vs oracle (FP64): %.2e" % (which, ferr))
    assert ferr <= 1e-5
    ids, ide = int(args[-18]), int(args[-17])
    for n in range(3):
        d = torch.randn(tuple(u0.shape), generator=gen, dtype=torch.float64)
        eps = 1e-4 * float(u0.abs().max() + 1.0) / float(d.abs().max())
        # no step across the upwind switch: the u that decides the edge faces
        # (columns ids+1 and ide-2) is not moved where it is near zero
        jds, jde = int(args[-16]), int(args[-15])
        if which == "u":
            for c in (ids + 1, ide - 2):
                near = u0[..., c].abs() < 10.0 * eps * d[..., c].abs()
                d[..., c][near] = 0.0
        elif which == "v":
            for r in (jds + 1, jde - 2):
                near = u0[..., r, :].abs() < 10.0 * eps * d[..., r, :].abs()
                d[..., r, :][near] = 0.0
        ref = float(((A(u0 + eps * d) - A(u0 - eps * d)) * lam).sum()) / (2.0 * eps)
        terms = x.grad.double().cpu() * d
        got = float(terms.sum())
        err = abs(got - ref) / max(abs(ref), float(terms.norm()))
        print("direction %d: <A^T lam, d> %.6e  difference %.6e  rel %.2e" % (n, got, ref, err))
        assert err <= 1e-5
