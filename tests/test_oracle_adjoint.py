"""The gradient checker of the adjoint tests, on the CPU: torch autograd
through the oracle's advect_scalar and rk_update_scalar (which update their
arrays in place) gives the derivative that FP64 central differences give.
The maps are linear in the field, so the differences are exact up to
rounding and a wrong gradient (a dropped face, a wrong sign) shows at once.
"""
import torch

N = None


def _setup():
    import oracle
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.step import RunConfig
    torch.set_num_threads(4)
    s = synthetic_state("config1c", dtype=torch.float64)
    cfg = RunConfig()
    gen = torch.Generator().manual_seed(0)
    # smooth, non-trivial mass fluxes of both signs
    for name in ("ru_m", "rv_m", "ww_m"):
        getattr(s, name).copy_(torch.randn(s.t.shape, generator=gen, dtype=torch.float64))
    return oracle.ops(), s, cfg, gen


def _advect(O, s, cfg, q):
    return O.advect_scalar(q, q, torch.zeros_like(q), s.ru_m, s.rv_m, s.ww_m, N, N, N, N, N, N, N, N, s.msftx, N,
                           s.fnm, s.fnp, cfg.rdx, cfg.rdy, s.rdnw, *s.grid.ix18)


def _stage(O, s, cfg, q1, q2, rk_step):
    t = _advect(O, s, cfg, q2[0])
    O.rk_update_scalar(0, 0, q1, q2, torch.zeros_like(q2), t.unsqueeze(0), N, N, N, s.msfty, N, N, s.mu_1, s.mu,
                       s.mub, rk_step, 18.0, 1, False, *s.grid.ix18)


def test_autograd_through_the_oracle_equals_central_differences():
    O, s, cfg, gen = _setup()
    lam = torch.randn(s.t.shape, generator=gen, dtype=torch.float64)

    def J(q):
        q1 = q.unsqueeze(0).clone()
        q2 = q.unsqueeze(0).clone()
        for rk in (1, 2, 3):
            _stage(O, s, cfg, q1, q2, rk)
        return (q2[0] * lam).sum() + 0.5 * (_advect(O, s, cfg, q) ** 2).sum()

    q = s.moist[1].clone().requires_grad_()
    J(q).backward()
    with torch.no_grad():
        for _ in range(5):
            d = torch.randn(q.shape, generator=gen, dtype=torch.float64)
            eps = 1e-3
            fd = float(J(q.detach() + eps * d) - J(q.detach() - eps * d)) / (2 * eps)
            ad = float((q.grad * d).sum())
            assert abs(ad - fd) <= 1e-8 * abs(fd), (ad, fd)
