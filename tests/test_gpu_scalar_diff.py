"""First-stage diffusion of the transported species (rk_scalar_tend with
rk_step 1, wrf_dynamics.py:4018-4046) with nonzero coefficients, against the
oracle on identical inputs.  The step's own configuration has khdif = kvdif =
0, where the diffusion terms vanish; here kvdif is set and xkmhd is a random
positive field, so every face factor, the 'p' kind of water vapour in the
moist call and the kts rule of kind 'm' enter the result."""
import pytest
import torch

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _args(s, cfg, scs, sce, moist_step, xkmhd, kvdif, dev):
    from pywrf_amd.step import DIFF_6TH_OPT, DIFF_6TH_FACTOR
    d = (lambda t: t.to(dev).clone()) if dev != "cpu" else (lambda t: t.clone())
    arr = s.moist if moist_step else torch.rand_like(s.scalar)   # the tracers are zero in this state
    old = s.moist_old if moist_step else s.scalar_old
    tend = torch.randn_like(arr)
    adv = torch.zeros_like(arr)
    return (scs, sce, False, 1, cfg.dt / 3, d(s.ru_m), d(s.rv_m), d(s.ww_m), d(s.muts), d(s.mub), d(s.mu_1),
            d(s.c1h), d(s.c2h), d(s.alt), d(old), d(arr), d(tend), d(adv), d(s.h_tendency), d(s.z_tendency),
            d(s.RQVFTEN), d(s.qv_base), moist_step, d(s.fnm), d(s.fnp), d(s.msfux), d(s.msfuy), d(s.msfvx),
            d(s.msfvx_inv), d(s.msfvy), d(s.msftx), d(s.msfty), cfg.rdx, cfg.rdy, d(s.rdn), d(s.rdnw), 0.0, kvdif,
            d(xkmhd), DIFF_6TH_OPT, DIFF_6TH_FACTOR, 0, *s.grid.ix18)


@pytest.mark.parametrize("moist_step", [True, False])
def test_scalar_diffusion_matches_oracle_with_nonzero_coefficients(moist_step):
    import oracle
    import pywrf_amd
    from parity import state, run
    from pywrf_amd.step import RunConfig
    cfg = RunConfig()
    s = state("config1c")
    run(s, oracle.ops(), 1)
    torch.manual_seed(11)
    xkmhd = torch.rand_like(s.xkhh) * 3.0e4 + 1.0e3
    scs, sce = (1, 6) if moist_step else (1, 2)
    torch.manual_seed(12)
    a_cpu = _args(s, cfg, scs, sce, moist_step, xkmhd, 5.0, "cpu")
    torch.manual_seed(12)
    a_gpu = _args(s, cfg, scs, sce, moist_step, xkmhd, 5.0, "cuda")
    t0 = a_cpu[16].clone()
    ref = oracle.ops().rk_scalar_tend(*a_cpu)
    got = pywrf_amd.ops().rk_scalar_tend(*a_gpu)
    torch.cuda.synchronize()
    r, g = ref[0].double(), got[0].double().cpu()
    change = (r - t0.double()).abs().max().item()
    err = (g - r).abs().max().item() / r.abs().max().item()
    print("diffusion change %.3e, normwise difference %.2e" % (change, err))
    assert change > 1e-6 * r.abs().max().item()     # the diffusion did something
    assert err <= TOL
    a_r, a_g = ref[1].double(), got[1].double().cpu()
    assert (a_g - a_r).abs().max().item() <= TOL * max(a_r.abs().max().item(), 1e-30)
