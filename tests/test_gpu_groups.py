"""The batched entry points (pywrf_*_group, pywrf_coriolis_curvature) give
exactly what the individual calls they replace give."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_zero_group_zeroes_exactly_its_tensors():
    import pywrf_amd.wrf_dynamics as dyn
    torch.manual_seed(0)
    shapes = [(41, 40, 40), (40, 40), (7, 41, 40, 40), (7,), (1,), (3, 5, 11)]
    ts = [torch.randn(sh, device="cuda") + 3.0 for sh in shapes]
    keep = torch.randn(41, 40, 40, device="cuda")
    keep0 = keep.clone()
    dyn.zero_group(ts)
    torch.cuda.synchronize()
    for t in ts:
        assert int(torch.count_nonzero(t)) == 0
    assert torch.equal(keep, keep0)
    dyn.zero_group([])                       # no-op


@pytest.mark.parametrize("n", [40, 120])
def test_bc_group_equals_individual_calls(n):
    import pywrf_amd.wrf_dynamics as dyn
    from pywrf_amd.step import Grid
    g = Grid(41, n, n)
    torch.manual_seed(1)
    items = [(torch.randn(41, n, n, device="cuda"), v, 0) for v in "uvwtpef"]
    items += [(torch.randn(n, n, device="cuda"), v, 1) for v in "uvt"]
    ref = [t.clone() for t, _, _ in items]
    for r, (_, v, two) in zip(ref, items):
        if two:
            dyn.set_physical_bc2d(r, v, *g.ix2d)
        else:
            dyn.set_physical_bc3d(r, v, *g.ix24)
    dyn.set_physical_bc_group(items, g.ix24, g.ix2d)
    torch.cuda.synchronize()
    for r, (t, v, two) in zip(ref, items):
        assert torch.equal(r, t), (v, two)


def _ix2d(g):
    (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps, kpe,
     its, ite, jts, jte, kts, kte) = g.ix24
    return (ids, ide, jds, jde, 0, 0, ims, ime, jms, jme, 0, 0, ips, ipe, jps, jpe, 0, 0, its, ite, jts, jte, 0, 0)


@pytest.mark.parametrize("n", [40, 120])
def test_spec_bdyupdate_group_equals_individual_calls(n):
    import pywrf_amd.wrf_boundary as bdy
    from pywrf_amd.step import Grid
    g = Grid(41, n, n)
    ix2 = _ix2d(g)
    torch.manual_seed(2)
    items = [(torch.randn(41, n, n, device="cuda"), torch.randn(41, n, n, device="cuda"), v, 0) for v in "uvth"]
    items += [(torch.randn(n, n, device="cuda"), torch.randn(n, n, device="cuda"), "m", 1) for _ in range(2)]
    ref = [f.clone() for f, _, _, _ in items]
    for r, (_, t, v, two) in zip(ref, items):
        bdy.spec_bdyupdate(r, t, 13.5, v, 1, *(ix2 if two else g.ix24))
    bdy.spec_bdyupdate_group(items, 13.5, g.ix24, ix2)
    torch.cuda.synchronize()
    for r, (f, _, v, two) in zip(ref, items):
        assert torch.equal(r, f), (v, two)


@pytest.mark.parametrize("n", [40, 120])
def test_flow_dep_group_equals_individual_calls(n):
    import pywrf_amd.wrf_boundary as bdy
    from pywrf_amd.step import Grid
    g = Grid(41, n, n)
    torch.manual_seed(3)
    q = torch.rand(7, 41, n, n, device="cuda")
    u, v = torch.randn(41, n, n, device="cuda"), torch.randn(41, n, n, device="cuda")
    ref = q.clone()
    for im in range(2, 7):
        bdy.flow_dep_bdy(ref[im], u, v, 1, *g.ix24)
    bdy.flow_dep_bdy_group([q[im] for im in range(2, 7)], u, v, g.ix24)
    torch.cuda.synchronize()
    assert torch.equal(ref, q)


def test_coriolis_curvature_equals_the_two_calls():
    """The fused entry point against coriolis followed by curvature, on the
    state after one RK3 step of the 40x40 moist workload (bit for bit)."""
    import pywrf_amd
    from pywrf_amd.step import RunConfig
    from parity import state, run
    cfg = RunConfig()
    o = pywrf_amd.ops()
    s = state("config1c", device="cuda")
    run(s, o, 1)
    I18 = s.grid.ix18
    torch.manual_seed(4)
    t0 = [torch.randn_like(s.ru_tend) for _ in range(3)]
    ref = [t.clone() for t in t0]
    o.coriolis(s.ru, s.rv, s.rw, *ref, s.msftx, s.msfty, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.f, s.e, s.sina,
               s.cosa, s.fnm, s.fnp, *I18)
    o.curvature(s.ru, s.rv, s.rw, s.u, s.v, s.w, *ref, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty, s.clat,
                s.fnm, s.fnp, cfg.rdx, cfg.rdy, *I18)
    got = [t.clone() for t in t0]
    o.coriolis_curvature(s.ru, s.rv, s.rw, s.u, s.v, s.w, *got, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty,
                         s.f, s.e, s.sina, s.cosa, s.clat, s.fnm, s.fnp, cfg.rdx, cfg.rdy, *I18)
    torch.cuda.synchronize()
    for a, b in zip(got, ref):
        assert torch.equal(a, b)


def _bdy_arrays(kz, n, gen):
    return tuple(torch.randn((kz, n, 5) if m % 4 < 2 else (kz, 5, n), generator=gen).cuda() for m in range(8))


@pytest.mark.parametrize("n", [40, 120])
def test_spec_bdytend_group_equals_individual_calls(n):
    import pywrf_amd.wrf_boundary as bdy
    from pywrf_amd.step import Grid
    g = Grid(41, n, n)
    (ids, ide, jds, jde, _, _, ims, ime, jms, jme, _, _, ips, ipe, jps, jpe, _, _, its, ite, jts, jte, _, _) = g.ix24
    ix_mu = (ids, ide, jds, jde, 0, 1, ims, ime, jms, jme, 0, 1, ips, ipe, jps, jpe, 0, 1, its, ite, jts, jte, 0, 1)
    gen = torch.Generator().manual_seed(5)
    items = [(torch.randn(41, n, n, generator=gen).cuda(), _bdy_arrays(41, n, gen), v, False) for v in "uvht"]
    items.append((torch.randn(1, n, n, generator=gen).cuda(), _bdy_arrays(1, n, gen), "m", True))
    ref = [t.clone() for t, _, _, _ in items]
    for r, (_, arrs, v, two) in zip(ref, items):
        bdy.spec_bdytend(r, *arrs, v, 5, 1, *(ix_mu if two else g.ix24))
    bdy.spec_bdytend_group([(t, arrs[4:], v, two) for t, arrs, v, two in items], g.ix24, ix_mu)
    torch.cuda.synchronize()
    for r, (t, _, v, two) in zip(ref, items):
        assert torch.equal(r, t), (v, two)


@pytest.mark.parametrize("n", [40, 120])
def test_spec_bdy_final_group_equals_individual_calls(n):
    import pywrf_amd.wrf_boundary as bdy
    from pywrf_amd.step import Grid
    g = Grid(41, n, n)
    gen = torch.Generator().manual_seed(6)
    c1 = torch.rand(41, generator=gen).cuda()
    c2 = torch.rand(41, generator=gen).cuda()
    mus = [torch.rand(n, n, generator=gen).cuda() + 1.0 for _ in range(3)]
    msfs = [torch.rand(n, n, generator=gen).cuda() + 0.5 for _ in range(3)]
    q = torch.randn(7, 41, n, n, generator=gen).cuda()
    items = [(torch.randn(41, n, n, generator=gen).cuda(), mus[0], msfs[0], _bdy_arrays(41, n, gen), "u"),
             (torch.randn(41, n, n, generator=gen).cuda(), mus[1], msfs[1], _bdy_arrays(41, n, gen), "v"),
             (torch.randn(41, n, n, generator=gen).cuda(), mus[2], msfs[2], _bdy_arrays(41, n, gen), "t"),
             (torch.randn(41, n, n, generator=gen).cuda(), mus[2], msfs[2], _bdy_arrays(41, n, gen), "h"),
             (q[0], mus[2], msfs[2], _bdy_arrays(41, n, gen), "t")]
    ref = [f.clone() for f, _, _, _, _ in items]
    q0 = q.clone()
    for r, (_, mu, msf, arrs, v) in zip(ref, items):
        bdy.spec_bdy_final(r, mu, c1, c2, msf, *arrs, v, 5, 1, 36.0, *g.ix24)
    bdy.spec_bdy_final_group(items, 36.0, g.ix24)
    torch.cuda.synchronize()
    for r, (f, _, _, _, v) in zip(ref, items):
        assert torch.equal(r, f), v
    assert torch.equal(q[1:], q0[1:])


def _transport_args(s, rk_step, moist, arrs):
    from pywrf_amd.step import (RunConfig, KHDIF, DIFF_6TH_OPT, DIFF_6TH_FACTOR, MOIST_ADV_OPT,
                                SCALAR_ADV_OPT)
    cfg = RunConfig()
    old, cur, tend, adv, h, z = arrs
    mf = (s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty)
    sct = (cfg.rdx, cfg.rdy, s.rdn, s.rdnw, KHDIF, 5.0, s.xkhh, DIFF_6TH_OPT, DIFF_6TH_FACTOR)
    n = 6 if moist else 2
    return (1, n, False, rk_step, cfg.dt / 3, s.ru_m, s.rv_m, s.ww_m, s.muts, s.mub, s.mu_1, s.c1h, s.c2h, s.alt,
            old, cur, tend, adv, h, z, s.RQVFTEN, s.qv_base, moist, s.fnm, s.fnp, *mf, *sct,
            MOIST_ADV_OPT if moist else SCALAR_ADV_OPT, *s.grid.ix18)


def _update_args(s, rk_step, n, arrs, h, z):
    s1, s2, tend, adv = arrs
    return (1, n, s1, s2, tend, adv, h, z, s.msftx, s.msfty, s.c1h, s.c2h, s.mu_1, s.mu, s.mub, rk_step, 18.0, 1,
            False, *s.grid.ix18)


@pytest.mark.parametrize("rk_step", [1, 2])
def test_transport_pairs_equal_the_two_calls(rk_step):
    """rk_scalar_tend_pair and rk_update_scalar_pair against the moist call
    followed by the tracer call (bit for bit), on the state after one step,
    with random tracer species and tendencies and a nonzero vertical
    diffusion coefficient."""
    s, o = _stepped()
    torch.manual_seed(9)
    sc = torch.rand_like(s.scalar)
    base = {"m": [s.moist_old, s.moist, torch.randn_like(s.moist), torch.randn_like(s.moist)],
            "s": [torch.rand_like(s.scalar), sc, torch.randn_like(sc), torch.randn_like(sc)]}
    hz = [torch.randn_like(s.h_tendency), torch.randn_like(s.z_tendency)]
    ref = {k: [x.clone() for x in v] for k, v in base.items()}
    got = {k: [x.clone() for x in v] for k, v in base.items()}
    rh, gh = [x.clone() for x in hz], [x.clone() for x in hz]
    o.rk_scalar_tend(*_transport_args(s, rk_step, True, ref["m"] + rh))
    o.rk_scalar_tend(*_transport_args(s, rk_step, False, ref["s"] + rh))
    o.rk_scalar_tend_pair(_transport_args(s, rk_step, True, got["m"] + gh),
                          _transport_args(s, rk_step, False, got["s"] + gh))
    torch.cuda.synchronize()
    for k in ("m", "s"):
        for n, (a, b) in enumerate(zip(ref[k], got[k])):
            assert torch.equal(a, b), (k, n)
    assert torch.equal(rh[0], gh[0]) and torch.equal(rh[1], gh[1])
    o.rk_update_scalar(*_update_args(s, rk_step, 6, ref["m"], *rh))
    o.rk_update_scalar(*_update_args(s, rk_step, 2, ref["s"], *rh))
    o.rk_update_scalar_pair(_update_args(s, rk_step, 6, got["m"], *gh), _update_args(s, rk_step, 2, got["s"], *gh))
    torch.cuda.synchronize()
    for k in ("m", "s"):
        for n, (a, b) in enumerate(zip(ref[k], got[k])):
            assert torch.equal(a, b), (k, n)


def test_transport_pair_refuses_calls_that_differ_in_shared_arguments():
    s, o = _stepped()
    a = _transport_args(s, 2, True, [s.moist_old, s.moist, s.moist_tend.clone(), s.advect_tend.clone(),
                                     s.h_tendency, s.z_tendency])
    b = list(_transport_args(s, 2, False, [s.scalar_old, s.scalar, s.scalar_tend.clone(),
                                           s.advect_tend_sca.clone(), s.h_tendency, s.z_tendency]))
    b[5] = s.ru_m.clone()                          # a different ru
    with pytest.raises(ValueError):
        o.rk_scalar_tend_pair(a, tuple(b))


def _stepped(config="config1c"):
    import pywrf_amd
    from parity import state, run
    o = pywrf_amd.ops()
    s = state(config, device="cuda")
    run(s, o, 1)
    torch.cuda.synchronize()
    return s, o


def test_fresh_first_writers_equal_zeroing_then_the_call():
    """advect_u/v/w_fresh, advect_scalar_fresh and rhs_ph_fresh on a tendency
    holding random values against zeroing it and making the plain call (the
    rk_tendency sequence), bit for bit over the whole array."""
    from pywrf_amd.step import RunConfig, TIME_STEP
    cfg = RunConfig()
    s, o = _stepped()
    I18 = s.grid.ix18
    adv = (s.c1h, s.c2h, s.mut, TIME_STEP, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty, s.fnm, s.fnp,
           cfg.rdx, cfg.rdy)
    calls = {
        "advect_u": lambda f, t: f(s.u, s.u, t, s.ru, s.rv, s.ww, *adv, s.rdnw, *I18),
        "advect_v": lambda f, t: f(s.v, s.v, t, s.ru, s.rv, s.ww, *adv, s.rdnw, *I18),
        "advect_w": lambda f, t: f(s.w, s.w, t, s.ru, s.rv, s.ww, *adv, s.rdn, *I18),
        "advect_scalar": lambda f, t: f(s.t, s.t, t, s.ru, s.rv, s.ww, *adv, s.rdnw, *I18),
        "rhs_ph": lambda f, t: f(t, s.u, s.v, s.ww, s.ph, s.ph, s.phb, s.w, s.mut, s.muu, s.muv, s.c1f, s.c2f,
                                 s.fnm, s.fnp, s.rdnw, s.cfn, s.cfn1, cfg.rdx, cfg.rdy, s.msfux, s.msfuy, s.msfvx,
                                 s.msfvx_inv, s.msfvy, s.msftx, s.msfty, True, *I18),
    }
    torch.manual_seed(10)
    for name, run_call in calls.items():
        junk = torch.randn_like(s.u) * 1e3
        ref = torch.zeros_like(junk)
        run_call(getattr(o, name), ref)
        got = junk.clone()
        run_call(getattr(o, name + "_fresh"), got)
        torch.cuda.synchronize()
        assert int(torch.count_nonzero(ref)) > 0, name
        assert torch.equal(ref, got), name


@pytest.mark.parametrize("config", ["config1c", "config4"])
def test_pgf_buoy_coriolis_curvature_equals_the_three_calls(config):
    """The fused pressure-gradient / buoyancy / Coriolis-curvature entry point
    against horizontal_pressure_gradient, pg_buoy_w and coriolis_curvature in
    sequence, bit for bit (tendencies and cqw), on the state after one step
    with random incoming tendencies."""
    from pywrf_amd.step import RunConfig, G
    cfg = RunConfig()
    s, o = _stepped(config)
    I18 = s.grid.ix18
    torch.manual_seed(13)
    t0 = [torch.randn_like(s.ru_tend) for _ in range(3)]
    cq0 = s.cqw.clone()

    def args(ru_t, rv_t, rw_t, cqw):
        return ((ru_t, rv_t, s.ph, s.alt, s.p, s.pb, s.al, s.php, s.cqu, s.cqv, s.muu, s.muv, s.mu, s.c1h, s.c2h,
                 s.fnm, s.fnp, s.rdnw, s.cf1, s.cf2, s.cf3, s.cfn, s.cfn1, cfg.rdx, cfg.rdy, s.msfux, s.msfuy,
                 s.msfvx, s.msfvy, s.msftx, s.msfty, True, False, *I18),
                (rw_t, s.p, cqw, s.mu, s.mub, s.c1f, s.c2f, s.rdnw, s.rdn, G, s.msftx, s.msfty, *I18),
                (s.ru, s.rv, s.rw, s.u, s.v, s.w, ru_t, rv_t, rw_t, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.msftx,
                 s.msfty, s.f, s.e, s.sina, s.cosa, s.clat, s.fnm, s.fnp, cfg.rdx, cfg.rdy, *I18))
    ref = [t.clone() for t in t0] + [cq0.clone()]
    ha, ba, ca = args(*ref)
    o.horizontal_pressure_gradient(*ha)
    o.pg_buoy_w(*ba)
    o.coriolis_curvature(*ca)
    got = [t.clone() for t in t0] + [cq0.clone()]
    o.pgf_buoy_coriolis_curvature(*args(*got))
    torch.cuda.synchronize()
    for n, (a, b) in enumerate(zip(ref, got)):
        assert torch.equal(a, b), n


@pytest.mark.parametrize("config", ["config1c", "config4"])
def test_diffusion_dry_equals_the_eight_calls(config):
    """The first stage's fused diffusion entry point against the eight
    horizontal/vertical diffusion calls in the order solver.py:1079-1141 makes
    them, bit for bit, with random tendencies, positive random eddy
    coefficients and non-zero vertical coefficients."""
    from pywrf_amd.step import RunConfig
    cfg = RunConfig()
    s, o = _stepped(config)
    I18 = s.grid.ix18
    gen = torch.Generator(device="cuda").manual_seed(29)
    t0 = [torch.randn(x.shape, device="cuda", generator=gen) for x in (s.ru_tendf, s.rv_tendf, s.rw_tendf, s.t_tendf)]
    xkmh = torch.rand(s.xkmh.shape, device="cuda", generator=gen) * 50.
    xkhh = torch.rand(s.xkhh.shape, device="cuda", generator=gen) * 150.
    hd = (s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty)
    rdx, rdy, kv = cfg.rdx, cfg.rdy, 0.7

    def args(ru, rv, rw, tt):
        return (('u', s.u, ru, s.mut, s.c1h, s.c2h, *hd, 0., xkmh, rdx, rdy, *I18),
                ('v', s.v, rv, s.mut, s.c1h, s.c2h, *hd, 0., xkmh, rdx, rdy, *I18),
                ('w', s.w, rw, s.mut, s.c1f, s.c2f, *hd, 0., xkmh, rdx, rdy, *I18),
                ('m', s.t, tt, s.mut, s.c1h, s.c2h, s.t_init, *hd, 0., xkhh, rdx, rdy, *I18),
                (s.u, ru, s.u_base, s.c1h, s.c2h, s.alt, s.muu, s.rdn, s.rdnw, kv, *I18),
                (s.v, rv, s.v_base, s.c1h, s.c2h, s.alt, s.muv, s.rdn, s.rdnw, kv, *I18),
                ('w', s.w, rw, s.c1f, s.c2f, s.alt, s.mut, s.rdn, s.rdnw, kv, *I18),
                (s.t, tt, s.t_init, s.c1h, s.c2h, s.alt, s.mut, s.rdn, s.rdnw, 3. * kv, *I18))
    ref = [t.clone() for t in t0]
    hu, hv, hw, ht, vu, vv, vw, vt = args(*ref)
    o.horizontal_diffusion(*hu)
    o.horizontal_diffusion(*hv)
    o.horizontal_diffusion(*hw)
    o.horizontal_diffusion_3dmp(*ht)
    o.vertical_diffusion_u(*vu)
    o.vertical_diffusion_v(*vv)
    o.vertical_diffusion(*vw)
    o.vertical_diffusion_3dmp(*vt)
    got = [t.clone() for t in t0]
    o.diffusion_dry(*args(*got))
    torch.cuda.synchronize()
    for n, (a, b) in enumerate(zip(ref, got)):
        assert not torch.equal(a, t0[n]), n
        assert torch.equal(a, b), n
    bad = list(args(*got))
    bad[5] = (s.u,) + bad[5][1:]  # the v vertical call on another field
    with pytest.raises(ValueError):
        o.diffusion_dry(*bad)


def _epilogue_items(s):
    """The end-of-step fills and spec_bdy_final calls of the step (solver.py:1776-1904)."""
    b = s.bdy
    bcs = [(s.u_1, 'U', 0), (s.u, 'U', 0), (s.v_1, 'V', 0), (s.v, 'V', 0), (s.w_1, 'w', 0), (s.w, 'w', 0),
           (s.t_1, 'p', 0), (s.t, 'p', 0), (s.ph_1, 'w', 0), (s.ph, 'w', 0), (s.mu_1, 't', 1), (s.mu, 't', 1),
           (s.tke_1, 'p', 0), (s.tke, 'p', 0)] + [(s.moist[im], 'p', 0) for im in range(1, 7)] + \
        [(s.scalar[im], 'p', 0) for im in range(1, 3)]
    finals = [(s.u, s.muus, s.msfuy, b['u'], 'u'), (s.v, s.muvs, s.msfvx, b['v'], 'v'),
              (s.t, s.muts, s.msfty, b['t'], 't'), (s.ph, s.muts, s.msfty, b['ph'], 'h'),
              (s.moist[1], s.muts, s.msfty, b['qv'], 't')]
    return bcs, finals


@pytest.mark.parametrize("on_device", [False, True])
def test_boundary_epilogue_equals_the_grouped_calls(on_device):
    """One launch for the end-of-step fills and boundary rows equals
    set_physical_bc_group then spec_bdy_final_group, bit for bit, on a state
    with live boundary tendencies (24 x 24 x 41 after two steps), every field
    and its halo randomised first so that each copy and each row is seen."""
    import pywrf_amd
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.state import clone_state
    from pywrf_amd.step import rk3_step, RunConfig
    s0 = synthetic_state("bdy24", device="cuda")
    for _ in range(2):
        rk3_step(s0, pywrf_amd.ops(), RunConfig())
    gen = torch.Generator(device="cuda").manual_seed(9)
    bcs0, _ = _epilogue_items(s0)
    for t, _, _ in bcs0:
        t.add_(torch.randn(t.shape, device="cuda", generator=gen) * (t.abs().max() * 0.1 + 1e-3))
    a, c = clone_state(s0, device="cuda"), clone_state(s0, device="cuda")
    g = s0.grid
    dtbc = torch.full((1,), 108.0, device="cuda") if on_device else 108.0
    bcs_a, fin_a = _epilogue_items(a)
    pywrf_amd.ops().set_physical_bc_group(bcs_a, g.ix24, g.ix2d)
    pywrf_amd.ops().spec_bdy_final_group(fin_a, dtbc, g.ix24)
    bcs_c, fin_c = _epilogue_items(c)
    pywrf_amd.ops().boundary_epilogue(bcs_c, fin_c, dtbc, g.ix24, g.ix2d)
    torch.cuda.synchronize()
    for (ta, v, _), (tc, _, _) in zip(bcs_a, bcs_c):
        assert torch.equal(ta, tc), v
    changed = sum(int((ta != t0).sum()) for (ta, _, _), (t0, _, _) in zip(bcs_a, bcs0))
    assert changed > 0


@pytest.mark.parametrize("n", [40, 120])
def test_spec_bdyupdate_ph_zero_grad_equals_the_two_calls(n):
    """The paired launch against spec_bdyupdate_ph on ph then zero_grad_bdy on
    w, bit for bit on both whole arrays, with random fields (positive column
    masses)."""
    import pywrf_amd.wrf_boundary as bdy
    from pywrf_amd.step import Grid
    g = Grid(41, n, n)
    torch.manual_seed(4)
    r3 = lambda: torch.randn(41, n, n, device="cuda")
    ph_save, ph_tend, w0, ph0 = r3(), r3(), r3(), r3()
    mu_tend = torch.randn(n, n, device="cuda")
    muts = torch.rand(n, n, device="cuda") * 1e4 + 5e4
    c1f, c2f = torch.rand(41, device="cuda"), torch.rand(41, device="cuda")
    args = lambda ph: (ph_save, ph, ph_tend, mu_tend, muts, c1f, c2f, 13.5, "h", 1, *g.ix24)
    ph_a, w_a = ph0.clone(), w0.clone()
    bdy.spec_bdyupdate_ph(*args(ph_a))
    bdy.zero_grad_bdy(w_a, "w", 1, *g.ix24)
    ph_b, w_b = ph0.clone(), w0.clone()
    bdy.spec_bdyupdate_ph_zero_grad(args(ph_b), (w_b, "w", 1, *g.ix24))
    torch.cuda.synchronize()
    assert not torch.equal(ph_a, ph0) and not torch.equal(w_a, w0)
    assert torch.equal(ph_a, ph_b)
    assert torch.equal(w_a, w_b)
