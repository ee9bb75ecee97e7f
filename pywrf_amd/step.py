"""One RK3 step of the ARW core: the body of the reference's time loop
(pywrf/solver.py:454-1910) restated as a function.

`rk3_step(s, ops, cfg)` calls the hot-path functions through `ops`, a
namespace exposing them under the reference's names and positional
signatures.  With `pywrf_amd.ops()` every call lands in the HIP kernels
(the drop-in path); the same driver also runs any other implementation with
the same signatures.  The driver reproduces the reference's rebinding and
in-place semantics exactly (which functions return fresh tensors and which
update their arguments, the `mu_tend_e` view aliasing at solver.py:983,
1195-1203 and 1248-1255, and the stage/step boundary-condition calls).

Deviations from the reference loop that do not change any computed value:
probe prints, the unused argmax diagnostics (solver.py:752-757, 1049-1054)
and `moist_last_step` are dropped; the NaN exits (solver.py:1648-1654,
1768-1774) raise `FloatingPointError` instead of calling sys.exit, and are
only evaluated when `check_nan=True` (they force a device sync).
"""
from dataclasses import dataclass


@dataclass
class RunConfig:
    dt: float = 54.0
    soundsteps: int = 4
    bdy_interval: int = 360
    rk_order: int = 3
    p_top: float = 5000.0
    dx: float = 9000.0
    dy: float = 9000.0

    @property
    def rdx(self):
        return 1. / self.dx

    @property
    def rdy(self):
        return 1. / self.dy


# namelist values the reference driver passes literally (config_params.py:29-49, 114-118)
SPEC_BDY_WIDTH, SPEC_ZONE, RELAX_ZONE = 5, 1, 4
SMDIV, EMDIV, KVDIF, KHDIF = 0.1, 0.01, 0., 0.
EPSSM, T0, P0, G = 0.1, 300.0, 100000.0, 9.810000
DAMPCOEF, ZDAMP, DAMP_OPT = 0.2, 5000., 3
DIFF_6TH_OPT, DIFF_6TH_FACTOR, MOIST_ADV_OPT, SCALAR_ADV_OPT = 0, 0.25, 0, 0
HYPSOMETRIC_OPT, N_MOIST, N_SCALAR, TIME_STEP = 1, 7, 3, 60
P_QV, P_QC = 1, 2


class Grid:
    """WRF index sets for one patch of an (nz, ny, nx) allocation with a
    5-point halo (config_params.py:67-99 at ny = nx = 240)."""

    def __init__(self, nz, ny, nx, halo=5):
        self.nz, self.ny, self.nx = nz, ny, nx
        self.ids = self.its = self.ips = halo
        self.jds = self.jts = self.jps = halo
        self.ide = self.ite = self.ipe = nx - halo
        self.jde = self.jte = self.jpe = ny - halo
        self.kds = self.kts = self.kps = self.kms = 0
        self.kde = self.kte = self.kpe = self.kme = nz
        self.ims, self.ime, self.jms, self.jme = 0, nx, 0, ny

    @property
    def ix18(self):
        return (self.ids, self.ide, self.jds, self.jde, self.kds, self.kde,
                self.ims, self.ime, self.jms, self.jme, self.kms, self.kme,
                self.its, self.ite, self.jts, self.jte, self.kts, self.kte)

    @property
    def ix24(self):
        return (self.ids, self.ide, self.jds, self.jde, self.kds, self.kde,
                self.ims, self.ime, self.jms, self.jme, self.kms, self.kme,
                self.ips, self.ipe, self.jps, self.jpe, self.kps, self.kpe,
                self.its, self.ite, self.jts, self.jte, self.kts, self.kte)

    @property
    def ix2d(self):
        return (self.ids, self.ide, self.jds, self.jde, self.ims, self.ime, self.jms, self.jme,
                self.ips, self.ipe, self.jps, self.jpe, self.its, self.ite, self.jts, self.jte)

    def phys18(self):
        """Index set the driver passes to the physics (one in from each edge)."""
        return (self.ids + 1, self.ide - 1, self.jds + 1, self.jde - 1, self.kds, self.kde,
                self.ims, self.ime, self.jms, self.jme, self.kms, self.kme,
                self.its + 1, self.ite - 1, self.jts + 1, self.jte - 1, self.kts, self.kte)


class DeviceClock:
    """The step counter, the boundary record index it_bdy and the boundary
    time dtbc = dt (step - it_bdy bdy_interval) (solver.py:457-458, 713) held
    as device tensors and advanced by one kernel (pywrf_clock_advance) at the
    start of each step, so that a captured graph of steps forms each replay's
    dtbc itself.
    dtbc is formed in double and rounded to float once, the value the
    kernels take for the host's Python float."""

    def __init__(self, s, device="cuda"):
        import torch
        self.step = torch.tensor(s.step, dtype=torch.int64, device=device)
        self.it_bdy = torch.tensor(s.it_bdy, dtype=torch.int64, device=device)
        self.dtbc = torch.zeros((), dtype=torch.float32, device=device)

    def sync_from(self, s):
        """Set the device counters to the host's (asynchronous fills)."""
        self.step.fill_(s.step)
        self.it_bdy.fill_(s.it_bdy)

    def advance(self, cfg):
        from ._lib import call, stream
        call("pywrf_clock_advance", self.step.data_ptr(), self.it_bdy.data_ptr(), self.dtbc.data_ptr(),
             int(cfg.bdy_interval), float(cfg.dt), stream())


def _nan_check(p, where):
    import torch
    if bool(torch.isnan(p).any()):
        raise FloatingPointError("NaN in p (%s)" % where)


class NanFlag:
    """The reference's NaN exits (solver.py:1648-1654, 1768-1774) kept on the
    device: rk3_step(..., nan_flag=f) scans p after each RK stage (codes 1-3)
    and after the physics (code 4) with pywrf_nan_scan, which records the
    first check that saw a NaN, one NaN point of it and the NaN count; no
    synchronisation inside the step, so the scans can be captured in a HIP
    graph.  check() reads the record (one small copy) and raises
    FloatingPointError naming the check and the point."""

    WHERE = {1: "rk_step 1", 2: "rk_step 2", 3: "rk_step 3", 4: "physics"}

    def __init__(self, device="cuda"):
        import torch
        self.flag = torch.zeros(4, dtype=torch.int32, device=device)

    def _ptr(self):
        import ctypes
        return ctypes.c_void_p(self.flag.data_ptr())

    def clear(self):
        from ._lib import call, stream
        call("pywrf_nan_flag_clear", self._ptr(), stream())

    def scan(self, field, code, idx):
        from ._lib import call, grid, ptr, stream
        call("pywrf_nan_scan", grid(field.shape, idx), ptr(field), int(code), self._ptr(), stream())

    def record(self):
        code, first, count, _ = (int(x) for x in self.flag.cpu())
        return code, first, count

    def check(self, shape):
        code, first, count = self.record()
        if code:
            nz, ny, nx = shape
            k, rem = divmod(first, ny * nx)
            j, i = divmod(rem, nx)
            raise FloatingPointError("NaN in p after %s: %d points, one at (k, j, i) = (%d, %d, %d)"
                                     % (self.WHERE.get(code, code), count, k, j, i))


def rk3_step(s, ops, cfg, check_nan=False, clock=None, nan_flag=None):
    """Advance state `s` (a namespace of tensors, see pywrf_amd.state) by one
    RK3 step using the functions in `ops`.  solver.py:454-1910.  With a
    DeviceClock the boundary time dtbc is its device tensor (the form a
    captured graph needs); otherwise it is computed on the host."""
    gr = s.grid
    I18, I24, I2 = gr.ix18, gr.ix24, gr.ix2d
    ids, ide, jds, jde, kds, kde = gr.ids, gr.ide, gr.jds, gr.jde, gr.kds, gr.kde
    ims, ime, jms, jme, kms, kme = gr.ims, gr.ime, gr.jms, gr.jme, gr.kms, gr.kme
    ips, ipe, jps, jpe, kps, kpe = gr.ips, gr.ipe, gr.jps, gr.jpe, gr.kps, gr.kpe
    its, ite, jts, jte, kts, kte = gr.its, gr.ite, gr.jts, gr.jte, gr.kts, gr.kte
    o = ops
    b = s.bdy
    bc3 = lambda a, v: o.set_physical_bc3d(a, v, *I24)
    bc2 = lambda a, v: o.set_physical_bc2d(a, v, *I2)
    grouped = hasattr(o, "set_physical_bc_group")
    grouped_upd = hasattr(o, "spec_bdyupdate_group")

    def bcs(*items):
        """A run of consecutive open-boundary fills, items (tensor, flag, is_2d);
        one batched launch when the implementation offers it."""
        if grouped:
            o.set_physical_bc_group(list(items), I24, I2)
            return [t for t, _, _ in items]
        return [bc2(t, v) if two else bc3(t, v) for t, v, two in items]

    def zeros(*arrays):
        """A run of zero_() calls; one launch when the implementation offers it."""
        if hasattr(o, "zero_group"):
            o.zero_group(list(arrays))
        else:
            for a in arrays:
                a.zero_()

    s.step += 1
    if s.step % cfg.bdy_interval == 1:
        s.it_bdy = int(s.step / cfg.bdy_interval)
    dtm = cfg.dt
    dts = cfg.dt / cfg.soundsteps
    if clock is not None:
        clock.advance(cfg)
        dtbc = clock.dtbc
    else:
        dtbc = cfg.dt * (s.step - s.it_bdy * cfg.bdy_interval)
    rdx, rdy = cfg.rdx, cfg.rdy

    for rk_step in range(1, 4):
        if rk_step == 1:
            dt_rk, dts_rk, nsmall = dtm / 3., dtm / 3., 1
        elif rk_step == 2:
            dt_rk, dts_rk, nsmall = 0.5 * dtm, dts, int(cfg.soundsteps / 2)
        else:
            dt_rk, dts_rk, nsmall = dtm, dts, cfg.soundsteps
        # rk_step_prep (solver.py:736-778)
        s.mut = o.calculate_full(s.mut, s.mub, s.mu, *I18)
        s.muu, s.muv = o.calc_mu_uv(s.mu, s.mub, s.muu, s.muv, *I18)
        s.ru, s.rv, s.rw = o.couple_momentum(s.muu, s.ru, s.u, s.msfu, s.muv, s.rv, s.v, s.msfv, s.msfvx_inv,
                                             s.mut, s.rw, s.w, s.msft, s.c1h, s.c2h, s.c1f, s.c2f, *I18)
        s.ww = o.calc_ww_cp(s.u, s.v, s.mu, s.mub, s.muu, s.muv, s.c1h, s.c2h, s.ww, rdx, rdy, s.msftx, s.msfty,
                            s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.dnw, *I18)
        s.cqu, _cuv, s.cqw = o.calc_cq(s.moist, s.cqu, s.cqv, s.cqw, N_MOIST, *I18)
        s.alt = o.calc_alt(s.alt, s.al, s.alb, *I18)
        s.php = o.calc_php(s.php, s.ph, s.phb, *I18)
        # rk_phys_bc_dry_1 (solver.py:781-843)
        (s.ru, s.rv, s.rw, s.ww, s.php, s.alt, s.p, s.muu, s.muv, s.mut, s.al, s.ph) = bcs(
            (s.ru, 'u', 0), (s.rv, 'v', 0), (s.rw, 'w', 0), (s.ww, 'w', 0), (s.php, 'w', 0), (s.alt, 't', 0),
            (s.p, 'p', 0), (s.muu, 'u', 1), (s.muv, 'v', 1), (s.mut, 't', 1), (s.al, 'p', 0), (s.ph, 'w', 0))
        if rk_step == 1:
            # first_rk_step_part1/2 (solver.py:846-968)
            zeros(s.ru_tendf, s.rv_tendf, s.rw_tendf, s.ph_tendf, s.t_tendf, s.tke_tend, s.mu_tendf,
                  s.moist_tend, s.scalar_tend)
            (s.th_phy, s.p_phy, s.pi_phy, s.t_phy, s.rho, s.u_phy, s.v_phy, s.z_at_w, s.dz8w, s.z, s.p8w, s.t8w,
             s.p_hyd_w, s.p_hyd) = o.phy_prep(
                s.mut, s.muu, s.muv, s.c1h, s.c2h, s.c1f, s.c2f, s.u, s.v, s.p, s.pb, s.alt, s.ph, s.phb, s.t,
                s.moist, N_MOIST, s.rho, s.th_phy, s.p_phy, s.pi_phy, s.u_phy, s.v_phy, s.p8w, s.t_phy, s.t8w,
                s.z, s.z_at_w, s.dz8w, s.p_hyd, s.p_hyd_w, s.dnw, s.fnm, s.fnp, s.znw, cfg.p_top, *I18)
            (s.rublten, s.rvblten, s.rthblten, s.rqvblten, s.rqcblten, _rqi, s.scalar_tend) = o.calculate_phy_tend(
                s.c1h, s.c2h, s.mut, s.muu, s.muv, s.pi_phy, s.rublten, s.rvblten, s.rthblten, s.rqvblten,
                s.rqcblten, s.rqiblten, s.scalar, s.scalar_tend, N_SCALAR, *I18)
            s.rdzw, s.rdz, s.zx, s.zy, s.z = o.compute_diff_metrics(s.ph, s.phb, s.z, s.rdz, s.rdzw, s.zx, s.zy,
                                                                    rdx, rdy, *I18)
            s.rdzw, s.rdz, s.z, s.zx, s.zy = bcs((s.rdzw, 'w', 0), (s.rdz, 'w', 0), (s.z, 'w', 0), (s.zx, 'e', 0),
                                                 (s.zy, 'f', 0))
            (s.defor11, s.defor12, s.defor13, s.defor22, s.defor23, s.defor33, s.div) = o.cal_deform_and_div(
                s.u, s.v, s.w, s.div, s.defor11, s.defor22, s.defor33, s.defor12, s.defor13, s.defor23,
                s.u_base, s.v_base, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty, rdx, rdy, s.dn, s.dnw,
                s.rdz, s.rdzw, s.fnm, s.fnp, s.cf1, s.cf2, s.cf3, s.zx, s.zy, *I18)
            s.BN2, s.xkmh, s.xkmv, s.xkhh, s.xkhv = o.calculate_km_kh(
                dtm, DAMPCOEF, ZDAMP, DAMP_OPT, s.xkmh, s.xkmv, s.xkhh, s.xkhv, s.BN2, KHDIF, KVDIF, s.div,
                s.defor11, s.defor22, s.defor33, s.defor12, s.defor13, s.defor23, s.tke, s.p8w, s.t8w,
                s.th_phy, s.t_phy, s.p_phy, s.moist, s.dn, s.dnw, cfg.dx, cfg.dy, s.rdz, s.rdzw, 0, N_MOIST,
                s.cf1, s.cf2, s.cf3, 0, 0.1, s.msftx, s.msfty, s.zx, s.zy, *I18)
            s.rublten, s.rvblten, s.xkmh, s.xkhh = bcs((s.rublten, 't', 0), (s.rvblten, 't', 0), (s.xkmh, 't', 0),
                                                       (s.xkhh, 't', 0))
            s.t_tendf, s.ru_tendf, s.rv_tendf, s.moist_tend = o.update_phy_ten(
                s.ph_tendf, s.t_tendf, s.ru_tendf, s.rv_tendf, s.moist_tend, s.scalar_tend, s.mu_tendf,
                s.rthblten, s.rublten, s.rvblten, s.rqvblten, s.rqcblten, s.rqiblten, N_MOIST, N_SCALAR,
                rk_step, True, *I18)
        # rk_tendency (solver.py:971-1216)
        # with the *_fresh calls the five tendencies are zeroed by their first
        # writer (advect_u/v/w, advect_scalar, rhs_ph below), which nothing before reads
        fresh = hasattr(o, "advect_u_fresh")
        if fresh:
            zeros(s.u_save, s.v_save, s.w_save, s.ph_save, s.t_save, s.mu_tend, s.mu_save)
        else:
            zeros(s.ru_tend, s.rv_tend, s.rw_tend, s.t_tend, s.ph_tend, s.u_save, s.v_save, s.w_save,
                  s.ph_save, s.t_save, s.mu_tend, s.mu_save)
        mu_tend_e = s.mu_tend.unsqueeze(0)
        mu_e = s.mu.unsqueeze(0)
        adv = (s.c1h, s.c2h, s.mut, TIME_STEP, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty, s.fnm, s.fnp,
               rdx, rdy)
        adv_u, adv_v, adv_w, adv_s, rhs = ((o.advect_u_fresh, o.advect_v_fresh, o.advect_w_fresh,
                                            o.advect_scalar_fresh, o.rhs_ph_fresh) if fresh else
                                           (o.advect_u, o.advect_v, o.advect_w, o.advect_scalar, o.rhs_ph))
        s.ru_tend = adv_u(s.u, s.u, s.ru_tend, s.ru, s.rv, s.ww, *adv, s.rdnw, *I18)
        s.rv_tend = adv_v(s.v, s.v, s.rv_tend, s.ru, s.rv, s.ww, *adv, s.rdnw, *I18)
        s.rw_tend = adv_w(s.w, s.w, s.rw_tend, s.ru, s.rv, s.ww, *adv, s.rdn, *I18)
        s.t_tend = adv_s(s.t, s.t, s.t_tend, s.ru, s.rv, s.ww, *adv, s.rdnw, *I18)
        s.ph_tend = rhs(s.ph_tend, s.u, s.v, s.ww, s.ph, s.ph, s.phb, s.w, s.mut, s.muu, s.muv, s.c1f, s.c2f,
                        s.fnm, s.fnp, s.rdnw, s.cfn, s.cfn1, rdx, rdy, s.msfux, s.msfuy, s.msfvx, s.msfvx_inv,
                        s.msfvy, s.msftx, s.msfty, True, *I18)
        hpg_args = (s.ru_tend, s.rv_tend, s.ph, s.alt, s.p, s.pb, s.al, s.php, s.cqu, s.cqv, s.muu, s.muv, s.mu,
                    s.c1h, s.c2h, s.fnm, s.fnp, s.rdnw, s.cf1, s.cf2, s.cf3, s.cfn, s.cfn1, rdx, rdy, s.msfux,
                    s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty, True, False, *I18)
        buoy_args = (s.rw_tend, s.p, s.cqw, s.mu, s.mub, s.c1f, s.c2f, s.rdnw, s.rdn, G, s.msftx, s.msfty, *I18)
        if hasattr(o, "pgf_buoy_coriolis_curvature"):
            # the three calls in one launch, each point in the same order
            o.pgf_buoy_coriolis_curvature(
                hpg_args, buoy_args,
                (s.ru, s.rv, s.rw, s.u, s.v, s.w, s.ru_tend, s.rv_tend, s.rw_tend, s.msfux, s.msfuy, s.msfvx,
                 s.msfvy, s.msftx, s.msfty, s.f, s.e, s.sina, s.cosa, s.clat, s.fnm, s.fnp, rdx, rdy, *I18))
        else:
            s.ru_tend, s.rv_tend = o.horizontal_pressure_gradient(*hpg_args)
            s.rw_tend, s.cqw = o.pg_buoy_w(*buoy_args)
            if hasattr(o, "coriolis_curvature"):
                o.coriolis_curvature(s.ru, s.rv, s.rw, s.u, s.v, s.w, s.ru_tend, s.rv_tend, s.rw_tend, s.msfux,
                                     s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty, s.f, s.e, s.sina, s.cosa, s.clat,
                                     s.fnm, s.fnp, rdx, rdy, *I18)
            else:
                s.ru_tend, s.rv_tend, s.rw_tend = o.coriolis(s.ru, s.rv, s.rw, s.ru_tend, s.rv_tend, s.rw_tend,
                                                             s.msftx, s.msfty, s.msfux, s.msfuy, s.msfvx, s.msfvy,
                                                             s.f, s.e, s.sina, s.cosa, s.fnm, s.fnp, *I18)
                s.ru_tend, s.rv_tend, s.rw_tend = o.curvature(s.ru, s.rv, s.rw, s.u, s.v, s.w, s.ru_tend,
                                                              s.rv_tend, s.rw_tend, s.msfux, s.msfuy, s.msfvx,
                                                              s.msfvy, s.msftx, s.msfty, s.clat, s.fnm, s.fnp, rdx,
                                                              rdy, *I18)
        if rk_step == 1:
            hd = (s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty)
            diff_args = (
                ('u', s.u, s.ru_tendf, s.mut, s.c1h, s.c2h, *hd, KHDIF, s.xkmh, rdx, rdy, *I18),
                ('v', s.v, s.rv_tendf, s.mut, s.c1h, s.c2h, *hd, KHDIF, s.xkmh, rdx, rdy, *I18),
                ('w', s.w, s.rw_tendf, s.mut, s.c1f, s.c2f, *hd, KHDIF, s.xkmh, rdx, rdy, *I18),
                ('m', s.t, s.t_tendf, s.mut, s.c1h, s.c2h, s.t_init, *hd, 3. * KHDIF, s.xkhh, rdx, rdy, *I18),
                (s.u, s.ru_tendf, s.u_base, s.c1h, s.c2h, s.alt, s.muu, s.rdn, s.rdnw, KVDIF, *I18),
                (s.v, s.rv_tendf, s.v_base, s.c1h, s.c2h, s.alt, s.muv, s.rdn, s.rdnw, KVDIF, *I18),
                ('w', s.w, s.rw_tendf, s.c1f, s.c2f, s.alt, s.mut, s.rdn, s.rdnw, KVDIF, *I18),
                (s.t, s.t_tendf, s.t_init, s.c1h, s.c2h, s.alt, s.mut, s.rdn, s.rdnw, 3. * KVDIF, *I18))
            if hasattr(o, "diffusion_dry"):
                o.diffusion_dry(*diff_args)
            else:
                hu, hv, hw, ht, vu, vv, vw, vt = diff_args
                s.ru_tendf = o.horizontal_diffusion(*hu)
                s.rv_tendf = o.horizontal_diffusion(*hv)
                s.rw_tendf = o.horizontal_diffusion(*hw)
                s.t_tendf = o.horizontal_diffusion_3dmp(*ht)
                s.ru_tendf = o.vertical_diffusion_u(*vu)
                s.rv_tendf = o.vertical_diffusion_v(*vv)
                s.rw_tendf = o.vertical_diffusion(*vw)
                s.t_tendf = o.vertical_diffusion_3dmp(*vt)
            # relax_bdy_dry (solver.py:1146-1203)
            rb = lambda fld, tend, key, var, ix=I24: o.relax_bdytend(
                fld, tend, *b[key], var, SPEC_BDY_WIDTH, SPEC_ZONE, RELAX_ZONE, dtbc, s.fcx, s.gcx, *ix)
            s.u_save = rb(s.ru, s.u_save, 'u', 'u')
            s.v_save = rb(s.rv, s.v_save, 'v', 'v')
            i_start, i_end = max(its - 1, ids), min(ite + 1, ide - 1)
            j_start, j_end = max(jts - 1, jds), min(jte + 1, jde - 1)
            mw = (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme,
                  its - 1, ite + 1, jts - 1, jte + 1, kts, kte, i_start, i_end, j_start, j_end, kts, kte)
            s.rfield = o.mass_weight(s.ph, s.mut, s.rfield, s.c1f, s.c2f, *mw)
            s.ph_save = rb(s.rfield, s.ph_save, 'ph', 'h')
            s.rfield = o.mass_weight(s.t, s.mut, s.rfield, s.c1h, s.c2h, *mw)
            s.t_save = rb(s.rfield, s.t_save, 't', 't')
            ix_mu = (ids, ide, jds, jde, 0, 1, ims, ime, jms, jme, 0, 1, ips, ipe, jps, jpe, 0, 1,
                     its, ite, jts, jte, 0, 1)
            mu_tend_e = rb(mu_e, mu_tend_e, 'mu', 'm', ix_mu)
            s.mu_tend = mu_tend_e[0, :, :] + 0.
        s.ru_tend, s.rv_tend, s.rw_tend, s.ph_tend, s.t_tend, s.mu_tend = o.rk_addtend_dry(
            s.ru_tend, s.rv_tend, s.rw_tend, s.ph_tend, s.t_tend, s.ru_tendf, s.rv_tendf, s.rw_tendf, s.ph_tendf,
            s.t_tendf, s.u_save, s.v_save, s.w_save, s.ph_save, s.t_save, s.mu_tend, s.mu_tendf, rk_step, s.c1h,
            s.c2h, s.h_diabatic, s.mut, s.msftx, s.msfty, s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, *I24)
        # spec_bdy_dry (solver.py:1220-1255)
        ix_mu = (ids, ide, jds, jde, 0, 1, ims, ime, jms, jme, 0, 1, ips, ipe, jps, jpe, 0, 1, its, ite, jts, jte, 0, 1)
        if hasattr(o, "spec_bdytend_group"):
            o.spec_bdytend_group([(s.ru_tend, b['u'][4:], 'u', False), (s.rv_tend, b['v'][4:], 'v', False),
                                  (s.ph_tend, b['ph'][4:], 'h', False), (s.t_tend, b['t'][4:], 't', False),
                                  (mu_tend_e, b['mu'][4:], 'm', True)], I24, ix_mu)
        else:
            sb = lambda tend, key, var, ix=I24: o.spec_bdytend(tend, *b[key], var, SPEC_BDY_WIDTH, SPEC_ZONE, *ix)
            s.ru_tend = sb(s.ru_tend, 'u', 'u')
            s.rv_tend = sb(s.rv_tend, 'v', 'v')
            s.ph_tend = sb(s.ph_tend, 'ph', 'h')
            s.t_tend = sb(s.t_tend, 't', 't')
            mu_tend_e = sb(mu_tend_e, 'mu', 'm', ix_mu)
        s.mu_tend = mu_tend_e[0, :, :] + 0.
        # small_step_prep, calc_p_rho, calc_coef_w (solver.py:1258-1294)
        (s.u_1, s.v_1, s.w_1, s.t_1, s.ph_1, s.u_save, s.v_save, s.w_save, s.t_save, s.ph_save, s.u, s.v, s.w,
         s.t, s.ph, s.c2a, s.ww1, s.mu_1, s.mu, s.muus, s.muvs, s.muts, s.mudf, s.mu_save) = o.small_step_prep(
            s.u_1, s.u, s.v_1, s.v, s.w_1, s.w, s.t_1, s.t, s.ph_1, s.ph, s.mub, s.mu_1, s.mu, s.muu, s.muus,
            s.muv, s.muvs, s.mut, s.muts, s.mudf, s.c1h, s.c2h, s.c1f, s.c2f, s.c3h, s.c4h, s.c3f, s.c4f,
            s.u_save, s.v_save, s.w_save, s.t_save, s.ph_save, s.mu_save, s.ww, s.ww1, s.c2a, s.pb, s.p, s.alt,
            s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty, rdx, rdy, rk_step, *I18)
        cpr = lambda step: o.calc_p_rho(s.al, s.p, s.ph, s.alt, s.t, s.t_save, s.c2a, s.pm1, s.mu, s.muts, s.c1h,
                                        s.c2h, s.c1f, s.c2f, s.c3h, s.c4h, s.c3f, s.c4f, s.znu, T0, s.rdnw, s.dnw,
                                        SMDIV, True, step, *I18)
        s.al, s.p, s.pm1 = cpr(0)
        s.a, s.alpha, s.gamma = o.calc_coef_w(s.a, s.alpha, s.gamma, s.mut, s.c1h, s.c2h, s.c1f, s.c2f, s.c3h,
                                              s.c4h, s.c3f, s.c4f, s.cqw, s.rdn, s.rdnw, s.c2a, dts_rk, G, EPSSM,
                                              False, *I18)
        # set_phys_bc2_tim (solver.py:1296-1345)
        (s.ru_tend, s.rv_tend, s.ph, s.al, s.p, s.t_1, s.t_save, s.mu_1, s.mu, s.mudf) = bcs(
            (s.ru_tend, 'u', 0), (s.rv_tend, 'v', 0), (s.ph, 'w', 0), (s.al, 'p', 0), (s.p, 'p', 0), (s.t_1, 't', 0),
            (s.t_save, 't', 0), (s.mu_1, 't', 1), (s.mu, 't', 1), (s.mudf, 't', 1))
        ix_2d = (ids, ide, jds, jde, 0, 0, ims, ime, jms, jme, 0, 0, ips, ipe, jps, jpe, 0, 0,
                 its, ite, jts, jte, 0, 0)
        # acoustic small steps (solver.py:1349-1493)
        for iteration in range(1, nsmall + 1):
            s.u, s.v = o.advance_uv(s.u, s.ru_tend, s.v, s.rv_tend, s.p, s.pb, s.ph, s.php, s.alt, s.al, s.mu,
                                    s.muu, s.cqu, s.muv, s.cqv, s.mudf, s.c1h, s.c2h, s.c1f, s.c2f, s.c3h, s.c4h,
                                    s.c3f, s.c4f, s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, rdx, rdy, dts_rk,
                                    s.cf1, s.cf2, s.cf3, s.fnm, s.fnp, EMDIV, s.rdnw, SPEC_ZONE, True, False, *I18)
            if grouped_upd:
                o.spec_bdyupdate_group([(s.u, s.ru_tend, 'u', 0), (s.v, s.rv_tend, 'v', 0)], dts_rk, I24, ix_2d)
            else:
                s.u = o.spec_bdyupdate(s.u, s.ru_tend, dts_rk, 'u', SPEC_ZONE, *I24)
                s.v = o.spec_bdyupdate(s.v, s.rv_tend, dts_rk, 'v', SPEC_ZONE, *I24)
            (s.ww, s.ww1, s.t, s.t_2save, s.ru_m, s.rv_m, s.ww_m, s.muave, s.muts, s.mudf, s.mu) = o.advance_mu_t(
                s.ww, s.ww1, s.u, s.u_save, s.v, s.v_save, s.mu, s.mut, s.muave, s.muts, s.muu, s.muv, s.mudf,
                s.c1h, s.c2h, s.c1f, s.c2f, s.c3h, s.c4h, s.c3f, s.c4f, s.ru_m, s.rv_m, s.ww_m, s.t, s.t_save,
                s.t_2save, s.t_tend, s.mu_tend, rdx, rdy, dts_rk, EPSSM, s.dnw, s.fnm, s.fnp, s.rdnw, s.msfux,
                s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty, iteration, *I18)
            if grouped_upd:
                o.spec_bdyupdate_group([(s.t, s.t_tend, 't', 0), (s.mu, s.mu_tend, 'm', 1),
                                        (s.muts, s.mu_tend, 'm', 1)], dts_rk, I24, ix_2d)
            else:
                s.t = o.spec_bdyupdate(s.t, s.t_tend, dts_rk, 't', SPEC_ZONE, *I24)
                s.mu = o.spec_bdyupdate(s.mu, s.mu_tend, dts_rk, 'm', SPEC_ZONE, *ix_2d)
                s.muts = o.spec_bdyupdate(s.muts, s.mu_tend, dts_rk, 'm', SPEC_ZONE, *ix_2d)
            s.t_2save, s.w, s.ph = o.advance_w(
                s.w, s.rw_tend, s.ww, s.w_save, s.u, s.v, s.mu, s.mut, s.muave, s.muts, s.c1h, s.c2h, s.c1f, s.c2f,
                s.c3h, s.c4h, s.c3f, s.c4f, s.t_2save, s.t, s.t_save, s.ph, s.ph_save, s.phb, s.ph_tend, s.ht, s.c2a,
                s.cqw, s.alt, s.alb, s.a, s.alpha, s.gamma, rdx, rdy, dts_rk, T0, EPSSM, s.dnw, s.fnm, s.fnp, s.rdnw,
                s.rdn, s.cf1, s.cf2, s.cf3, s.msftx, s.msfty, False, *I18)
            s.ru_m, s.rv_m, s.ww_m = o.sumflux(s.u, s.v, s.ww, s.u_save, s.v_save, s.ww1, s.muu, s.muv, s.c1h, s.c2h,
                                               s.c1f, s.c2f, s.c3h, s.c4h, s.c3f, s.c4f, s.ru_m, s.rv_m, s.ww_m, EPSSM,
                                               s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, iteration, nsmall,
                                               *I18)
            ph_args = (s.ph_save, s.ph, s.ph_tend, s.mu_tend, s.muts, s.c1f, s.c2f, dts_rk, 'h', SPEC_ZONE, *I24)
            if hasattr(o, "spec_bdyupdate_ph_zero_grad"):
                s.ph, s.w = o.spec_bdyupdate_ph_zero_grad(ph_args, (s.w, 'w', SPEC_ZONE, *I24))
            else:
                s.ph = o.spec_bdyupdate_ph(*ph_args)
                s.w = o.zero_grad_bdy(s.w, 'w', SPEC_ZONE, *I24)
            s.al, s.p, s.pm1 = cpr(iteration)
            s.ph, s.al, s.p, s.muts, s.mu, s.mudf = bcs((s.ph, 'w', 0), (s.al, 'p', 0), (s.p, 'p', 0),
                                                        (s.muts, 't', 1), (s.mu, 't', 1), (s.mudf, 't', 1))
        # back to full variables (solver.py:1496-1543)
        s.muus, s.muvs = o.calc_mu_uv_1(s.muts, s.muus, s.muvs, *I18)
        s.u, s.v, s.w, s.ph, s.ww, s.t, s.mu = o.small_step_finish(
            s.u, s.u_1, s.v, s.v_1, s.w, s.w_1, s.t, s.t_1, s.ph, s.ph_1, s.ww, s.ww1, s.mu, s.mu_1, s.mut, s.muts,
            s.muu, s.muus, s.muv, s.muvs, s.c1h, s.c2h, s.c1f, s.c2f, s.c3h, s.c4h, s.c3f, s.c4f, s.u_save, s.v_save,
            s.w_save, s.t_save, s.ph_save, s.mu_save, s.msfux, s.msfuy, s.msfvx, s.msfvy, s.msftx, s.msfty,
            s.h_diabatic, nsmall, dts_rk, rk_step, cfg.rk_order, *I18)
        if rk_step == cfg.rk_order:
            s.ru_m, s.rv_m, s.ww_m, s.mut, s.muts = bcs((s.ru_m, 'u', 0), (s.rv_m, 'v', 0), (s.ww_m, 'w', 0),
                                                        (s.mut, 't', 1), (s.muts, 't', 1))
        # moist and scalar transport (solver.py:1548-1636)
        sct = (rdx, rdy, s.rdn, s.rdnw, KHDIF, KVDIF, s.xkhh, DIFF_6TH_OPT, DIFF_6TH_FACTOR)
        mf = (s.msfux, s.msfuy, s.msfvx, s.msfvx_inv, s.msfvy, s.msftx, s.msfty)
        paired = hasattr(o, "rk_scalar_tend_pair") and hasattr(o, "rk_update_scalar_pair")
        if paired:
            # the moist and the tracer calls in the same launches (they share the
            # fluxes, the index set and h/z_tendency, which both only zero); the
            # water-vapour boundary terms come between tendency and update as in
            # the reference, and the tracer calls touch neither
            (s.moist_tend, s.advect_tend, s.h_tendency, s.z_tendency), (s.scalar_tend, s.advect_tend_sca, _h, _z) = (
                o.rk_scalar_tend_pair(
                    (1, 6, False, rk_step, dt_rk, s.ru_m, s.rv_m, s.ww_m, s.muts, s.mub, s.mu_1, s.c1h, s.c2h, s.alt,
                     s.moist_old, s.moist, s.moist_tend, s.advect_tend, s.h_tendency, s.z_tendency, s.RQVFTEN,
                     s.qv_base, True, s.fnm, s.fnp, *mf, *sct, MOIST_ADV_OPT, *I18),
                    (1, 2, False, rk_step, dt_rk, s.ru_m, s.rv_m, s.ww_m, s.muts, s.mub, s.mu_1, s.c1h, s.c2h, s.alt,
                     s.scalar_old, s.scalar, s.scalar_tend, s.advect_tend_sca, s.h_tendency, s.z_tendency,
                     s.RQVFTEN, s.qv_base, False, s.fnm, s.fnp, *mf, *sct, SCALAR_ADV_OPT, *I18)))
            if rk_step == 1:
                s.moist_tend[P_QV, :, :, :] = o.relax_bdy_scalar(
                    s.moist_tend[P_QV, :, :, :], s.moist[P_QV, :, :, :], s.mut, s.c1h, s.c2h, *b['qv'],
                    SPEC_BDY_WIDTH, SPEC_ZONE, RELAX_ZONE, dtbc, s.fcx, s.gcx, *I24)
                s.moist_tend[P_QV, :, :, :] = o.spec_bdy_scalar(s.moist_tend[P_QV, :, :, :], *b['qv'],
                                                               SPEC_BDY_WIDTH, SPEC_ZONE, *I24)
            (s.moist_old, s.moist), (s.scalar_old, s.scalar) = o.rk_update_scalar_pair(
                (1, 6, s.moist_old, s.moist, s.moist_tend, s.advect_tend, s.h_tendency, s.z_tendency, s.msftx,
                 s.msfty, s.c1h, s.c2h, s.mu_1, s.mu, s.mub, rk_step, dt_rk, SPEC_ZONE, False, *I18),
                (1, 2, s.scalar_old, s.scalar, s.scalar_tend, s.advect_tend_sca, s.h_tendency, s.z_tendency,
                 s.msftx, s.msfty, s.c1h, s.c2h, s.mu_1, s.mu, s.mub, rk_step, dt_rk, SPEC_ZONE, False, *I18))
            o.flow_dep_bdy_group([s.moist[im] for im in range(2, 7)], s.ru_m, s.rv_m, I24)
        else:
            s.moist_tend, s.advect_tend, s.h_tendency, s.z_tendency = o.rk_scalar_tend(
                1, 6, False, rk_step, dt_rk, s.ru_m, s.rv_m, s.ww_m, s.muts, s.mub, s.mu_1, s.c1h, s.c2h, s.alt,
                s.moist_old, s.moist, s.moist_tend, s.advect_tend, s.h_tendency, s.z_tendency, s.RQVFTEN, s.qv_base,
                True, s.fnm, s.fnp, *mf, *sct, MOIST_ADV_OPT, *I18)
            if rk_step == 1:
                s.moist_tend[P_QV, :, :, :] = o.relax_bdy_scalar(
                    s.moist_tend[P_QV, :, :, :], s.moist[P_QV, :, :, :], s.mut, s.c1h, s.c2h, *b['qv'],
                    SPEC_BDY_WIDTH, SPEC_ZONE, RELAX_ZONE, dtbc, s.fcx, s.gcx, *I24)
                s.moist_tend[P_QV, :, :, :] = o.spec_bdy_scalar(s.moist_tend[P_QV, :, :, :], *b['qv'], SPEC_BDY_WIDTH,
                                                               SPEC_ZONE, *I24)
            s.moist_old, s.moist = o.rk_update_scalar(1, 6, s.moist_old, s.moist, s.moist_tend, s.advect_tend,
                                                      s.h_tendency, s.z_tendency, s.msftx, s.msfty, s.c1h, s.c2h,
                                                      s.mu_1, s.mu, s.mub, rk_step, dt_rk, SPEC_ZONE, False, *I18)
            if hasattr(o, "flow_dep_bdy_group"):
                o.flow_dep_bdy_group([s.moist[im] for im in range(2, 7)], s.ru_m, s.rv_m, I24)
            else:
                for im in range(2, 7):
                    s.moist[im, :, :, :] = o.flow_dep_bdy(s.moist[im, :, :, :], s.ru_m, s.rv_m, 1, *I24)
            s.scalar_tend, s.advect_tend_sca, s.h_tendency, s.z_tendency = o.rk_scalar_tend(
                1, 2, False, rk_step, dt_rk, s.ru_m, s.rv_m, s.ww_m, s.muts, s.mub, s.mu_1, s.c1h, s.c2h, s.alt,
                s.scalar_old, s.scalar, s.scalar_tend, s.advect_tend_sca, s.h_tendency, s.z_tendency, s.RQVFTEN,
                s.qv_base, False, s.fnm, s.fnp, *mf, *sct, SCALAR_ADV_OPT, *I18)
            s.scalar_old, s.scalar = o.rk_update_scalar(1, 2, s.scalar_old, s.scalar, s.scalar_tend, s.advect_tend_sca,
                                                        s.h_tendency, s.z_tendency, s.msftx, s.msfty, s.c1h, s.c2h,
                                                        s.mu_1, s.mu, s.mub, rk_step, dt_rk, SPEC_ZONE, False, *I18)
        # EOS (solver.py:1639-1654)
        s.al, s.p, s.ph = o.calc_p_rho_phi(s.moist, N_MOIST, HYPSOMETRIC_OPT, s.al, s.alb, s.mu, s.muts, s.c1h,
                                           s.c2h, s.c3h, s.c4h, s.c3f, s.c4f, s.ph, s.phb, s.p, s.pb, s.t, P0, T0,
                                           cfg.p_top, s.znu, s.znw, s.dnw, s.rdnw, s.rdn, True, *I18)
        if check_nan:
            _nan_check(s.p, "rk_step %d" % rk_step)
        if nan_flag is not None:
            nan_flag.scan(s.p, rk_step, I18)
        if rk_step < cfg.rk_order:
            (s.u, s.v, s.w, s.t, s.ph, s.mu) = bcs(
                (s.u, 'U', 0), (s.v, 'V', 0), (s.w, 'w', 0), (s.t, 'p', 0), (s.ph, 'w', 0), (s.mu, 't', 1),
                *[(s.moist[im, :, :, :], 'p', 0) for im in range(1, 7)],
                *[(s.scalar[im, :, :, :], 'p', 0) for im in range(1, 3)])[:6]
    # end of the RK loop: phy_prep_part2 (solver.py:1702-1709)
    (s.rublten, s.rvblten, s.rthblten, s.rqvblten, s.rqcblten, s.rqiblten) = o.phy_prep_part2(
        s.mut, s.muu, s.muv, s.c1h, s.c2h, s.c1f, s.c2f, s.rthblten, s.rublten, s.rvblten, s.rqvblten,
        s.rqcblten, s.rqiblten, *I18)
    # moist physics (solver.py:1714-1766)
    P18 = gr.phys18()
    (s.rho, s.th_phy, s.pi_phy, s.p_phy, s.z, s.z_at_w, s.dz8w, s.p8w, s.h_diabatic, s.qv_diabatic,
     s.qc_diabatic, s.t, s.t_1) = o.moist_physics_prep_em(
        s.t, s.t_1, T0, s.rho, s.al, s.alb, s.p, s.p8w, P0, s.pb, s.ph, s.phb, s.th_phy, s.pi_phy, s.p_phy, s.z,
        s.z_at_w, s.dz8w, dtm, s.h_diabatic, s.moist[P_QV, :, :, :], s.qv_diabatic, s.moist[P_QC, :, :, :],
        s.qc_diabatic, s.fnm, s.fnp, *P18)
    m = s.moist
    out = o.microphysics_driver(s.th_phy, s.rho, s.pi_phy, s.p_phy, s.ht, s.dz8w, s.p8w, dtm, cfg.dx, cfg.dy,
                                m[1], m[2], m[3], m[4], m[5], m[6], s.rainnc, s.rainncv, s.snownc, s.snowncv,
                                s.sr, s.refl_10cm, 1, 1, s.graupelnc, s.graupelncv, 0, 0, 0, 0, 0, 0, *P18)
    (s.th_phy, m[1], m[2], m[3], m[4], m[5], m[6], s.rainnc, s.rainncv, s.snownc, s.snowncv, s.graupelnc,
     s.graupelncv, s.sr, s.refl_10cm) = out
    (s.t, s.t_1, s.th_phy, s.h_diabatic, s.qv_diabatic, s.qc_diabatic) = o.moist_physics_finish_em(
        s.t, s.t_1, T0, s.muts, s.th_phy, s.h_diabatic, dtm, m[P_QV], s.qv_diabatic, m[P_QC], s.qc_diabatic, *P18)
    s.al, s.p, s.ph = o.calc_p_rho_phi(s.moist, N_MOIST, HYPSOMETRIC_OPT, s.al, s.alb, s.mu, s.muts, s.c1h, s.c2h,
                                       s.c3h, s.c4h, s.c3f, s.c4f, s.ph, s.phb, s.p, s.pb, s.t, P0, T0, cfg.p_top,
                                       s.znu, s.znw, s.dnw, s.rdnw, s.rdn, True, *I18)
    if check_nan:
        _nan_check(s.p, "physics")
    if nan_flag is not None:
        nan_flag.scan(s.p, 4, I18)
    # set_phys_bc_dry2 (solver.py:1776-1858), then spec_bdy_final (solver.py:1863-1904)
    end_bcs = ((s.u_1, 'U', 0), (s.u, 'U', 0), (s.v_1, 'V', 0), (s.v, 'V', 0), (s.w_1, 'w', 0), (s.w, 'w', 0),
               (s.t_1, 'p', 0), (s.t, 'p', 0), (s.ph_1, 'w', 0), (s.ph, 'w', 0), (s.mu_1, 't', 1), (s.mu, 't', 1),
               (s.tke_1, 'p', 0), (s.tke, 'p', 0),
               *[(s.moist[im, :, :, :], 'p', 0) for im in range(1, 7)],
               *[(s.scalar[im, :, :, :], 'p', 0) for im in range(1, 3)])
    finals = [(s.u, s.muus, s.msfuy, b['u'], 'u'), (s.v, s.muvs, s.msfvx, b['v'], 'v'),
              (s.t, s.muts, s.msfty, b['t'], 't'), (s.ph, s.muts, s.msfty, b['ph'], 'h'),
              (s.moist[P_QV, :, :, :], s.muts, s.msfty, b['qv'], 't')]
    if hasattr(o, "boundary_epilogue"):
        # both in one launch: each (field, level) fills its halo, then its boundary rows
        o.boundary_epilogue(list(end_bcs), finals, dtbc, I24, I2)
    else:
        (s.u_1, s.u, s.v_1, s.v, s.w_1, s.w, s.t_1, s.t, s.ph_1, s.ph, s.mu_1, s.mu, s.tke_1, s.tke) = \
            bcs(*end_bcs)[:14]
        if hasattr(o, "spec_bdy_final_group"):
            o.spec_bdy_final_group(finals, dtbc, I24)
        else:
            sf = lambda fld, mu, msf, key, var: o.spec_bdy_final(fld, mu, s.c1h, s.c2h, msf, *b[key], var,
                                                                  SPEC_BDY_WIDTH, SPEC_ZONE, dtbc, *I24)
            s.u = sf(s.u, s.muus, s.msfuy, 'u', 'u')
            s.v = sf(s.v, s.muvs, s.msfvx, 'v', 'v')
            s.t = sf(s.t, s.muts, s.msfty, 't', 't')
            s.ph = sf(s.ph, s.muts, s.msfty, 'ph', 'h')
            s.moist[P_QV, :, :, :] = sf(s.moist[P_QV, :, :, :], s.muts, s.msfty, 'qv', 't')
    # surface w (solver.py:1906-1910)
    s.w = o.set_w_surface(s.znw, False, s.w, s.ht, s.u, s.v, s.cf1, s.cf2, s.cf3, rdx, rdy, s.msftx, s.msfty,
                          *I18)
    return s
