"""Drop-in for pywrf/wrf_dynamics.py on the ARW step's path: the same function
names, positional parameters (including the ones the reference ignores) and
return tuples, each forwarding to one C-ABI entry point of libpywrf_amd.so.

Semantics of the reference are kept: arguments the reference updates in place
are updated in place and returned as the same tensor objects; where it
returns a fresh tensor (clone, torch.where, new arithmetic result) a fresh
tensor is allocated here and filled by the kernel.  All work is enqueued on
torch's current stream; nothing here synchronises the device.

Tensors are contiguous FP32 CUDA tensors, 3-D (nz, ny, nx), 2-D (ny, nx),
level arrays (nz,); 0-d tensors for cf1..cf3, cfn, cfn1.
"""
import inspect
from types import SimpleNamespace

import torch

from ._lib import call, grid, ptr, flag, fresh, stream


def _cf(cf1, cf2, cf3):
    """Pointer to (cf1, cf2, cf3) as three consecutive floats."""
    if (cf1.dim() == 0 and cf2.dim() == 0 and cf3.dim() == 0 and cf1.is_cuda
            and cf2.data_ptr() == cf1.data_ptr() + 4 and cf3.data_ptr() == cf1.data_ptr() + 8):
        return ptr(cf1), None
    t = torch.stack([cf1, cf2, cf3]).contiguous()
    return ptr(t), t


# ----------------------------------------------------------------- RK-stage prep
def calculate_full(mut, mub, mu,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:30-38"""
    g = grid(mut.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calculate_full", g, ptr(mut), ptr(mub), ptr(mu), stream())
    return mut


def calc_mu_uv(mu, mub, muu, muv,
               ids, ide, jds, jde, kds, kde,
               ims, ime, jms, jme, kms, kme,
               its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:41-61"""
    g = grid(mu.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_mu_uv", g, ptr(mu), ptr(mub), ptr(muu), ptr(muv), stream())
    return muu, muv


def couple_momentum(muu, ru, u, msfu,
                    muv, rv, v, msfv, msfvx_inv,
                    mut, rw, w, msft,
                    c1h, c2h, c1f, c2f,
                    ids, ide, jds, jde, kds, kde,
                    ims, ime, jms, jme, kms, kme,
                    its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:64-86"""
    g = grid(u.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_couple_momentum", g, ptr(muu), ptr(ru), ptr(u), ptr(msfu), ptr(muv), ptr(rv), ptr(v), ptr(msfv),
         ptr(msfvx_inv), ptr(mut), ptr(rw), ptr(w), ptr(msft), stream())
    return ru, rv, rw


def calc_ww_cp(u, v, mu, mub, muu, muv, c1h, c2h, ww,
               rdx, rdy, msftx, msfty,
               msfux, msfuy, msfvx, msfvx_inv,
               msfvy, dnw,
               ids, ide, jds, jde, kds, kde,
               ims, ime, jms, jme, kms, kme,
               its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:89-118"""
    g = grid(u.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_ww_cp", g, ptr(u), ptr(v), ptr(muu), ptr(muv), ptr(c1h), ptr(ww), float(rdx), float(rdy),
         ptr(msftx), ptr(msfuy), ptr(msfvx_inv), ptr(dnw), stream())
    return ww


def calc_cq(moist, cqu, cqv, cqw, n_moist,
            ids, ide, jds, jde, kds, kde,
            ims, ime, jms, jme, kms, kme,
            its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:121-136"""
    g = grid(cqu.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_cq", g, ptr(moist), int(moist.shape[0]), ptr(cqu), ptr(cqv), ptr(cqw), stream())
    return cqu, cqv, cqw


def calc_alt(alt, al, alb,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:139-147"""
    g = grid(alt.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_alt", g, ptr(alt), ptr(al), ptr(alb), stream())
    return alt


def calc_php(php, ph, phb,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:150-159"""
    g = grid(php.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_php", g, ptr(php), ptr(ph), ptr(phb), stream())
    return php


def set_physical_bc2d(dat, variable_in,
                      ids, ide, jds, jde,
                      ims, ime, jms, jme,
                      ips, ipe, jps, jpe,
                      its, ite, jts, jte):
    """wrf_dynamics.py:164-212"""
    g = grid(dat.shape, (ids, ide, jds, jde, ims, ime, jms, jme, ips, ipe, jps, jpe, its, ite, jts, jte))
    call("pywrf_set_physical_bc2d", g, ptr(dat), flag(variable_in), stream())
    return dat


def set_physical_bc3d(dat, variable_in,
                      ids, ide, jds, jde, kds, kde,
                      ims, ime, jms, jme, kms, kme,
                      ips, ipe, jps, jpe, kps, kpe,
                      its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:216-278"""
    g = grid(dat.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps, kpe,
                         its, ite, jts, jte, kts, kte))
    call("pywrf_set_physical_bc3d", g, ptr(dat), flag(variable_in), stream())
    return dat


# ----------------------------------------------------------------- first-stage physics prep
def phy_prep(mut, muu, muv,
             c1h, c2h, c1f, c2f,
             u, v, p, pb, alt, ph,
             phb, t, moist, n_moist,
             rho, th_phy, p_phy, pi_phy,
             u_phy, v_phy, p8w, t_phy, t8w,
             z, z_at_w, dz8w,
             p_hyd, p_hyd_w, dnw,
             fzm, fzp, znw, p_top,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:314-411 (pi_phy, t_phy returned as fresh arrays)"""
    g = grid(p.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    pi_new, t_new = fresh(p_phy), fresh(p_phy)
    call("pywrf_phy_prep", g, ptr(mut), ptr(u), ptr(v), ptr(p), ptr(pb), ptr(alt), ptr(ph), ptr(phb), ptr(t),
         ptr(moist), int(moist.shape[0]), ptr(rho), ptr(th_phy), ptr(p_phy), ptr(u_phy), ptr(v_phy), ptr(p8w),
         ptr(t8w), ptr(z), ptr(z_at_w), ptr(dz8w), ptr(p_hyd), ptr(p_hyd_w), ptr(dnw), ptr(fzm), ptr(fzp),
         float(p_top), ptr(pi_new), ptr(t_new), stream())
    return th_phy, p_phy, pi_new, t_new, rho, u_phy, v_phy, z_at_w, dz8w, z, p8w, t8w, p_hyd_w, p_hyd


def calculate_phy_tend(c1, c2,
                       mut, muu, muv, pi3d,
                       RUBLTEN, RVBLTEN, RTHBLTEN,
                       RQVBLTEN, RQCBLTEN, RQIBLTEN,
                       scalar, scalar_tend, num_scalar,
                       ids, ide, jds, jde, kds, kde,
                       ims, ime, jms, jme, kms, kme,
                       its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:449-488"""
    g = grid(RUBLTEN.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts,
                             kte))
    call("pywrf_calculate_phy_tend", g, ptr(mut), ptr(RUBLTEN), ptr(RVBLTEN), ptr(RTHBLTEN), ptr(RQVBLTEN),
         ptr(RQCBLTEN), ptr(RQIBLTEN), ptr(scalar_tend), stream())
    return RUBLTEN, RVBLTEN, RTHBLTEN, RQVBLTEN, RQCBLTEN, RQIBLTEN, scalar_tend


def compute_diff_metrics(ph, phb, z, rdz, rdzw,
                         zx, zy, rdx, rdy,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:491-536"""
    g = grid(ph.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_compute_diff_metrics", g, ptr(ph), ptr(phb), ptr(z), ptr(rdz), ptr(rdzw), ptr(zx), ptr(zy),
         float(rdx), float(rdy), stream())
    return rdzw, rdz, zx, zy, z


def cal_deform_and_div(u, v, w, div,
                       defor11, defor22, defor33,
                       defor12, defor13, defor23,
                       u_base, v_base, msfux, msfuy,
                       msfvx, msfvy, msftx, msfty,
                       rdx, rdy, dn, dnw, rdz, rdzw,
                       fnm, fnp, cf1, cf2, cf3, zx, zy,
                       ids, ide, jds, jde, kds, kde,
                       ims, ime, jms, jme, kms, kme,
                       its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:539-813"""
    g = grid(u.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    cf, _keep = _cf(cf1, cf2, cf3)
    call("pywrf_cal_deform_and_div", g, ptr(u), ptr(v), ptr(w), ptr(div), ptr(defor11), ptr(defor22),
         ptr(defor33), ptr(defor12), ptr(defor13), ptr(defor23), ptr(msfux), ptr(msfuy), ptr(msfvx), ptr(msfvy),
         ptr(msftx), ptr(msfty), float(rdx), float(rdy), ptr(dn), ptr(dnw), ptr(rdz), ptr(rdzw), ptr(fnm),
         ptr(fnp), cf, ptr(zx), ptr(zy), stream())
    return defor11, defor12, defor13, defor22, defor23, defor33, div


def calculate_km_kh(dt,
                    dampcoef, zdamp, damp_opt,
                    xkmh, xkmv, xkhh, xkhv,
                    BN2, khdif, kvdif, div,
                    defor11, defor22, defor33,
                    defor12, defor13, defor23,
                    tke, p8w, t8w, theta, t, p, moist,
                    dn, dnw, dx, dy, rdz, rdzw, isotropic,
                    n_moist, cf1, cf2, cf3, warm_rain,
                    mix_upper_bound,
                    msftx, msfty,
                    zx, zy,
                    ids, ide, jds, jde, kds, kde,
                    ims, ime, jms, jme, kms, kme,
                    its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:816-995 (BN2 and xkmh returned as fresh arrays)"""
    g = grid(t.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    cf, _keep = _cf(cf1, cf2, cf3)
    bn2_new, xkmh_new = fresh(t), fresh(t)
    call("pywrf_calculate_km_kh", g, ptr(xkmh), ptr(xkmv), ptr(xkhh), ptr(xkhv), ptr(defor11), ptr(defor22),
         ptr(defor12), ptr(p8w), ptr(t8w), ptr(theta), ptr(t), ptr(p), ptr(moist), float(dx), float(dy), ptr(rdz),
         ptr(rdzw), cf, ptr(msftx), ptr(msfty), ptr(bn2_new), ptr(xkmh_new), stream())
    return bn2_new, xkmh_new, xkmv, xkhh, xkhv


def update_phy_ten(rph_tendf, rt_tendf, ru_tendf, rv_tendf, moist_tendf,
                   scalar_tendf, mu_tendf,
                   RTHBLTEN, RUBLTEN, RVBLTEN,
                   RQVBLTEN, RQCBLTEN, RQIBLTEN,
                   n_moist, n_scalar, rk_step, adv_moist_cond,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:998-1076"""
    g = grid(rt_tendf.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts,
                              kte))
    call("pywrf_update_phy_ten", g, ptr(rt_tendf), ptr(ru_tendf), ptr(rv_tendf), ptr(moist_tendf), ptr(RTHBLTEN),
         ptr(RUBLTEN), ptr(RVBLTEN), ptr(RQVBLTEN), ptr(RQCBLTEN), ptr(RQIBLTEN), stream())
    return rt_tendf, ru_tendf, rv_tendf, moist_tendf


# ----------------------------------------------------------------- advection
def advect_u(u, u_old, tendency,
             ru, rv, rom,
             c1, c2,
             mut, time_step,
             msfux, msfuy, msfvx, msfvy,
             msftx, msfty,
             fzm, fzp,
             rdx, rdy, rdzw,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:1112-1307"""
    return _advect_u("pywrf_advect_u", **locals())


def _advect_u(entry, u, u_old, tendency, ru, rv, rom, c1, c2, mut, time_step, msfux, msfuy, msfvx, msfvy, msftx,
              msfty, fzm, fzp, rdx, rdy, rdzw, ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite,
              jts, jte, kts, kte):
    g = grid(u.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call(entry, g, ptr(u), ptr(tendency), ptr(ru), ptr(rv), ptr(rom), ptr(msfux), ptr(fzm), ptr(fzp),
         float(rdx), float(rdy), ptr(rdzw), stream())
    return tendency


def advect_u_fresh(*args):
    """advect_u (same positional arguments) into a tendency the caller would have
    zeroed first (the zeroing at solver.py:971-975, then advect_u): the
    kernel writes 0 + the tendency inside its index box and 0 everywhere else,
    so the array needs no separate zeroing and is not read.  Same result as
    zeroing and calling advect_u.  Not part of the reference API."""
    return _advect_u("pywrf_advect_u_fresh", **vars(_bind(advect_u, args)))


def advect_v(v, v_old, tendency,
             ru, rv, rom,
             c1, c2,
             mut, time_step,
             msfux, msfuy, msfvx, msfvy,
             msftx, msfty,
             fzm, fzp,
             rdx, rdy, rdzw,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:1332-1532"""
    return _advect_v("pywrf_advect_v", **locals())


def _advect_v(entry, v, v_old, tendency, ru, rv, rom, c1, c2, mut, time_step, msfux, msfuy, msfvx, msfvy, msftx,
              msfty, fzm, fzp, rdx, rdy, rdzw, ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite,
              jts, jte, kts, kte):
    g = grid(v.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call(entry, g, ptr(v), ptr(tendency), ptr(ru), ptr(rv), ptr(rom), ptr(msfvx), ptr(msfvy), ptr(fzm),
         ptr(fzp), float(rdx), float(rdy), ptr(rdzw), stream())
    return tendency


def advect_v_fresh(*args):
    """advect_v (same positional arguments) into a tendency the caller would have
    zeroed first (the zeroing at solver.py:971-975, then advect_v): the
    kernel writes 0 + the tendency inside its index box and 0 everywhere else,
    so the array needs no separate zeroing and is not read.  Same result as
    zeroing and calling advect_v.  Not part of the reference API."""
    return _advect_v("pywrf_advect_v_fresh", **vars(_bind(advect_v, args)))


def advect_w(w, w_old, tendency,
             ru, rv, rom,
             c1, c2,
             mut, time_step,
             msfux, msfuy, msfvx, msfvy,
             msftx, msfty,
             fzm, fzp,
             rdx, rdy, rdzu,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:1535-1822"""
    return _advect_w("pywrf_advect_w", **locals())


def _advect_w(entry, w, w_old, tendency, ru, rv, rom, c1, c2, mut, time_step, msfux, msfuy, msfvx, msfvy, msftx,
              msfty, fzm, fzp, rdx, rdy, rdzu, ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite,
              jts, jte, kts, kte):
    g = grid(w.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call(entry, g, ptr(w), ptr(tendency), ptr(ru), ptr(rv), ptr(rom), ptr(msftx), ptr(fzm), ptr(fzp),
         float(rdx), float(rdy), ptr(rdzu), stream())
    return tendency


def advect_w_fresh(*args):
    """advect_w (same positional arguments) into a tendency the caller would have
    zeroed first (the zeroing at solver.py:971-975, then advect_w): the
    kernel writes 0 + the tendency inside its index box and 0 everywhere else,
    so the array needs no separate zeroing and is not read.  Same result as
    zeroing and calling advect_w.  Not part of the reference API."""
    return _advect_w("pywrf_advect_w_fresh", **vars(_bind(advect_w, args)))


def advect_scalar(field, field_old, tendency,
                  ru, rv, rom,
                  c1, c2,
                  mut, time_step,
                  msfux, msfuy, msfvx, msfvy,
                  msftx, msfty,
                  fzm, fzp,
                  rdx, rdy, rdzw,
                  ids, ide, jds, jde, kds, kde,
                  ims, ime, jms, jme, kms, kme,
                  its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:1841-2028"""
    return _advect_scalar("pywrf_advect_scalar", **locals())


def _advect_scalar(entry, field, field_old, tendency, ru, rv, rom, c1, c2, mut, time_step, msfux, msfuy, msfvx, msfvy,
                   msftx, msfty, fzm, fzp, rdx, rdy, rdzw, ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call(entry, g, ptr(field), ptr(tendency), ptr(ru), ptr(rv), ptr(rom), ptr(msftx), ptr(fzm),
         ptr(fzp), float(rdx), float(rdy), ptr(rdzw), stream())
    return tendency


def advect_scalar_fresh(*args):
    """advect_scalar (same positional arguments) into a tendency the caller would have
    zeroed first (the zeroing at solver.py:971-975, then advect_scalar): the
    kernel writes 0 + the tendency inside its index box and 0 everywhere else,
    so the array needs no separate zeroing and is not read.  Same result as
    zeroing and calling advect_scalar.  Not part of the reference API."""
    return _advect_scalar("pywrf_advect_scalar_fresh", **vars(_bind(advect_scalar, args)))


# ----------------------------------------------------------------- other large-step tendencies
def rhs_ph(ph_tend, u, v, ww,
           ph, ph_old, phb, w,
           mut, muuf, muvf,
           c1f, c2f,
           fnm, fnp,
           rdnw, cfn, cfn1, rdx, rdy,
           msfux, msfuy, msfvx,
           msfvx_inv, msfvy,
           msftx, msfty,
           non_hydrostatic,
           ids, ide, jds, jde, kds, kde,
           ims, ime, jms, jme, kms, kme,
           its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2031-2297"""
    return _rhs_ph("pywrf_rhs_ph", **locals())


def _rhs_ph(entry, ph_tend, u, v, ww, ph, ph_old, phb, w, mut, muuf, muvf, c1f, c2f, fnm, fnp, rdnw, cfn, cfn1, rdx,
            rdy, msfux, msfuy, msfvx, msfvx_inv, msfvy, msftx, msfty, non_hydrostatic, ids, ide, jds, jde, kds, kde,
            ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte):
    g = grid(ph.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call(entry, g, ptr(ph_tend), ptr(u), ptr(v), ptr(ww), ptr(ph), ptr(phb), ptr(w), ptr(mut), ptr(muuf),
         ptr(muvf), ptr(fnm), ptr(fnp), ptr(rdnw), ptr(cfn), ptr(cfn1), float(rdx), float(rdy), ptr(msfux),
         ptr(msfvy), ptr(msfty), stream())
    return ph_tend


def rhs_ph_fresh(*args):
    """rhs_ph (same positional arguments) into a tendency the caller would have
    zeroed first (the zeroing at solver.py:971-975, then rhs_ph): the
    kernel writes 0 + the tendency inside its index box and 0 everywhere else,
    so the array needs no separate zeroing and is not read.  Same result as
    zeroing and calling rhs_ph.  Not part of the reference API."""
    return _rhs_ph("pywrf_rhs_ph_fresh", **vars(_bind(rhs_ph, args)))


def horizontal_pressure_gradient(ru_tend, rv_tend,
                                 ph, alt, p, pb, al, php, cqu, cqv,
                                 muu, muv, mu, c1h, c2h, fnm, fnp, rdnw,
                                 cf1, cf2, cf3, cfn, cfn1,
                                 rdx, rdy, msfux, msfuy,
                                 msfvx, msfvy, msftx, msfty,
                                 non_hydrostatic, top_lid,
                                 ids, ide, jds, jde, kds, kde,
                                 ims, ime, jms, jme, kms, kme,
                                 its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2300-2421"""
    g = grid(p.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    cf, _keep = _cf(cf1, cf2, cf3)
    call("pywrf_horizontal_pressure_gradient", g, ptr(ru_tend), ptr(rv_tend), ptr(ph), ptr(alt), ptr(p), ptr(pb),
         ptr(al), ptr(php), ptr(cqu), ptr(cqv), ptr(muu), ptr(muv), ptr(mu), ptr(fnm), ptr(fnp), ptr(rdnw), cf,
         float(rdx), float(rdy), ptr(msfux), ptr(msfuy), ptr(msfvx), ptr(msfvy), stream())
    return ru_tend, rv_tend


def pg_buoy_w(rw_tend, p, cqw, muf, mubf,
              c1f, c2f,
              rdnw, rdn, g_, msftx, msfty,
              ids, ide, jds, jde, kds, kde,
              ims, ime, jms, jme, kms, kme,
              its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2424-2460"""
    g = grid(p.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_pg_buoy_w", g, ptr(rw_tend), ptr(p), ptr(cqw), ptr(muf), ptr(mubf), ptr(rdnw), ptr(rdn), float(g_),
         ptr(msfty), stream())
    return rw_tend, cqw


def coriolis(ru, rv, rw,
             ru_tend, rv_tend, rw_tend,
             msftx, msfty, msfux, msfuy,
             msfvx, msfvy,
             f, e, sina, cosa, fzm, fzp,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2463-2532"""
    g = grid(ru.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_coriolis", g, ptr(ru), ptr(rv), ptr(rw), ptr(ru_tend), ptr(rv_tend), ptr(rw_tend), ptr(msftx),
         ptr(msfty), ptr(msfux), ptr(msfuy), ptr(msfvx), ptr(msfvy), ptr(f), ptr(e), ptr(sina), ptr(cosa), ptr(fzm),
         ptr(fzp), stream())
    return ru_tend, rv_tend, rw_tend


def curvature(ru, rv, rw, u, v, w, ru_tend, rv_tend, rw_tend,
              msfux, msfuy, msfvx, msfvy, msftx, msfty,
              xlat, fzm, fzp, rdx, rdy,
              ids, ide, jds, jde, kds, kde,
              ims, ime, jms, jme, kms, kme,
              its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2535-2620"""
    g = grid(ru.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_curvature", g, ptr(ru), ptr(rv), ptr(rw), ptr(u), ptr(v), ptr(ru_tend), ptr(rv_tend), ptr(rw_tend),
         ptr(msfuy), ptr(msfvx), ptr(msfvy), ptr(msftx), ptr(msfty), ptr(fzm), ptr(fzp), float(rdx), float(rdy),
         stream())
    return ru_tend, rv_tend, rw_tend


def coriolis_curvature(ru, rv, rw, u, v, w, ru_tend, rv_tend, rw_tend,
                       msfux, msfuy, msfvx, msfvy, msftx, msfty,
                       f, e, sina, cosa, xlat, fzm, fzp, rdx, rdy,
                       ids, ide, jds, jde, kds, kde,
                       ims, ime, jms, jme, kms, kme,
                       its, ite, jts, jte, kts, kte):
    """coriolis (wrf_dynamics.py:2463-2532) then curvature (:2535-2620) in one
    launch; the tendencies are updated in place.  Not part of the reference
    API."""
    g = grid(ru.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_coriolis_curvature", g, ptr(ru), ptr(rv), ptr(rw), ptr(u), ptr(v), ptr(ru_tend), ptr(rv_tend),
         ptr(rw_tend), ptr(msftx), ptr(msfty), ptr(msfux), ptr(msfuy), ptr(msfvx), ptr(msfvy), ptr(f), ptr(e),
         ptr(sina), ptr(cosa), ptr(fzm), ptr(fzp), float(rdx), float(rdy), stream())
    return ru_tend, rv_tend, rw_tend


def pgf_buoy_coriolis_curvature(hpg_args, buoy_args, cc_args):
    """horizontal_pressure_gradient, pg_buoy_w and coriolis_curvature, each
    given as the tuple of its positional arguments, in one launch (the
    sequence of solver.py:1038-1080).  The calls must act on the same
    tendencies (ru_tend, rv_tend of the first and the third, rw_tend of the
    second and the third), pressure, mu, rdnw, map factors, index set and grid
    spacing, with the third's fzm, fzp the first's fnm, fnp.  Returns the
    three calls' results.  Not part of the reference API."""
    h = _bind(horizontal_pressure_gradient, hpg_args)
    b = _bind(pg_buoy_w, buoy_args)
    c = _bind(coriolis_curvature, cc_args)
    same = [(c.ru_tend, h.ru_tend), (c.rv_tend, h.rv_tend), (c.rw_tend, b.rw_tend), (b.p, h.p), (b.muf, h.mu),
            (b.rdnw, h.rdnw), (b.msfty, c.msfty), (c.fzm, h.fnm), (c.fzp, h.fnp), (c.msfux, h.msfux),
            (c.msfuy, h.msfuy), (c.msfvx, h.msfvx), (c.msfvy, h.msfvy)]
    idx = ("ids", "ide", "jds", "jde", "kds", "kde", "ims", "ime", "jms", "jme", "kms", "kme", "its", "ite", "jts",
           "jte", "kts", "kte")
    if (not all(x is y for x, y in same) or h.rdx != c.rdx or h.rdy != c.rdy
            or any(getattr(h, n) != getattr(b, n) or getattr(h, n) != getattr(c, n) for n in idx)):
        raise ValueError("pgf_buoy_coriolis_curvature: the three calls do not share their common arguments")
    g = grid(h.p.shape, tuple(getattr(h, n) for n in idx))
    cf, _keep = _cf(h.cf1, h.cf2, h.cf3)
    call("pywrf_pgf_buoy_coriolis_curvature", g, ptr(h.ru_tend), ptr(h.rv_tend), ptr(b.rw_tend), ptr(h.ph),
         ptr(h.alt), ptr(h.p), ptr(h.pb), ptr(h.al), ptr(h.php), ptr(h.cqu), ptr(h.cqv), ptr(b.cqw), ptr(h.muu),
         ptr(h.muv), ptr(h.mu), ptr(b.mubf), ptr(h.fnm), ptr(h.fnp), ptr(h.rdnw), ptr(b.rdn), cf, float(b.g_),
         ptr(c.ru), ptr(c.rv), ptr(c.rw), ptr(c.u), ptr(c.v), ptr(c.msftx), ptr(c.msfty), ptr(c.msfux),
         ptr(c.msfuy), ptr(c.msfvx), ptr(c.msfvy), ptr(c.f), ptr(c.e), ptr(c.sina), ptr(c.cosa), float(h.rdx),
         float(h.rdy), stream())
    return (h.ru_tend, h.rv_tend), (b.rw_tend, b.cqw), (c.ru_tend, c.rv_tend, c.rw_tend)


def horizontal_diffusion(name, field, tendency, mut, c1, c2,
                         msfux, msfuy, msfvx, msfvx_inv,
                         msfvy, msftx, msfty,
                         khdif, xkmhd, rdx, rdy,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2623-2838"""
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    nm = name if name in ("u", "v", "w") else "m"
    call("pywrf_horizontal_diffusion", g, flag(nm), ptr(field), ptr(tendency), ptr(mut), ptr(msfux), ptr(msfuy),
         ptr(msfvx), ptr(msfvx_inv), ptr(msfvy), ptr(msftx), ptr(msfty), ptr(xkmhd), float(rdx), float(rdy),
         stream())
    return tendency


def horizontal_diffusion_3dmp(name, field, tendency, mut, c1, c2,
                              base_3d,
                              msfux, msfuy, msfvx, msfvx_inv,
                              msfvy, msftx, msfty,
                              khdif, xkmhd, rdx, rdy,
                              ids, ide, jds, jde, kds, kde,
                              ims, ime, jms, jme, kms, kme,
                              its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2841-2915"""
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_horizontal_diffusion_3dmp", g, ptr(field), ptr(tendency), ptr(mut), ptr(base_3d), ptr(msfux),
         ptr(msfuy), ptr(msfvx), ptr(msfvx_inv), ptr(msfvy), ptr(msftx), ptr(msfty), ptr(xkmhd), float(rdx),
         float(rdy), stream())
    return tendency


def vertical_diffusion(name, field, tendency,
                       c1, c2,
                       alt, MUT, rdn, rdnw, kvdif,
                       ids, ide, jds, jde, kds, kde,
                       ims, ime, jms, jme, kms, kme,
                       its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2917-2959"""
    if name not in ("w", "m"):
        return tendency
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_vertical_diffusion", g, flag(name), ptr(field), ptr(tendency), ptr(alt), ptr(MUT), ptr(rdn),
         ptr(rdnw), float(kvdif), stream())
    return tendency


def vertical_diffusion_u(field, tendency,
                         u_base, c1h, c2h,
                         alt, muu, rdn, rdnw, kvdif,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2962-2991"""
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_vertical_diffusion_u", g, ptr(field), ptr(tendency), ptr(u_base), ptr(alt), ptr(muu), ptr(rdn),
         ptr(rdnw), float(kvdif), stream())
    return tendency


def vertical_diffusion_v(field, tendency,
                         v_base, c1h, c2h,
                         alt, muv, rdn, rdnw, kvdif,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:2994-3024"""
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_vertical_diffusion_v", g, ptr(field), ptr(tendency), ptr(v_base), ptr(alt), ptr(muv), ptr(rdn),
         ptr(rdnw), float(kvdif), stream())
    return tendency


def vertical_diffusion_mp(field, tendency,
                          base, c1, c2,
                          alt, MUT, rdn, rdnw, kvdif,
                          ids, ide, jds, jde, kds, kde,
                          ims, ime, jms, jme, kms, kme,
                          its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3026-3055"""
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_vertical_diffusion_mp", g, ptr(field), ptr(tendency), ptr(base), ptr(alt), ptr(MUT), ptr(rdn),
         ptr(rdnw), float(kvdif), stream())
    return tendency


def vertical_diffusion_3dmp(field, tendency,
                            base_3d, c1, c2,
                            alt, MUT, rdn, rdnw, kvdif,
                            ids, ide, jds, jde, kds, kde,
                            ims, ime, jms, jme, kms, kme,
                            its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3058-3086"""
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_vertical_diffusion_3dmp", g, ptr(field), ptr(tendency), ptr(base_3d), ptr(alt), ptr(MUT), ptr(rdn),
         ptr(rdnw), float(kvdif), stream())
    return tendency


def diffusion_dry(hu_args, hv_args, hw_args, ht_args, vu_args, vv_args, vw_args, vt_args):
    """The first stage's eight diffusion calls (solver.py:1079-1141), each given
    as the tuple of its positional arguments: horizontal_diffusion 'u', 'v',
    'w', horizontal_diffusion_3dmp, vertical_diffusion_u, vertical_diffusion_v,
    vertical_diffusion 'w' and vertical_diffusion_3dmp, in one launch.  Each
    vertical call acts on its horizontal call's field and tendency; the calls
    share mut, map factors, alt, rdn, rdnw, the index set and grid spacing,
    the momentum calls their xkmhd and kvdif; the four tendencies are distinct
    and alias no input.  Returns the four tendencies.  Not part of the
    reference API."""
    hs = [_bind(horizontal_diffusion, a) for a in (hu_args, hv_args, hw_args)] + [
        _bind(horizontal_diffusion_3dmp, ht_args)]
    vu, vv = _bind(vertical_diffusion_u, vu_args), _bind(vertical_diffusion_v, vv_args)
    vw, vt = _bind(vertical_diffusion, vw_args), _bind(vertical_diffusion_3dmp, vt_args)
    vs = [vu, vv, vw, vt]
    hu = hs[0]
    idx = ("ids", "ide", "jds", "jde", "kds", "kde", "ims", "ime", "jms", "jme", "kms", "kme", "its", "ite", "jts",
           "jte", "kts", "kte")
    shared = ("mut", "msfux", "msfuy", "msfvx", "msfvx_inv", "msfvy", "msftx", "msfty")
    same = [(h.field, v.field) for h, v in zip(hs, vs)] + [(h.tendency, v.tendency) for h, v in zip(hs, vs)]
    same += [(getattr(h, n), getattr(hu, n)) for h in hs[1:] for n in shared]
    same += [(h.xkmhd, hu.xkmhd) for h in hs[1:3]] + [(hs[3].base_3d, vt.base_3d), (vw.MUT, hu.mut),
                                                      (vt.MUT, hu.mut)]
    same += [(getattr(v, n), getattr(vu, n)) for v in vs[1:] for n in ("alt", "rdn", "rdnw")]
    tends = [h.tendency for h in hs]
    inputs = [h.field for h in hs] + [hu.mut, hs[3].base_3d, hu.xkmhd, hs[3].xkmhd, vu.u_base, vv.v_base, vu.alt,
                                      vu.muu, vv.muv] + [getattr(hu, n) for n in shared]
    if (hs[0].name != "u" or hs[1].name != "v" or hs[2].name != "w" or vw.name != "w"
            or not all(x is y for x, y in same)
            or any(h.rdx != hu.rdx or h.rdy != hu.rdy for h in hs)
            or vv.kvdif != vu.kvdif or vw.kvdif != vu.kvdif
            or any(getattr(c, n) != getattr(hu, n) for c in hs[1:] + vs for n in idx)):
        raise ValueError("diffusion_dry: the eight calls do not share their common arguments")
    ptrs = {t.data_ptr() for t in tends}
    if len(ptrs) != 4 or any(x.data_ptr() in ptrs for x in inputs):
        raise ValueError("diffusion_dry: the tendencies must be four distinct arrays that alias no input")
    g = grid(hu.field.shape, tuple(getattr(hu, n) for n in idx))
    call("pywrf_diffusion_dry", g, ptr(hs[0].field), ptr(tends[0]), ptr(hs[1].field), ptr(tends[1]),
         ptr(hs[2].field), ptr(tends[2]), ptr(hs[3].field), ptr(tends[3]), ptr(hs[3].base_3d), ptr(hu.mut),
         ptr(hu.msfux), ptr(hu.msfuy), ptr(hu.msfvx), ptr(hu.msfvx_inv), ptr(hu.msfvy), ptr(hu.msftx),
         ptr(hu.msfty), ptr(hu.xkmhd), ptr(hs[3].xkmhd), ptr(vu.u_base), ptr(vv.v_base), ptr(vu.alt), ptr(vu.muu),
         ptr(vv.muv), ptr(vu.rdn), ptr(vu.rdnw), float(vu.kvdif), float(vt.kvdif), float(hu.rdx), float(hu.rdy),
         stream())
    return tuple(tends)


def rk_addtend_dry(ru_tend, rv_tend, rw_tend, ph_tend, t_tend,
                   ru_tendf, rv_tendf, rw_tendf, ph_tendf, t_tendf,
                   u_save, v_save, w_save, ph_save, t_save,
                   mu_tend, mu_tendf, rk_step, c1, c2,
                   h_diabatic, mut, msftx, msfty, msfux, msfuy,
                   msfvx, msfvx_inv, msfvy,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   ips, ipe, jps, jpe, kps, kpe,
                   its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3100-3148"""
    g = grid(ru_tend.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps, kpe,
                             its, ite, jts, jte, kts, kte))
    call("pywrf_rk_addtend_dry", g, ptr(ru_tend), ptr(rv_tend), ptr(rw_tend), ptr(ph_tend), ptr(t_tend),
         ptr(ru_tendf), ptr(rv_tendf), ptr(rw_tendf), ptr(ph_tendf), ptr(t_tendf), ptr(u_save), ptr(v_save),
         ptr(w_save), ptr(ph_save), ptr(t_save), ptr(mu_tend), ptr(mu_tendf), int(rk_step), ptr(h_diabatic),
         ptr(mut), ptr(msfty), ptr(msfuy), ptr(msfvx), ptr(msfvx_inv), stream())
    return ru_tend, rv_tend, rw_tend, ph_tend, t_tend, mu_tend


# ----------------------------------------------------------------- acoustic small step
def small_step_prep(u_1, u_2, v_1, v_2, w_1, w_2,
                    t_1, t_2, ph_1, ph_2,
                    mub, mu_1, mu_2,
                    muu, muus, muv, muvs,
                    mut, muts, mudf,
                    c1h, c2h, c1f, c2f,
                    c3h, c4h, c3f, c4f,
                    u_save, v_save, w_save,
                    t_save, ph_save, mu_save,
                    ww, ww_save,
                    c2a, pb, p, alt,
                    msfux, msfuy, msfvx,
                    msfvx_inv,
                    msfvy, msftx, msfty,
                    rdx, rdy,
                    rk_step,
                    ids, ide, jds, jde, kds, kde,
                    ims, ime, jms, jme, kms, kme,
                    its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3152-3266 (the five perturbation fields are fresh arrays)"""
    g = grid(u_2.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    u2s, v2s, w2s, t2s, ph2s = fresh(u_2), fresh(v_2), fresh(w_2), fresh(t_2), fresh(ph_2)
    call("pywrf_small_step_prep", g, ptr(u_1), ptr(u_2), ptr(v_1), ptr(v_2), ptr(w_1), ptr(w_2), ptr(t_1), ptr(t_2),
         ptr(ph_1), ptr(ph_2), ptr(mub), ptr(mu_1), ptr(mu_2), ptr(muu), ptr(muus), ptr(muv), ptr(muvs), ptr(mut),
         ptr(muts), ptr(mudf), ptr(u_save), ptr(v_save), ptr(w_save), ptr(t_save), ptr(ph_save), ptr(mu_save),
         ptr(ww), ptr(ww_save), ptr(c2a), ptr(pb), ptr(p), ptr(alt), ptr(msfuy), ptr(msfvx_inv), ptr(msfty),
         int(rk_step), ptr(u2s), ptr(v2s), ptr(w2s), ptr(t2s), ptr(ph2s), stream())
    return (u_1, v_1, w_1, t_1, ph_1, u_save, v_save, w_save, t_save, ph_save, u2s, v2s, w2s, t2s, ph2s,
            c2a, ww_save, mu_1, mu_2, muus, muvs, muts, mudf, mu_save)


def calc_p_rho(al_ori, p_ori, ph,
               alt, t_2, t_1, c2a, pm1,
               mu, mut,
               c1h, c2h, c1f, c2f,
               c3h, c4h, c3f, c4f,
               znu, t0,
               rdnw, dnw, smdiv,
               non_hydrostatic, step,
               ids, ide, jds, jde, kds, kde,
               ims, ime, jms, jme, kms, kme,
               its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3269-3313 (al, p fresh; pm1 in place, or a fresh copy of p at step 0)"""
    g = grid(p_ori.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    al, p = fresh(al_ori), fresh(p_ori)
    if step == 0:
        pm1 = fresh(p_ori)
    call("pywrf_calc_p_rho", g, ptr(al_ori), ptr(p_ori), ptr(ph), ptr(alt), ptr(t_2), ptr(t_1), ptr(c2a), ptr(pm1),
         ptr(mu), ptr(mut), float(t0), ptr(rdnw), float(smdiv), int(step), ptr(al), ptr(p), stream())
    return al, p, pm1


def calc_coef_w(a, alpha, gamma,
                mut,
                c1h, c2h, c1f, c2f,
                c3h, c4h, c3f, c4f,
                cqw,
                rdn, rdnw, c2a,
                dts, g_, epssm, top_lid,
                ids, ide, jds, jde, kds, kde,
                ims, ime, jms, jme, kms, kme,
                its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3316-3388"""
    g = grid(a.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_coef_w", g, ptr(a), ptr(alpha), ptr(gamma), ptr(mut), ptr(cqw), ptr(rdn), ptr(rdnw), ptr(c2a),
         float(dts), float(g_), float(epssm), int(bool(top_lid)), stream())
    return a, alpha, gamma


def advance_uv(u_ori, ru_tend, v_ori, rv_tend,
               p, pb,
               ph, php, alt, al, mu,
               muu, cqu, muv, cqv, mudf,
               c1h, c2h, c1f, c2f,
               c3h, c4h, c3f, c4f,
               msfux, msfuy, msfvx,
               msfvx_inv, msfvy,
               rdx, rdy, dts,
               cf1, cf2, cf3, fnm, fnp,
               emdiv,
               rdnw, spec_zone,
               non_hydrostatic, top_lid,
               ids, ide, jds, jde, kds, kde,
               ims, ime, jms, jme, kms, kme,
               its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3391-3525 (u, v fresh)"""
    g = grid(u_ori.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    cf, _keep = _cf(cf1, cf2, cf3)
    u, v = fresh(u_ori), fresh(v_ori)
    call("pywrf_advance_uv", g, ptr(u_ori), ptr(ru_tend), ptr(v_ori), ptr(rv_tend), ptr(p), ptr(pb), ptr(ph),
         ptr(php), ptr(alt), ptr(al), ptr(mu), ptr(muu), ptr(cqu), ptr(muv), ptr(cqv), ptr(mudf), ptr(msfux),
         ptr(msfuy), ptr(msfvx), ptr(msfvx_inv), ptr(msfvy), float(rdx), float(rdy), float(dts), cf, ptr(fnm),
         ptr(fnp), float(emdiv), ptr(rdnw), int(spec_zone), ptr(u), ptr(v), stream())
    return u, v


def advance_mu_t(ww_ori, ww_1, u, u_1, v, v_1,
                 mu_ori, mut, muave_ori, muts_ori, muu, muv, mudf_ori,
                 c1h, c2h, c1f, c2f,
                 c3h, c4h, c3f, c4f,
                 uam, vam, wwam, t_ori, t_1,
                 t_ave_ori, ft, mu_tend,
                 rdx, rdy, dts, epssm,
                 dnw, fnm, fnp, rdnw,
                 msfux, msfuy, msfvx, msfvx_inv,
                 msfvy, msftx, msfty,
                 step,
                 ids, ide, jds, jde, kds, kde,
                 ims, ime, jms, jme, kms, kme,
                 its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3528-3638 (returns the reference's 11-tuple; ww, t, t_ave, muave, muts,
    mudf, mu are fresh)"""
    g = grid(u.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    ww, t, t_ave = fresh(ww_ori), fresh(t_ori), fresh(t_ave_ori)
    muave, muts, mudf, mu = fresh(muave_ori), fresh(muts_ori), fresh(mudf_ori), fresh(mu_ori)
    call("pywrf_advance_mu_t", g, ptr(ww_ori), ptr(ww_1), ptr(u), ptr(u_1), ptr(v), ptr(v_1), ptr(mu_ori), ptr(mut),
         ptr(muave_ori), ptr(muts_ori), ptr(muu), ptr(muv), ptr(mudf_ori), ptr(t_ori), ptr(t_1), ptr(t_ave_ori),
         ptr(ft), ptr(mu_tend), float(rdx), float(rdy), float(dts), float(epssm), ptr(dnw), ptr(fnm), ptr(fnp),
         ptr(rdnw), ptr(msfuy), ptr(msfvx_inv), ptr(msftx), ptr(msfty), ptr(ww), ptr(t), ptr(t_ave), ptr(muave),
         ptr(muts), ptr(mudf), ptr(mu), stream())
    return ww, ww_1, t, t_ave, uam, vam, wwam, muave, muts, mudf, mu


def advance_w(w_ori, rw_tend, ww, w_save, u, v,
              mu1, mut, muave, muts,
              c1h, c2h, c1f, c2f,
              c3h, c4h, c3f, c4f,
              t_2ave_ori, t_2, t_1,
              ph_ori, ph_1, phb, ph_tend,
              ht, c2a, cqw, alt, alb,
              a, alpha, gamma,
              rdx, rdy, dts, t0, epssm,
              dnw, fnm, fnp, rdnw, rdn,
              cf1, cf2, cf3, msftx, msfty,
              top_lid,
              ids, ide, jds, jde, kds, kde,
              ims, ime, jms, jme, kms, kme,
              its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3641-3796 (t_2ave, w, ph fresh)"""
    g = grid(w_ori.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    cf, _keep = _cf(cf1, cf2, cf3)
    t_2ave, w, ph = fresh(t_2ave_ori), fresh(w_ori), fresh(ph_ori)
    call("pywrf_advance_w", g, ptr(w_ori), ptr(rw_tend), ptr(ww), ptr(w_save), ptr(u), ptr(v), ptr(mut), ptr(muave),
         ptr(muts), ptr(t_2ave_ori), ptr(t_2), ptr(t_1), ptr(ph_ori), ptr(ph_1), ptr(phb), ptr(ph_tend), ptr(ht),
         ptr(c2a), ptr(cqw), ptr(alt), ptr(a), ptr(alpha), ptr(gamma), float(rdx), float(rdy), float(dts),
         float(t0), float(epssm), ptr(fnm), ptr(fnp), ptr(rdnw), ptr(rdn), cf, ptr(msftx), ptr(msfty), ptr(t_2ave),
         ptr(w), ptr(ph), stream())
    return t_2ave, w, ph


def sumflux(ru, rv, ww,
            u_lin, v_lin, ww_lin,
            muu, muv,
            c1h, c2h, c1f, c2f,
            c3h, c4h, c3f, c4f,
            ru_m, rv_m, ww_m, epssm,
            msfux, msfuy, msfvx, msfvx_inv, msfvy,
            iteration, number_of_small_timesteps,
            ids, ide, jds, jde, kds, kde,
            ims, ime, jms, jme, kms, kme,
            its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3799-3839"""
    g = grid(ru.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_sumflux", g, ptr(ru), ptr(rv), ptr(ww), ptr(u_lin), ptr(v_lin), ptr(ww_lin), ptr(muu), ptr(muv),
         ptr(ru_m), ptr(rv_m), ptr(ww_m), ptr(msfuy), ptr(msfvx_inv), int(iteration), int(number_of_small_timesteps),
         stream())
    return ru_m, rv_m, ww_m


def calc_mu_uv_1(mu, muu, muv,
                 ids, ide, jds, jde, kds, kde,
                 ims, ime, jms, jme, kms, kme,
                 its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3844-3860"""
    g = grid(mu.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_mu_uv_1", g, ptr(mu), ptr(muu), ptr(muv), stream())
    return muu, muv


def small_step_finish(u_2, u_1, v_2, v_1, w_2, w_1,
                      t_2, t_1, ph_2, ph_1, ww, ww1,
                      mu_2, mu_1,
                      mut, muts, muu, muus, muv, muvs,
                      c1h, c2h, c1f, c2f,
                      c3h, c4h, c3f, c4f,
                      u_save, v_save, w_save,
                      t_save, ph_save, mu_save,
                      msfux, msfuy, msfvx, msfvy,
                      msftx, msfty,
                      h_diabatic,
                      number_of_small_timesteps, dts,
                      rk_step, rk_order,
                      ids, ide, jds, jde, kds, kde,
                      ims, ime, jms, jme, kms, kme,
                      its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3863-3930 (u, v, w, ph, t fresh; ww, mu_2 in place)"""
    g = grid(u_2.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    u, v, w, ph, t = fresh(u_2), fresh(v_2), fresh(w_2), fresh(ph_2), fresh(t_2)
    call("pywrf_small_step_finish", g, ptr(u_2), ptr(v_2), ptr(w_2), ptr(t_2), ptr(ph_2), ptr(ww), ptr(ww1),
         ptr(mu_2), ptr(mut), ptr(muts), ptr(muu), ptr(muus), ptr(muv), ptr(muvs), ptr(u_save), ptr(v_save),
         ptr(w_save), ptr(t_save), ptr(ph_save), ptr(mu_save), ptr(msfuy), ptr(msfvx), ptr(msfty), ptr(h_diabatic),
         int(number_of_small_timesteps), float(dts), int(rk_step), int(rk_order), ptr(u), ptr(v), ptr(w), ptr(ph),
         ptr(t), stream())
    return u, v, w, ph, ww, t, mu_2


# ----------------------------------------------------------------- scalars, EOS, surface w
def rk_scalar_tend(scs, sce,
                   tenddec,
                   rk_step, dt,
                   ru, rv, ww, mut, mub, mu_old,
                   c1h, c2h, alt,
                   scalar_old, scalar,
                   scalar_tends, advect_tend,
                   h_tendency, z_tendency,
                   RQVFTEN,
                   base, moist_step, fnm, fnp,
                   msfux, msfuy, msfvx, msfvx_inv,
                   msfvy, msftx, msfty,
                   rdx, rdy, rdn, rdnw,
                   khdif, kvdif, xkmhd,
                   diff_6th_opt, diff_6th_factor,
                   adv_opt,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:3982-4051"""
    g = grid(ru.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_rk_scalar_tend", g, int(scs), int(sce), int(rk_step), ptr(ru), ptr(rv), ptr(ww), ptr(mut), ptr(alt),
         ptr(scalar), ptr(scalar_tends), ptr(advect_tend), ptr(h_tendency), ptr(z_tendency), ptr(base),
         int(bool(moist_step)), ptr(fnm), ptr(fnp), ptr(msfux), ptr(msfuy), ptr(msfvx), ptr(msfvx_inv), ptr(msfvy),
         ptr(msftx), ptr(msfty), float(rdx), float(rdy), ptr(rdn), ptr(rdnw), float(kvdif), ptr(xkmhd), stream())
    return scalar_tends, advect_tend, h_tendency, z_tendency


def _bind(fn, args):
    """The positional arguments of a call of `fn`, by parameter name."""
    names = list(inspect.signature(fn).parameters)
    if len(args) != len(names):
        raise TypeError("%s takes %d positional arguments, got %d" % (fn.__name__, len(names), len(args)))
    return SimpleNamespace(**dict(zip(names, args)))


def _shared(fn, a, b, own):
    """Raise unless the two bound calls of `fn` agree on every argument outside `own`."""
    for n in inspect.signature(fn).parameters:
        if n in own:
            continue
        x, y = getattr(a, n), getattr(b, n)
        same = x is y if isinstance(x, torch.Tensor) or isinstance(y, torch.Tensor) else x == y
        if not same:
            raise ValueError("%s pair: argument %r differs between the two calls" % (fn.__name__, n))


def rk_scalar_tend_pair(args1, args2):
    """Two rk_scalar_tend calls, each given as the tuple of its positional
    arguments, that share every argument but the species range, the species
    arrays, the two tendency arrays, the moist flag and the advection option
    (the moist and the tracer call of solver.py:1548-1610), in the same
    launches.  Returns the two calls' results.  Not part of the reference API."""
    a, b = _bind(rk_scalar_tend, args1), _bind(rk_scalar_tend, args2)
    _shared(rk_scalar_tend, a, b, ("scs", "sce", "scalar_old", "scalar", "scalar_tends", "advect_tend",
                                   "moist_step", "adv_opt"))
    if a.scalar is b.scalar or a.scalar_tends is b.scalar_tends or a.advect_tend is b.advect_tend:
        raise ValueError("rk_scalar_tend pair: the two calls must have distinct species and tendency arrays")
    g = grid(a.ru.shape, (a.ids, a.ide, a.jds, a.jde, a.kds, a.kde, a.ims, a.ime, a.jms, a.jme, a.kms, a.kme,
                          a.its, a.ite, a.jts, a.jte, a.kts, a.kte))
    call("pywrf_rk_scalar_tend_pair", g, int(a.rk_step), int(a.scs), int(a.sce), ptr(a.scalar), ptr(a.scalar_tends),
         ptr(a.advect_tend), int(bool(a.moist_step)), int(b.scs), int(b.sce), ptr(b.scalar), ptr(b.scalar_tends),
         ptr(b.advect_tend), int(bool(b.moist_step)), ptr(a.ru), ptr(a.rv), ptr(a.ww), ptr(a.mut), ptr(a.alt),
         ptr(a.h_tendency), ptr(a.z_tendency), ptr(a.base), ptr(a.fnm), ptr(a.fnp), ptr(a.msfux), ptr(a.msfuy),
         ptr(a.msfvx), ptr(a.msfvx_inv), ptr(a.msfvy), ptr(a.msftx), ptr(a.msfty), float(a.rdx), float(a.rdy),
         ptr(a.rdn), ptr(a.rdnw), float(a.kvdif), ptr(a.xkmhd), stream())
    return ((a.scalar_tends, a.advect_tend, a.h_tendency, a.z_tendency),
            (b.scalar_tends, b.advect_tend, b.h_tendency, b.z_tendency))


def rk_update_scalar(scs, sce,
                     scalar_1, scalar_2, sc_tend,
                     advect_tend,
                     h_tendency, z_tendency,
                     msftx, msfty, c1, c2,
                     mu_old, mu_new, mu_base,
                     rk_step, dt, spec_zone,
                     tenddec,
                     ids, ide, jds, jde, kds, kde,
                     ims, ime, jms, jme, kms, kme,
                     its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:4055-4119"""
    g = grid(scalar_2.shape[1:], (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte,
                                  kts, kte))
    call("pywrf_rk_update_scalar", g, int(scs), int(sce), ptr(scalar_1), ptr(scalar_2), ptr(sc_tend),
         ptr(advect_tend), ptr(msfty), ptr(mu_old), ptr(mu_new), ptr(mu_base), int(rk_step), float(dt),
         int(spec_zone), stream())
    return scalar_1, scalar_2


def rk_update_scalar_pair(args1, args2):
    """Two rk_update_scalar calls, each given as the tuple of its positional
    arguments, that share every argument but the species range, the two
    species arrays and their tendencies (the moist and the tracer update of
    solver.py:1548-1636), in one launch.  Returns the two calls' results.  Not
    part of the reference API."""
    a, b = _bind(rk_update_scalar, args1), _bind(rk_update_scalar, args2)
    _shared(rk_update_scalar, a, b, ("scs", "sce", "scalar_1", "scalar_2", "sc_tend", "advect_tend"))
    if a.scalar_1 is b.scalar_1 or a.scalar_2 is b.scalar_2:
        raise ValueError("rk_update_scalar pair: the two calls must have distinct species arrays")
    if a.scalar_2.shape[1:] != b.scalar_2.shape[1:]:
        raise ValueError("rk_update_scalar pair: the two species arrays differ in shape per species")
    g = grid(a.scalar_2.shape[1:], (a.ids, a.ide, a.jds, a.jde, a.kds, a.kde, a.ims, a.ime, a.jms, a.jme, a.kms,
                                    a.kme, a.its, a.ite, a.jts, a.jte, a.kts, a.kte))
    call("pywrf_rk_update_scalar_pair", g, int(a.rk_step), int(a.scs), int(a.sce), ptr(a.scalar_1), ptr(a.scalar_2),
         ptr(a.sc_tend), ptr(a.advect_tend), int(b.scs), int(b.sce), ptr(b.scalar_1), ptr(b.scalar_2),
         ptr(b.sc_tend), ptr(b.advect_tend), ptr(a.msfty), ptr(a.mu_old), ptr(a.mu_new), ptr(a.mu_base),
         float(a.dt), int(a.spec_zone), stream())
    return (a.scalar_1, a.scalar_2), (b.scalar_1, b.scalar_2)


def calc_p_rho_phi(moist, n_moist, hypsometric_opt,
                   al, alb, mu, muts,
                   c1, c2, c3h, c4h, c3f, c4f,
                   ph, phb, p, pb,
                   t, p0, t0, ptop, znu, znw, dnw, rdnw,
                   rdn, non_hydrostatic,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:4124-4164"""
    g = grid(p.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_calc_p_rho_phi", g, ptr(moist), ptr(al), ptr(alb), ptr(mu), ptr(muts), ptr(ph), ptr(p), ptr(pb),
         ptr(t), float(p0), float(t0), ptr(rdnw), stream())
    return al, p, ph


def phy_prep_part2(mut, muu, muv,
                   c1h, c2h, c1f, c2f,
                   RTHBLTEN, RUBLTEN, RVBLTEN,
                   RQVBLTEN, RQCBLTEN, RQIBLTEN,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:4179-4213"""
    g = grid(RUBLTEN.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts,
                             kte))
    call("pywrf_phy_prep_part2", g, ptr(mut), ptr(RTHBLTEN), ptr(RUBLTEN), ptr(RVBLTEN), ptr(RQVBLTEN),
         ptr(RQCBLTEN), ptr(RQIBLTEN), stream())
    return RUBLTEN, RVBLTEN, RTHBLTEN, RQVBLTEN, RQCBLTEN, RQIBLTEN


def set_w_surface(znw, fill_w_flag,
                  w, ht, u, v, cf1, cf2, cf3, rdx, rdy,
                  msftx, msfty,
                  ids, ide, jds, jde, kds, kde,
                  ims, ime, jms, jme, kms, kme,
                  its, ite, jts, jte, kts, kte):
    """wrf_dynamics.py:4218-4324"""
    g = grid(w.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    cf, _keep = _cf(cf1, cf2, cf3)
    call("pywrf_set_w_surface", g, ptr(w), ptr(ht), ptr(u), ptr(v), cf, float(rdx), float(rdy), ptr(msftx),
         ptr(msfty), stream())
    return w


# ----------------------------------------------------------------- batched open boundaries
def set_physical_bc_group(items, idx24, idx2d):
    """A run of consecutive set_physical_bc3d / set_physical_bc2d calls on
    distinct fields, in one launch.  items: [(tensor, variable, is_2d)];
    idx24 / idx2d: the index tuples the individual calls would take.  Same
    result as calling the two functions in order (the fields are distinct, so
    the fills are independent).  Not part of the reference API."""
    import ctypes
    n = len(items)
    if n == 0:
        return
    if len({t.data_ptr() for t, _, _ in items}) != n:
        raise ValueError("set_physical_bc_group: a field appears twice")
    t3 = next((t for t, _, two in items if not two), None)
    t2 = next((t for t, _, two in items if two), None)
    shape3 = t3.shape if t3 is not None else (idx24[5],) + tuple(items[0][0].shape[-2:])
    shape2 = t2.shape if t2 is not None else tuple(items[0][0].shape[-2:])
    g3, g2 = grid(shape3, idx24), grid(shape2, idx2d)
    ptrs = (ctypes.c_void_p * n)(*[ptr(t).value for t, _, _ in items])
    flags = b"".join(v.encode() for _, v, _ in items)
    kinds = (ctypes.c_int * n)(*[1 if two else 0 for _, _, two in items])
    call("pywrf_set_physical_bc_group", g3, g2, n, ptrs, flags, kinds, stream())


def zero_group(tensors):
    """tensor.zero_() on each of up to 32 distinct contiguous FP32 device
    tensors, in one launch (the driver's runs of zero_ calls,
    solver.py:848-851, 975-981).  Not part of the reference API."""
    import ctypes
    n = len(tensors)
    if n == 0:
        return
    if n > 32:
        raise ValueError("zero_group: at most 32 tensors per call")
    if len({t.data_ptr() for t in tensors}) != n:
        raise ValueError("zero_group: a tensor appears twice")
    ptrs = (ctypes.c_void_p * n)(*[ptr(t).value for t in tensors])
    counts = (ctypes.c_longlong * n)(*[t.numel() for t in tensors])
    call("pywrf_zero_group", n, ptrs, counts, stream())


# ----------------------------------------------------------------- physics hooks
def pbl_driver():
    """The boundary-layer hook of the first RK stage (wrf_dynamics.py:413-415,
    marked at solver.py:868): no work in the reference, none here; the PBL
    tendencies rublten .. rqiblten stay as the driver set them."""
    return None


def sixth_order_diffusion(name, field, tendency, MUT, dt,
                          config_flags, c1, c2,
                          diff_6th_opt, diff_6th_factor,
                          ids, ide, jds, jde, kds, kde,
                          ims, ime, jms, jme, kms, kme,
                          its, ite, jts, jte, kts, kte):
    """The 6th-order diffusion hook (wrf_dynamics.py:3088-3097): the reference
    forms diff_6th_coef and returns None without touching the tendency; so
    does this, with no device work."""
    return None
