"""Oracle: scalar transport, equation of state and surface w
(test infrastructure only).

CPU restatement of pywrf/wrf_dynamics.py:3982-4324.
"""
import torch

from .constants import P_QV, prandtl, rvovrd, r_d, cpovcv, time_step
from ._box import Box, zeros_like3
from .dyn_adv import advect_scalar
from .dyn_tend import horizontal_diffusion, vertical_diffusion, vertical_diffusion_mp


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
    """Advection (+ first-stage diffusion) tendencies of scalar species.
    wrf_dynamics.py:3982-4051."""
    khdq, kvdq = khdif / prandtl, kvdif / prandtl
    ix = (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte)
    for im in range(scs, sce + 1):
        advect_tend[im, kts:kte, jts:jte, its:ite] = 0.
        h_tendency[:, :, :] = 0.
        z_tendency[:, :, :] = 0.
        advect_tend[im] = advect_scalar(scalar[im], scalar[im], advect_tend[im], ru, rv, ww, c1h, c2h,
                                        mut, time_step, msfux, msfuy, msfvx, msfvy, msftx, msfty, fnm, fnp,
                                        rdx, rdy, rdnw, *ix)
        if rk_step == 1:
            scalar_tends[im] = horizontal_diffusion('m', scalar[im], scalar_tends[im], mut, c1h, c2h,
                                                    msfux, msfuy, msfvx, msfvx_inv, msfvy, msftx, msfty,
                                                    khdq, xkmhd, rdx, rdy, *ix)
            if moist_step and im == P_QV:
                scalar_tends[im] = vertical_diffusion_mp(scalar[im], scalar_tends[im], base, c1h, c2h,
                                                         alt, mut, rdn, rdnw, kvdq, *ix)
            else:
                scalar_tends[im] = vertical_diffusion('m', scalar[im], scalar_tends[im], c1h, c2h,
                                                      alt, mut, rdn, rdnw, kvdq, *ix)
    return scalar_tends, advect_tend, h_tendency, z_tendency


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
    """Mass-weighted RK update of scalar species.  wrf_dynamics.py:4055-4119."""
    bs = Box(kts, kte - 1, jts, min(jte, jde - 1), its, min(ite, ide - 1))
    bi = Box(kts, min(kte, kde - 1), max(jts, jds + spec_zone), min(jte, jde - spec_zone - 1),
             max(its, ids + spec_zone), min(ite, ide - spec_zone - 1))
    bf = Box(kts, kde - 1, jts, jde - 1, its, ide - 1)
    muold = torch.zeros(mu_base.shape, dtype=mu_base.dtype)
    munew = torch.zeros(mu_base.shape, dtype=mu_base.dtype)
    muold[jts:jde - 1, its:ide - 1] = mu_old[jts:jde - 1, its:ide - 1] + mu_base[jts:jde - 1, its:ide - 1]
    munew[jts:jde - 1, its:ide - 1] = mu_new[jts:jde - 1, its:ide - 1] + mu_base[jts:jde - 1, its:ide - 1]
    tendency = zeros_like3(scalar_2[0])
    for im in range(scs, sce + 1):
        bf.put(tendency, 0.)
        bi.put(tendency, bi(advect_tend[im]) * bi.m(msfty))
        bs.put(tendency, bs(tendency) + bs(sc_tend[im]))
        if rk_step == 1:
            scalar_1[im, :, :, :] = scalar_2[im, :, :, :] + 0.0
        bf.put(scalar_2[im], (bf.m(muold) * bf(scalar_1[im]) + dt * bf(tendency)) / bf.m(munew))
    return scalar_1, scalar_2


def calc_p_rho_phi(moist, n_moist, hypsometric_opt,
                   al, alb, mu, muts,
                   c1, c2, c3h, c4h, c3f, c4f,
                   ph, phb, p, pb,
                   t, p0, t0, ptop, znu, znw, dnw, rdnw,
                   rdn, non_hydrostatic,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """Moist equation of state for al and p (phi is not re-integrated).
    wrf_dynamics.py:4124-4164."""
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    b.put(al, -1. / b.m(muts) * (b(alb) * b.m(mu) + b.c(rdnw) * (b(ph, 1) - b(ph))))
    qvf = 1. + rvovrd * b(moist[P_QV])
    temp = (r_d * (t0 + b(t)) * qvf) / (p0 * (b(al) + b(alb)))
    b.put(p, temp ** cpovcv)
    b.put(p, b(p) * p0 - b(pb))
    return al, p, ph


def phy_prep_part2(mut, muu, muv,
                   c1h, c2h, c1f, c2f,
                   RTHBLTEN, RUBLTEN, RVBLTEN,
                   RQVBLTEN, RQCBLTEN, RQIBLTEN,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """Decouple PBL tendencies from mut.  wrf_dynamics.py:4179-4213."""
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    for a in (RUBLTEN, RVBLTEN, RTHBLTEN, RQVBLTEN, RQCBLTEN, RQIBLTEN):
        b.put(a, b(a) / b.m(mut))
    return RUBLTEN, RVBLTEN, RTHBLTEN, RQVBLTEN, RQCBLTEN, RQIBLTEN


def _wsfc(w, ht, u, v, cf1, cf2, cf3, rdx, rdy, msftx, msfty, js, is_, ydiff, xdiff, mx=None):
    """One statement of set_w_surface over rows js x columns is_.

    ydiff / xdiff pick which of the two one-sided slope terms appear:
    'both', 'up' (j+1 - j), 'dn' (j - j-1); mx overrides the msftx sample."""
    J = lambda d: slice(js.start + d, js.stop + d)
    I = lambda d: slice(is_.start + d, is_.stop + d)
    cv = lambda dj: cf1 * v[0, J(dj), is_] + cf2 * v[1, J(dj), is_] + cf3 * v[2, J(dj), is_]
    cu = lambda di: cf1 * u[0, js, I(di)] + cf2 * u[1, js, I(di)] + cf3 * u[2, js, I(di)]
    if ydiff == 'both':
        ty = (ht[J(1), is_] - ht[js, is_]) * cv(1) + (ht[js, is_] - ht[J(-1), is_]) * cv(0)
    elif ydiff == 'up':
        ty = (ht[J(1), is_] - ht[js, is_]) * cv(1)
    elif ydiff == 'zero':
        ty = (ht[J(1), is_] - ht[J(1), is_]) * cv(1)
    else:
        ty = (ht[js, is_] - ht[J(-1), is_]) * cv(0)
    if xdiff == 'both':
        tx = (ht[js, I(1)] - ht[js, is_]) * cu(1) + (ht[js, is_] - ht[js, I(-1)]) * cu(0)
    elif xdiff == 'up':
        tx = (ht[js, I(1)] - ht[js, is_]) * cu(1)
    else:
        tx = (ht[js, is_] - ht[js, I(-1)]) * cu(0)
    msx = msftx[js, is_] if mx is None else mx
    w[0, js, is_] = msfty[js, is_] * 0.5 * rdy * ty + msx * 0.5 * rdx * tx


def set_w_surface(znw, fill_w_flag,
                  w, ht, u, v, cf1, cf2, cf3, rdx, rdy,
                  msftx, msfty,
                  ids, ide, jds, jde, kds, kde,
                  ims, ime, jms, jme, kms, kme,
                  its, ite, jts, jte, kts, kte):
    """Kinematic lower boundary w = V . grad(h) with one-sided edge stencils,
    including the reference's corner quirks.  wrf_dynamics.py:4218-4324."""
    a = (w, ht, u, v, cf1, cf2, cf3, rdx, rdy, msftx, msfty)
    S = lambda x0, x1: slice(x0, x1)
    _wsfc(*a, S(jts + 1, jte - 2), S(its + 1, ite - 2), 'both', 'both')
    _wsfc(*a, S(jts, jts + 1), S(its + 1, ite - 2), 'up', 'both')
    _wsfc(*a, S(jte - 2, jte - 1), S(its + 1, ite - 2), 'dn', 'both')
    _wsfc(*a, S(jts + 1, jte - 2), S(its, its + 1), 'both', 'up')
    _wsfc(*a, S(jts + 1, jte - 2), S(ite - 2, ite - 1), 'both', 'dn')
    _wsfc(*a, S(jts, jts + 1), S(its, its + 1), 'up', 'up')
    _wsfc(*a, S(jte - 2, jte - 1), S(its, its + 1), 'dn', 'up')
    # y slope is ht[jts+1]-ht[jts+1] (identically zero) in the reference
    _wsfc(*a, S(jts, jts + 1), S(ite - 2, ite - 1), 'zero', 'dn')
    # x term samples msftx at (jts, its) in the reference
    _wsfc(*a, S(jte - 2, jte - 1), S(ite - 2, ite - 1), 'dn', 'dn', mx=msftx[jts, its])
    return w
