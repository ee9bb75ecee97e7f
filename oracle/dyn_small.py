"""Oracle: acoustic small-step machinery (test infrastructure only).

CPU restatement of pywrf/wrf_dynamics.py:3152-3930: small_step_prep,
calc_p_rho, calc_coef_w, advance_uv, advance_mu_t (with the reference's
no-grad mass update), advance_w (tridiagonal Thomas solve and Rayleigh
damping), sumflux, calc_mu_uv_1 and small_step_finish.
"""
import torch

from .constants import cpovcv, g
from ._box import Box, zeros_like3


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
    """Save the stage state and form the small-step perturbation variables.
    wrf_dynamics.py:3152-3266; the five perturbation fields are new tensors."""
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    k0, k1 = kts, min(kte, kde - 1)
    i1u, j1v = ite, jte
    u2s, v2s, w2s, t2s, ph2s = u_2.clone(), v_2.clone(), w_2.clone(), t_2.clone(), ph_2.clone()
    h = (slice(j0, j1), slice(i0, i1))
    if rk_step == 1:
        mu_1[h] = mu_2[h] + 0.0
        ww_save[(kde - 1,) + h] = 0.0
        ww_save[(0,) + h] = 0.0
        mudf[h] = 0.0
        u_1[k0:k1, j0:j1, i0:i1u] = u_2[k0:k1, j0:j1, i0:i1u] + 0.0
        v_1[k0:k1, j0:j1v, i0:i1] = v_2[k0:k1, j0:j1v, i0:i1] + 0.0
        t_1[k0:k1, j0:j1, i0:i1] = t_2[k0:k1, j0:j1, i0:i1] + 0.0
        w_1[k0:kde, j0:j1, i0:i1] = w_2[k0:kde, j0:j1, i0:i1] + 0.0
        ph_1[k0:kde, j0:j1, i0:i1] = ph_2[k0:kde, j0:j1, i0:i1] + 0.0
        muts[h] = mub[h] + mu_2[h]
        muus[j0:j1, i0:i1u] = muu[j0:j1, i0:i1u] + 0.0
        muvs[j0:j1v, i0:i1] = muv[j0:j1v, i0:i1] + 0.0
        mu_save[h] = mu_2[h] + 0.0
        mu_2[h] = 0.0
    else:
        muts[h] = mub[h] + mu_1[h]
        muus[j0:j1, i0:i1u] = 0.5 * (mub[j0:j1, i0:i1u] + mu_1[j0:j1, i0:i1u]
                                     + mub[j0:j1, i0 - 1:i1u - 1] + mu_1[j0:j1, i0 - 1:i1u - 1])
        muvs[j0:j1v, i0:i1u] = 0.5 * (mub[j0:j1v, i0:i1u] + mu_1[j0:j1v, i0:i1u]
                                      + mub[j0 - 1:j1v - 1, i0:i1u] + mu_1[j0 - 1:j1v - 1, i0:i1u])
        mu_save[h] = mu_2[h] + 0.0
        mu_2[h] = mu_1[h] - mu_2[h]
    ww_save[(kde - 1,) + h] = 0.0
    ww_save[(0,) + h] = 0.0
    b = Box(k0, k1, j0, j1, i0, i1)
    b.put(c2a, cpovcv * (b(pb) + b(p)) / b(alt))
    bu = Box(k0, k1, j0, j1, i0, i1u)
    bu.put(u_save, bu(u_2) + 0.0)
    bu.put(u2s, (bu.m(muus) * bu(u_1) - bu.m(muu) * bu(u_2)) / bu.m(msfuy))
    bv = Box(k0, k1, j0, j1v, i0, i1)
    bv.put(v_save, bv(v_2) + 0.0)
    bv.put(v2s, (bv.m(muvs) * bv(v_1) - bv.m(muv) * bv(v_2)) * bv.m(msfvx_inv))
    b.put(t_save, b(t_2) + 0.0)
    b.put(t2s, b.m(muts) * b(t_1) - b.m(mut) * b(t_2))
    bw = Box(k0, kde, j0, j1, i0, i1)
    bw.put(w_save, bw(w_2) + 0.0)
    bw.put(w2s, (bw.m(muts) * bw(w_1) - bw.m(mut) * bw(w_2)) / bw.m(msfty))
    bw.put(ph_save, bw(ph_2) + 0.0)
    bw.put(ph2s, bw(ph_1) - bw(ph_2))
    bw.put(ww_save, bw(ww))
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
    """Small-step inverse density and pressure with divergence damping.
    wrf_dynamics.py:3269-3313; al and p are new tensors, pm1 is updated in
    place (or replaced by a copy of p at step 0)."""
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    al, p = al_ori.clone(), p_ori.clone()
    b.put(al, -1. / b.m(mut) * (b(alt) * b.m(mu) + b.c(rdnw) * (b(ph, 1) - b(ph))))
    b.put(p, b(c2a) * (b(alt) * (b(t_2) - b.m(mu) * b(t_1)) / (b.m(mut) * (t0 + b(t_1))) - b(al)))
    if step == 0:
        pm1 = p + 0.0
    else:
        ptmp = b(p) + 0.0
        b.put(p, b(p) + smdiv * (b(p) - b(pm1)))
        b.put(pm1, ptmp)
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
    """Tridiagonal coefficients and forward-elimination factors for the
    implicit w-phi system.  wrf_dynamics.py:3316-3388."""
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    h = (slice(j0, j1), slice(i0, i1))
    lid = 0 if top_lid else 1
    C = (0.5 * dts * g_ * (1. + epssm)) ** 2
    bb, cc = zeros_like3(c2a), zeros_like3(c2a)
    mm = mut[h] * mut[h]
    a[(1,) + h] = 0.0
    a[(kde - 1,) + h] = -2. * C * rdnw[kde - 2] ** 2 * c2a[(kde - 2,) + h] * lid / (mut[h] * mut[h])
    gamma[(0,) + h] = 0.0
    b = Box(2, kde - 1, j0, j1, i0, i1)
    b.put(a, -b(cqw) * C * b.c(rdn) * b.c(rdnw, -1) * b(c2a, -1) / (b.m(mut) * b.m(mut)))
    b = Box(1, kde - 1, j0, j1, i0, i1)
    b.put(bb, 1. + b(cqw) * C * b.c(rdn) * (b.c(rdnw) * b(c2a) / (b.m(mut) * b.m(mut))
                                           + b.c(rdnw, -1) * b(c2a, -1) / (b.m(mut) * b.m(mut))))
    b.put(cc, -b(cqw) * C * b.c(rdn) * b.c(rdnw) * b(c2a) / (b.m(mut) * b.m(mut)))
    for k in range(1, kde - 1):
        alpha[(k,) + h] = 1. / (bb[(k,) + h] - a[(k,) + h] * gamma[(k - 1,) + h])
        gamma[(k,) + h] = cc[(k,) + h] * alpha[(k,) + h]
    bb[(kde - 1,) + h] = 1. + 2. * C * (rdnw[kde - 2] ** 2) * c2a[(kde - 2,) + h] / (mut[h] * mut[h])
    alpha[(kde - 1,) + h] = 1. / (bb[(kde - 1,) + h] - a[(kde - 1,) + h] * gamma[(kde - 2,) + h])
    gamma[(kde - 1,) + h] = 0.0
    del mm
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
    """Explicit small-step u and v update (PGF + external-mode damping).
    wrf_dynamics.py:3391-3525; u and v are new tensors."""
    i0, i1 = max(its, ids + spec_zone), min(ite, ide - spec_zone - 1)
    j0, j1 = max(jts, jds + spec_zone), min(jte, jde - spec_zone - 1)
    k0, k1 = kts, min(kte, kde - 1)
    i1u, j1v = min(ite, ide - spec_zone), min(jte, jde - spec_zone)
    dx, dy = 1. / rdx, 1. / rdy
    u, v = u_ori.clone(), v_ori.clone()
    dpn = zeros_like3(p)
    for (bj0, bj1, bi0, bi1, d, vel, tend, mv, mx, my, rd, cq, mf, dfac) in (
            (j0, j1, i0, i1u, (0, -1), u, ru_tend, muu, msfux, msfuy, rdx, cqu, msfuy, -emdiv * dx),
            (j0, j1v, i0, i1, (-1, 0), v, rv_tend, muv, msfvy, msfvx, rdy, cqv, msfvx_inv, -emdiv * dy)):
        b = Box(k0, k1, bj0, bj1, bi0, bi1)
        b.put(vel, b(vel) + dts * b(tend))
        if d == (0, -1):
            mudf_xy = dfac * (b.m(mudf) - b.m(mudf, *d)) / b.m(mf)
        else:
            mudf_xy = dfac * (b.m(mudf) - b.m(mudf, *d)) * b.m(mf)
        dpxy = (b.m(mx) / b.m(my)) * 0.5 * rd * b.m(mv) * (
            ((b(ph, 1) - b(ph, 1, *d)) + (b(ph) - b(ph, 0, *d))) + (b(alt) + b(alt, 0, *d)) * (b(p) - b(p, 0, *d))
            + (b(al) + b(al, 0, *d)) * (b(pb) - b(pb, 0, *d)))
        h = (slice(bj0, bj1), slice(bi0, bi1))
        hd = (slice(bj0 + d[0], bj1 + d[0]), slice(bi0 + d[1], bi1 + d[1]))
        dpn[(0,) + h] = 0.5 * (cf1 * (p[(0,) + h] + p[(0,) + hd]) + cf2 * (p[(1,) + h] + p[(1,) + hd])
                               + cf3 * (p[(2,) + h] + p[(2,) + hd]))
        dpn[(kde - 1,) + h] = 0.
        bi = Box(k0 + 1, k1, bj0, bj1, bi0, bi1)
        bi.put(dpn, 0.5 * (bi.c(fnm) * (bi(p) + bi(p, 0, *d)) + bi.c(fnp) * (bi(p, -1) + bi(p, -1, *d))))
        dpxy = dpxy + (b.m(mx) / b.m(my)) * rd * (b(php) - b(php, 0, *d)) * (
            b.c(rdnw) * (b(dpn, 1) - b(dpn)) - 0.5 * (b.m(mu, *d) + b.m(mu)))
        b.put(vel, b(vel) - dts * b(cq) * dpxy + mudf_xy)
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
    """Column mass and theta small-step update, eta-velocity recurrence.
    wrf_dynamics.py:3528-3638 (returns an 11-tuple of new tensors)."""
    k1 = kte - 1
    j0, j1, i0, i1 = max(jts, jds + 1), min(jte, jde - 2), max(its, ids + 1), min(ite, ide - 2)
    ww, t, t_ave = ww_ori.clone(), t_ori.clone(), t_ave_ori.clone()
    muave, muts, mudf, mu = muave_ori.clone(), muts_ori.clone(), mudf_ori.clone(), mu_ori.clone()
    dvdxi, dmdt_k, wdtn = zeros_like3(u), zeros_like3(u), zeros_like3(u)
    b = Box(kts, k1, j0, j1, i0, i1)
    b.put(dvdxi, b.m(msftx) * b.m(msfty) * (
        rdy * ((b(v, 0, 1) + b.m(muv, 1) * b(v_1, 0, 1) * b.m(msfvx_inv, 1)) - (b(v) + b.m(muv) * b(v_1) * b.m(msfvx_inv)))
        + rdx * ((b(u, 0, 0, 1) + b.m(muu, 0, 1) * b(u_1, 0, 0, 1) / b.m(msfuy, 0, 1))
                 - (b(u) + b.m(muu) * b(u_1) / b.m(msfuy)))))
    b.put(dmdt_k, b.c(dnw) * b(dvdxi))
    dmdt = dmdt_k.sum(dim=0)
    h = (slice(j0, j1), slice(i0, i1))
    with torch.no_grad():
        muave[h] = mu[h] + 0.0
        mu[h] = mu[h] + dts * (dmdt[h] + mu_tend[h])
        mudf[h] = (dmdt[h] + mu_tend[h])
        muts[h] = mut[h] + mu[h]
        muave[h] = 0.5 * ((1. + epssm) * mu[h] + (1. - epssm) * muave[h])
    for k in range(1, k1):
        ww[(k,) + h] = ww[(k - 1,) + h] - dnw[k - 1] * (dmdt[h] + dvdxi[(k - 1,) + h] + mu_tend[h]) / msfty[h]
    ww[(slice(0, k1),) + h] = ww[(slice(0, k1),) + h] - ww_1[(slice(0, k1),) + h]
    b = Box(0, k1, j0, j1, i0, i1)
    b.put(t_ave, b(t) + 0.0)
    b.put(t, b(t) + b.m(msfty) * dts * b(ft))
    wdtn[(0,) + h] = 0.
    wdtn[(kde - 1,) + h] = 0.
    bi = Box(1, k1, j0, j1, i0, i1)
    bi.put(wdtn, bi(ww) * (bi.c(fnm) * bi(t_1) + bi.c(fnp) * bi(t_1, -1)))
    b.put(t, b(t) - dts * b.m(msfty) * (b.m(msftx) * (
        0.5 * rdy * (b(v, 0, 1) * (b(t_1, 0, 1) + b(t_1)) - b(v) * (b(t_1) + b(t_1, 0, -1)))
        + 0.5 * rdx * (b(u, 0, 0, 1) * (b(t_1, 0, 0, 1) + b(t_1)) - b(u) * (b(t_1) + b(t_1, 0, 0, -1))))
        + b.c(rdnw) * (b(wdtn, 1) - b(wdtn))))
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
    """Implicit w / phi small step: RHS, surface w, explicit predictor,
    Thomas sweeps, upper Rayleigh damping, phi update.
    wrf_dynamics.py:3641-3796 (returns new t_2ave, w, ph)."""
    k1 = kte - 1
    j0, j1, i0, i1 = max(jts, jds + 1), min(jte, jde - 2), max(its, ids + 1), min(ite, ide - 2)
    pi_ = 3.14159265359
    dampmag = dts * 0.2
    hdepth = 5000.
    t_2ave, w, ph = t_2ave_ori.clone(), w_ori.clone(), ph_ori.clone()
    h = (slice(j0, j1), slice(i0, i1))
    rhs, wdwn = zeros_like3(w), zeros_like3(w)
    msft_inv = torch.zeros(msfty.shape, dtype=w.dtype)
    msft_inv[h] = 1. / msfty[h]
    b = Box(0, k1, j0, j1, i0, i1)
    b.put(t_2ave, 0.5 * ((1. + epssm) * b(t_2) + (1. - epssm) * b(t_2ave)))
    b.put(t_2ave, (b(t_2ave) + b.m(muave) * t0) / (b.m(muts) * (t0 + b(t_1))))
    bw = Box(1, k1 + 1, j0, j1, i0, i1)
    bw.put(wdwn, 0.5 * (bw(ww) + bw(ww, -1)) * bw.c(rdnw, -1) * (bw(ph_1) - bw(ph_1, -1) + bw(phb) - bw(phb, -1)))
    bw.put(rhs, dts * (bw(ph_tend) + 0.5 * g * (1. - epssm) * bw(w)))
    bi = Box(1, k1, j0, j1, i0, i1)
    bi.put(rhs, bi(rhs) - dts * (bi.c(fnm) * bi(wdwn, 1) + bi.c(fnp) * bi(wdwn)))
    bw.put(rhs, bw(ph) + bw.m(msfty) * bw(rhs) / bw.m(mut))
    s = Box(0, 1, j0, j1, i0, i1)
    w[(0,) + h] = (msfty[h] * 0.5 * rdy * (
        (s.m(ht, 1) - s.m(ht)) * (cf1 * v[(0,) + _sh(h, 1, 0)] + cf2 * v[(1,) + _sh(h, 1, 0)] + cf3 * v[(2,) + _sh(h, 1, 0)])
        + (s.m(ht) - s.m(ht, -1)) * (cf1 * v[(0,) + h] + cf2 * v[(1,) + h] + cf3 * v[(2,) + h]))
        + msftx[h] * 0.5 * rdx * (
        (s.m(ht, 0, 1) - s.m(ht)) * (cf1 * u[(0,) + _sh(h, 0, 1)] + cf2 * u[(1,) + _sh(h, 0, 1)] + cf3 * u[(2,) + _sh(h, 0, 1)])
        + (s.m(ht) - s.m(ht, 0, -1)) * (cf1 * u[(0,) + h] + cf2 * u[(1,) + h] + cf3 * u[(2,) + h])))
    mi = lambda bx: bx.m(msft_inv)
    bi.put(w, bi(w) + dts * bi(rw_tend) + mi(bi) * bi(cqw) * (0.5 * dts * g * bi.c(rdn) * (
        bi(c2a) * bi.c(rdnw) / bi.m(mut) * ((1. + epssm) * (bi(rhs, 1) - bi(rhs)) + (1. - epssm) * (bi(ph, 1) - bi(ph)))
        - bi(c2a, -1) * bi.c(rdnw, -1) / bi.m(mut) * ((1. + epssm) * (bi(rhs) - bi(rhs, -1)) + (1. - epssm) * (bi(ph) - bi(ph, -1)))))
        + dts * g * mi(bi) * (bi.c(rdn) * (bi(c2a) * bi(alt) * bi(t_2ave) - bi(c2a, -1) * bi(alt, -1) * bi(t_2ave, -1))
                              - bi.m(muave)))
    bt = Box(k1, k1 + 1, j0, j1, i0, i1)
    bt.put(w, bt(w) + dts * bt(rw_tend) + mi(bt) * (
        -0.5 * dts * g / bt.m(mut) * bt.c(rdnw, -1) ** 2 * 2. * bt(c2a, -1) * (
            (1. + epssm) * (bt(rhs) - bt(rhs, -1)) + (1. - epssm) * (bt(ph) - bt(ph, -1)))
        - dts * g * (2. * bt.c(rdnw, -1) * bt(c2a, -1) * bt(alt, -1) * bt(t_2ave, -1) + bt.m(muave))))
    for k in range(1, k1 + 1):
        w[(k,) + h] = (w[(k,) + h] - a[(k,) + h] * w[(k - 1,) + h]) * alpha[(k,) + h]
    for k in range(k1 - 1, 0, -1):
        w[(k,) + h] = w[(k,) + h] - gamma[(k,) + h] * w[(k + 1,) + h]
    htop = torch.zeros(msfty.shape, dtype=w.dtype)
    hbot = torch.zeros(msfty.shape, dtype=w.dtype)
    dampwt, hk = zeros_like3(w), zeros_like3(w)
    htop[h] = (ph_1[(k1,) + h] + phb[(k1,) + h]) / g
    hbot[h] = htop[h] - hdepth
    bw.put(hk, (bw(ph_1) + bw(phb)) / g)
    cond = hk >= hbot
    bw.put(dampwt, dampmag * (torch.sin(0.5 * pi_ * (bw(hk) - bw.m(hbot)) / hdepth)) ** 2)
    dampwt = torch.where(cond, dampwt, torch.zeros((), dtype=w.dtype))
    bw.put(w, (bw(w) - bw(dampwt) * bw.m(mut) * bw(w_save)) / (1. + bw(dampwt)))
    bw.put(ph, bw(rhs) + bw.m(msfty) * 0.5 * dts * g * (1. + epssm) * bw(w) / bw.m(muts))
    return t_2ave, w, ph


def _sh(h, dj, di):
    return (slice(h[0].start + dj, h[0].stop + dj), slice(h[1].start + di, h[1].stop + di))


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
    """Time-averaged mass fluxes for scalar transport.  wrf_dynamics.py:3799-3839."""
    if iteration == 1:
        for a in (ru_m, rv_m, ww_m):
            a[kts:kte, jts:jte, its:ite] = 0.
    mini, minj, mink = min(ide - 1, ite), min(jde - 1, jte), min(kde - 1, kte)
    bu, bv, bw = Box(kts, mink, jts, minj, its, ite), Box(kts, mink, jts, jte, its, mini), Box(kts, kte, jts, minj, its, mini)
    bu.put(ru_m, bu(ru_m) + bu(ru))
    bv.put(rv_m, bv(rv_m) + bv(rv))
    bw.put(ww_m, bw(ww_m) + bw(ww))
    if iteration == number_of_small_timesteps:
        n = number_of_small_timesteps
        bu.put(ru_m, bu(ru_m) / n + bu.m(muu) * bu(u_lin) / bu.m(msfuy))
        bv.put(rv_m, bv(rv_m) / n + bv.m(muv) * bv(v_lin) * bv.m(msfvx_inv))
        bw.put(ww_m, bw(ww_m) / n + bw(ww_lin))
    return ru_m, rv_m, ww_m


def calc_mu_uv_1(mu, muu, muv,
                 ids, ide, jds, jde, kds, kde,
                 ims, ime, jms, jme, kms, kme,
                 its, ite, jts, jte, kts, kte):
    """muus, muvs from the small-step muts.  wrf_dynamics.py:3844-3860."""
    jtf = min(jte, jde - 1)
    muu[jts:jtf, its + 1:ite - 1] = 0.5 * (mu[jts:jtf, its + 1:ite - 1] + mu[jts:jtf, its:ite - 2])
    muu[jts:jtf, its] = mu[jts:jtf, its]
    muu[jts:jtf, ite - 1] = mu[jts:jtf, ite - 2]
    itf = min(ite, ide - 1)
    muv[jts + 1:jte - 1, its:itf] = 0.5 * (mu[jts + 1:jte - 1, its:itf] + mu[jts:jte - 2, its:itf])
    muv[jts, its:itf] = mu[jts, its:itf]
    muv[jte - 1, its:itf] = mu[jte - 2, its:itf]
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
    """Back from perturbation to full variables.  wrf_dynamics.py:3863-3930
    (u, v, w, ph, t are new tensors; ww and mu_2 are updated in place)."""
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    u2b, v2b, w2b, t2b, ph2b = u_2.clone(), v_2.clone(), w_2.clone(), t_2.clone(), ph_2.clone()
    b = Box(kds, kde - 1, j0, jte, i0, i1)
    b.put(v2b, (b.m(msfvx) * b(v_2) + b(v_save) * b.m(muv)) / b.m(muvs))
    b = Box(kds, kde - 1, j0, j1, i0, ite)
    b.put(u2b, (b.m(msfuy) * b(u_2) + b(u_save) * b.m(muu)) / b.m(muus))
    b = Box(kds, kde, j0, j1, i0, i1)
    b.put(w2b, (b.m(msfty) * b(w_2) + b(w_save) * b.m(mut)) / b.m(muts))
    b.put(ph2b, b(ph_2) + b(ph_save))
    b.put(ww, b(ww) + b(ww1))
    b = Box(kds, kde - 1, j0, j1, i0, i1)
    if rk_step < rk_order:
        b.put(t2b, (b(t_2) + b(t_save) * b.m(mut)) / b.m(muts))
    else:
        b.put(t2b, (b(t_2) - dts * number_of_small_timesteps * b.m(mut) * b(h_diabatic)
                    + b(t_save) * b.m(mut)) / b.m(muts))
    mu_2[j0:j1, i0:i1] = mu_2[j0:j1, i0:i1] + mu_save[j0:j1, i0:i1]
    return u2b, v2b, w2b, ph2b, ww, t2b, mu_2
