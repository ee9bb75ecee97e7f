"""Oracle: WSM6 microphysics and the moist-physics wrappers (test
infrastructure only -- never imported by the product).

CPU restatement of pywrf/wrf_physics.py:27-1584 and 1988-2137.  Like the
reference, the sedimentation helpers work on the domain tile taken from the
array shape (halo 5: rows/columns [5, n-5), levels [0, nz)), while wsm6 itself
works on the tile it is called with (one row/column inside the mass interior).
The quirks that fix the reference's numbers are kept: the rain slope caps used
for snow and graupel, the 5.387e7 / 5.38e7 ice-number constants, paacw counted
twice, pgacs identically zero, and the piecewise-constant overlap remap with
the top source cell dropped and a 3800 m cap on the top target cell.
"""
import torch

from .constants import (P_QV, P_QC, P_QR, P_QI, P_QS, P_QG, g, Cp, Cpv, r_d, r_v, SVPT0, ep_1, ep_2,
                        epsilon, XLS, XLV, XLF, rhoair0, rhowater, cliq, cice, psat, rcp,
                        dtcldcr, n0r, avtr, qcrmin, dens, qs0, qc0, qck1, pvtr, pacrr, g6pbr, precr1,
                        precr2, roqimax, pvts, pvtg, precs1, precs2, pidn0r, pidn0s, pidn0g, xlv1,
                        pacrc, pi, precg1, precg2, n0s, n0g, n0smax, alpha, pfrz1, pfrz2, xncr,
                        dicon, dimax, bvtr, bvts, bvtg, rslopermax, rsloperbmax, rsloper2max,
                        rsloper3max)
from ._box import Box

HALO = 5


def _tile(a):
    """Global tile (jts:jte, its:ite) the reference's helpers take from globals."""
    return HALO, a.shape[-2] - HALO, HALO, a.shape[-1] - HALO


def moist_physics_prep_em(t_new, t_old, t0, rho, al, alb,
                          p, p8w, p0, pb, ph, phb,
                          th_phy, pii, pf,
                          z, z_at_w, dz8w,
                          dt, h_diabatic,
                          qv, qv_diabatic,
                          qc, qc_diabatic,
                          fzm, fzp,
                          ids, ide, jds, jde, kds, kde,
                          ims, ime, jms, jme, kms, kme,
                          its, ite, jts, jte, kts, kte):
    """Physics inputs for WSM6 on the physics tile.  wrf_physics.py:27-91."""
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    k0, k1 = kts, min(kte, kde - 1)
    h = (slice(j0, j1), slice(i0, i1))
    bw = Box(k0, kte, j0, j1, i0, i1)
    bw.put(z_at_w, (bw(ph) + bw(phb)) / g)
    bd = Box(k0, kte - 1, j0, j1, i0, i1)
    bd.put(dz8w, bd(z_at_w, 1) - bd(z_at_w))
    dz8w[(kte - 1,) + h] = 0.
    b = Box(k0, k1, j0, j1, i0, i1)
    b.put(th_phy, b(t_new) + t0)
    b.put(h_diabatic, b(th_phy) + 0.0)
    b.put(qv_diabatic, b(qv) + 0.0)
    b.put(qc_diabatic, b(qc) + 0.0)
    b.put(rho, 1. / (b(al) + b(alb)))
    b.put(pii, ((b(p) + b(pb)) / p0) ** rcp)
    b.put(z, 0.5 * (b(z_at_w) + b(z_at_w, 1)))
    b.put(pf, b(p) + b(pb))
    bi = Box(1, k1, j0, j1, i0, i1)
    bi.put(p8w, bi.c(fzm) * bi(pf) + bi.c(fzp) * bi(pf, -1))
    w1 = (z_at_w[(0,) + h] - z[(1,) + h]) / (z[(0,) + h] - z[(1,) + h])
    w2 = 1. - w1
    p8w[(0,) + h] = w1 * pf[(0,) + h] + w2 * pf[(1,) + h]
    w1 = (z_at_w[(kte - 1,) + h] - z[(k1 - 2,) + h]) / (z[(k1 - 1,) + h] - z[(k1 - 2,) + h])
    w2 = 1. - w1
    p8w[(kde - 1,) + h] = torch.exp(w1 * torch.log(pf[(kde - 2,) + h]) + w2 * torch.log(pf[(kde - 3,) + h]))
    return rho, th_phy, pii, pf, z, z_at_w, dz8w, p8w, h_diabatic, qv_diabatic, qc_diabatic, t_new, t_old


def microphysics_driver(th, rho, pi_phy, p, ht,
                        dz8w, p8w, dt, dx, dy,
                        qv_curr, qc_curr, qr_curr, qi_curr, qs_curr, qg_curr,
                        rainnc, rainncv, snownc, snowncv, sr,
                        refl_10cm, diagflag, do_radar_ref,
                        graupelnc, graupelncv,
                        has_reqc, has_reqi, has_reqs,
                        re_cloud, re_ice, re_snow,
                        ids, ide, jds, jde, kds, kde,
                        ims, ime, jms, jme, kms, kme,
                        its, ite, jts, jte, kts, kte):
    """Dispatch to WSM6.  wrf_physics.py:94-134."""
    return wsm6(th, qv_curr, qc_curr, qr_curr, qi_curr, qs_curr, qg_curr,
                rho, pi_phy, p, dz8w, dt, g, Cp, Cpv, r_d, r_v, SVPT0, ep_1, ep_2, epsilon, XLS, XLV, XLF,
                rhoair0, rhowater, cliq, cice, psat, rainnc, rainncv, snownc, snowncv, sr, refl_10cm,
                diagflag, do_radar_ref, graupelnc, graupelncv, has_reqc, has_reqi, has_reqs,
                re_cloud, re_ice, re_snow,
                ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte)


# --------------------------------------------------------------------------- helpers
def _slope3(q0, q1, q2, den, denfac, t):
    """slope_wsm6 on a box: returns rslope, rslopeb, rslope2, rslope3, vt as
    3-tuples (rain, snow, graupel).  wrf_physics.py:2048-2103 (all three
    species are capped with the rain maxima, as in the reference)."""
    supcol = t0c_ - t
    n0sfac = torch.maximum(torch.minimum(torch.exp(alpha * supcol), _c(n0smax / n0s, t)), _c(1., t))
    lam = (lambda x, y: (pidn0r / (x * y)) ** 0.25,
           lambda x, y: (pidn0s * n0sfac / (x * y)) ** 0.25,
           lambda x, y: (pidn0g / (x * y)) ** 0.25)
    bexp, pv = (bvtr, bvts, bvtg), (pvtr, pvts, pvtg)
    rs, rb, r2, r3, vt = [], [], [], [], []
    for n, q in enumerate((q0, q1, q2)):
        small = q <= qcrmin
        r = torch.where(small, _c(rslopermax, t), 1. / lam[n](q, den))
        rs.append(r)
        rb.append(torch.where(small, _c(rsloperbmax, t), r ** bexp[n]))
        r2.append(torch.where(small, _c(rsloper2max, t), r ** 2))
        r3.append(torch.where(small, _c(rsloper3max, t), r ** 3))
        v = pv[n] * rb[n] * denfac
        vt.append(torch.where(q <= 0.0, _c(0.0, t), v))
    return rs, rb, r2, r3, vt


def _c(x, like):
    return torch.tensor(x, dtype=like.dtype)


t0c_ = SVPT0  # t0c == SVPT0 in config_params.py:130-131


def _slope1(kind, qr, den, denfac, t):
    """slope_rain / slope_snow / slope_graup terminal velocity on a box.
    wrf_physics.py:1988-2046."""
    if kind == 'r':
        r = torch.where(qr <= qcrmin, _c(rslopermax, t), 1. / (pidn0r / (qr * den)) ** 0.25)
        rb = torch.where(qr <= qcrmin, _c(rsloperbmax, t), r ** bvtr)
        v = pvtr * rb * denfac
    elif kind == 's':
        supcol = t0c_ - t
        n0sfac = torch.maximum(torch.minimum(torch.exp(alpha * supcol), _c(n0smax / n0s, t)), _c(1., t))
        r = torch.where(qr <= qcrmin, _c(rslopermax, t), 1. / (pidn0s * n0sfac / (qr * den)) ** 0.25)
        rb = torch.where(qr <= qcrmin, _c(rsloperbmax, t), r ** bvts)
        v = pvts * rb * denfac
    else:
        r = torch.where(qr <= qcrmin, _c(rslopermax, t), 1. / (pidn0g / (qr * den)) ** 0.25)
        rb = torch.where(qr <= qcrmin, _c(rsloperbmax, t), r ** bvtg)
        v = pvtg * rb * denfac
    return torch.where(qr <= 0.0, _c(0.0, t), v)


def _sediment_setup(km, dz, ww, dt):
    """Shared head of the nislfv_* schemes: heights, arrival-interface
    velocities with the CFL limiter, departure heights.  Inputs are (km, ...)
    columns; returns zi, wi (km+1) after the limiter."""
    zi = torch.zeros((km + 1,) + dz.shape[1:], dtype=dz.dtype)
    for k in range(km):
        zi[k + 1] = zi[k] + dz[k]
    return zi


def _interface_w(km, dz, ww, dt, dzlim):
    wi = torch.zeros((km + 1,) + dz.shape[1:], dtype=dz.dtype)
    wi[0] = ww[0] + 0.0
    wi[km] = ww[km - 1] + 0.0
    for k in range(1, km):
        wi[k] = (ww[k] * dz[k - 1] + ww[k - 1] * dz[k]) / (dz[k - 1] + dz[k])
    fa1, fa2 = 9. / 16., 1. / 16.
    wi[0] = ww[0] + 0.0
    wi[1] = 0.5 * (ww[1] + ww[0])
    for k in range(2, km - 1):
        wi[k] = fa1 * (ww[k] + ww[k - 1]) - fa2 * (ww[k + 1] + ww[k - 2])
    wi[km - 1] = 0.5 * (ww[km - 1] + ww[km - 2])
    wi[km] = ww[km - 1] + 0.0
    wi[1:km] = torch.where(ww[1:km] == 0.0, ww[0:km - 1], wi[1:km])
    for k in range(km - 1, -1, -1):
        decfl = (wi[k + 1] - wi[k]) * dt / dz[k]
        wi[k] = torch.where(decfl > 0.05, wi[k + 1] - 0.05 * dzlim[k] / dt, wi[k])
    return wi


def _remap(km, zi, za, qa, dz):
    """Piecewise-constant interval-overlap remap onto the fixed grid
    (wrf_physics.py:1339-1358): the top source cell is dropped and both top
    cells are capped at 3800 m above their lower interface."""
    zi_high = torch.cat([zi[1:km], zi[km - 1:km] + 3800.], dim=0)
    za_high = torch.cat([za[1:km], za[km - 1:km] + 3800.], dim=0)
    min_high = torch.minimum(zi_high.unsqueeze(1), za_high.unsqueeze(0))
    max_low = torch.maximum(zi[0:km].unsqueeze(1), za[0:km].unsqueeze(0))
    overlap = torch.maximum(torch.tensor(0.0, dtype=zi.dtype), min_high - max_low)
    qa[km - 1] = 0.0
    qq_new = (qa[0:km].unsqueeze(0) * overlap).sum(dim=1)
    return qq_new[0:km] / dz[0:km]


def _precip(km, za, qa, dza, precip):
    for k in range(km):
        c1 = (za[k] < 0.0) & (za[k + 1] < 0.0)
        precip = torch.where(c1, precip + qa[k] * dza[k], precip)
        c3 = (za[k] < 0.0) & (za[k + 1] > 0.0)
        precip = torch.where(c3, precip + qa[k] * (0.0 - za[k]), precip)
    return precip


def nislfv_rain_plm(km, denl, denfacl, tkl, dzl, wwl, rql, precip, dt, id, iter):
    """Semi-Lagrangian sedimentation of one species.  wrf_physics.py:1216-1381.
    precip is accumulated in place on the tile and returned with rql."""
    j0, j1, i0, i1 = _tile(rql)
    nz = rql.shape[0]
    T = lambda a: a[0:nz, j0:j1, i0:i1]
    dz, qq, ww = T(dzl) + 0.0, T(rql) + 0.0, T(wwl) + 0.0
    den, denfac, tk = T(denl) + 0.0, T(denfacl) + 0.0, T(tkl) + 0.0
    allold = qq.sum(dim=0)
    zi = _sediment_setup(km, dz, ww, dt)
    wd = ww.clone()
    was = torch.zeros_like(ww)
    n = 0 if iter == 0 else 1
    while n <= iter + 1:
        wi = _interface_w(km, dz, ww, dt, dz)
        za = zi - wi * dt
        dza = torch.zeros_like(za)
        dza[0:km] = za[1:km + 1] - za[0:km]
        dza[km] = zi[km] - za[km]
        qa = torch.zeros_like(za)
        qa[0:km] = qq[0:km] * dz[0:km] / dza[0:km]
        qr = torch.zeros_like(qq)
        qr[0:km] = qa[0:km] / den[0:km]
        qa[km] = 0.0
        if n <= iter:
            wa = torch.zeros_like(ww)
            wa[0:km] = _slope1('r', qr[0:km], den[0:km], denfac[0:km], tk[0:km])
            if n >= 2:
                wa[0:km] = 0.5 * (wa[0:km] + was[0:km])
            ww = 0.5 * (wd + wa)
            was = wa + 0.0
        n = n + 1
        if iter == 0:
            n = 2
    qn = _remap(km, zi, za, qa, dz)
    pr = _precip(km, za, qa, dza, precip[j0:j1, i0:i1])
    cond = allold > 0.
    precip[j0:j1, i0:i1] = torch.where(cond, pr, _c(0., pr))
    rql[0:km, j0:j1, i0:i1] = torch.where(cond, qn, _c(0., qn))
    return precip, rql


def nislfv_rain_plm6(km, denl, denfacl, tkl, dzl, wwl, rql, rql2, precip1, precip2, dt, id, iter):
    """Joint sedimentation of snow and graupel with a mass-weighted fall
    speed.  wrf_physics.py:1383-1584 (precip1/precip2 are new tensors)."""
    j0, j1, i0, i1 = _tile(rql)
    nz = rql.shape[0]
    T = lambda a: a[0:nz, j0:j1, i0:i1]
    dz, qq, qq2, ww = T(dzl) + 0.0, T(rql) + 0.0, T(rql2) + 0.0, T(wwl) + 0.0
    den, denfac, tk = T(denl) + 0.0, T(denfacl) + 0.0, T(tkl) + 0.0
    allold = qq.sum(dim=0) + qq2.sum(dim=0)
    zi = _sediment_setup(km, dz, ww, dt)
    wd = ww.clone()
    was = torch.zeros_like(ww)
    n = 0 if iter == 0 else 1
    while n <= iter + 1:
        wi = _interface_w(km, dz, ww, dt, T(dzl))
        za = zi - wi * dt
        dza = torch.zeros_like(za)
        dza[0:km] = za[1:km + 1] - za[0:km]
        dza[km] = zi[km] - za[km]
        qa, qa2 = torch.zeros_like(za), torch.zeros_like(za)
        qr, qr2 = torch.zeros_like(qq), torch.zeros_like(qq)
        qa[0:km] = qq[0:km] * dz[0:km] / dza[0:km]
        qr[0:km] = qa[0:km] / den[0:km]
        qa2[0:km] = qq2[0:km] * dz[0:km] / dza[0:km]
        qr2[0:km] = qa2[0:km] / den[0:km]
        qa[km] = 0.0
        qa2[km] = 0.0
        if n <= iter:
            wa, wa2 = torch.zeros_like(ww), torch.zeros_like(ww)
            wa[0:km] = _slope1('s', qr[0:km], den[0:km], denfac[0:km], tk[0:km])
            wa2[0:km] = _slope1('g', qr2[0:km], den[0:km], denfac[0:km], tk[0:km])
            tmp = torch.maximum(qr + qr2, _c(1.e-15, qr))
            wa = torch.where(tmp > 1.e-15, (wa * qr + wa2 * qr2) / tmp, _c(0., qr))
            if n >= 2:
                wa[0:km] = 0.5 * (wa[0:km] + was[0:km])
            ww = 0.5 * (wd + wa)
            was = wa + 0.0
        n = n + 1
        if iter == 0:
            n = 2
    cond = allold > 0.
    outs = []
    for ist, src in enumerate((qa, qa2)):
        q = src.clone()
        qn = _remap(km, zi, za, q, dz)
        pr = _precip(km, za, q, dza, torch.zeros_like(allold))
        outs.append((torch.where(cond, pr, _c(0., pr)), torch.where(cond, qn, _c(0., qn))))
    precip1 = torch.zeros(rql.shape[-2:], dtype=rql.dtype)
    precip2 = torch.zeros(rql.shape[-2:], dtype=rql.dtype)
    precip1[j0:j1, i0:i1] = outs[0][0]
    rql[0:km, j0:j1, i0:i1] = outs[0][1]
    precip2[j0:j1, i0:i1] = outs[1][0]
    rql2[0:km, j0:j1, i0:i1] = outs[1][1]
    return precip1, precip2, rql, rql2


# --------------------------------------------------------------------------- WSM6
def wsm6(th, q, qc, qr, qi, qs, qg,
         den, pii, p, delz,
         delt, g_, cpd, cpv, rd, rv, t0c,
         ep1, ep2, qmin,
         xls, xlv0, xlf0, den0, denr,
         cliq_, cice_, psat_,
         rain, rainncv,
         snow, snowncv,
         sr,
         refl_10cm, diagflag, do_radar_ref,
         graupel, graupelncv,
         has_reqc, has_reqi, has_reqs,
         re_cloud, re_ice, re_snow,
         ids, ide, jds, jde, kds, kde,
         ims, ime, jms, jme, kms, kme,
         its, ite, jts, jte, kts, kte):
    """WSM6 single-moment six-class microphysics, one sub-cycle per call at
    delt <= dtcldcr.  wrf_physics.py:137-1214."""
    kte, ite, jte = kte - 1, ite - 1, jte - 1
    B = Box(kts, kte, jts, jte, its, ite)
    H = (slice(jts, jte), slice(its, ite))
    full = lambda: torch.zeros(th.shape, dtype=th.dtype)
    C = lambda x: _c(x, th)
    mx, mn, where = torch.maximum, torch.minimum, torch.where

    t = B(th) * B(pii)
    qci0, qci1 = mx(B(qc) + 0.0, C(0.)), mx(B(qi) + 0.0, C(0.))
    qrs0, qrs1, qrs2 = mx(B(qr) + 0.0, C(0.)), mx(B(qs) + 0.0, C(0.)), mx(B(qg) + 0.0, C(0.))
    qv = B(q)
    cpm = where(qv >= qmin, cpv * qv + (1 - qv) * cpd, C(cpd * (1 - qmin) + cpv * qmin))
    xl = xlv0 - xlv1 * (t - t0c)
    d, dz_, pp = B(den), B(delz), B(p)
    for a in (rainncv, snowncv, graupelncv, sr):
        a[H] = 0.0
    tstepsnow = torch.zeros(th.shape[-2:], dtype=th.dtype)
    tstepgraup = torch.zeros(th.shape[-2:], dtype=th.dtype)
    dtcld = delt + 0.0 if delt <= dtcldcr else 1.0 * delt / max(delt / dtcldcr, 1)
    kdim = kte - kts

    denfac = torch.sqrt(1.0 / d * den0)
    ttp = t0c + 0.01
    xa = -(cpv - cliq_) / rv
    xb = xa + xlv0 / (rv * ttp)
    xai = -(cpv - cice_) / rv
    xbi = xai + xls / (rv * ttp)

    def fpvs(t_):
        tr = ttp / t_
        s0 = psat_ * torch.exp(torch.log(tr) * xa) * torch.exp(xb * (1. - tr))
        s0 = mn(s0, 0.99 * pp)
        s0 = ep2 * s0 / (pp - s0)
        s0 = mx(s0, C(qmin))
        ice = psat_ * torch.exp(torch.log(tr) * xai) * torch.exp(xbi * (1. - tr))
        liq = psat_ * torch.exp(torch.log(tr) * xa) * torch.exp(xb * (1. - tr))
        s1 = where(t_ < ttp, ice, liq)
        s1 = mn(s1, 0.99 * pp)
        s1 = ep2 * s1 / (pp - s1)
        s1 = mx(s1, C(qmin))
        return s0, s1

    qs0_, qs1_ = fpvs(t)
    rh0 = mx(qv / qs0_, C(qmin))
    rh1 = mx(qv / qs1_, C(qmin))
    zeros = lambda: torch.zeros_like(t)
    temp = (d * mx(qci1, C(qmin))) ** 0.75
    xni = mn(mx(5.387e7 * temp, C(1.e3)), C(1.e6))
    rs, rb, r2, r3, vt = _slope3(qrs0, qrs1, qrs2, d, denfac, t)
    workr = vt[0]
    qsum = mx(qrs1 + qrs2, C(1.e-15))
    worka = where(qsum > 1.e-15, (vt[1] * qrs1 + vt[2] * qrs2) / qsum, C(0.))
    workr = where(qrs0 <= 0.0, C(0.0), workr)
    # sedimentation works on full-size arrays over the domain tile
    den_tmp, delz_tmp, denfac_f, t_f = full(), full(), full(), full()
    B.put(den_tmp, d)
    B.put(delz_tmp, dz_)
    B.put(denfac_f, denfac)
    B.put(t_f, t)
    F = lambda x: (lambda a: (B.put(a, x), a)[1])(full())
    denqrs1, denqrs2, denqrs3 = F(d * qrs0), F(d * qrs1), F(d * qrs2)
    delqrs1 = torch.zeros(th.shape[-2:], dtype=th.dtype)
    delqrs1, denqrs1 = nislfv_rain_plm(kdim, den_tmp, denfac_f, t_f, delz_tmp, F(workr), denqrs1, delqrs1, dtcld, 1, 1)
    delqrs2, delqrs3, denqrs2, denqrs3 = nislfv_rain_plm6(kdim, den_tmp, denfac_f, t_f, delz_tmp, F(worka),
                                                          denqrs2, denqrs3, None, None, dtcld, 1, 1)
    qrs0 = mx(B(denqrs1) / d, C(0))
    qrs1 = mx(B(denqrs2) / d, C(0))
    qrs2 = mx(B(denqrs3) / d, C(0))
    fall0, fall1, fall2 = B(denqrs1) * workr / dz_, B(denqrs2) * worka / dz_, B(denqrs3) * worka / dz_
    fall0[0] = delqrs1[H] / dz_[0] / dtcld
    fall1[0] = delqrs2[H] / dz_[0] / dtcld
    fall2[0] = delqrs3[H] / dz_[0] / dtcld
    rs, rb, r2, r3, vt = _slope3(qrs0, qrs1, qrs2, d, denfac, t)

    xlf = xlf0 + 0.
    viscos = lambda x, y: 1.496e-6 * (x * torch.sqrt(x)) / (x + 120.) / y
    diffus = lambda x, y: 8.794e-5 * torch.exp(torch.log(x) * 1.81) / y
    xka = lambda x, y: 1.414e3 * viscos(x, y) * y
    diffac = lambda a, b, c, dd, e: dd * a * a / (xka(c, dd) * rv * c * c) + 1. / (e * diffus(c, b))
    venfac = lambda a, b, c: (torch.exp(torch.log((viscos(b, c) / diffus(b, a))) * 0.3333333)
                              / torch.sqrt(viscos(b, c)) * torch.sqrt(torch.sqrt(den0 / c)))
    supcol = t0c - t
    n0sfac = mx(mn(torch.exp(alpha * supcol), C(n0smax / n0s)), C(1))
    # psmlt / pgmlt
    cond = t > t0c
    work2 = where(cond, venfac(pp, t, d), zeros())
    c2 = (qrs1 > 0.) & cond
    coeres = where(c2, r2[1] * (rs[1] * rb[1]) ** 0.5, zeros())
    psmlt = where(c2, xka(t, d) / xlf * (t0c - t) * pi / 2. * n0sfac * (precs1 * r2[1] + precs2 * work2 * coeres) / d, zeros())
    psmlt = where(c2, mn(mx(psmlt * dtcld / 1., -qrs1 / 1.), C(0)), psmlt)
    qrs1 = where(c2, qrs1 + psmlt, qrs1)
    qrs0 = where(c2, qrs0 - psmlt, qrs0)
    t = where(c2, t + xlf / cpm * psmlt, t)
    c2 = (qrs2 > 0.) & cond
    coeres = where(c2, r2[2] * (rs[2] * rb[2]) ** 0.5, zeros())
    pgmlt = where(c2, xka(t, d) / xlf * (t0c - t) * (precg1 * r2[2] + precg2 * work2 * coeres) / d, zeros())
    pgmlt = where(c2, mn(mx(pgmlt * dtcld / 1., -qrs2 / 1.), C(0)), pgmlt)
    qrs2 = where(c2, qrs2 + pgmlt, qrs2)
    qrs0 = where(c2, qrs0 - pgmlt, qrs0)
    t = where(c2, t + xlf / cpm * pgmlt, t)
    # ice fall
    cond = qci1 <= 0.
    wt = mx(mn(dicon * (d * qci1 / xni) ** 0.5, C(dimax)), C(1.e-25))
    wt = 1.49e4 * torch.exp(torch.log(wt) * 1.31)
    work1c = where(cond, C(0), wt)
    B.put(t_f, t)
    delqi = torch.zeros(th.shape[-2:], dtype=th.dtype)
    delqi, denqci = nislfv_rain_plm(kdim, den_tmp, denfac_f, t_f, delz_tmp, F(work1c), F(d * qci1), delqi, dtcld, 1, 0)
    qci1 = mx(B(denqci) / d, C(0))
    fallc0 = delqi[H] / dz_[0] / dtcld
    fallsum = fall0[0] + fall1[0] + fall2[0] + fallc0
    fallsum_qsi = fall1[0] + fallc0
    fallsum_qg = fall2[0] + 0.0
    dz0 = dz_[0]
    acc = lambda fs: fs * dz0 / denr * dtcld * 1000.
    cnd = fallsum > 0.
    rainncv[H] = where(cnd, acc(fallsum) + rainncv[H], rainncv[H])
    rain[H] = where(cnd, acc(fallsum) + rain[H], rain[H])
    cnd = fallsum_qsi > 0.
    tstepsnow[H] = where(cnd, acc(fallsum_qsi) + tstepsnow[H], tstepsnow[H])
    snowncv[H] = where(cnd, acc(fallsum_qsi) + snowncv[H], snowncv[H])
    snow[H] = where(cnd, acc(fallsum_qsi) + snow[H], snow[H])
    cnd = fallsum_qg > 0.
    tstepgraup[H] = where(cnd, acc(fallsum_qg) + tstepgraup[H], tstepgraup[H])
    graupelncv[H] = where(cnd, acc(fallsum_qg) + graupelncv[H], graupelncv[H])
    graupel[H] = where(cnd, acc(fallsum_qg) + graupel[H], graupel[H])
    cnd = fallsum > 0.
    sr[H] = where(cnd, (snowncv[H] + graupelncv[H]) / (rainncv[H] + 1.e-12), sr[H])
    # instantaneous melting / freezing
    supcol = t0c - t
    xlf3 = xls - xl
    c = (supcol < 0.) & (qci1 > 0.)
    qci0 = where(c, qci0 + qci1, qci0)
    t = where(c, t - xlf3 / cpm * qci1, t)
    qci1 = where(c, C(0), qci1)
    c = (supcol > 40.) & (qci0 > 0.)
    qci1 = where(c, qci1 + qci0, qci1)
    t = where(c, t + xlf3 / cpm * qci0, t)
    qci0 = where(c, C(0), qci0)
    c = (supcol > 0.) & (qci0 > qmin)
    pfrzdtc = mn(pfrz1 * (torch.exp(pfrz2 * mn(supcol, C(50.))) - 1.) * d / denr / xncr * qci0 * qci0 * dtcld, qci0)
    qci1 = where(c, qci1 + pfrzdtc, qci1)
    t = where(c, t + xlf3 / cpm * pfrzdtc, t)
    qci0 = where(c, qci0 - pfrzdtc, qci0)
    c = (supcol > 0.) & (qrs0 > 0.)
    pfrzdtr = mn(20. * pi * pi * pfrz1 * n0r * denr / d * (torch.exp(pfrz2 * mn(supcol, C(50.))) - 1.)
                 * r3[0] * r3[0] * rs[0] * dtcld, qrs0)
    qrs2 = where(c, qrs2 + pfrzdtr, qrs2)
    t = where(c, t + xlf3 / cpm * pfrzdtr, t)
    qrs0 = where(c, qrs0 - pfrzdtr, qrs0)
    rs, rb, r2, r3, vt = _slope3(qrs0, qrs1, qrs2, d, denfac, t)
    work1_0 = diffac(xl, pp, t, d, qs0_)
    work1_1 = diffac(xls, pp, t, d, qs1_)
    work2 = venfac(pp, t, d)
    # warm rain
    supsat = mx(qv, C(qmin)) - qs0_
    praut = qck1 * qci0 ** (7. / 3.)
    praut = where(qci0 > qc0, mn(praut, qci0 / dtcld), C(0))
    c = (qrs0 > qcrmin) & (qci0 > qmin)
    pracw = where(c, mn(pacrr * r3[0] * rb[0] * qci0 * denfac, qci0 / dtcld), zeros())
    c = qrs0 > 0.
    coeres = r2[0] * (rs[0] * rb[0]) ** 0.5
    prevp = where(c, (rh0 - 1.) * (precr1 * r2[0] + precr2 * work2 * coeres) / work1_0, zeros())
    c2, c3 = (prevp < 0.) & c, (prevp >= 0.) & c
    prevp = where(c2, mx(prevp, -qrs0 / dtcld), prevp)
    prevp = where(c2, mx(prevp, supsat / dtcld / 2.), prevp)
    prevp = where(c3, mn(prevp, supsat / dtcld / 2.), prevp)
    # cold rain
    supcol = t0c - t
    n0sfac = mx(mn(torch.exp(alpha * supcol), C(n0smax / n0s)), C(1.))
    supsat = mx(qv, C(qmin)) - qs1_
    satdt = supsat / dtcld
    ifsat = torch.zeros(t.shape, dtype=torch.int32)
    xni = mn(mx(5.38e7 * (d * mx(qci1, C(qmin))) ** 0.75, C(1.e3)), C(1.e6))
    eacrs = torch.exp(0.07 * (-supcol))
    xmi = d * qci1 / xni
    diameter = mn(dicon * xmi ** 0.5, C(dimax))
    vt2i = 1.49e4 * diameter ** 1.31
    vt2r, vt2s, vt2g = pvtr * rb[0] * denfac, pvts * rb[1] * denfac, pvtg * rb[2] * denfac
    qsum = mx(qrs1 + qrs2, C(1.e-15))
    vt2ave = where(qsum > 1.e-15, (vt2s * qrs1 + vt2g * qrs2) / qsum, C(0))
    cc = (supcol > 0.) & (qci1 > qmin)
    c1 = cc & (qrs0 > qcrmin)
    acr = 2. * r3[0] + 2. * diameter * r2[0] + diameter ** 2 * rs[0]
    praci = pi * qci1 * n0r * torch.abs(vt2r - vt2i) * acr
    praci = where(c1, mn(praci, qci1 / dtcld), C(0))
    piacr = pi ** 2 * avtr * n0r * denr * xni * denfac * g6pbr * r3[0] ** 2 * rb[0] / 24. / d
    piacr = where(c1, mn(piacr, qrs0 / dtcld), C(0))
    c1 = (qrs1 > qcrmin) & cc
    acr = 2. * r3[1] + 2. * diameter * r2[1] + diameter ** 2 * rs[1]
    psaci = pi * qci1 * eacrs * n0s * n0sfac * torch.abs(vt2ave - vt2i) * acr / 4.
    psaci = where(c1, mn(psaci, qci1 / dtcld), C(0))
    c1 = (qrs2 > qcrmin) & cc
    egi = torch.exp(-0.07 * supcol)
    acr = 2. * r3[2] + 2. * diameter * r2[2] + diameter ** 2 * rs[2]
    pgaci = pi * egi * qci1 * n0g * torch.abs(vt2ave - vt2i) * acr / 4.
    pgaci = where(c1, torch.min(pgaci, qci1 / dtcld), C(0))
    c = (qrs1 > qcrmin) & (qci0 > qmin)
    psacw = where(c, mn(pacrc * n0sfac * r3[1] * rb[1] * qci0 * denfac, qci0 / dtcld), C(0))
    c = (qrs2 > qcrmin) & (qci0 > qmin)
    pgacw = where(c, mn(pacrc * r3[2] * rb[2] * qci0 * denfac, qci0 / dtcld), C(0))
    paacw = where(qsum > 1.e-15, (qrs1 * psacw + qrs2 * pgacw) / qsum, C(0))
    c = (qrs1 > qcrmin) & (qrs0 > qcrmin)
    c1 = (supcol > 0) & c
    acr = 5. * r3[1] ** 2 * rs[0] + 2. * r3[1] * r2[1] * r2[0] + 0.5 * r2[1] ** 2 * r3[0]
    pracs = pi ** 2 * n0r * n0s * n0sfac * torch.abs(vt2r - vt2ave) * (dens / d) * acr
    pracs = where(c1, mn(pracs, qrs1 / dtcld), C(0))
    acr = 5. * r3[0] ** 2 * rs[1] + 2. * r3[0] * r2[0] * r2[1] + 0.5 * r2[0] ** 2 * r3[1]
    psacr = pi ** 2 * n0r * n0s * n0sfac * torch.abs(vt2r - vt2ave) * (denr / d) * acr
    psacr = where(c, mn(psacr, qrs0 / dtcld), C(0))
    c = (qrs2 > qcrmin) & (qrs0 > qcrmin)
    acr = 5. * r3[0] ** 2 * rs[2] + 2. * r3[0] * r2[0] * r2[2] + 0.5 * r2[0] ** 2 * r3[2]
    pgacr = pi ** 2 * n0r * n0g * torch.abs(vt2r - vt2ave) * (denr / d) * acr
    pgacr = where(c, mn(pgacr, qrs0 / dtcld), C(0))
    pgacs = zeros()
    cc = supcol <= 0.
    c1 = cc & (qrs1 > 0.)
    pseml = where(c1, mn(mx(cliq_ * supcol * (paacw + psacr) / xlf0, -qrs1 / dtcld), C(0.)), C(0))
    c1 = cc & (qrs2 > 0.)
    pgeml = where(c1, mn(mx(cliq_ * supcol * (paacw + psacr) / xlf0, -qrs2 / dtcld), C(0.)), C(0))
    cc = supcol > 0.
    # deposition / sublimation with the ifsat chain
    c1 = (qci1 > 0.) & (ifsat != 1) & cc
    pidep = where(c1, 4. * diameter * xni * (rh1 - 1.) / work1_1, C(0))
    supice = where(c1, satdt - prevp, C(0))
    c2, c3 = (pidep < 0.) & c1, (pidep >= 0.) & c1
    pidep = where(c2, mx(mx(mx(pidep, satdt / 2), supice), -qci1 / dtcld), pidep)
    pidep = where(c3, mn(mn(pidep, satdt / 2), supice), pidep)
    ifsat = where((torch.abs(prevp + pidep) >= torch.abs(satdt)) & c1, torch.tensor(1, dtype=torch.int32), ifsat)
    c1 = (qrs1 > 0.) & (ifsat != 1) & cc
    coeres = r2[1] * (rs[1] * rb[1]) ** 0.5
    psdep = where(c1, (rh1 - 1.) * n0sfac * (precs1 * r2[1] + precs2 * work2 * coeres) / work1_1, C(0))
    supice = where(c1, satdt - prevp - pidep, C(0))
    c2, c3 = (psdep < 0.) & c1, (psdep >= 0.) & c1
    psdep = where(c2, mx(mx(mx(psdep, -qrs1 / dtcld), satdt / 2), supice), psdep)
    psdep = where(c3, mn(mn(psdep, satdt / 2), supice), psdep)
    ifsat = where((torch.abs(prevp + pidep + psdep) >= torch.abs(satdt)) & c1, torch.tensor(1, dtype=torch.int32), ifsat)
    c1 = (qrs2 > 0.) & (ifsat != 1) & cc
    coeres = r2[2] * (rs[2] * rb[2]) ** 0.5
    pgdep = where(c1, (rh1 - 1.) * (precg1 * r2[2] + precg2 * work2 * coeres) / work1_1, C(0))
    supice = where(c1, satdt - prevp - pidep - psdep, C(0))
    c2, c3 = (pgdep < 0.) & c1, (pgdep >= 0.) & c1
    pgdep = where(c2, mx(mx(mx(pgdep, -qrs2 / dtcld), satdt / 2), supice), pgdep)
    pgdep = where(c3, mn(mn(pgdep, satdt / 2), supice), pgdep)
    ifsat = where((torch.abs(prevp + pidep + psdep + pgdep) >= torch.abs(satdt)) & c1,
                  torch.tensor(1, dtype=torch.int32), ifsat)
    c1 = (supsat > 0.) & (ifsat != 1) & cc
    supice = satdt - prevp - pidep - psdep - pgdep
    roqi0 = 4.92e-11 * (1.e3 * torch.exp(0.1 * supcol)) ** 1.33
    pigen = mx(C(0.), (roqi0 / d - mx(qci1, C(0.))) / dtcld)
    pigen = where(c1, mn(mn(pigen, satdt), supice), C(0.))
    c1 = (qci1 > 0.) & cc
    psaut = where(c1, mx(C(0.), (qci1 - roqimax / d) / dtcld), C(0.))
    c1 = (qrs1 > 0.) & cc
    pgaut = where(c1, mn(mx(C(0.), 1.e-3 * torch.exp(-0.09 * supcol) * (qrs1 - qs0)), qrs1 / dtcld), C(0.))
    cc = supcol < 0.
    c1 = (qrs1 > 0.) & (rh0 < 1.) & cc
    coeres = r2[1] * (rs[1] * rb[1]) ** 0.5
    psevp = (rh0 - 1.) * n0sfac * (precs1 * r2[1] + precs2 * work2 * coeres) / work1_0
    psevp = where(c1, mn(mx(psevp, -qrs1 / dtcld), C(0.)), C(0.))
    c1 = (qrs2 > 0.) & (rh0 < 1.) & cc
    coeres = r2[2] * (rs[2] * rb[2]) ** 0.5
    pgevp = (rh0 - 1.) * (precg1 * r2[2] + precg2 * work2 * coeres) / work1_0
    pgevp = where(c1, mn(mx(pgevp, -qrs2 / dtcld), C(0.)), C(0.))
    # mass-conservation limiters and update, T <= T0
    delta2 = where((qrs0 < 1.e-4) & (qrs1 < 1.e-4), C(1.), C(0.))
    delta3 = where(qrs0 < 1.e-4, C(1.), C(0.))
    cc = t <= t0c

    def limit(value, source, names, rates, cond):
        c1 = (source > value) & cond
        f = value / source
        return [where(c1, r * f, r) for r in rates]

    praut, pracw, paacw = limit(mx(C(qmin), qci0), (praut + pracw + paacw + paacw) * dtcld, None,
                                [praut, pracw, paacw], cc)
    psaut, pigen, pidep, praci, psaci, pgaci = limit(
        mx(C(qmin), qci1), (psaut - pigen - pidep + praci + psaci + pgaci) * dtcld, None,
        [psaut, pigen, pidep, praci, psaci, pgaci], cc)
    praut, prevp, pracw, piacr, psacr, pgacr = limit(
        mx(C(qmin), qrs0), (-praut - prevp - pracw + piacr + psacr + pgacr) * dtcld, None,
        [praut, prevp, pracw, piacr, psacr, pgacr], cc)
    psdep, psaut, pgaut, paacw, piacr, praci, psaci, pracs, psacr, pgacs = limit(
        mx(C(qmin), qrs1), -(psdep + psaut - pgaut + paacw + piacr * delta3 + praci * delta3
                            - pracs * (1. - delta2) + psacr * delta2 + psaci - pgacs) * dtcld, None,
        [psdep, psaut, pgaut, paacw, piacr, praci, psaci, pracs, psacr, pgacs], cc)
    pgdep, pgaut, piacr, praci, psacr, pracs, paacw, pgaci, pgacr, pgacs = limit(
        mx(C(qmin), qrs2), -(pgdep + pgaut + piacr * (1. - delta3) + praci * (1. - delta3) + psacr * (1. - delta2)
                            + pracs * (1. - delta2) + pgaci + paacw + pgacr + pgacs) * dtcld, None,
        [pgdep, pgaut, piacr, praci, psacr, pracs, paacw, pgaci, pgacr, pgacs], cc)
    work2 = where(cc, -(prevp + psdep + pgdep + pigen + pidep), work2)
    qv_new = where(cc, qv + work2 * dtcld, qv)
    qci0 = where(cc, mx(qci0 - (praut + pracw + paacw + paacw) * dtcld, C(0.)), qci0)
    qrs0 = where(cc, mx(qrs0 + (praut + pracw + prevp - piacr - pgacr - psacr) * dtcld, C(0.)), qrs0)
    qci1 = where(cc, mx(qci1 - (psaut + praci + psaci + pgaci - pigen - pidep) * dtcld, C(0.)), qci1)
    qrs1 = where(cc, mx(qrs1 + (psdep + psaut + paacw - pgaut + piacr * delta3 + praci * delta3 + psaci - pgacs
                               - pracs * (1. - delta2) + psacr * delta2) * dtcld, C(0.)), qrs1)
    qrs2 = where(cc, mx(qrs2 + (pgdep + pgaut + piacr * (1. - delta3) + praci * (1. - delta3)
                               + psacr * (1. - delta2) + pracs * (1. - delta2) + pgaci + paacw + pgacr + pgacs) * dtcld,
                        C(0.)), qrs2)
    xlf3 = xls - xl
    xlwork2 = (-xls * (psdep + pgdep + pidep + pigen) - xl * prevp
               - xlf3 * (piacr + paacw + paacw + pgacr + psacr))
    t = where(cc, t - xlwork2 / cpm * dtcld, t)
    # T > T0
    cc = t > t0c
    praut, pracw, paacw = limit(mx(C(qmin), qci0), (praut + pracw + paacw + paacw) * dtcld, None,
                                [praut, pracw, paacw], cc)
    praut, prevp, pracw, paacw, pseml, pgeml = limit(
        mx(C(qmin), qrs0), (-paacw - praut + pseml + pgeml - pracw - paacw - prevp) * dtcld, None,
        [praut, prevp, pracw, paacw, pseml, pgeml], cc)
    pgacs, psevp, pseml = limit(mx(C(qcrmin), qrs1), (pgacs - pseml - psevp) * dtcld, None,
                                [pgacs, psevp, pseml], cc)
    pgacs, pgevp, pgeml = limit(mx(C(qcrmin), qrs2), -(pgacs + pgevp + pgeml) * dtcld, None,
                                [pgacs, pgevp, pgeml], cc)
    work2 = where(cc, -(prevp + psevp + pgevp), work2)
    qv_new = where(cc, qv_new + work2 * dtcld, qv_new)
    qci0 = where(cc, mx(qci0 - (praut + pracw + paacw + paacw) * dtcld, C(0.)), qci0)
    qrs0 = where(cc, mx(qrs0 + (praut + pracw + prevp + paacw + paacw - pseml - pgeml) * dtcld, C(0.)), qrs0)
    qrs1 = where(cc, mx(qrs1 + (psevp - pgacs + pseml) * dtcld, C(0.)), qrs1)
    qrs2 = where(cc, mx(qrs2 + (pgacs + pgevp + pgeml) * dtcld, C(0.)), qrs2)
    xlf3 = xls - xl
    xlwork2 = -xl * (prevp + psevp + pgevp) - xlf3 * (pseml + pgeml)
    t = where(cc, t - xlwork2 / cpm * dtcld, t)
    # saturation adjustment
    qs0_, qs1_ = fpvs(t)
    w1 = (mx(qv_new, C(qmin)) - qs0_) / (1. + xl * xl / (rv * cpm) * qs0_ / (t * t))
    pcond = mn(mx(w1 / dtcld, C(0.)), mx(qv_new, C(0.)) / dtcld)
    c = (qci0 > 0.) & (w1 < 0.)
    pcond = where(c, mx(w1, -qci0) / dtcld, pcond)
    qv_new = qv_new - pcond * dtcld
    qci0 = mx(qci0 + pcond * dtcld, C(0.))
    t = t + pcond * xl / cpm * dtcld
    qci0 = where(qci0 <= qmin, C(0.), qci0)
    qci1 = where(qci1 <= qmin, C(0.), qci1)
    B.put(q, qv_new)
    B.put(th, t / B(pii))
    B.put(qc, qci0)
    B.put(qi, qci1)
    B.put(qr, qrs0)
    B.put(qs, qrs1)
    B.put(qg, qrs2)
    return th, q, qc, qr, qi, qs, qg, rain, rainncv, snow, snowncv, graupel, graupelncv, sr, refl_10cm


def moist_physics_finish_em(t_new, t_old, t0, mut,
                            th_phy, h_diabatic, dt,
                            qv, qv_diabatic,
                            qc, qc_diabatic,
                            ids, ide, jds, jde, kds, kde,
                            ims, ime, jms, jme, kms, kme,
                            its, ite, jts, jte, kts, kte):
    """theta increment from microphysics and diabatic rates.
    wrf_physics.py:2106-2137."""
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    mpten = b(th_phy) - b(h_diabatic)
    qvten = b(qv) - b(qv_diabatic)
    qcten = b(qc) - b(qc_diabatic)
    b.put(t_new, b(t_new) + mpten)
    b.put(h_diabatic, mpten / dt)
    b.put(qv_diabatic, qvten / dt)
    b.put(qc_diabatic, qcten / dt)
    return t_new, t_old, th_phy, h_diabatic, qv_diabatic, qc_diabatic
