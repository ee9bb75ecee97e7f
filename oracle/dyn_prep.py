"""Oracle: RK-stage preparation, physical BCs, physics prep and Smagorinsky
coefficients (test infrastructure only -- never imported by the product).

CPU restatement of pywrf/wrf_dynamics.py:30-1076 with the reference's
positional signatures, index boxes and left-to-right expression order, so that
on the same inputs it reproduces the reference bit for bit on the CPU.
"""
import torch

from .constants import (P_QV, P_QC, P_QI, t0, g, rcp, p1000mb, SVP1, SVP2, SVP3,
                        SVPT0, EP_2, XLV, r_d, Cp, r_v, prandtl, c_s)
from ._box import Box, zeros_like3


def calculate_full(mut, mub, mu,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """mut = mub + mu on the mass interior.  wrf_dynamics.py:30-38."""
    itf, jtf = min(ite, ide - 1), min(jte, jde - 1)
    mut[jts:jtf, its:itf] = mub[jts:jtf, its:itf] + mu[jts:jtf, its:itf]
    return mut


def calc_mu_uv(mu, mub, muu, muv,
               ids, ide, jds, jde, kds, kde,
               ims, ime, jms, jme, kms, kme,
               its, ite, jts, jte, kts, kte):
    """Dry column mass at u and v points.  wrf_dynamics.py:41-61."""
    jtf = min(jte, jde - 1)
    a, b = slice(its + 1, ite - 1), slice(its, ite - 2)
    muu[jts:jtf, a] = 0.5 * (mu[jts:jtf, a] + mu[jts:jtf, b] + mub[jts:jtf, a] + mub[jts:jtf, b])
    muu[jts:jtf, its] = mu[jts:jtf, its] + mub[jts:jtf, its]
    muu[jts:jtf, ite - 1] = mu[jts:jtf, ite - 2] + mub[jts:jtf, ite - 2]
    itf = min(ite, ide - 1)
    a, b = slice(jts + 1, jte - 1), slice(jts, jte - 2)
    muv[a, its:itf] = 0.5 * (mu[a, its:itf] + mu[b, its:itf] + mub[a, its:itf] + mub[b, its:itf])
    muv[jts, its:itf] = mu[jts, its:itf] + mub[jts, its:itf]
    muv[jte - 1, its:itf] = mu[jte - 2, its:itf] + mub[jte - 2, its:itf]
    return muu, muv


def couple_momentum(muu, ru, u, msfu,
                    muv, rv, v, msfv, msfvx_inv,
                    mut, rw, w, msft,
                    c1h, c2h, c1f, c2f,
                    ids, ide, jds, jde, kds, kde,
                    ims, ime, jms, jme, kms, kme,
                    its, ite, jts, jte, kts, kte):
    """Mass-coupled momenta ru, rv, rw.  wrf_dynamics.py:64-86."""
    ktf = min(kte, kde - 1)
    b = Box(kts, ktf, jts, min(jte, jde - 1), its, ite)
    b.put(ru, b(u) * b.m(muu) / b.m(msfu))
    b = Box(kts, ktf, jts, jte, its, min(ite, ide - 1))
    b.put(rv, b(v) * b.m(muv) * b.m(msfvx_inv))
    b = Box(kts, kte, jts, min(jte, jde - 1), its, min(ite, ide - 1))
    b.put(rw, b(w) * b.m(mut) / b.m(msft))
    return ru, rv, rw


def calc_ww_cp(u, v, mu, mub, muu, muv, c1h, c2h, ww,
               rdx, rdy, msftx, msfty,
               msfux, msfuy, msfvx, msfvx_inv,
               msfvy, dnw,
               ids, ide, jds, jde, kds, kde,
               ims, ime, jms, jme, kms, kme,
               its, ite, jts, jte, kts, kte):
    """Coupled eta velocity by column integration.  wrf_dynamics.py:89-118."""
    jtf, ktf, itf = min(jte, jde - 1), min(kte, kde - 1), min(ite, ide - 1)
    divv = zeros_like3(u)
    b = Box(kts, ktf, jts, jtf, its, itf)
    b.put(divv, b.m(msftx) * b.c(dnw) * (
        rdx * (b.m(muu, 0, 1) * b(u, 0, 0, 1) / b.m(msfuy, 0, 1) - b.m(muu) * b(u) / b.m(msfuy))
        + rdy * (b.m(muv, 1, 0) * b(v, 0, 1, 0) * b.m(msfvx_inv, 1, 0) - b.m(muv) * b(v) * b.m(msfvx_inv))))
    dmdt = divv.sum(dim=0)
    for k in range(ktf - 1):
        ww[k + 1, jts:jtf, its:itf] = (ww[k, jts:jtf, its:itf] - dnw[k] * c1h[k] * dmdt[jts:jtf, its:itf]
                                       - divv[k, jts:jtf, its:itf])
    return ww


def calc_cq(moist, cqu, cqv, cqw, n_moist,
            ids, ide, jds, jde, kds, kde,
            ims, ime, jms, jme, kms, kme,
            its, ite, jts, jte, kts, kte):
    """Moisture coupling factors at u, v and w points.  wrf_dynamics.py:121-136."""
    ktf = min(kte, kde - 1)
    jtf = min(jte, jde - 1)
    s = lambda k0, k1, j0, j1, i0, i1: moist[:, k0:k1, j0:j1, i0:i1].sum(dim=0)
    cqu[kts:ktf, jts:jtf, its:ite] = 1.0 / (1.0 + 0.5 * (s(kts, ktf, jts, jtf, its, ite)
                                                       + s(kts, ktf, jts, jtf, its - 1, ite - 1)))
    itf = min(ite, ide - 1)
    cqv[kts:ktf, jts:jte, its:itf] = 1.0 / (1.0 + 0.5 * (s(kts, ktf, jts, jte, its, itf)
                                                       + s(kts, ktf, jts - 1, jte - 1, its, itf)))
    cqw[kts + 1:ktf, jts:jtf, its:itf] = 0.5 * (s(kts + 1, ktf, jts, jtf, its, itf)
                                                + s(kts, ktf - 1, jts, jtf, its, itf))
    return cqu, cqv, cqw


def calc_alt(alt, al, alb,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """alt = al + alb.  wrf_dynamics.py:139-147."""
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    b.put(alt, b(al) + b(alb))
    return alt


def calc_php(php, ph, phb,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """Full geopotential on mass levels.  wrf_dynamics.py:150-159."""
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    b.put(php, 0.5 * (b(phb) + b(phb, 1) + b(ph) + b(ph, 1)))
    return php


def set_physical_bc2d(dat, variable_in,
                      ids, ide, jds, jde,
                      ims, ime, jms, jme,
                      ips, ipe, jps, jpe,
                      its, ite, jts, jte):
    """Open-boundary halo copy for a 2-D field.  wrf_dynamics.py:164-212."""
    v = variable_in.lower()
    if v not in ('u', 'v', 'w', 't', 'x', 'y', 'r', 'p'):
        raise NameError("open_bc_copy undefined for variable %r" % variable_in)
    istag = 0 if v in ('u', 'x', 'd', 'e') else -1
    jstag = 0 if v in ('v', 'y', 'd', 'f') else -1
    i0, i1 = max(ids, its - 1), min(ite + 1, ide + istag)
    rows = slice(jds, jde + jstag)
    for h in (1, 2, 3):
        dat[rows, ids - h] = dat[rows, ids]
    if v not in ('u', 'x'):
        for c in (ide - 1, ide, ide + 1):
            dat[rows, c] = dat[rows, ide - 2]
    else:
        for c in (ide, ide + 1, ide + 2):
            dat[rows, c] = dat[rows, ide - 1]
    for h in (1, 2, 3):
        dat[jds - h, i0:i1] = dat[jds, i0:i1]
    if v not in ('v', 'y'):
        for r in (jde - 1, jde, jde + 1):
            dat[r, i0:i1] = dat[jde - 2, i0:i1]
    else:
        for r in (jde, jde + 1, jde + 2):
            dat[r, i0:i1] = dat[jde - 1, i0:i1]
    return dat


def set_physical_bc3d(dat, variable_in,
                      ids, ide, jds, jde, kds, kde,
                      ims, ime, jms, jme, kms, kme,
                      ips, ipe, jps, jpe, kps, kpe,
                      its, ite, jts, jte, kts, kte):
    """Open-boundary halo copy for a 3-D field.  wrf_dynamics.py:216-278."""
    v = {'U': 'u', 'V': 'v', 'M': 'm', 'H': 'h'}.get(variable_in, variable_in)
    if v not in ('u', 'v', 'w', 't', 'd', 'e', 'x', 'y', 'f', 'r', 'p'):
        raise NameError("open_bc_copy undefined for variable %r" % variable_in)
    istag = 0 if v in ('u', 'x', 'd', 'e') else -1
    jstag = 0 if v in ('v', 'y', 'd', 'f') else -1
    k1 = min(kde, kte) if v in ('e', 'f', 'w') else min(kde - 1, kte)
    ks = slice(kts, k1)
    rows = slice(jts - 4, jde + jstag + 4)
    for h in (1, 2, 3):
        dat[ks, rows, ids - h] = dat[ks, rows, ids]
    if v not in ('u', 'x'):
        for c in (ide - 1, ide, ide + 1):
            dat[ks, rows, c] = dat[ks, rows, ide - 2]
    else:
        rows5 = slice(jts - 5, jde + jstag + 4)
        for c in (ide, ide + 1, ide + 2):
            dat[ks, rows5, c] = dat[ks, rows5, ide - 1]
    cols = slice(ids, ide + istag)
    for h in (1, 2, 3):
        dat[ks, jds - h, cols] = dat[ks, jds, cols]
    if v not in ('v', 'y'):
        for r in (jde - 1, jde, jde + 1):
            dat[ks, r, cols] = dat[ks, jde - 2, cols]
    else:
        for r in (jde, jde + 1, jde + 2):
            dat[ks, r, cols] = dat[ks, jde - 1, cols]
    return dat


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
    """Physics scratch (theta, Exner, rho, heights, w-level p and T, hydrostatic
    pressure).  wrf_dynamics.py:314-411.  pi_phy and t_phy are returned as new
    whole-array tensors, the rest are written in place."""
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    k0, k1 = kts, min(kte, kde - 1)
    b = Box(k0, k1, j0, j1, i0, i1)
    b.put(th_phy, b(t) + t0)
    b.put(p_phy, b(p) + b(pb))
    pi_phy = (p_phy / p1000mb) ** rcp
    t_phy = th_phy * pi_phy
    b.put(rho, 1. / b(alt) * (1 + b(moist[P_QV])))
    b.put(u_phy, 0.5 * (b(u) + b(u, 0, 0, 1)))
    b.put(v_phy, 0.5 * (b(v) + b(v, 0, 1, 0)))
    b.put(z_at_w, (b(phb) + b(ph)) / g)
    bw = Box(k0, kte - 1, j0, j1, i0, i1)
    bw.put(dz8w, bw(z_at_w, 1) - bw(z_at_w))
    dz8w[kte - 1, j0:j1, i0:i1] = 0.0
    b.put(z, 0.5 * (b(z_at_w) + b(z_at_w, 1)))
    bi = Box(k0 + 1, k1, j0, j1, i0, i1)
    bi.put(p8w, bi.c(fzm) * bi(p_phy) + bi.c(fzp) * bi(p_phy, -1))
    bi.put(t8w, bi.c(fzm) * bi(t_phy) + bi.c(fzp) * bi(t_phy, -1))
    h = (slice(j0, j1), slice(i0, i1))
    z0, z1, z2 = z_at_w[(0,) + h], z[(0,) + h], z[(1,) + h]
    w1 = (z0 - z2) / (z1 - z2)
    w2 = 1. - w1
    p8w[(0,) + h] = w1 * p_phy[(0,) + h] + w2 * p_phy[(1,) + h]
    t8w[(0,) + h] = w1 * t_phy[(0,) + h] + w2 * t_phy[(1,) + h]
    z0, z1, z2 = z_at_w[(kte - 1,) + h], z[(k1 - 1,) + h], z[(k1 - 2,) + h]
    w1 = (z0 - z2) / (z1 - z2)
    w2 = 1. - w1
    p8w[(kde - 1,) + h] = torch.exp(w1 * torch.log(p_phy[(kde - 2,) + h]) + w2 * torch.log(p_phy[(kde - 3,) + h]))
    t8w[(kde - 1,) + h] = w1 * t_phy[(kde - 2,) + h] + w2 * t_phy[(kde - 3,) + h]
    p_hyd_w[(kte - 1,) + h] = p_top
    q_tot = zeros_like3(p)
    q_tot[(slice(k0, k1 - 1),) + h] = moist[(slice(None), slice(k0, k1 - 1)) + h].sum(dim=0)
    for k in range(kte - 2, k0 - 1, -1):
        p_hyd_w[(k,) + h] = p_hyd_w[(k + 1,) + h] - (1. + q_tot[(k,) + h]) * mut[h] * dnw[k]
    b.put(p_hyd, 0.5 * (b(p_hyd_w) + b(p_hyd_w, 1)))
    return th_phy, p_phy, pi_phy, t_phy, rho, u_phy, v_phy, z_at_w, dz8w, z, p8w, t8w, p_hyd_w, p_hyd


def calculate_phy_tend(c1, c2,
                       mut, muu, muv, pi3d,
                       RUBLTEN, RVBLTEN, RTHBLTEN,
                       RQVBLTEN, RQCBLTEN, RQIBLTEN,
                       scalar, scalar_tend, num_scalar,
                       ids, ide, jds, jde, kds, kde,
                       ims, ime, jms, jme, kms, kme,
                       its, ite, jts, jte, kts, kte):
    """Couple PBL tendencies with mut.  wrf_dynamics.py:449-488."""
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    for a in (RUBLTEN, RVBLTEN, RTHBLTEN, RQVBLTEN, RQCBLTEN, RQIBLTEN):
        b.put(a, b.m(mut) * b(a))
    for im in range(3):
        b.put(scalar_tend[im], b.m(mut) * b(scalar_tend[im]))
    return RUBLTEN, RVBLTEN, RTHBLTEN, RQVBLTEN, RQCBLTEN, RQIBLTEN, scalar_tend


def compute_diff_metrics(ph, phb, z, rdz, rdzw,
                         zx, zy, rdx, rdy,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """1/dz on w and mass levels, terrain slopes zx, zy and z.  wrf_dynamics.py:491-536."""
    ktf = min(kte, kde - 1)
    zw = zeros_like3(ph)
    j0, j1, i0, i1 = jts - 1, jte, its, ide - 1
    Box(kts, kte, j0, j1, i0, i1).put(zw, (Box(kts, kte, j0, j1, i0, i1)(ph) + Box(kts, kte, j0, j1, i0, i1)(phb)) / g)
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(rdzw, 1.0 / (b(zw, 1) - b(zw)))
    b = Box(kts + 1, ktf, j0, j1, i0, i1)
    b.put(rdz, 2.0 / (b(zw, 1) - b(zw, -1)))
    rdz[0, j0:j1, i0:i1] = 2.0 / (zw[1, j0:j1, i0:i1] - zw[0, j0:j1, i0:i1])
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    b = Box(kts, kte, j0, j1, ids + 1, i1)
    b.put(zx, rdx * (b(phb) - b(phb, 0, 0, -1)) / g + rdx * (b(ph) - b(ph, 0, 0, -1)) / g)
    b = Box(kts, kte, jds + 1, j1, i0, i1)
    b.put(zy, rdy * (b(phb) - b(phb, 0, -1)) / g + rdy * (b(ph) - b(ph, 0, -1)) / g)
    zx[0:ktf, j0:j1, ide - 1] = 0.0
    zx[0:ktf, j0:j1, ids - 1] = 0.0
    zy[0:ktf, jde - 1, i0:i1] = 0.0
    zy[0:ktf, jds - 1, i0:i1] = 0.0
    b = Box(0, ktf, j0, j1, i0, i1)
    b.put(z, 0.5 * (b(ph) + b(phb) + b(ph, 1) + b(phb, 1)))
    return rdzw, rdz, zx, zy, z


def _hatavg_pair(hatavg, hat, b, fnm, fnp, cf1, cf2, cf3, cft1, cft2, ktes1, ktes2, kte, dj):
    """Vertical average of hat and its (dj, di) neighbour onto w levels
    (the three hatavg statements repeated four times in cal_deform_and_div)."""
    j0, j1, i0, i1 = b.j0, b.j1, b.i0, b.i1
    bi = Box(b.k0 + 1, b.k1, j0, j1, i0, i1)
    bi.put(hatavg, 0.5 * (bi.c(fnm) * (bi(hat, 0, *dj[0]) + bi(hat, 0, *dj[1]))
                          + bi.c(fnp) * (bi(hat, -1, *dj[0]) + bi(hat, -1, *dj[1]))))
    h = lambda k, o: hat[k, j0 + o[0]:j1 + o[0], i0 + o[1]:i1 + o[1]]
    hatavg[0, j0:j1, i0:i1] = 0.5 * (cf1 * h(0, dj[0]) + cf2 * h(1, dj[0]) + cf3 * h(2, dj[0])
                                     + cf1 * h(0, dj[1]) + cf2 * h(1, dj[1]) + cf3 * h(2, dj[1]))
    hatavg[kte - 1, j0:j1, i0:i1] = 0.5 * (cft1 * (h(ktes1, dj[0]) + h(ktes1, dj[1]))
                                           + cft2 * (h(ktes2, dj[0]) + h(ktes2, dj[1])))


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
    """Deformation tensor and divergence in terrain-following form.
    wrf_dynamics.py:539-813."""
    ktes1, ktes2 = kte - 2, kte - 3
    cft2 = - 0.5 * dnw[ktes1] / dn[ktes1]
    cft1 = 1.0 - cft2
    ktf = kde - 1
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    mm = torch.zeros(msftx.shape, dtype=u.dtype)
    hat, hatavg, tmp1 = zeros_like3(u), zeros_like3(u), zeros_like3(u)
    mm[j0:j1, i0:i1] = msftx[j0:j1, i0:i1] * msfty[j0:j1, i0:i1]
    args = (fnm, fnp, cf1, cf2, cf3, cft1, cft2, ktes1, ktes2, kte)
    b = Box(kts, ktf, j0, j1, i0, i1)
    # D11 and the first part of the divergence
    Box(kts, ktf, j0, j1, i0, i1 + 1).put(hat, Box(kts, ktf, j0, j1, i0, i1 + 1)(u)
                                          / Box(kts, ktf, j0, j1, i0, i1 + 1).m(msfuy))
    _hatavg_pair(hatavg, hat, b, *args, dj=((0, 0), (0, 1)))
    b.put(tmp1, (b(hatavg, 1) - b(hatavg)) * (0.25 * (b(zx) + b(zx, 0, 0, 1) + b(zx, 1) + b(zx, 1, 0, 1)) * b(rdzw)))
    b.put(tmp1, b.m(mm) * (rdx * (b(hat, 0, 0, 1) - b(hat)) - b(tmp1)))
    b.put(defor11, 2.0 * b(tmp1))
    b.put(div, b(tmp1) + 0.0)
    # D22
    Box(kts, ktf, j0, j1 + 1, i0, i1).put(hat, Box(kts, ktf, j0, j1 + 1, i0, i1)(v)
                                          / Box(kts, ktf, j0, j1 + 1, i0, i1).m(msfvx))
    _hatavg_pair(hatavg, hat, b, *args, dj=((0, 0), (1, 0)))
    b.put(tmp1, b.m(mm) * (rdy * (b(hat, 0, 1) - b(hat)) - (b(hatavg, 1) - b(hatavg)) * (
        0.25 * (b(zy) + b(zy, 0, 1) + b(zy, 1) + b(zy, 1, 1)) * b(rdzw))))
    b.put(defor22, 2.0 * b(tmp1))
    b.put(div, b(tmp1) + b(div))
    # D33
    b.put(tmp1, (b(w, 1) - b(w)) * b(rdzw))
    b.put(defor33, 2.0 * b(tmp1))
    b.put(div, b(tmp1) + b(div))
    # D12
    j0, j1, i0, i1 = jts + 1, jte - 1, its + 1, ite - 1
    mm[j0:j1, i0:i1] = 0.25 * (msfux[j0 - 1:j1 - 1, i0:i1] + msfux[j0:j1, i0:i1]) * \
        (msfvy[j0:j1, i0 - 1:i1 - 1] + msfvy[j0:j1, i0:i1])
    b = Box(kts, ktf, j0, j1, i0, i1)
    bh = Box(kts, ktf, j0 - 1, j1, i0, i1)
    bh.put(hat, bh(u) / bh.m(msfux))
    _hatavg_pair(hatavg, hat, b, *args, dj=((-1, 0), (0, 0)))
    b.put(tmp1, (b(hatavg, 1) - b(hatavg)) * (0.25 * (b(zy, 0, 0, -1) + b(zy) + b(zy, 1, 0, -1) + b(zy, 1)) * 0.25 * (
        b(rdzw) + b(rdzw, 0, 0, -1) + b(rdzw, 0, -1, -1) + b(rdzw, 0, -1, 0))))
    b.put(defor12, b.m(mm) * (rdy * (b(hat) - b(hat, 0, -1)) - b(tmp1)))
    bh = Box(kts, ktf, j0, j1, i0 - 1, i1)
    bh.put(hat, bh(v) / bh.m(msfvy))
    _hatavg_pair(hatavg, hat, b, *args, dj=((0, -1), (0, 0)))
    b.put(tmp1, (b(hatavg, 1) - b(hatavg)) * (0.25 * (b(zx, 0, -1) + b(zx) + b(zx, 1, -1) + b(zx, 1)) * 0.25 * (
        b(rdzw) + b(rdzw, 0, -1) + b(rdzw, 0, -1, -1) + b(rdzw, 0, 0, -1))))
    b.put(defor12, b(defor12) + b.m(mm) * (rdx * (b(hat) - b(hat, 0, 0, -1)) - b(tmp1)))
    defor12[kts:kte, jts:jte, ids] = defor12[kts:kte, jts:jte, ids + 1]
    defor12[kts:kte, jts:jte, ide - 1] = defor12[kts:kte, jts:jte, ide - 2]
    defor12[kts:kte, jds, its:ite] = defor12[kts:kte, jds + 1, its:ite]
    defor12[kts:kte, jde - 1, its:ite] = defor12[kts:kte, jde - 2, its:ite]
    # D13
    j0, j1, i0, i1 = jts, jde - 1, its, ide - 1
    mm[j0:j1, i0:i1] = msfux[j0:j1, i0:i1] * msfuy[j0:j1, i0:i1]
    Box(kts, kte, j0, j1, i0, i1).put(hat, Box(kts, kte, j0, j1, i0, i1)(w) / Box(kts, kte, j0, j1, i0, i1).m(msfty))
    hat[kts:kte, j0:j1, i0 - 1] = w[kts:kte, j0:j1, i0 - 1] / msfty[j0:j1, i0 - 1]
    hat[kts:kte, j0 - 1, i0:i1] = w[kts:kte, j0 - 1, i0:i1] / msfty[j0 - 1, i0:i1]
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(hatavg, 0.25 * (b(hat) + b(hat, 1) + b(hat, 0, 0, -1) + b(hat, 1, 0, -1)))
    bi = Box(kts + 1, ktf, j0, j1, i0, i1)
    bi.put(tmp1, (bi(hatavg) - bi(hatavg, -1)) * bi(zx) * 0.5 * (bi(rdz) + bi(rdz, 0, 0, -1)))
    bi.put(defor13, bi.m(mm) * (rdx * (bi(hat) - bi(hat, 0, 0, -1)) - bi(tmp1)))
    defor13[kts, j0:j1, i0:i1] = 0.
    defor13[ktf, j0:j1, i0:i1] = 0.
    bi.put(tmp1, (bi(u) - bi(u, -1)) * 0.5 * (bi(rdz) + bi(rdz, 0, 0, -1)))
    bi.put(defor13, bi(defor13) + bi(tmp1))
    # D23
    mm[j0:j1, i0:i1] = msfvx[j0:j1, i0:i1] * msfvy[j0:j1, i0:i1]
    Box(kts, kte, j0, j1, i0, i1).put(hat, Box(kts, kte, j0, j1, i0, i1)(w) / Box(kts, kte, j0, j1, i0, i1).m(msftx))
    hat[kts:kte, j0:j1, i0 - 1] = w[kts:kte, j0:j1, i0 - 1] / msftx[j0:j1, i0 - 1]
    hat[kts:kte, j0 - 1, i0:i1] = w[kts:kte, j0 - 1, i0:i1] / msftx[j0 - 1, i0:i1]
    b.put(hatavg, 0.25 * (b(hat) + b(hat, 1) + b(hat, 0, -1) + b(hat, 1, -1)))
    bi.put(tmp1, (bi(hatavg) - bi(hatavg, -1)) * bi(zy) * 0.5 * (bi(rdz) + bi(rdz, 0, -1)))
    bi.put(defor23, bi.m(mm) * (rdy * (bi(hat) - bi(hat, 0, -1)) - bi(tmp1)))
    defor23[kts, j0:j1, i0:i1] = 0.
    defor23[ktf, j0:j1, i0:i1] = 0.
    bi.put(tmp1, (bi(v) - bi(v, -1)) * 0.5 * (bi(rdz) + bi(rdz, 0, -1)))
    bi.put(defor23, bi(defor23) + bi(tmp1))
    for d in (defor13, defor23):
        d[kts:kte, jts:jte, ids] = d[kts:kte, jts:jte, ids + 1]
        d[kts:kte, jts:jte, ide - 1] = d[kts:kte, jts:jte, ide - 2]
        d[kts:kte, jds, its:ite] = d[kts:kte, jds + 1, its:ite]
        d[kts:kte, jde - 1, its:ite] = d[kts:kte, jde - 2, its:ite]
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
    """Brunt-Vaisala frequency and 2-D Smagorinsky eddy viscosity.
    wrf_dynamics.py:816-995.  BN2 and xkmh are returned as new tensors."""
    ktf = min(kte, kde - 1)
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    ktes1, ktes2 = kte - 2, kte - 3
    qc_cr = 0.00001
    Z3 = lambda: zeros_like3(t)
    Z2 = lambda: torch.zeros(msftx.shape, dtype=t.dtype)
    tmp1sfc, qctmp, tmp1, es, qvs, tmpdz, xlvqv, coefa = Z2(), Z3(), Z3(), Z3(), Z3(), Z3(), Z3(), Z3()
    thetaep1, thetaem1, BN2_1, BN2_0, thetasfc, thetaesfc, qvsfc = Z3(), Z3(), Z3(), Z3(), Z2(), Z2(), Z2()
    def2, tmp, mlen_h = Z3(), Z3(), Z2()
    qv, qc, qi = moist[P_QV], moist[P_QC], moist[P_QI]
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(qctmp, b(qc) + 0.0)
    b.put(tmp1, b(qv) + b(qc) + b(qi))
    h = (slice(j0, j1), slice(i0, i1))
    L = lambda a, k: a[(k,) + h]
    tmp1sfc[h] = (cf1 * (L(qv, 0) + L(qc, 0) + L(qi, 0)) + cf2 * (L(qv, 1) + L(qc, 1) + L(qi, 1))
                  + cf3 * (L(qv, 2) + L(qc, 2) + L(qi, 2)))
    b.put(es, 1000.0 * SVP1 * torch.exp(SVP2 * (b(t) - SVPT0) / (b(t) - SVP3)))
    b.put(qvs, EP_2 * b(es) / (b(p) - b(es)))
    bm = Box(kts + 1, ktf - 1, j0, j1, i0, i1)
    bm.put(tmpdz, 1.0 / bm(rdz) + 1.0 / bm(rdz, 1))
    bm.put(xlvqv, XLV * bm(qv))
    bm.put(coefa, (1.0 + bm(xlvqv) / r_d / bm(t)) / (1.0 + XLV * bm(xlvqv) / Cp / r_v / bm(t) / bm(t)) / bm(theta))
    bm.put(thetaep1, bm(theta, 1) * (1.0 + XLV * bm(qvs, 1)) / Cp / bm(t, 1))
    bm.put(thetaem1, bm(theta, -1) * (1.0 + XLV * bm(qvs, -1)) / Cp / bm(t, -1))
    bm.put(BN2_1, g * (bm(coefa) * (bm(thetaep1) - bm(thetaem1)) / bm(tmpdz)
                       - (bm(tmp1, 1) - bm(tmp1, -1)) / bm(tmpdz)))
    bm.put(BN2_0, g * ((bm(theta, 1) - bm(theta, -1)) / bm(theta) / bm(tmpdz)
                       + 1.61 * (bm(qv, 1) - bm(qv, -1)) / bm(tmpdz)
                       - (bm(tmp1, 1) - bm(tmp1, -1)) / bm(tmpdz)))
    cond = torch.logical_or(qv >= qvs, qctmp >= torch.tensor(qc_cr, dtype=t.dtype))
    BN2 = torch.where(cond, BN2_1, BN2_0)
    tmpdz[(kts,) + h] = 1.0 / L(rdz, kts + 1) + 0.5 / L(rdzw, kts)
    thetasfc[h] = L(t8w, kts) / ((L(p8w, kts) / p1000mb) ** (r_d / Cp))
    qvsfc[h] = cf1 * L(qvs, 0) + cf2 * L(qvs, 1) + cf3 * L(qvs, 2)
    xlvqv[(kts,) + h] = XLV * L(qv, kts)
    coefa[(kts,) + h] = ((1.0 + L(xlvqv, kts) / r_d / L(t, kts))
                         / (1.0 + XLV * L(xlvqv, kts) / Cp / r_v / L(t, kts) / L(t, kts)) / L(theta, kts))
    thetaep1[(kts,) + h] = L(theta, kts + 1) * (1.0 + XLV * L(qvs, kts + 1)) / Cp / L(t, kts + 1)
    thetaesfc[h] = thetasfc[h] * (1.0 + XLV * qvsfc[h] / Cp / L(t8w, kts))
    BN2_1[(kts,) + h] = g * (L(coefa, kts) * (L(thetaep1, kts) - thetaesfc[h]) / L(tmpdz, kts)
                             - (L(tmp1, kts + 1) - tmp1sfc[h]) / L(tmpdz, kts))
    qvsfc[h] = cf1 * L(qv, 0) + cf2 * L(qv, 1) + cf3 * L(qv, 2)
    tmpdz[(kts,) + h] = 1. / L(rdzw, kts)
    BN2_0[(kts,) + h] = g * ((L(theta, kts + 1) - L(theta, kts)) / L(theta, kts) / L(tmpdz, kts)
                             + 1.61 * (L(qv, kts + 1) - qvsfc[h]) / L(tmpdz, kts)
                             - (L(tmp1, kts + 1) - tmp1sfc[h]) / L(tmpdz, kts))
    cond = torch.logical_or(qv[kts] >= qvs[kts], qctmp[kts] >= qc_cr)
    BN2[kts] = torch.where(cond, BN2_1[kts], BN2_0[kts])
    BN2[ktf - 1] = BN2[ktf - 2]
    # 2-D Smagorinsky
    j0, j1, i0, i1 = jts + 1, min(jte, jde - 2), its + 1, min(ite, ide - 2)
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(def2, 0.25 * ((b(defor11) - b(defor22)) * (b(defor11) - b(defor22))))
    b.put(tmp, 0.25 * (b(defor12) + b(defor12, 0, 1) + b(defor12, 0, 0, 1) + b(defor12, 0, 1, 1)))
    b.put(def2, b(def2) + b(tmp) * b(tmp))
    mlen_h[j0:j1, i0:i1] = (dx / msftx[j0:j1, i0:i1] * dy / msfty[j0:j1, i0:i1]) ** 0.5
    b.put(tmp, b(def2) ** 0.5)
    b.put(xkmh, c_s * c_s * b.m(mlen_h) * b.m(mlen_h) * b(tmp))
    cap = 10. * mlen_h
    xkmh = torch.where(xkmh < cap, xkmh, cap.expand_as(xkmh))
    xkmv[:, :, :] = 0.
    b.put(xkhh, b(xkmh) / prandtl)
    xkhv[:, :, :] = 0.
    return BN2, xkmh, xkmv, xkhh, xkhv


def _add_a2a(lvar, rvar, ide, jde, kde, its, ite, jts, jte, kts, kte):
    b = Box(kts, min(kte, kde - 1), jts, min(jte, jde - 1), its, min(ite, ide - 1))
    b.put(lvar, b(lvar) + b(rvar))
    return lvar


def update_phy_ten(rph_tendf, rt_tendf, ru_tendf, rv_tendf, moist_tendf,
                   scalar_tendf, mu_tendf,
                   RTHBLTEN, RUBLTEN, RVBLTEN,
                   RQVBLTEN, RQCBLTEN, RQIBLTEN,
                   n_moist, n_scalar, rk_step, adv_moist_cond,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   its, ite, jts, jte, kts, kte):
    """Add PBL tendencies to the physics tendencies.  wrf_dynamics.py:998-1076."""
    ix = (ide, jde, kde, its, ite, jts, jte, kts, kte)
    rt_tendf = _add_a2a(rt_tendf, RTHBLTEN, *ix)
    ktf = min(kte, kde - 1)
    b = Box(kts, ktf, jts, min(jte, jde - 1), its, ite)
    b.put(ru_tendf, b(ru_tendf) + 0.5 * b(RUBLTEN) + 0.5 * b(RUBLTEN, 0, 0, -1))
    b = Box(kts, ktf, jts, jte, its, min(ite, ide - 1))
    b.put(rv_tendf, b(rv_tendf) + 0.5 * b(RVBLTEN) + 0.5 * b(RVBLTEN, 0, -1, 0))
    moist_tendf[P_QV] = _add_a2a(moist_tendf[P_QV], RQVBLTEN, *ix)
    moist_tendf[P_QC] = _add_a2a(moist_tendf[P_QC], RQCBLTEN, *ix)
    moist_tendf[P_QI] = _add_a2a(moist_tendf[P_QI], RQIBLTEN, *ix)
    return rt_tendf, ru_tendf, rv_tendf, moist_tendf
