"""Oracle: geopotential RHS, pressure gradient, buoyancy, Coriolis,
curvature, diffusion and tendency accumulation (test infrastructure only).

CPU restatement of pywrf/wrf_dynamics.py:2031-3148.
"""
from .constants import g, reradius
from ._box import Box, zeros_like3


def _ph_flux(ph_tend, ph, phb, vel, muf, msf, msfty, rd, cfn, cfn1, box_lo, box_top, d, order):
    """One direction of the horizontal phi advection in rhs_ph.

    d is the unit offset (dj, di); order is 6, 4 or 2 as in
    wrf_dynamics.py:2082-2295.  Interior levels use (v[k]+v[k-1]) and 0.25,
    the top level the cfn/cfn1 extrapolation and 0.5."""
    def sh(b, a, n, dk=0):
        return b(a, dk, n * d[0], n * d[1])

    def sh2(b, a, n):
        return b.m(a, n * d[0], n * d[1])

    for b, top in ((box_lo, False), (box_top, True)):
        if b.j1 <= b.j0 or b.i1 <= b.i0:
            continue
        if top:
            c = (0.5 * rd) / b.m(msfty)
            V = lambda n: cfn * sh(b, vel, n, -1) + cfn1 * sh(b, vel, n, -2)
        else:
            c = (0.25 * rd) / b.m(msfty)
            V = lambda n: sh(b, vel, n) + sh(b, vel, n, -1)
        F1 = sh2(b, muf, 1) * V(1) * sh2(b, msf, 1)
        F0 = sh2(b, muf, 0) * V(0) * sh2(b, msf, 0)
        if order == 6:
            s = (45. * (sh(b, ph, 1) - sh(b, ph, -1)) - 9. * (sh(b, ph, 2) - sh(b, ph, -2)) + (sh(b, ph, 3) - sh(b, ph, -3))
                 + 45. * (sh(b, phb, 1) - sh(b, phb, -1)) - 9. * (sh(b, phb, 2) - sh(b, phb, -2))
                 + (sh(b, phb, 3) - sh(b, phb, -3)))
            b.put(ph_tend, b(ph_tend) - c * ((F1 + F0) * (1. / 60.) * s))
        elif order == 4:
            s = (8. * (sh(b, ph, 1) - sh(b, ph, -1)) - (sh(b, ph, 2) - sh(b, ph, -2))
                 + 8. * (sh(b, phb, 1) - sh(b, phb, -1)) - (sh(b, phb, 2) - sh(b, phb, -2)))
            b.put(ph_tend, b(ph_tend) - c * ((F1 + F0) * (1. / 12.) * s))
        else:
            d1 = sh(b, phb, 1) - sh(b, phb, 0) + sh(b, ph, 1) - sh(b, ph, 0)
            d0 = sh(b, phb, 0) - sh(b, phb, -1) + sh(b, ph, 0) - sh(b, ph, -1)
            b.put(ph_tend, b(ph_tend) - c * (F1 * d1 + F0 * d0))


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
    """Geopotential tendency.  wrf_dynamics.py:2031-2297."""
    itf, jtf = min(ite, ide - 1), min(jte, jde - 1)
    wdwn = zeros_like3(ph)
    b = Box(1, kte, jts, jtf, its, itf)
    b.put(wdwn, 0.5 * (b(ww) + b(ww, -1)) * b.c(rdnw, -1) * (b(ph) - b(ph, -1) + b(phb) - b(phb, -1)))
    b = Box(1, kte - 1, jts, jtf, its, itf)
    b.put(ph_tend, b(ph_tend) - (b.c(fnm) * b(wdwn, 1) + b.c(fnp) * b(wdwn)))
    ph_tend[kte - 1, jts:jtf, its:itf] = 0.
    b = Box(1, kte, jts, jtf, its, itf)
    b.put(ph_tend, b(ph_tend) + b.m(mut) * g * b(w) / b.m(msfty))
    # y direction
    args = (ph_tend, ph, phb, v, muvf, msfvy, msfty, rdy, cfn, cfn1)
    lohi = lambda j0, j1, i0, i1: (Box(1, kte - 1, j0, j1, i0, i1), Box(kte - 1, kte, j0, j1, i0, i1))
    j0, j1 = max(jts, jds + 3), min(jtf, jde - 4)
    _ph_flux(*args, *lohi(j0, j1, its, itf), (1, 0), 6)
    _ph_flux(*args, *lohi(jds + 2, jds + 3, its, itf), (1, 0), 4)
    _ph_flux(*args, *lohi(jde - 4, jde - 3, its, itf), (1, 0), 4)
    _ph_flux(*args, *lohi(jds + 1, jds + 2, its, itf), (1, 0), 2)
    _ph_flux(*args, *lohi(jde - 3, jde - 2, its, itf), (1, 0), 2)
    # x direction (no 4th-order columns, as in the reference)
    args = (ph_tend, ph, phb, u, muuf, msfux, msfty, rdx, cfn, cfn1)
    i0, i1 = max(its, ids + 3), min(itf, ide - 4)
    _ph_flux(*args, *lohi(jts, jtf, i0, i1), (0, 1), 6)
    _ph_flux(*args, *lohi(jts, jtf, ids + 1, ids + 2), (0, 1), 2)
    _ph_flux(*args, *lohi(jts, jtf, ide - 3, ide - 2), (0, 1), 2)
    return ph_tend


def _dpn(dpn, p, fnm, fnp, cf1, cf2, cf3, kde, ktf, j0, j1, i0, i1, d):
    """Pressure on w levels averaged to the u or v face (d = (dj, di) = the
    neighbour on the other side of the face)."""
    h = (slice(j0, j1), slice(i0, i1))
    hd = (slice(j0 + d[0], j1 + d[0]), slice(i0 + d[1], i1 + d[1]))
    dpn[(0,) + h] = 0.5 * (cf1 * (p[(0,) + hd] + p[(0,) + h]) + cf2 * (p[(1,) + hd] + p[(1,) + h])
                           + cf3 * (p[(2,) + hd] + p[(2,) + h]))
    dpn[(kde - 1,) + h] = 0.
    b = Box(1, ktf, j0, j1, i0, i1)
    b.put(dpn, 0.5 * (b.c(fnm) * (b(p, 0, *d) + b(p)) + b.c(fnp) * (b(p, -1, *d) + b(p, -1))))


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
    """Terrain-following horizontal PGF on ru/rv tendencies.  wrf_dynamics.py:2300-2421."""
    ktf = min(kte, kde - 1)
    dpn = zeros_like3(p)
    for (j0, j1, i0, i1, d, mv, mx, my, rd, cq, tend) in (
            (jts + 1, jte - 1, its, min(ite, ide - 1), (-1, 0), muv, msfvy, msfvx, rdy, cqv, rv_tend),
            (jts, min(jte, jde - 1), its + 1, ite - 1, (0, -1), muu, msfux, msfuy, rdx, cqu, ru_tend)):
        _dpn(dpn, p, fnm, fnp, cf1, cf2, cf3, kde, ktf, j0, j1, i0, i1, d)
        b = Box(0, ktf, j0, j1, i0, i1)
        dpx = (b.m(mx) / b.m(my)) * 0.5 * rd * b.m(mv) * (
            (b(ph, 1) - b(ph, 1, *d) + b(ph) - b(ph, 0, *d)) + (b(alt) + b(alt, 0, *d)) * (b(p) - b(p, 0, *d))
            + (b(al) + b(al, 0, *d)) * (b(pb) - b(pb, 0, *d)))
        dpx = dpx + (b.m(mx) / b.m(my)) * rd * (b(php) - b(php, 0, *d)) * (
            b.c(rdnw) * (b(dpn, 1) - b(dpn)) - 0.5 * (b.m(mu, *d) + b.m(mu)))
        b.put(tend, b(tend) - b(cq) * dpx)
    return ru_tend, rv_tend


def pg_buoy_w(rw_tend, p, cqw, muf, mubf,
              c1f, c2f,
              rdnw, rdn, g_, msftx, msfty,
              ids, ide, jds, jde, kds, kde,
              ims, ime, jms, jme, kms, kme,
              its, ite, jts, jte, kts, kte):
    """Vertical PGF and buoyancy on w; overwrites cqw[1:kde-1] with 1/(1+cqw).
    wrf_dynamics.py:2424-2460."""
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    h = (slice(j0, j1), slice(i0, i1))
    k = kde - 1
    cq1 = 1. / (1. + cqw[(kde - 2,) + h])
    cq2 = cqw[(kde - 2,) + h] * cq1
    rw_tend[(k,) + h] = rw_tend[(k,) + h] + (1. / msfty[h]) * g_ * (
        cq1 * 2. * rdnw[kde - 2] * (-p[(kde - 2,) + h]) - muf[h] - cq2 * mubf[h])
    b = Box(1, kde - 1, j0, j1, i0, i1)
    cq1 = 1. / (1. + b(cqw))
    cq2 = b(cqw) * cq1
    b.put(cqw, cq1)
    b.put(rw_tend, b(rw_tend) + (1. / b.m(msfty)) * g_ * (
        cq1 * b.c(rdn) * (b(p) - b(p, -1)) - b.m(muf) - cq2 * b.m(mubf)))
    return rw_tend, cqw


def coriolis(ru, rv, rw,
             ru_tend, rv_tend, rw_tend,
             msftx, msfty, msfux, msfuy,
             msfvx, msfvy,
             f, e, sina, cosa, fzm, fzp,
             ids, ide, jds, jde, kds, kde,
             ims, ime, jms, jme, kms, kme,
             its, ite, jts, jte, kts, kte):
    """Coriolis terms.  wrf_dynamics.py:2463-2532."""
    ktf = min(kte, kde - 1)
    b = Box(kts, ktf, jts, jte - 1, max(ids + 1, its), min(ide - 1, ite))
    b.put(ru_tend, b(ru_tend) + (b.m(msfux) / b.m(msfuy)) * 0.5 * (b.m(f) + b.m(f, 0, -1)) * 0.25 * (
        b(rv, 0, 1, -1) + b(rv, 0, 1, 0) + b(rv, 0, 0, -1) + b(rv))
        - 0.5 * (b.m(e) + b.m(e, 0, -1)) * 0.5 * (b.m(cosa) + b.m(cosa, 0, -1)) * 0.25 * (
        b(rw, 1, 0, -1) + b(rw, 0, 0, -1) + b(rw, 1) + b(rw)))
    b = Box(kts, ktf, max(jds + 1, jts), min(jde - 1, jte), its, ite - 1)
    b.put(rv_tend, b(rv_tend) - (b.m(msfvy) / b.m(msfvx)) * 0.5 * (b.m(f) + b.m(f, -1)) * 0.25 * (
        b(ru) + b(ru, 0, 0, 1) + b(ru, 0, -1) + b(ru, 0, -1, 1))
        + (b.m(msfvy) / b.m(msfvx)) * 0.5 * (b.m(e) + b.m(e, -1)) * 0.5 * (b.m(sina) + b.m(sina, -1)) * 0.25 * (
        b(rw, 1, -1) + b(rw, 0, -1) + b(rw, 1) + b(rw)))
    b = Box(kts + 1, ktf, jts, jte - 1, its, ite - 1)
    b.put(rw_tend, b(rw_tend) + b.m(e) * (b.m(cosa) * 0.5 * (
        b.c(fzm) * (b(ru) + b(ru, 0, 0, 1)) + b.c(fzp) * (b(ru, -1) + b(ru, -1, 0, 1)))
        - (b.m(msftx) / b.m(msfty)) * b.m(sina) * 0.5 * (
        b.c(fzm) * (b(rv) + b(rv, 0, 1)) + b.c(fzp) * (b(rv, -1) + b(rv, -1, 1)))))
    return ru_tend, rv_tend, rw_tend


def curvature(ru, rv, rw, u, v, w, ru_tend, rv_tend, rw_tend,
              msfux, msfuy, msfvx, msfvy, msftx, msfty,
              xlat, fzm, fzp, rdx, rdy,
              ids, ide, jds, jde, kds, kde,
              ims, ime, jms, jme, kms, kme,
              its, ite, jts, jte, kts, kte):
    """Map-factor curvature terms.  wrf_dynamics.py:2535-2620."""
    ktf = min(kte, kde - 1)
    vxgm = zeros_like3(u)
    b = Box(kts, ktf, jts, jte - 1, its, ite - 1)
    b.put(vxgm, 0.5 * (b(u) + b(u, 0, 0, 1)) * (b.m(msfvx, 1, 0) - b.m(msfvx)) * rdy
          - 0.5 * (b(v) + b(v, 0, 1)) * (b.m(msfuy, 0, 1) - b.m(msfuy)) * rdx)
    vxgm[kts:ktf, jts:jte - 1, its - 1] = vxgm[kts:ktf, jts:jte - 1, its]
    vxgm[kts:ktf, jts:jte - 1, ite - 1] = vxgm[kts:ktf, jts:jte - 1, ite - 2]
    vxgm[kts:ktf, jts - 1, its - 1:ite] = vxgm[kts:ktf, jts, its - 1:ite]
    vxgm[kts:ktf, jte - 1, its - 1:ite] = vxgm[kts:ktf, jte - 2, its - 1:ite]
    b = Box(kts, ktf, jts, jte - 1, max(ids + 1, its), min(ide - 1, ite))
    b.put(ru_tend, b(ru_tend) + 0.5 * (b(vxgm) + b(vxgm, 0, 0, -1)) * 0.25 * (
        b(rv, 0, 1, -1) + b(rv, 0, 1) + b(rv, 0, 0, -1) + b(rv))
        - b(u) * reradius * 0.25 * (b(rw, 1, 0, -1) + b(rw, 0, 0, -1) + b(rw, 1) + b(rw)))
    b = Box(kts, ktf, max(jds + 1, jts), min(jde - 1, jte), its, ite - 1)
    b.put(rv_tend, b(rv_tend) - 0.5 * (b(vxgm) + b(vxgm, 0, -1)) * 0.25 * (
        b(ru) + b(ru, 0, 0, 1) + b(ru, 0, -1) + b(ru, 0, -1, 1))
        - (b.m(msfvy) / b.m(msfvx)) * b(v) * reradius * 0.25 * (b(rw, 1, -1) + b(rw, 0, -1) + b(rw, 1) + b(rw)))
    b = Box(kts + 1, ktf, jts, jte - 1, its, ite - 1)
    fz = lambda a, o: b.c(fzm) * (b(a) + b(a, 0, *o)) + b.c(fzp) * (b(a, -1) + b(a, -1, *o))
    b.put(rw_tend, b(rw_tend) + reradius * (0.5 * fz(ru, (0, 1)) * 0.5 * fz(u, (0, 1))
                                            + (b.m(msftx) / b.m(msfty)) * 0.5 * fz(rv, (1, 0)) * 0.5 * fz(v, (1, 0))))
    return ru_tend, rv_tend, rw_tend


def horizontal_diffusion(name, field, tendency, mut, c1, c2,
                         msfux, msfuy, msfvx, msfvx_inv,
                         msfvy, msftx, msfty,
                         khdif, xkmhd, rdx, rdy,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """2nd-order flux-form horizontal diffusion.  wrf_dynamics.py:2623-2838."""
    ktf = min(kte, kde - 1)
    f, xk = field, xkmhd
    if name == "u":
        b = Box(kts, ktf, max(jds + 1, jts), min(jde - 2, jte), max(ids + 1, its), min(ide - 1, ite))
        mkxm = (b.m(msftx, 0, -1) / b.m(msfty, 0, -1)) * b.m(mut, 0, -1) * b(xk, 0, 0, -1) * rdx
        mkxp = (b.m(msftx) / b.m(msfty)) * b.m(mut) * b(xk) * rdx
        mrdx = b.m(msfux) * b.m(msfuy) * rdx
        mkym = ((b.m(msfuy) + b.m(msfuy, -1)) / (b.m(msfux) + b.m(msfux, -1))) * 0.25 * (
            b.m(mut) + b.m(mut, -1) + b.m(mut, -1, -1) + b.m(mut, 0, -1)) * 0.25 * (
            b(xk) + b(xk, 0, -1) + b(xk, 0, -1, -1) + b(xk, 0, 0, -1)) * rdy
        mkyp = ((b.m(msfuy) + b.m(msfuy, 1)) / (b.m(msfux) + b.m(msfux, 1))) * 0.25 * (
            b.m(mut) + b.m(mut, 1) + b.m(mut, 1, -1) + b.m(mut, 0, -1)) * 0.25 * (
            b(xk) + b(xk, 0, 1) + b(xk, 0, 1, -1) + b(xk, 0, 0, -1)) * rdy
        mrdy = b.m(msfux) * b.m(msfuy) * rdy
    elif name == "v":
        b = Box(kts, ktf, max(jds + 1, jts), min(jde - 1, jte), max(ids + 1, its), min(ide - 2, ite))
        mkxm = ((b.m(msfvx) + b.m(msfvx, 0, -1)) / (b.m(msfvy) + b.m(msfvy, 0, -1))) * 0.25 * (
            b.m(mut) + b.m(mut, -1) + b.m(mut, -1, -1) + b.m(mut, 0, -1)) * 0.25 * (
            b(xk) + b(xk, 0, -1) + b(xk, 0, -1, -1) + b(xk, 0, 0, -1)) * rdx
        mkxp = ((b.m(msfvx) + b.m(msfvx, 0, 1)) / (b.m(msfvy) + b.m(msfvy, 0, 1))) * 0.25 * (
            b.m(mut) + b.m(mut, -1) + b.m(mut, -1, 1) + b.m(mut, 0, 1)) * 0.25 * (
            b(xk) + b(xk, 0, -1) + b(xk, 0, -1, 1) + b(xk, 0, 0, 1)) * rdx
        mrdx = b.m(msfvx) * b.m(msfvy) * rdx
        mkym = (b.m(msfty, -1) / b.m(msftx, -1)) * b(xk, 0, -1) * rdy
        mkyp = (b.m(msfty) / b.m(msftx)) * b(xk) * rdy
        mrdy = b.m(msfvx) * b.m(msfvy) * rdy
    elif name == "w":
        b = Box(kts + 1, ktf, max(jds + 1, jts), min(jde - 2, jte), max(ids + 1, its), min(ide - 2, ite))
        mkxm = (b.m(msfux) / b.m(msfuy)) * 0.25 * (
            b.m(mut) + b.m(mut, 0, -1) + b.m(mut) + b.m(mut, 0, -1)) * 0.25 * (
            b(xk) + b(xk, 0, 0, -1) + b(xk, -1) + b(xk, -1, 0, -1)) * rdx
        mkxp = (b.m(msfux, 0, 1) / b.m(msfuy, 0, 1)) * 0.25 * (
            b.m(mut, 0, 1) + b.m(mut) + b.m(mut, 0, 1) + b.m(mut)) * 0.25 * (
            b(xk, 0, 0, 1) + b(xk) + b(xk, -1, 0, 1) + b(xk, -1)) * rdx
        mrdx = b.m(msftx) * b.m(msfty) * rdx
        mkym = (b.m(msfvy) / b.m(msfvx)) * 0.25 * (
            b.m(mut) + b.m(mut, -1) + b.m(mut) + b.m(mut, -1)) * 0.25 * (
            b(xk) + b(xk, 0, -1) + b(xk, -1) + b(xk, -1, -1)) * rdy
        mkyp = (b.m(msfvy, 1) / b.m(msfvx, 1)) * 0.25 * (
            b.m(mut, 1) + b.m(mut) + b.m(mut, 1) + b.m(mut)) * 0.25 * (
            b(xk, 0, 1) + b(xk) + b(xk, -1, 1) + b(xk, -1)) * rdy
        mrdy = b.m(msftx) * b.m(msfty) * rdy
    else:
        b = Box(kts, ktf, max(jds + 1, jts), min(jde - 2, jte), max(ids + 1, its), min(ide - 2, ite))
        mkxm, mkxp, mrdx, mkym, mkyp, mrdy = _scalar_coefs(b, mut, xk, msfux, msfuy, msfvx_inv, msfvy,
                                                           msftx, msfty, rdx, rdy)
    b.put(tendency, b(tendency) + (mrdx * (mkxp * (b(f, 0, 0, 1) - b(f)) - mkxm * (b(f) - b(f, 0, 0, -1)))
                                   + mrdy * (mkyp * (b(f, 0, 1) - b(f)) - mkym * (b(f) - b(f, 0, -1)))))
    return tendency


def _scalar_coefs(b, mut, xk, msfux, msfuy, msfvx_inv, msfvy, msftx, msfty, rdx, rdy):
    mkxm = (b.m(msfux) / b.m(msfuy)) * 0.5 * (b(xk) + b(xk, 0, 0, -1)) * 0.5 * (b.m(mut) + b.m(mut, 0, -1)) * rdx
    mkxp = (b.m(msfux, 0, 1) / b.m(msfuy, 0, 1)) * 0.5 * (b(xk, 0, 0, 1) + b(xk)) * 0.5 * (
        b.m(mut, 0, 1) + b.m(mut)) * rdx
    mrdx = b.m(msftx) * b.m(msfty) * rdx
    mkym = (b.m(msfvy) * b.m(msfvx_inv)) * 0.5 * (b(xk) + b(xk, 0, -1)) * 0.5 * (b.m(mut) + b.m(mut, -1)) * rdy
    mkyp = (b.m(msfvy, 1) * b.m(msfvx_inv, 1)) * 0.5 * (b(xk, 0, 1) + b(xk)) * 0.5 * (
        b.m(mut, 1) + b.m(mut)) * rdy
    mrdy = b.m(msftx) * b.m(msfty) * rdy
    return mkxm, mkxp, mrdx, mkym, mkyp, mrdy


def horizontal_diffusion_3dmp(name, field, tendency, mut, c1, c2,
                              base_3d,
                              msfux, msfuy, msfvx, msfvx_inv,
                              msfvy, msftx, msfty,
                              khdif, xkmhd, rdx, rdy,
                              ids, ide, jds, jde, kds, kde,
                              ims, ime, jms, jme, kms, kme,
                              its, ite, jts, jte, kts, kte):
    """Horizontal diffusion of field - base_3d.  wrf_dynamics.py:2841-2915."""
    ktf = min(kte, kde - 1)
    b = Box(kts, ktf, max(jds + 1, jts), min(jde - 2, jte), max(ids + 1, its), min(ide - 2, ite))
    mkxm, mkxp, mrdx, mkym, mkyp, mrdy = _scalar_coefs(b, mut, xkmhd, msfux, msfuy, msfvx_inv, msfvy,
                                                       msftx, msfty, rdx, rdy)
    f, s = field, base_3d
    b.put(tendency, b(tendency) + (
        mrdx * (mkxp * (b(f, 0, 0, 1) - b(f) - b(s, 0, 0, 1) + b(s)) - mkxm * (b(f) - b(f, 0, 0, -1) - b(s) + b(s, 0, 0, -1)))
        + mrdy * (mkyp * (b(f, 0, 1) - b(f) - b(s, 0, 1) + b(s)) - mkym * (b(f) - b(f, 0, -1) - b(s) + b(s, 0, -1)))))
    return tendency


def vertical_diffusion(name, field, tendency,
                       c1, c2,
                       alt, MUT, rdn, rdnw, kvdif,
                       ids, ide, jds, jde, kds, kde,
                       ims, ime, jms, jme, kms, kme,
                       its, ite, jts, jte, kts, kte):
    """Vertical diffusion of w ('w') or a scalar ('m').  wrf_dynamics.py:2917-2959."""
    ktf = min(kte, kde - 1)
    vflux = zeros_like3(field)
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    f = field
    if name == "w":
        b = Box(kts, ktf - 1, j0, j1, i0, i1)
        b.put(vflux, (kvdif / b(alt)) * b.c(rdnw) * (b(f, 1) - b(f)))
        vflux[ktf - 1, j0:j1, i0:i1] = 0.
        b = Box(kts + 1, ktf, j0, j1, i0, i1)
        b.put(tendency, b(tendency) + b.c(rdn) * g * g / b.m(MUT) / (0.5 * (b(alt) + b(alt, -1))) * (b(vflux) - b(vflux, -1)))
    elif name == "m":
        b = Box(kts, ktf - 1, j0, j1, i0, i1)
        b.put(vflux, kvdif * b.c(rdn, 1) / (0.5 * (b(alt) + b(alt, 1))) * (b(f, 1) - b(f)))
        vflux[kts, j0:j1, i0:i1] = vflux[kts + 1, j0:j1, i0:i1]
        vflux[ktf - 1, j0:j1, i0:i1] = 0.
        b = Box(kts + 1, ktf, j0, j1, i0, i1)
        b.put(tendency, b(tendency) + g * g / b.m(MUT) / b(alt) * b.c(rdnw) * (b(vflux) - b(vflux, -1)))
    return tendency


def vertical_diffusion_u(field, tendency,
                         u_base, c1h, c2h,
                         alt, muu, rdn, rdnw, kvdif,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """Vertical diffusion of u - u_base.  wrf_dynamics.py:2962-2991."""
    return _vdiff_uv(field, tendency, u_base, alt, muu, rdn, rdnw, kvdif,
                     jts, min(jte, jde - 1), max(ids + 1, its), min(ide - 1, ite), kte, kde, kts, (0, -1))


def vertical_diffusion_v(field, tendency,
                         v_base, c1h, c2h,
                         alt, muv, rdn, rdnw, kvdif,
                         ids, ide, jds, jde, kds, kde,
                         ims, ime, jms, jme, kms, kme,
                         its, ite, jts, jte, kts, kte):
    """Vertical diffusion of v - v_base.  wrf_dynamics.py:2994-3024."""
    return _vdiff_uv(field, tendency, v_base, alt, muv, rdn, rdnw, kvdif,
                     jts, min(jte, jde - 1), its, min(ite, ide - 1), kte, kde, kts, (-1, 0))


def _vdiff_uv(f, tendency, base, alt, mus, rdn, rdnw, kvdif, j0, j1, i0, i1, kte, kde, kts, d):
    ktf = min(kte, kde - 1)
    vflux = zeros_like3(f)
    b = Box(kts, ktf - 1, j0, j1, i0, i1)
    b.put(vflux, kvdif * b.c(rdn, 1) / (0.25 * (b(alt) + b(alt, 0, *d) + b(alt, 1) + b(alt, 1, *d)))
          * (b(f, 1) - b(f) - b.c(base, 1) + b.c(base)))
    vflux[ktf - 1, j0:j1, i0:i1] = 0.
    b = Box(kts + 1, ktf - 1, j0, j1, i0, i1)
    b.put(tendency, b(tendency) + g * g * b.c(rdnw) / b.m(mus) / (0.5 * (b(alt, 0, *d) + b(alt))) * (b(vflux) - b(vflux, -1)))
    return tendency


def vertical_diffusion_mp(field, tendency,
                          base, c1, c2,
                          alt, MUT, rdn, rdnw, kvdif,
                          ids, ide, jds, jde, kds, kde,
                          ims, ime, jms, jme, kms, kme,
                          its, ite, jts, jte, kts, kte):
    """Vertical diffusion of a scalar minus a 1-D base.  wrf_dynamics.py:3026-3055."""
    ktf = min(kte, kde - 1)
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    vflux = zeros_like3(field)
    b = Box(kts, ktf - 1, j0, j1, i0, i1)
    b.put(vflux, kvdif * b.c(rdn, 1) / (0.5 * (b(alt) + b(alt, 1)))
          * (b(field, 1) - b(field) - b.c(base, 1) + b.c(base)))
    vflux[ktf - 1, j0:j1, i0:i1] = 0.
    b = Box(kts + 1, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) + g * g * b.m(MUT) / b(alt) * b.c(rdnw) * (b(vflux) - b(vflux, -1)))
    return tendency


def vertical_diffusion_3dmp(field, tendency,
                            base_3d, c1, c2,
                            alt, MUT, rdn, rdnw, kvdif,
                            ids, ide, jds, jde, kds, kde,
                            ims, ime, jms, jme, kms, kme,
                            its, ite, jts, jte, kts, kte):
    """Vertical diffusion of a scalar minus a 3-D base.  wrf_dynamics.py:3058-3086."""
    ktf = min(kte, kde - 1)
    j0, j1, i0, i1 = jts, min(jte, jde - 1), its, min(ite, ide - 1)
    vflux = zeros_like3(field)
    b = Box(kts, ktf - 1, j0, j1, i0, i1)
    b.put(vflux, kvdif * b.c(rdn, 1) / (0.5 * (b(alt) + b(alt, 1)))
          * (b(field, 1) - b(field) - b(base_3d, 1) + b(base_3d)))
    vflux[ktf - 1, j0:j1, i0:i1] = 0.
    b = Box(kts + 1, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) + g * g * b.m(MUT) / b(alt) * b.c(rdnw) * (b(vflux) - b(vflux, -1)))
    return tendency


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
    """Fold the first-stage tendencies and the boundary relaxation (held in the
    *_save arrays) into the RK tendencies.  wrf_dynamics.py:3100-3148."""
    b = Box(kts, kte - 1, jts, jde - 1, its, ite)
    if rk_step == 1:
        b.put(ru_tendf, b(ru_tendf) + b(u_save) * b.m(msfuy))
    b.put(ru_tend, b(ru_tend) + b(ru_tendf) / b.m(msfuy))
    b = Box(kts, kte - 1, jts, jte, its, ide - 1)
    if rk_step == 1:
        b.put(rv_tendf, b(rv_tendf) + b(v_save) * b.m(msfvx))
    b.put(rv_tend, b(rv_tend) + b(rv_tendf) * b.m(msfvx_inv))
    b = Box(kts, kte, jts, jde - 1, its, ide - 1)
    if rk_step == 1:
        b.put(rw_tendf, b(rw_tendf) + b(w_save) * b.m(msfty))
    b.put(rw_tend, b(rw_tend) + b(rw_tendf) / b.m(msfty))
    if rk_step == 1:
        b.put(ph_tendf, b(ph_tendf) + b(ph_save))
    b.put(ph_tend, b(ph_tend) + b(ph_tendf) / b.m(msfty))
    b = Box(kts, kte - 1, jts, jde - 1, its, ide - 1)
    if rk_step == 1:
        b.put(t_tendf, b(t_tendf) + b(t_save))
    b.put(t_tend, b(t_tend) + b(t_tendf) / b.m(msfty) + b.m(mut) * b(h_diabatic) / b.m(msfty))
    mu_tend[jts:jde - 1, its:ide - 1] = mu_tend[jts:jde - 1, its:ide - 1] + mu_tendf[jts:jde - 1, its:ide - 1]
    return ru_tend, rv_tend, rw_tend, ph_tend, t_tend, mu_tend


def pbl_driver():
    """The boundary-layer hook: no arithmetic.  wrf_dynamics.py:413-415."""
    return None


def sixth_order_diffusion(name, field, tendency, MUT, dt,
                          config_flags, c1, c2,
                          diff_6th_opt, diff_6th_factor,
                          ids, ide, jds, jde, kds, kde,
                          ims, ime, jms, jme, kms, kme,
                          its, ite, jts, jte, kts, kte):
    """The 6th-order diffusion hook: forms its coefficient and returns None,
    leaving the tendency unchanged.  wrf_dynamics.py:3088-3097."""
    return None
