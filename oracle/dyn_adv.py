"""Oracle: flux-form 5th/3rd-order advection (test infrastructure only).

CPU restatement of pywrf/wrf_dynamics.py:1112-2028 (advect_u, advect_v,
advect_w, advect_scalar) including the boundary-row order reductions and the
sign-source quirks of the 3rd-order boundary fluxes.
"""
from ._box import Box, zeros_like3, flux5, flux3


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
    """u advection tendency.  wrf_dynamics.py:1112-1307."""
    ktf = min(kte, kde - 1)
    vel, fq = zeros_like3(u), zeros_like3(u)
    # y flux at faces j (between rows j-1 and j), u columns [ids+1, ide-1)
    i0, i1 = max(ids + 1, its), min(ide - 1, ite)
    b = Box(kts, ktf, jds + 3, jde - 3, i0, i1)
    b.put(vel, 0.5 * (b(rv) + b(rv, 0, 0, -1)))
    b.put(fq, b(vel) * flux5(b(u, 0, -3), b(u, 0, -2), b(u, 0, -1), b(u), b(u, 0, 1), b(u, 0, 2), b(vel)))
    r = Box(kts, ktf, jds + 1, jds + 2, i0, i1)
    r.put(fq, 0.25 * (r(rv) + r(rv, 0, 0, -1)) * (r(u) + r(u, 0, -1)))
    r = Box(kts, ktf, jds + 2, jds + 3, i0, i1)
    r.put(vel, 0.5 * (r(rv) + r(rv, 0, 0, -1)))
    r.put(fq, r(vel) * flux3(r(u, 0, -2), r(u, 0, -1), r(u), r(u, 0, 1), r(vel)))
    r = Box(kts, ktf, jde - 2, jde - 1, i0, i1)
    r.put(fq, 0.25 * (r(rv) + r(rv, 0, 0, -1)) * (r(u) + r(u, 0, -1)))
    r = Box(kts, ktf, jde - 3, jde - 2, i0, i1)
    r.put(vel, 0.5 * (r(rv) + r(rv, 0, 0, -1)))
    # sign taken from vel at row jde-2, which this routine never sets
    r.put(fq, r(vel) * flux3(r(u, 0, -2), r(u, 0, -1), r(u), r(u, 0, 1), r(vel, 0, 1)))
    j0, j1 = max(jts, jds + 1), min(jte, jde - 2)
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - (b.m(msfux) * rdy) * (b(fq, 0, 1) - b(fq)))
    # x flux at faces i (between columns i-1 and i)
    j0, j1 = jts, min(jte, jde - 1)
    fq = zeros_like3(u)
    b = Box(kts, ktf, j0, j1, ids + 3, ide - 2)
    b.put(vel, 0.5 * (b(ru) + b(ru, 0, 0, -1)))
    b.put(fq, b(vel) * flux5(b(u, 0, 0, -3), b(u, 0, 0, -2), b(u, 0, 0, -1), b(u), b(u, 0, 0, 1), b(u, 0, 0, 2), b(vel)))
    c = Box(kts, ktf, j0, j1, ids + 1, ids + 2)
    ub = c(u).where(c(u) < 0., c(u, 0, 0, -1))
    c.put(fq, 0.25 * (c(ru) + c(ru, 0, 0, -1)) * (c(u) + ub))
    c = Box(kts, ktf, j0, j1, ids + 2, ids + 3)
    c.put(vel, 0.5 * (c(ru) + c(ru, 0, 0, -1)))
    c.put(fq, c(vel) * flux3(c(u, 0, 0, -2), c(u, 0, 0, -1), c(u), c(u, 0, 0, 1), c(vel)))
    c = Box(kts, ktf, j0, j1, ide - 1, ide)
    ub = c(u, 0, 0, -1).where(c(u, 0, 0, -1) > 0., c(u))
    c.put(fq, 0.25 * (c(ru) + c(ru, 0, 0, -1)) * (c(u, 0, 0, -1) + ub))
    c = Box(kts, ktf, j0, j1, ide - 2, ide - 1)
    c.put(vel, 0.5 * (c(ru) + c(ru, 0, 0, -1)))
    c.put(fq, c(vel) * flux3(c(u, 0, 0, -2), c(u, 0, 0, -1), c(u), c(u, 0, 0, 1), c(vel)))
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - (b.m(msfux) * rdx) * (b(fq, 0, 0, 1) - b(fq)))
    # z flux at w levels k
    fq = zeros_like3(u)
    b = Box(kts + 3, ktf - 2, j0, j1, i0, i1)
    b.put(vel, 0.5 * (b(rom, 0, 0, -1) + b(rom)))
    b.put(fq, b(vel) * flux5(b(u, -3), b(u, -2), b(u, -1), b(u), b(u, 1), b(u, 2), -b(vel)))
    k = Box(kts + 1, kts + 2, j0, j1, i0, i1)
    k.put(fq, 0.5 * (k(rom) + k(rom, 0, 0, -1)) * (k.c(fzm) * k(u) + k.c(fzp) * k(u, -1)))
    k = Box(kts + 2, kts + 3, j0, j1, i0, i1)
    k.put(vel, 0.5 * (k(rom) + k(rom, 0, 0, -1)))
    k.put(fq, k(vel) * flux3(k(u, -2), k(u, -1), k(u), k(u, 1), -k(vel)))
    k = Box(ktf - 2, ktf - 1, j0, j1, i0, i1)
    k.put(vel, 0.5 * (k(rom) + k(rom, 0, 0, -1)))
    k.put(fq, k(vel) * flux3(k(u, -2), k(u, -1), k(u), k(u, 1), -k(vel)))
    k = Box(ktf - 1, ktf, j0, j1, i0, i1)
    k.put(fq, 0.5 * (k(rom) + k(rom, 0, 0, -1)) * (k.c(fzm) * k(u) + k.c(fzp) * k(u, -1)))
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - b.c(rdzw) * (b(fq, 1) - b(fq)))
    return tendency


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
    """v advection tendency.  wrf_dynamics.py:1332-1532."""
    ktf = min(kte, kde - 1)
    vel, fq = zeros_like3(v), zeros_like3(v)
    i0, i1 = its, min(ite, ide - 1)
    b = Box(kts, ktf, jds + 3, jde - 2, i0, i1)
    b.put(vel, 0.5 * (b(rv) + b(rv, 0, -1)))
    b.put(fq, b(vel) * flux5(b(v, 0, -3), b(v, 0, -2), b(v, 0, -1), b(v), b(v, 0, 1), b(v, 0, 2), b(vel)))
    r = Box(kts, ktf, jds + 1, jds + 2, i0, i1)
    vb = r(v).where(r(v) < 0., r(v, 0, -1))
    r.put(fq, 0.25 * (r(rv) + r(rv, 0, -1)) * (r(v) + vb))
    r = Box(kts, ktf, jds + 2, jds + 3, i0, i1)
    r.put(vel, 0.5 * (r(rv) + r(rv, 0, -1)))
    r.put(fq, r(vel) * flux3(r(v, 0, -2), r(v, 0, -1), r(v), r(v, 0, 1), r(vel)))
    r = Box(kts, ktf, jde - 1, jde, i0, i1)
    vb = r(v, 0, -1).where(r(v, 0, -1) > 0., r(v))
    r.put(fq, 0.25 * (r(rv) + r(rv, 0, -1)) * (r(v, 0, -1) + vb))
    r = Box(kts, ktf, jde - 2, jde - 1, i0, i1)
    r.put(vel, 0.5 * (r(rv) + r(rv, 0, -1)))
    r.put(fq, r(vel) * flux3(r(v, 0, -2), r(v, 0, -1), r(v), r(v, 0, 1), r(vel)))
    j0, j1 = max(jts, jds + 1), min(jte, jde - 1)
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - (b.m(msfvy) * rdy) * (b(fq, 0, 1) - b(fq)))
    # x flux
    fq = zeros_like3(v)
    b = Box(kts, ktf, j0, j1, ids + 3, ide - 3)
    b.put(vel, 0.5 * (b(ru) + b(ru, 0, -1)))
    b.put(fq, b(vel) * flux5(b(v, 0, 0, -3), b(v, 0, 0, -2), b(v, 0, 0, -1), b(v), b(v, 0, 0, 1), b(v, 0, 0, 2), b(vel)))
    c = Box(kts, ktf, j0, j1, ids + 1, ids + 2)
    c.put(fq, 0.25 * (c(ru) + c(ru, 0, -1)) * (c(v) + c(v, 0, 0, -1)))
    c = Box(kts, ktf, j0, j1, ids + 2, ids + 3)
    c.put(vel, 0.5 * (c(ru) + c(ru, 0, -1)))
    c.put(fq, c(vel) * flux3(c(v, 0, 0, -2), c(v, 0, 0, -1), c(v), c(v, 0, 0, 1), c(vel)))
    c = Box(kts, ktf, j0, j1, ide - 2, ide - 1)
    c.put(fq, 0.25 * (c(ru) + c(ru, 0, -1)) * (c(v) + c(v, 0, 0, -1)))
    c = Box(kts, ktf, j0, j1, ide - 3, ide - 2)
    c.put(vel, 0.5 * (c(ru) + c(ru, 0, -1)))
    c.put(fq, c(vel) * flux3(c(v, 0, 0, -2), c(v, 0, 0, -1), c(v), c(v, 0, 0, 1), c(vel)))
    b = Box(kts, ktf, j0, j1, max(ids + 1, its), min(ide - 2, ite))
    b.put(tendency, b(tendency) - (b.m(msfvy) * rdx) * (b(fq, 0, 0, 1) - b(fq)))
    # z flux
    fq = zeros_like3(v)
    b = Box(kts + 3, ktf - 2, j0, j1, i0, i1)
    b.put(vel, 0.5 * (b(rom) + b(rom, 0, -1)))
    b.put(fq, b(vel) * flux5(b(v, -3), b(v, -2), b(v, -1), b(v), b(v, 1), b(v, 2), -b(vel)))
    k = Box(kts + 1, kts + 2, j0, j1, i0, i1)
    k.put(fq, 0.5 * (k(rom) + k(rom, 0, -1)) * (k.c(fzm) * k(v) + k.c(fzp) * k(v, -1)))
    k = Box(kts + 2, kts + 3, j0, j1, i0, i1)
    k.put(vel, 0.5 * (k(rom) + k(rom, 0, -1)))
    k.put(fq, k(vel) * flux3(k(v, -2), k(v, -1), k(v), k(v, 1), -k(vel)))
    k = Box(ktf - 2, ktf - 1, j0, j1, i0, i1)
    k.put(vel, 0.5 * (k(rom) + k(rom, 0, -1)))
    k.put(fq, k(vel) * flux3(k(v, -2), k(v, -1), k(v), k(v, 1), k(vel)))     # +vel as in the reference
    k = Box(ktf - 1, ktf, j0, j1, i0, i1)
    k.put(fq, 0.5 * (k(rom) + k(rom, 0, -1)) * (k.c(fzm) * k(v) + k.c(fzp) * k(v, -1)))
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - (b.m(msfvy) / b.m(msfvx)) * b.c(rdzw) * (b(fq, 1) - b(fq)))
    return tendency


def _wtop(rx, fzm, fzp, ktf, box):
    """Mass flux extrapolated to the model top, (2-fzm)*r[ktf-1] - fzp*r[ktf-2]."""
    return (2. - fzm[ktf - 1]) * box(rx, -1) - fzp[ktf - 1] * box(rx, -2)


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
    """w advection tendency on full levels.  wrf_dynamics.py:1535-1822."""
    ktf = min(kte, kde - 1)
    vel, fq = zeros_like3(w), zeros_like3(w)
    i0, i1 = its, min(ite, ide - 1)

    def ybox(j0, j1):
        return Box(kts + 1, ktf, j0, j1, i0, i1), Box(ktf, ktf + 1, j0, j1, i0, i1)

    lo, top = ybox(jds + 3, jde - 3)
    lo.put(vel, lo.c(fzm) * lo(rv) + lo.c(fzp) * lo(rv, -1))
    lo.put(fq, lo(vel) * flux5(lo(w, 0, -3), lo(w, 0, -2), lo(w, 0, -1), lo(w), lo(w, 0, 1), lo(w, 0, 2), lo(vel)))
    top.put(vel, _wtop(rv, fzm, fzp, ktf, top))
    top.put(fq, top(vel) * flux5(top(w, 0, -3), top(w, 0, -2), top(w, 0, -1), top(w), top(w, 0, 1), top(w, 0, 2), top(vel)))
    for jr, jn in ((jds + 1, -1), (jde - 2, -1)):
        lo, top = ybox(jr, jr + 1)
        lo.put(fq, 0.5 * (lo.c(fzm) * lo(rv) + lo.c(fzp) * lo(rv, -1)) * (lo(w) + lo(w, 0, jn)))
        top.put(fq, 0.5 * _wtop(rv, fzm, fzp, ktf, top) * (top(w) + top(w, 0, jn)))
    for jr in (jds + 2, jde - 3):
        lo, top = ybox(jr, jr + 1)
        lo.put(vel, lo.c(fzm) * lo(rv) + lo.c(fzp) * lo(rv, -1))
        lo.put(fq, lo(vel) * flux3(lo(w, 0, -2), lo(w, 0, -1), lo(w), lo(w, 0, 1), lo(vel)))
        top.put(vel, _wtop(rv, fzm, fzp, ktf, top))
        top.put(fq, top(vel) * flux3(top(w, 0, -2), top(w, 0, -1), top(w), top(w, 0, 1), top(vel)))
    j0, j1 = max(jts, jds + 1), min(jte, jde - 2)
    b = Box(kts + 1, ktf + 1, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - (b.m(msftx) * rdy) * (b(fq, 0, 1) - b(fq)))
    # x flux
    j0, j1 = jts, min(jte, jde - 1)
    fq = zeros_like3(w)

    def xbox(c0, c1):
        return Box(kts + 1, ktf, j0, j1, c0, c1), Box(ktf, ktf + 1, j0, j1, c0, c1)

    lo, top = xbox(ids + 3, ide - 3)
    lo.put(vel, lo.c(fzm) * lo(ru) + lo.c(fzp) * lo(ru, -1))
    lo.put(fq, lo(vel) * flux5(lo(w, 0, 0, -3), lo(w, 0, 0, -2), lo(w, 0, 0, -1), lo(w), lo(w, 0, 0, 1), lo(w, 0, 0, 2), lo(vel)))
    top.put(vel, _wtop(ru, fzm, fzp, ktf, top))
    top.put(fq, top(vel) * flux5(top(w, 0, 0, -3), top(w, 0, 0, -2), top(w, 0, 0, -1), top(w), top(w, 0, 0, 1),
                                 top(w, 0, 0, 2), top(vel)))
    for ic in (ids + 1, ide - 2):
        lo, top = xbox(ic, ic + 1)
        lo.put(fq, 0.5 * (lo.c(fzm) * lo(ru) + lo.c(fzp) * lo(ru, -1)) * (lo(w) + lo(w, 0, 0, -1)))
        top.put(fq, 0.5 * _wtop(ru, fzm, fzp, ktf, top) * (top(w) + top(w, 0, 0, -1)))
    for ic in (ids + 2, ide - 3):
        lo, top = xbox(ic, ic + 1)
        lo.put(vel, lo.c(fzm) * lo(ru) + lo.c(fzp) * lo(ru, -1))
        lo.put(fq, lo(vel) * flux3(lo(w, 0, 0, -2), lo(w, 0, 0, -1), lo(w), lo(w, 0, 0, 1), lo(vel)))
        top.put(vel, _wtop(ru, fzm, fzp, ktf, top))
        top.put(fq, top(vel) * flux3(top(w, 0, 0, -2), top(w, 0, 0, -1), top(w), top(w, 0, 0, 1), top(vel)))
    b = Box(kts + 1, ktf + 1, j0, j1, max(ids + 1, its), min(ide - 2, ite))
    b.put(tendency, b(tendency) - (b.m(msftx) * rdx) * (b(fq, 0, 0, 1) - b(fq)))
    # z flux at mass levels (index k = level between w[k-1] and w[k])
    fq = zeros_like3(w)
    b = Box(kts + 3, ktf - 1, j0, j1, i0, i1)
    b.put(vel, 0.5 * (b(rom) + b(rom, -1)))
    b.put(fq, b(vel) * flux5(b(w, -3), b(w, -2), b(w, -1), b(w), b(w, 1), b(w, 2), -b(vel)))
    k = Box(kts + 1, kts + 2, j0, j1, i0, i1)
    k.put(fq, 0.25 * (k(rom) + k(rom, -1)) * (k(w) + k(w, -1)))
    k = Box(kts + 2, kts + 3, j0, j1, i0, i1)
    k.put(vel, 0.5 * (k(rom) + k(rom, -1)))
    k.put(fq, k(vel) * flux3(k(w, -2), k(w, -1), k(w), k(w, 1), -k(vel)))
    k = Box(ktf - 1, ktf, j0, j1, i0, i1)
    k.put(vel, 0.5 * (k(rom) + k(rom, -1)))
    k.put(fq, k(vel) * flux3(k(w, -2), k(w, -1), k(w), k(w, 1), k(vel)))     # +vel as in the reference
    k = Box(ktf, ktf + 1, j0, j1, i0, i1)
    k.put(fq, 0.25 * (k(rom) + k(rom, -1)) * (k(w) + k(w, -1)))
    b = Box(kts + 1, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - b.c(rdzu) * (b(fq, 1) - b(fq)))
    t = Box(ktf, ktf + 1, j0, j1, i0, i1)
    t.put(tendency, t(tendency) + 2. * rdzu[ktf - 1] * t(fq))
    return tendency


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
    """Mass-point scalar advection tendency.  wrf_dynamics.py:1841-2028."""
    ktf = min(kte, kde - 1)
    f = field
    fq = zeros_like3(f)
    i0, i1 = its, min(ite, ide - 1)
    b = Box(kts, ktf, jds + 3, jde - 3, i0, i1)
    b.put(fq, b(rv) * flux5(b(f, 0, -3), b(f, 0, -2), b(f, 0, -1), b(f), b(f, 0, 1), b(f, 0, 2), b(rv)))
    for jr in (jds + 1, jde - 2):
        r = Box(kts, ktf, jr, jr + 1, i0, i1)
        r.put(fq, 0.5 * r(rv) * (r(f) + r(f, 0, -1)))
    for jr in (jds + 2, jde - 3):
        r = Box(kts, ktf, jr, jr + 1, i0, i1)
        r.put(fq, r(rv) * flux3(r(f, 0, -2), r(f, 0, -1), r(f), r(f, 0, 1), r(rv)))
    j0, j1 = max(jts, jds + 1), min(jte, jde - 2)
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - (b.m(msftx) * rdy) * (b(fq, 0, 1) - b(fq)))
    j0, j1 = jts, min(jte, jde - 1)
    fq = zeros_like3(f)
    b = Box(kts, ktf, j0, j1, ids + 3, ide - 3)
    b.put(fq, b(ru) * flux5(b(f, 0, 0, -3), b(f, 0, 0, -2), b(f, 0, 0, -1), b(f), b(f, 0, 0, 1), b(f, 0, 0, 2), b(ru)))
    for ic in (ids + 1, ide - 2):
        c = Box(kts, ktf, j0, j1, ic, ic + 1)
        c.put(fq, 0.5 * c(ru) * (c(f) + c(f, 0, 0, -1)))
    c = Box(kts, ktf, j0, j1, ids + 2, ids + 3)
    c.put(fq, c(ru) * flux3(c(f, 0, 0, -2), c(f, 0, 0, -1), c(f), c(f, 0, 0, 1), c(ru)))
    c = Box(kts, ktf, j0, j1, ide - 3, ide - 2)
    # sign taken from ru at column ide-2, as in the reference
    c.put(fq, c(ru) * flux3(c(f, 0, 0, -2), c(f, 0, 0, -1), c(f), c(f, 0, 0, 1), c(ru, 0, 0, 1)))
    b = Box(kts, ktf, j0, j1, max(ids + 1, its), min(ide - 2, ite))
    b.put(tendency, b(tendency) - (b.m(msftx) * rdx) * (b(fq, 0, 0, 1) - b(fq)))
    fq = zeros_like3(f)
    b = Box(kts + 3, ktf - 2, j0, j1, i0, i1)
    b.put(fq, b(rom) * flux5(b(f, -3), b(f, -2), b(f, -1), b(f), b(f, 1), b(f, 2), -b(rom)))
    for kk in (kts + 1, ktf - 1):
        k = Box(kk, kk + 1, j0, j1, i0, i1)
        k.put(fq, k(rom) * (k.c(fzm) * k(f) + k.c(fzp) * k(f, -1)))
    for kk in (kts + 2, ktf - 2):
        k = Box(kk, kk + 1, j0, j1, i0, i1)
        k.put(fq, k(rom) * flux3(k(f, -2), k(f, -1), k(f), k(f, 1), -k(rom)))
    b = Box(kts, ktf, j0, j1, i0, i1)
    b.put(tendency, b(tendency) - b.c(rdzw) * (b(fq, 1) - b(fq)))
    return tendency
