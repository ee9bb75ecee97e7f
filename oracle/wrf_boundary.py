"""Oracle: lateral boundary routines (test infrastructure only).

CPU restatement of pywrf/wrf_boundary.py:52-587 with the reference's
positional signatures.  Boundary arrays are the permuted views the reference
driver builds (solver.py:468-703): x-boundaries (nz, ny, 5), y-boundaries
(nz, 5, nx).
"""
import torch

from ._box import Box, zeros_like3


def _bounds(variable, ide, jde, kde, kte, itf_=None):
    ibs, ibe, jbs, jbe, ktf = None, ide - 1, None, jde - 1, kde - 1
    if variable == 'u':
        ibe = ide
    if variable == 'v':
        jbe = jde
    if variable in ('m', 'h'):
        ktf = kte
    return ibe, jbe, ktf


def relax_bdytend(field, field_tend,
                  field_bdy_xs, field_bdy_xe,
                  field_bdy_ys, field_bdy_ye,
                  field_bdy_tend_xs, field_bdy_tend_xe,
                  field_bdy_tend_ys, field_bdy_tend_ye,
                  variable_in,
                  spec_bdy_width, spec_zone, relax_zone,
                  dtbc, fcx, gcx,
                  ids, ide, jds, jde, kds, kde,
                  ims, ime, jms, jme, kms, kme,
                  ips, ipe, jps, jpe, kps, kpe,
                  its, ite, jts, jte, kts, kte):
    """Relaxation-zone tendency fcx*F - gcx*(5-point Laplacian of F), with
    F = bdy + dtbc*bdy_tend - field, on rows/columns b_dist = 1..3.
    wrf_boundary.py:52-161."""
    ibs, jbs = ids, jds
    ibe, jbe, ktf = _bounds(variable_in, ide, jde, kde, kte)
    ks = slice(kts, ktf)
    for b in (1, 2, 3):
        cols = lambda d: slice(ibs + b + d, ibe - b + d)
        F = lambda r, j, d: (field_bdy_ys[ks, r, cols(d)] + dtbc * field_bdy_tend_ys[ks, r, cols(d)]
                             - field[ks, j, cols(d)])
        j = jbs + b
        f0, f1, f2 = F(b, j, 0), F(b, j, -1), F(b, j, 1)
        f3, f4 = F(b - 1, j - 1, 0), F(b + 1, j + 1, 0)
        field_tend[ks, j, cols(0)] = field_tend[ks, j, cols(0)] + fcx[b] * f0 - gcx[b] * (f1 + f2 + f3 + f4 - 4. * f0)
    for b in (1, 2, 3):
        cols = lambda d: slice(ibs + b + d, ibe - b + d)
        F = lambda r, j, d: (field_bdy_ye[ks, r, cols(d)] + dtbc * field_bdy_tend_ye[ks, r, cols(d)]
                             - field[ks, j, cols(d)])
        j = jbe - 1 - b
        f0, f1, f2 = F(b, j, 0), F(b, j, -1), F(b, j, 1)
        f3, f4 = F(b - 1, j + 1, 0), F(b + 1, j - 1, 0)
        field_tend[ks, j, cols(0)] = field_tend[ks, j, cols(0)] + fcx[b] * f0 - gcx[b] * (f1 + f2 + f3 + f4 - 4. * f0)
    for b in (1, 2, 3):
        rows = lambda d: slice(jbs + b + 1 + d, jbe - b - 1 + d)
        F = lambda c, i, d: (field_bdy_xs[ks, rows(d), c] + dtbc * field_bdy_tend_xs[ks, rows(d), c]
                             - field[ks, rows(d), i])
        i = ibs + b
        f0, f1, f2 = F(b, i, 0), F(b, i, -1), F(b, i, 1)
        f3, f4 = F(b - 1, i - 1, 0), F(b + 1, i + 1, 0)
        field_tend[ks, rows(0), i] = field_tend[ks, rows(0), i] + fcx[b] * f0 - gcx[b] * (f1 + f2 + f3 + f4 - 4. * f0)
    for b in (1, 2, 3):
        rows = lambda d: slice(jbs + b + 1 + d, jbe - b - 1 + d)
        F = lambda c, i, d: (field_bdy_xe[ks, rows(d), c] + dtbc * field_bdy_tend_xe[ks, rows(d), c]
                             - field[ks, rows(d), i])
        i = ibe - 1 - b
        f0, f1, f2 = F(b, i, 0), F(b, i, -1), F(b, i, 1)
        f3, f4 = F(b - 1, i + 1, 0), F(b + 1, i - 1, 0)
        field_tend[ks, rows(0), i] = field_tend[ks, rows(0), i] + fcx[b] * f0 - gcx[b] * (f1 + f2 + f3 + f4 - 4. * f0)
    return field_tend


def spec_bdytend(field_tend,
                 field_bdy_xs, field_bdy_xe,
                 field_bdy_ys, field_bdy_ye,
                 field_bdy_tend_xs, field_bdy_tend_xe,
                 field_bdy_tend_ys, field_bdy_tend_ye,
                 variable_in,
                 spec_bdy_width, spec_zone,
                 ids, ide, jds, jde, kds, kde,
                 ims, ime, jms, jme, kms, kme,
                 ips, ipe, jps, jpe, kps, kpe,
                 its, ite, jts, jte, kts, kte):
    """Outermost row/column tendency = boundary tendency.  wrf_boundary.py:188-225."""
    ibs, jbs = ids, jds
    ibe, jbe, ktf = _bounds(variable_in, ide, jde, kde, kte)
    field_tend[kts:ktf, jbs, ibs:ibe] = field_bdy_tend_ys[kts:ktf, 0, ibs:ibe]
    field_tend[kts:ktf, jbe - 1, ibs:ibe] = field_bdy_tend_ye[kts:ktf, 0, ibs:ibe]
    field_tend[kts:ktf, jbs:jbe, ibs] = field_bdy_tend_xs[kts:ktf, jbs:jbe, 0]
    field_tend[kts:ktf, jbs:jbe, ibe - 1] = field_bdy_tend_xe[kts:ktf, jbs:jbe, 0]
    return field_tend


def spec_bdyupdate(field,
                   field_tend, dt,
                   variable,
                   spec_zone,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   ips, ipe, jps, jpe, kps, kpe,
                   its, ite, jts, jte, kts, kte):
    """field += dt*tend on the outermost rows/columns.  wrf_boundary.py:228-284."""
    variable = {'U': 'u', 'V': 'v', 'M': 'm', 'H': 'h'}.get(variable, variable)
    jbs = jds
    itf, jtf, ktf = min(ite, ide - 1), min(jte, jde - 1), kde - 1
    jbe = jde - 1
    if variable == 'u':
        itf = min(ite, ide)
    if variable == 'v':
        jbe, jtf = jde, min(jte, jde)
    if variable in ('m', 'h'):
        ktf = kte
    two_d = kds == 0 and kde == 0
    f, t = (field.unsqueeze(0), field_tend.unsqueeze(0)) if two_d else (field, field_tend)
    if two_d:
        ktf = 1
    f[kts:ktf, jts, its:itf] = f[kts:ktf, jts, its:itf] + dt * t[kts:ktf, jts, its:itf]
    f[kts:ktf, jtf - 1, its:itf] = f[kts:ktf, jtf - 1, its:itf] + dt * t[kts:ktf, jtf - 1, its:itf]
    f[kts:ktf, jbs + 1:jbe - 1, its] = f[kts:ktf, jbs + 1:jbe - 1, its] + dt * t[kts:ktf, jbs + 1:jbe - 1, its]
    f[kts:ktf, jbs + 1:jbe - 1, itf - 1] = f[kts:ktf, jbs + 1:jbe - 1, itf - 1] + dt * t[kts:ktf, jbs + 1:jbe - 1, itf - 1]
    return field


def spec_bdyupdate_ph(ph_save, field,
                      field_tend, mu_tend, muts,
                      c1, c2, dt,
                      variable_in,
                      spec_zone,
                      ids, ide, jds, jde, kds, kde,
                      ims, ime, jms, jme, kms, kme,
                      ips, ipe, jps, jpe, kps, kpe,
                      its, ite, jts, jte, kts, kte):
    """phi boundary update with the old/new column-mass ratio.
    wrf_boundary.py:287-358."""
    variable = {'U': 'u', 'V': 'v', 'M': 'm', 'H': 'h'}.get(variable_in, variable_in)
    ibs, jbs = ids, jds
    ibe, jbe, ktf = _bounds(variable, ide, jde, kde, kte)
    ks = slice(kts, ktf)
    mu_old = torch.zeros(muts.shape, dtype=muts.dtype)

    def upd(js, is_):
        mu_old[js, is_] = muts[js, is_] - dt * mu_tend[js, is_]
        mo, mt = mu_old[js, is_], muts[js, is_]
        field[ks, js, is_] = (field[ks, js, is_] * mo / mt + dt * field_tend[ks, js, is_] / mt
                              + ph_save[ks, js, is_] * (mo / mt - 1.))
    upd(jts, slice(ibs, ibe))
    upd(jbe - 1, slice(ibs, ibe))
    upd(slice(jbs + 1, jbe - 1), ibs)
    upd(slice(jbs + 1, jbe - 1), ibe - 1)
    return field


def zero_grad_bdy(field,
                  variable_in,
                  spec_zone,
                  ids, ide, jds, jde, kds, kde,
                  ims, ime, jms, jme, kms, kme,
                  ips, ipe, jps, jpe, kps, kpe,
                  its, ite, jts, jte, kts, kte):
    """Zero-gradient copy onto the outermost rows/columns.  wrf_boundary.py:361-405."""
    variable = {'U': 'u', 'V': 'v'}.get(variable_in, variable_in)
    ibs, jbs = ids, jds
    ibe, jbe, ktf = ide - 1, jde - 1, kde - 1
    if variable == 'u':
        ibe = ide
    if variable == 'v':
        jbe = jde
    if variable == 'w':
        ktf = kde
    ks = slice(kts, ktf)
    field[ks, jbs, ibs] = field[ks, jbs + 1, ibs + 1]
    field[ks, jbs, ibe - 1] = field[ks, jbs + 1, ibe - 2]
    field[ks, jbs, ibs + 1:ibe - 1] = field[ks, jbs + 1, ibs + 1:ibe - 1]
    field[ks, jbe - 1, ibs] = field[ks, jbe - 2, ibs + 1]
    field[ks, jbe - 1, ibe - 1] = field[ks, jbe - 2, ibe - 2]
    field[ks, jbe - 1, ibs + 1:ibe - 1] = field[ks, jbe - 2, ibs + 1:ibe - 1]
    field[ks, jbs + 1:jbe - 1, ibs] = field[ks, jbs + 1:jbe - 1, ibs + 1]
    field[ks, jbs + 1:jbe - 1, ibe - 1] = field[ks, jbs + 1:jbe - 1, ibe - 2]
    return field


def mass_weight(field, mut, rfield, c1, c2,
                ids, ide, jds, jde, kds, kde,
                ims, ime, jms, jme, kms, kme,
                irs, ire, jrs, jre, krs, kre,
                its, ite, jts, jte, kts, kte):
    """rfield = field * mut on the given tile.  wrf_boundary.py:408-415."""
    rfield[kts:kte, jts:jte, its:ite] = field[kts:kte, jts:jte, its:ite] * mut[jts:jte, its:ite]
    return rfield


def relax_bdy_scalar(scalar_tend,
                     scalar, mu, c1h, c2h,
                     scalar_bxs, scalar_bxe, scalar_bys, scalar_bye,
                     scalar_btxs, scalar_btxe, scalar_btys, scalar_btye,
                     spec_bdy_width, spec_zone, relax_zone,
                     dtbc, fcx, gcx,
                     ids, ide, jds, jde, kds, kde,
                     ims, ime, jms, jme, kms, kme,
                     ips, ipe, jps, jpe, kps, kpe,
                     its, ite, jts, jte, kts, kte):
    """Relaxation of a moist scalar: mass-weight then relax_bdytend('q').
    wrf_boundary.py:418-451."""
    i0, i1 = max(its - 1, ids), min(ite + 1, ide - 1)
    j0, j1 = max(jts - 1, jds), min(jte + 1, jde - 1)
    rscalar = zeros_like3(scalar)
    rscalar = mass_weight(scalar, mu, rscalar, c1h, c2h, ids, ide, jds, jde, kds, kde,
                          ims, ime, jms, jme, kms, kme, ims, ime, jms, jme, kms, kme,
                          i0, i1, j0, j1, kts, kte - 1)
    return relax_bdytend(rscalar, scalar_tend, scalar_bxs, scalar_bxe, scalar_bys, scalar_bye,
                         scalar_btxs, scalar_btxe, scalar_btys, scalar_btye, 'q',
                         spec_bdy_width, spec_zone, relax_zone, dtbc, fcx, gcx,
                         ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme,
                         ips, ipe, jps, jpe, kps, kpe, its, ite, jts, jte, kts, kte)


def spec_bdy_scalar(scalar_tend,
                    scalar_bxs, scalar_bxe, scalar_bys, scalar_bye,
                    scalar_btxs, scalar_btxe, scalar_btys, scalar_btye,
                    spec_bdy_width, spec_zone,
                    ids, ide, jds, jde, kds, kde,
                    ims, ime, jms, jme, kms, kme,
                    ips, ipe, jps, jpe, kps, kpe,
                    its, ite, jts, jte, kts, kte):
    """Specified-zone tendency of a moist scalar.  wrf_boundary.py:454-472."""
    return spec_bdytend(scalar_tend, scalar_bxs, scalar_bxe, scalar_bys, scalar_bye,
                        scalar_btxs, scalar_btxe, scalar_btys, scalar_btye, 'q',
                        spec_bdy_width, spec_zone, ids, ide, jds, jde, kds, kde,
                        ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps, kpe,
                        its, ite, jts, jte, kts, kte)


def flow_dep_bdy(field,
                 u, v,
                 spec_zone,
                 ids, ide, jds, jde, kds, kde,
                 ims, ime, jms, jme, kms, kme,
                 ips, ipe, jps, jpe, kps, kpe,
                 its, ite, jts, jte, kts, kte):
    """Inflow zero / outflow zero-gradient on the outermost rows/columns.
    wrf_boundary.py:475-510."""
    itf, jtf, ktf = min(ite, ide - 1), min(jte, jde - 1), kde - 1
    ks = slice(kts, ktf)
    z = torch.zeros((), dtype=field.dtype)
    field[ks, jts, its:itf] = torch.where(v[ks, jts, its:itf] < 0., field[ks, jts + 1, its:itf], z)
    field[ks, jts, its] = field[ks, jts + 1, its + 1] + 0.0
    field[ks, jts, itf - 1] = field[ks, jts + 1, itf - 2] + 0.0
    field[ks, jtf - 1, its:itf] = torch.where(v[ks, jtf, its:itf] > 0., field[ks, jtf - 2, its:itf], z)
    field[ks, jtf - 1, its] = field[ks, jtf - 2, its + 1] + 0.0
    field[ks, jtf - 1, itf - 1] = field[ks, jtf - 2, itf - 2] + 0.0
    field[ks, jts:jtf, its] = torch.where(u[ks, jts:jtf, its] < 0., field[ks, jts:jtf, its + 1], z)
    field[ks, jts, its] = field[ks, jts + 1, its + 1] + 0.0
    field[ks, jtf - 1, its] = field[ks, jtf - 2, its + 1] + 0.0
    field[ks, jts:jtf, itf - 1] = torch.where(u[ks, jts:jtf, itf] > 0., field[ks, jts:jtf, itf - 2], z)
    field[ks, jts, itf - 1] = field[ks, jts + 1, itf - 2] + 0.0
    field[ks, jtf - 1, itf - 1] = field[ks, jtf - 2, itf - 2] + 0.0
    return field


def spec_bdy_final(field, mu, c1, c2, msf,
                   field_bdy_xs, field_bdy_xe,
                   field_bdy_ys, field_bdy_ye,
                   field_bdy_tend_xs, field_bdy_tend_xe,
                   field_bdy_tend_ys, field_bdy_tend_ye,
                   variable_in,
                   spec_bdy_width, spec_zone,
                   dtbc,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   ips, ipe, jps, jpe, kps, kpe,
                   its, ite, jts, jte, kts, kte):
    """End-of-step specified values msf*(bdy + dtbc*tend)/mu on the outermost
    rows/columns.  wrf_boundary.py:513-587."""
    variable = {'U': 'u', 'V': 'v', 'W': 'w', 'M': 'm', 'T': 't', 'H': 'h'}.get(variable_in, variable_in)
    jbs = jds
    ibe, jbe = ide - 1, jde - 1
    itf, jtf, ktf = min(ite, ide - 1), min(jte, jde - 1), kde - 1
    if variable == 'u':
        ibe, itf = ide, min(ite, ide)
    if variable == 'v':
        jbe, jtf = jde, min(jte, jde)
    if variable in ('m', 'h', 'w'):
        ktf = kde
    ks = slice(kts, ktf)
    one = torch.ones((), dtype=field.dtype)
    xmsf = msf if variable in ('u', 'v', 'w') else None
    xmu = mu if variable != 'm' else None
    S = lambda a, js, is_: one if a is None else a[js, is_]

    def fin(js, is_, bdy, tend):
        field[ks, js, is_] = S(xmsf, js, is_) * (bdy + dtbc * tend) / S(xmu, js, is_)
    fin(jts, slice(its, ibe), field_bdy_ys[ks, 0, its:ibe], field_bdy_tend_ys[ks, 0, its:ibe])
    fin(jtf - 1, slice(its, ibe), field_bdy_ye[ks, 0, its:ibe], field_bdy_tend_ye[ks, 0, its:ibe])
    fin(slice(jts, jbe), its, field_bdy_xs[ks, jts:jbe, 0], field_bdy_tend_xs[ks, jts:jbe, 0])
    fin(slice(jts, jbe), itf - 1, field_bdy_xe[ks, jts:jbe, 0], field_bdy_tend_xe[ks, jts:jbe, 0])
    return field
