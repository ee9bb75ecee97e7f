"""Drop-in for pywrf/wrf_boundary.py on the ARW step's path (same names,
positional parameters and returns), forwarding to libpywrf_amd.so.

Boundary arrays are the reference driver's: x-boundaries (kz, ny, 5) and
y-boundaries (kz, 5, nx), as the permuted views of (5, kz, ny) / (5, kz, nx)
arrays it builds (solver.py:468-703), which the kernels read in place; arrays
in another layout are copied into that one (see _lib.boundary_layout).  kz is
the level count of the field they belong to.  Fields are updated in place, as
in the reference.
"""
import torch

from ._lib import call, grid, ptr, flag, stream, bdy_ptrs, check_operands


def _width(spec_bdy_width):
    if int(spec_bdy_width) != 5:
        raise IndexError("spec_bdy_width %r: the boundary arrays of the kernels are 5 points wide"
                         % (spec_bdy_width,))


def _bdy(g, arrs):
    gc = g.contents
    return bdy_ptrs(arrs, gc.nz, gc.ny, gc.nx)


def _on_device(dtbc):
    """True when the boundary time is a device tensor (one float, read by the
    kernel when it runs: the form a captured graph of the step uses)."""
    if isinstance(dtbc, torch.Tensor):
        if dtbc.numel() != 1:
            raise TypeError("dtbc: expected one value, got shape %s" % (tuple(dtbc.shape),))
        return True
    return False


def _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps, kpe,
         its, ite, jts, jte, kts, kte):
    return (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps, kpe,
            its, ite, jts, jte, kts, kte)


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
    """wrf_boundary.py:52-161"""
    _width(spec_bdy_width)
    g = grid(field.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                               kpe, its, ite, jts, jte, kts, kte))
    bp = _bdy(g, (field_bdy_xs, field_bdy_xe, field_bdy_ys, field_bdy_ye, field_bdy_tend_xs, field_bdy_tend_xe,
                  field_bdy_tend_ys, field_bdy_tend_ye))
    if _on_device(dtbc):
        call("pywrf_relax_bdytend_dtbc_dev", g, ptr(field), ptr(field_tend), *bp, flag(variable_in), ptr(dtbc),
             ptr(fcx), ptr(gcx), stream())
    else:
        call("pywrf_relax_bdytend", g, ptr(field), ptr(field_tend), *bp, flag(variable_in), float(dtbc), ptr(fcx),
             ptr(gcx), stream())
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
    """wrf_boundary.py:188-225"""
    _width(spec_bdy_width)
    g = grid(field_tend.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe,
                                    kps, kpe, its, ite, jts, jte, kts, kte))
    _bdy(g, (field_bdy_xs, field_bdy_xe, field_bdy_ys, field_bdy_ye))           # shape checks only
    bp = _bdy(g, (field_bdy_tend_xs, field_bdy_tend_xe, field_bdy_tend_ys, field_bdy_tend_ye))
    call("pywrf_spec_bdytend", g, ptr(field_tend), *bp, flag(variable_in), stream())
    return field_tend


def spec_bdyupdate(field,
                   field_tend, dt,
                   variable,
                   spec_zone,
                   ids, ide, jds, jde, kds, kde,
                   ims, ime, jms, jme, kms, kme,
                   ips, ipe, jps, jpe, kps, kpe,
                   its, ite, jts, jte, kts, kte):
    """wrf_boundary.py:228-284"""
    g = grid(field.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                               kpe, its, ite, jts, jte, kts, kte))
    call("pywrf_spec_bdyupdate", g, ptr(field), ptr(field_tend), float(dt), flag(variable), stream())
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
    """wrf_boundary.py:287-358"""
    g = grid(field.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                               kpe, its, ite, jts, jte, kts, kte))
    call("pywrf_spec_bdyupdate_ph", g, ptr(ph_save), ptr(field), ptr(field_tend), ptr(mu_tend), ptr(muts), float(dt),
         flag(variable_in), stream())
    return field


def zero_grad_bdy(field,
                  variable_in,
                  spec_zone,
                  ids, ide, jds, jde, kds, kde,
                  ims, ime, jms, jme, kms, kme,
                  ips, ipe, jps, jpe, kps, kpe,
                  its, ite, jts, jte, kts, kte):
    """wrf_boundary.py:361-405"""
    g = grid(field.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                               kpe, its, ite, jts, jte, kts, kte))
    call("pywrf_zero_grad_bdy", g, ptr(field), flag(variable_in), stream())
    return field


def spec_bdyupdate_ph_zero_grad(ph_args, zg_args):
    """spec_bdyupdate_ph then zero_grad_bdy (wrf_boundary.py:287-358,
    361-405) on two different arrays, each given as the tuple of its
    positional arguments, in one launch; the index sets must agree.  Returns
    the two calls' results.  Not part of the reference API."""
    (ph_save, field, field_tend, mu_tend, muts, c1, c2, dt, variable_in, spec_zone, *ix) = ph_args
    (zg_field, zg_variable, zg_zone, *zix) = zg_args
    if tuple(ix) != tuple(zix) or spec_zone != zg_zone:
        raise ValueError("spec_bdyupdate_ph_zero_grad: the two calls must share the index set and zone")
    if zg_field is field:
        raise ValueError("spec_bdyupdate_ph_zero_grad: the two calls must act on different arrays")
    g = grid(field.shape, _I24(*ix))
    call("pywrf_spec_bdyupdate_ph_zero_grad", g, ptr(ph_save), ptr(field), ptr(field_tend), ptr(mu_tend), ptr(muts),
         float(dt), flag(variable_in), ptr(zg_field), flag(zg_variable), stream())
    return field, zg_field


def mass_weight(field, mut, rfield, c1, c2,
                ids, ide, jds, jde, kds, kde,
                ims, ime, jms, jme, kms, kme,
                irs, ire, jrs, jre, krs, kre,
                its, ite, jts, jte, kts, kte):
    """wrf_boundary.py:408-415"""
    g = grid(field.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, irs, ire, jrs, jre, krs, kre,
                           its, ite, jts, jte, kts, kte))
    call("pywrf_mass_weight", g, ptr(field), ptr(mut), ptr(rfield), stream())
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
    """wrf_boundary.py:418-451"""
    _width(spec_bdy_width)
    g = grid(scalar.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                                kpe, its, ite, jts, jte, kts, kte))
    bp = _bdy(g, (scalar_bxs, scalar_bxe, scalar_bys, scalar_bye, scalar_btxs, scalar_btxe, scalar_btys,
                  scalar_btye))
    if _on_device(dtbc):
        call("pywrf_relax_bdy_scalar_dtbc_dev", g, ptr(scalar_tend), ptr(scalar), ptr(mu), *bp, ptr(dtbc), ptr(fcx),
             ptr(gcx), stream())
    else:
        call("pywrf_relax_bdy_scalar", g, ptr(scalar_tend), ptr(scalar), ptr(mu), *bp, float(dtbc), ptr(fcx),
             ptr(gcx), stream())
    return scalar_tend


def spec_bdy_scalar(scalar_tend,
                    scalar_bxs, scalar_bxe, scalar_bys, scalar_bye,
                    scalar_btxs, scalar_btxe, scalar_btys, scalar_btye,
                    spec_bdy_width, spec_zone,
                    ids, ide, jds, jde, kds, kde,
                    ims, ime, jms, jme, kms, kme,
                    ips, ipe, jps, jpe, kps, kpe,
                    its, ite, jts, jte, kts, kte):
    """wrf_boundary.py:454-472"""
    _width(spec_bdy_width)
    g = grid(scalar_tend.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe,
                                     kps, kpe, its, ite, jts, jte, kts, kte))
    _bdy(g, (scalar_bxs, scalar_bxe, scalar_bys, scalar_bye))                   # shape checks only
    bp = _bdy(g, (scalar_btxs, scalar_btxe, scalar_btys, scalar_btye))
    call("pywrf_spec_bdy_scalar", g, ptr(scalar_tend), *bp, stream())
    return scalar_tend


def flow_dep_bdy(field,
                 u, v,
                 spec_zone,
                 ids, ide, jds, jde, kds, kde,
                 ims, ime, jms, jme, kms, kme,
                 ips, ipe, jps, jpe, kps, kpe,
                 its, ite, jts, jte, kts, kte):
    """wrf_boundary.py:475-510"""
    g = grid(field.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                               kpe, its, ite, jts, jte, kts, kte))
    call("pywrf_flow_dep_bdy", g, ptr(field), ptr(u), ptr(v), stream())
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
    """wrf_boundary.py:513-587"""
    _width(spec_bdy_width)
    if _on_device(dtbc):
        spec_bdy_final_group([(field, mu, msf, (field_bdy_xs, field_bdy_xe, field_bdy_ys, field_bdy_ye,
                                                field_bdy_tend_xs, field_bdy_tend_xe, field_bdy_tend_ys,
                                                field_bdy_tend_ye), variable_in)], dtbc,
                             (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                              kpe, its, ite, jts, jte, kts, kte))
        return field
    g = grid(field.shape, _I24(ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, ips, ipe, jps, jpe, kps,
                               kpe, its, ite, jts, jte, kts, kte))
    bp = _bdy(g, (field_bdy_xs, field_bdy_xe, field_bdy_ys, field_bdy_ye, field_bdy_tend_xs, field_bdy_tend_xe,
                  field_bdy_tend_ys, field_bdy_tend_ye))
    call("pywrf_spec_bdy_final", g, ptr(field), ptr(mu), ptr(msf), *bp, flag(variable_in), float(dtbc), stream())
    return field


# ----------------------------------------------------------------- batched boundary updates
def spec_bdyupdate_group(items, dt, idx24, idx2d):
    """A run of spec_bdyupdate calls on distinct fields in one launch.  items:
    [(field, field_tend, variable, is_2d)], all with time step dt; idx24 /
    idx2d: the index tuples the individual calls would take (idx2d with
    kds = kde = 0).  Same result as the calls in order.  Not part of the
    reference API."""
    import ctypes
    n = len(items)
    if n == 0:
        return
    if len({f.data_ptr() for f, _, _, _ in items}) != n:
        raise ValueError("spec_bdyupdate_group: a field appears twice")
    f3 = next((f for f, _, _, two in items if not two), None)
    f2 = next((f for f, _, _, two in items if two), None)
    shape3 = f3.shape if f3 is not None else (idx24[5],) + tuple(items[0][0].shape[-2:])
    shape2 = f2.shape if f2 is not None else tuple(items[0][0].shape[-2:])
    g3 = grid(shape3, _I24(*idx24))
    g2 = grid(shape2, _I24(*idx2d))
    fields = (ctypes.c_void_p * n)(*[ptr(f).value for f, _, _, _ in items])
    tends = (ctypes.c_void_p * n)(*[ptr(t).value for _, t, _, _ in items])
    flags = b"".join(v.encode() for _, _, v, _ in items)
    kinds = (ctypes.c_int * n)(*[1 if two else 0 for _, _, _, two in items])
    call("pywrf_spec_bdyupdate_group", g3, g2, n, fields, tends, flags, kinds, float(dt), stream())


def flow_dep_bdy_group(fields, u, v, idx24):
    """flow_dep_bdy on several distinct fields with the same u, v in one
    launch (the moisture species loop).  Not part of the reference API."""
    import ctypes
    n = len(fields)
    if n == 0:
        return
    if len({f.data_ptr() for f in fields}) != n:
        raise ValueError("flow_dep_bdy_group: a field appears twice")
    g = grid(fields[0].shape, _I24(*idx24))
    arr = (ctypes.c_void_p * n)(*[ptr(f).value for f in fields])
    call("pywrf_flow_dep_bdy_group", g, n, arr, ptr(u), ptr(v), stream())


def spec_bdytend_group(items, idx24, idx2d):
    """A run of spec_bdytend calls on distinct tendencies in one launch.
    items: [(field_tend, (btxs, btxe, btys, btye), variable, is_2d)]; idx24 /
    idx2d: the index tuples the individual calls would take (is_2d selects
    idx2d).  Same result as the calls in order.  Not part of the reference
    API."""
    import ctypes
    n = len(items)
    if n == 0:
        return
    if len({t.data_ptr() for t, _, _, _ in items}) != n:
        raise ValueError("spec_bdytend_group: a tendency appears twice")
    t3 = next((t for t, _, _, two in items if not two), None)
    t2 = next((t for t, _, _, two in items if two), None)
    shape3 = t3.shape if t3 is not None else (idx24[5],) + tuple(items[0][0].shape[-2:])
    shape2 = t2.shape if t2 is not None else (1,) + tuple(items[0][0].shape[-2:])
    g3, g2 = grid(shape3, _I24(*idx24)), grid(shape2, _I24(*idx2d))
    for t, _, _, two in items:
        check_operands("spec_bdytend_group", (g2 if two else g3).contents, ("tend",), (ptr(t),))
    tends = (ctypes.c_void_p * n)(*[ptr(t).value for t, _, _, _ in items])
    bps = [p for _, arrs, _, two in items for p in _bdy(g2 if two else g3, arrs)]
    bt = (ctypes.c_void_p * (4 * n))(*[p.value for p in bps])
    flags = b"".join(v.encode() for _, _, v, _ in items)
    kinds = (ctypes.c_int * n)(*[1 if two else 0 for _, _, _, two in items])
    call("pywrf_spec_bdytend_group", g3, g2, n, tends, bt, flags, kinds, stream())


def spec_bdy_final_group(items, dtbc, idx24):
    """A run of spec_bdy_final calls on distinct fields in one launch.
    items: [(field, mu, msf, (bxs, bxe, bys, bye, btxs, btxe, btys, btye),
    variable)]; dtbc a number or a one-value device tensor.  Same result as
    the calls in order.  Not part of the reference API."""
    import ctypes
    n = len(items)
    if n == 0:
        return
    if len({f.data_ptr() for f, _, _, _, _ in items}) != n:
        raise ValueError("spec_bdy_final_group: a field appears twice")
    g = grid(items[0][0].shape, _I24(*idx24))
    for f, mu, msf, _, _ in items:
        check_operands("spec_bdy_final_group", g.contents, ("field", "mu", "msf"), (ptr(f), ptr(mu), ptr(msf)))
    vp = lambda xs: (ctypes.c_void_p * len(xs))(*[ptr(x).value for x in xs])
    fields = vp([f for f, _, _, _, _ in items])
    mus = vp([m for _, m, _, _, _ in items])
    msfs = vp([m for _, _, m, _, _ in items])
    bps = [p for _, _, _, arrs, _ in items for p in _bdy(g, arrs)]
    bdys = (ctypes.c_void_p * len(bps))(*[p.value for p in bps])
    flags = b"".join(v.encode() for _, _, _, _, v in items)
    if _on_device(dtbc):
        check_operands("spec_bdy_final_group", g.contents, ("dtbc",), (ptr(dtbc),))
        call("pywrf_spec_bdy_final_group_dtbc_dev", g, n, fields, mus, msfs, bdys, flags, ptr(dtbc), stream())
    else:
        call("pywrf_spec_bdy_final_group", g, n, fields, mus, msfs, bdys, flags, float(dtbc), stream())


def boundary_epilogue(bc_items, final_items, dtbc, idx24, idx2d):
    """The end-of-step boundary calls (solver.py:1776-1910) in one launch:
    set_physical_bc3d / bc2d on bc_items [(tensor, variable, is_2d)] (distinct
    fields), then spec_bdy_final on final_items [(field, mu, msf, (bxs, bxe,
    bys, bye, btxs, btxe, btys, btye), variable)], each field of which must be
    one of the bc_items.  Same results as set_physical_bc_group followed by
    spec_bdy_final_group.  dtbc: a number or a one-value device tensor.  Not
    part of the reference API."""
    import ctypes
    n = len(bc_items)
    if n == 0:
        return
    if len({t.data_ptr() for t, _, _ in bc_items}) != n:
        raise ValueError("boundary_epilogue: a field appears twice")
    t3 = next((t for t, _, two in bc_items if not two), None)
    t2 = next((t for t, _, two in bc_items if two), None)
    shape3 = t3.shape if t3 is not None else (idx24[5],) + tuple(bc_items[0][0].shape[-2:])
    shape2 = t2.shape if t2 is not None else tuple(bc_items[0][0].shape[-2:])
    g3, g2 = grid(shape3, _I24(*idx24)), grid(shape2, idx2d)
    where = {t.data_ptr(): k for k, (t, _, _) in enumerate(bc_items)}
    final_of = []
    for f, mu, msf, _, _ in final_items:
        if f.data_ptr() not in where:
            raise ValueError("boundary_epilogue: a spec_bdy_final field is not among the filled fields")
        check_operands("boundary_epilogue", g3.contents, ("field", "mu", "msf"), (ptr(f), ptr(mu), ptr(msf)))
        final_of.append(where[f.data_ptr()])
    vp = lambda xs: (ctypes.c_void_p * max(len(xs), 1))(*[ptr(x).value for x in xs])
    fields = vp([t for t, _, _ in bc_items])
    flags = b"".join(v.encode() for _, v, _ in bc_items)
    kinds = (ctypes.c_int * n)(*[1 if two else 0 for _, _, two in bc_items])
    m = len(final_items)
    fin = (ctypes.c_int * max(m, 1))(*final_of)
    mus = vp([x for _, x, _, _, _ in final_items])
    msfs = vp([x for _, _, x, _, _ in final_items])
    bps = [p for _, _, _, arrs, _ in final_items for p in _bdy(g3, arrs)]
    bdys = (ctypes.c_void_p * max(len(bps), 1))(*[p.value for p in bps])
    fflags = b"".join(v.encode() for _, _, _, _, v in final_items)
    if _on_device(dtbc):
        check_operands("boundary_epilogue", g3.contents, ("dtbc",), (ptr(dtbc),))
        call("pywrf_boundary_epilogue_dtbc_dev", g3, g2, n, fields, flags, kinds, m, fin, mus, msfs, bdys, fflags,
             ptr(dtbc), stream())
    else:
        call("pywrf_boundary_epilogue", g3, g2, n, fields, flags, kinds, m, fin, mus, msfs, bdys, fflags,
             float(dtbc), stream())
