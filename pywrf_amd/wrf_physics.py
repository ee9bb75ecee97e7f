"""Drop-in for pywrf/wrf_physics.py on the ARW step's path: the moist-physics
prep and finish wrappers and WSM6 (with its sedimentation), same names,
positional parameters and return tuples, forwarding to libpywrf_amd.so.
"""
from ._lib import call, grid, ptr, stream


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
    """wrf_physics.py:27-91"""
    g = grid(p.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_moist_physics_prep_em", g, ptr(t_new), float(t0), ptr(rho), ptr(al), ptr(alb), ptr(p), ptr(p8w),
         float(p0), ptr(pb), ptr(ph), ptr(phb), ptr(th_phy), ptr(pii), ptr(pf), ptr(z), ptr(z_at_w), ptr(dz8w),
         ptr(h_diabatic), ptr(qv), ptr(qv_diabatic), ptr(qc), ptr(qc_diabatic), ptr(fzm), ptr(fzp), stream())
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
    """wrf_physics.py:94-134: WSM6 with the configured constants (15-tuple)."""
    g = grid(th.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_microphysics_driver", g, ptr(th), ptr(rho), ptr(pi_phy), ptr(p), ptr(dz8w), float(dt), ptr(qv_curr),
         ptr(qc_curr), ptr(qr_curr), ptr(qi_curr), ptr(qs_curr), ptr(qg_curr), ptr(rainnc), ptr(rainncv),
         ptr(snownc), ptr(snowncv), ptr(sr), ptr(graupelnc), ptr(graupelncv), stream())
    return (th, qv_curr, qc_curr, qr_curr, qi_curr, qs_curr, qg_curr, rainnc, rainncv, snownc, snowncv,
            graupelnc, graupelncv, sr, refl_10cm)


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
    """wrf_physics.py:137-1214 (the physical constants are the library's; the
    reference's microphysics_driver passes exactly those)."""
    g = grid(th.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_wsm6", g, ptr(th), ptr(q), ptr(qc), ptr(qr), ptr(qi), ptr(qs), ptr(qg), ptr(den), ptr(pii), ptr(p),
         ptr(delz), float(delt), ptr(rain), ptr(rainncv), ptr(snow), ptr(snowncv), ptr(sr), ptr(graupel),
         ptr(graupelncv), stream())
    return th, q, qc, qr, qi, qs, qg, rain, rainncv, snow, snowncv, graupel, graupelncv, sr, refl_10cm


def moist_physics_finish_em(t_new, t_old, t0, mut,
                            th_phy, h_diabatic, dt,
                            qv, qv_diabatic,
                            qc, qc_diabatic,
                            ids, ide, jds, jde, kds, kde,
                            ims, ime, jms, jme, kms, kme,
                            its, ite, jts, jte, kts, kte):
    """wrf_physics.py:2106-2137"""
    g = grid(t_new.shape, (ids, ide, jds, jde, kds, kde, ims, ime, jms, jme, kms, kme, its, ite, jts, jte, kts, kte))
    call("pywrf_moist_physics_finish_em", g, ptr(t_new), ptr(th_phy), ptr(h_diabatic), float(dt), ptr(qv),
         ptr(qv_diabatic), ptr(qc), ptr(qc_diabatic), stream())
    return t_new, t_old, th_phy, h_diabatic, qv_diabatic, qc_diabatic
