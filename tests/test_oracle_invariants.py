"""Pins of the oracle that do not depend on the reference's numbers (CPU):
closed forms of the flux operators, the equation-of-state inverse, the WSM6
column water budget, and a resting atmosphere that must stay at rest."""
import math

import numpy as np
import pytest
import torch

from parity import state, run


# ---- flux operators (wrf_dynamics.py:1313-1329) --------------------------
def _faces(q, flux, width):
    """Face values at faces i = 3 .. n-3 from point values q (float64)."""
    n = q.numel()
    ones = torch.ones(1, dtype=q.dtype)
    out = []
    for i in range(3, n - 2):
        if width == 6:
            out.append(flux(q[i - 3:i - 2], q[i - 2:i - 1], q[i - 1:i], q[i:i + 1], q[i + 1:i + 2], q[i + 2:i + 3], ones))
        else:
            out.append(flux(q[i - 2:i - 1], q[i - 1:i], q[i:i + 1], q[i + 1:i + 2], ones))
    return torch.cat(out)


@pytest.mark.parametrize("deg", range(0, 7))
def test_flux5_difference_is_sixth_order_derivative(deg):
    """With a uniform velocity the difference of the 6th-order centred part
    across a cell is the (45, -9, 1)/60 derivative stencil: exact for
    polynomials of degree <= 6; the upwind correction adds an undivided 6th
    difference, zero for degree <= 5."""
    from oracle._box import flux5
    x = torch.arange(-12, 13, dtype=torch.float64) * 0.37
    q = x ** deg
    f = _faces(q, flux5, 6)                       # face i sits between x[i-1] and x[i]
    dq = (f[1:] - f[:-1]) / 0.37                  # centred on x[i]
    xi = x[3:3 + dq.numel()]
    exact = deg * xi ** (deg - 1) if deg else torch.zeros_like(xi)
    central = (45 * (q[4:4 + dq.numel()] - q[2:2 + dq.numel()]) - 9 * (q[5:5 + dq.numel()] - q[1:1 + dq.numel()])
               + (q[6:6 + dq.numel()] - q[0:dq.numel()])) / 60 / 0.37
    scale = max(1.0, float(exact.abs().max()))
    if deg <= 5:
        assert float((dq - exact).abs().max()) < 1e-9 * scale
    else:
        # only the upwind term (a 6th difference) departs from the centred derivative
        assert float((central - exact).abs().max()) < 1e-9 * scale
        assert float((dq - exact).abs().max()) > 1e-6


@pytest.mark.parametrize("deg", range(0, 5))
def test_flux3_difference_is_fourth_order_derivative(deg):
    """flux3 = flux4 + upwind: flux4 differences give the (8, -1)/12 stencil
    (exact to degree 4); the upwind 4th difference vanishes to degree 3."""
    from oracle._box import flux3
    x = torch.arange(-12, 13, dtype=torch.float64) * 0.53
    q = x ** deg
    f = _faces(q, flux3, 4)
    dq = (f[1:] - f[:-1]) / 0.53
    xi = x[3:3 + dq.numel()]
    exact = deg * xi ** (deg - 1) if deg else torch.zeros_like(xi)
    scale = max(1.0, float(exact.abs().max()))
    if deg <= 3:
        assert float((dq - exact).abs().max()) < 1e-9 * scale
    else:
        assert float((dq - exact).abs().max()) > 1e-6


def test_upwind_sign_follows_velocity():
    from oracle._box import flux5, flux3
    q = [torch.tensor([float(v)], dtype=torch.float64) for v in (1, 4, 2, 8, 5, 7)]
    up = flux5(*q, torch.tensor([1.0], dtype=torch.float64))
    dn = flux5(*q, torch.tensor([-1.0], dtype=torch.float64))
    zero = flux5(*q, torch.tensor([0.0], dtype=torch.float64))
    assert float(up) != float(dn) and float(zero) == float(up)    # SIGN(1, 0) = +1
    q3 = q[1:5]
    assert float(flux3(*q3, torch.tensor([1.0], dtype=torch.float64))) != \
        float(flux3(*q3, torch.tensor([-1.0], dtype=torch.float64)))


# ---- equation of state (wrf_dynamics.py:4124-4164) -----------------------
def test_eos_inverse():
    """calc_p_rho_phi's p satisfies (p + pb)/p0 = (Rd theta_m / (p0 alpha))^(cp/cv)."""
    import oracle
    from oracle.constants import r_d, rvovrd, cpovcv
    s = state("config1", dtype=torch.float64)
    g = s.grid
    al, p, _ = oracle.ops().calc_p_rho_phi(s.moist, 7, 1, s.al.clone(), s.alb, s.mu, s.mub + s.mu, s.c1h, s.c2h,
                                           s.c3h, s.c4h, s.c3f, s.c4f, s.ph, s.phb, s.p.clone(), s.pb, s.t, 1e5,
                                           300.0, 5000.0, s.znu, s.znw, s.dnw, s.rdnw, s.rdn, True, *g.ix18)
    b = (slice(0, g.nz - 1), slice(g.jts, g.jte - 1), slice(g.its, g.ite - 1))
    theta_m = (300.0 + s.t[b]) * (1 + rvovrd * s.moist[1][b])
    lhs = ((p[b] + s.pb[b]) / 1e5) ** (1.0 / cpovcv)
    rhs = r_d * theta_m / (1e5 * (al[b] + s.alb[b]))
    assert float(((lhs - rhs) / rhs).abs().max()) < 1e-12


# ---- WSM6 column water budget (wrf_physics.py:137-1214) ------------------
def test_wsm6_conserves_column_water():
    """Per column, total water (vapour + 5 hydrometeors) times rho dz, plus
    this step's surface precipitation, is conserved by WSM6 (FP64)."""
    import oracle
    from oracle.constants import rhowater
    s = state("config1c", dtype=torch.float64)
    g = s.grid
    o = oracle.ops()
    P18 = g.phys18()
    (s.rho, s.th_phy, s.pi_phy, s.p_phy, s.z, s.z_at_w, s.dz8w, s.p8w, s.h_diabatic, s.qv_diabatic,
     s.qc_diabatic, s.t, s.t_1) = o.moist_physics_prep_em(
        s.t, s.t_1, 300.0, s.rho, s.al, s.alb, s.p, s.p8w, 1e5, s.pb, s.ph, s.phb, s.th_phy, s.pi_phy, s.p_phy, s.z,
        s.z_at_w, s.dz8w, 54.0, s.h_diabatic, s.moist[1], s.qv_diabatic, s.moist[2], s.qc_diabatic, s.fnm, s.fnp,
        *P18)
    k1, (j0, j1), (i0, i1) = g.nz - 1, (P18[14], P18[15] - 1), (P18[12], P18[13] - 1)
    col = (slice(0, k1), slice(j0, j1), slice(i0, i1))

    def water():
        q = sum(s.moist[m][col] for m in range(1, 7))
        return (q * s.rho[col] * s.dz8w[col]).sum(dim=0)

    before = water()
    hyd0 = sum(s.moist[m][col].abs().sum() for m in range(2, 7))
    m = s.moist
    o.microphysics_driver(s.th_phy, s.rho, s.pi_phy, s.p_phy, s.ht, s.dz8w, s.p8w, 54.0, 9000., 9000., m[1], m[2],
                          m[3], m[4], m[5], m[6], s.rainnc, s.rainncv, s.snownc, s.snowncv, s.sr, s.refl_10cm, 1, 1,
                          s.graupelnc, s.graupelncv, 0, 0, 0, 0, 0, 0, *P18)
    after = water()
    surf = s.rainncv[j0:j1, i0:i1] / 1000.0 * rhowater         # mm of liquid water -> kg m-2
    resid = (before - after - surf).abs()
    hyd1 = sum(s.moist[m][col].abs().sum() for m in range(2, 7))
    assert abs(float(hyd1 - hyd0)) > 1e-3 * float(hyd0)        # the processes did act
    assert float(resid.max()) < 1e-9 * float(before.max()), float(resid.max() / before.max())


# ---- composite: a resting atmosphere stays at rest -----------------------
def test_resting_atmosphere_stays_at_rest():
    import oracle
    s = run(state("rest"), oracle.ops(), 3)
    assert float(s.u.abs().max()) < 1e-3 and float(s.v.abs().max()) < 1e-3
    assert float(s.w.abs().max()) < 1e-4
