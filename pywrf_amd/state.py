"""Model state for one domain: every array the reference driver allocates
(pywrf/solver.py:140-447) under the reference's names, in (k, j, i) layout,
i fastest, FP32 (or FP64), with a 5-point halo.

Lateral-boundary arrays are held as the reference driver holds them
(solver.py:468-703): x-boundaries as the (nz, ny, 5) view permute(1, 2, 0) of
a (5, nz, ny) array, y-boundaries as the (nz, 5, nx) view permute(1, 0, 2) of
a (5, nz, nx) array, under `s.bdy[var] = (bxs, bxe, bys, bye, btxs, btxe,
btys, btye)` for var in u, v, ph, t, mu, qv (nz = 1 for mu).
"""
from types import SimpleNamespace

import torch

from .step import Grid

# fields of shape (nz, ny, nx)
FIELDS3 = (
    "u v w t ph p al pb phb t_init alb ru rv rw ww ww1 cqu cqv cqw alt php "
    "ru_tendf rv_tendf rw_tendf ph_tendf t_tendf tke_tend h_tendency z_tendency "
    "ru_tend rv_tend rw_tend ph_tend t_tend u_save v_save w_save ph_save t_save "
    "rho u_phy v_phy z_at_w dz8w z p8w t8w p_hyd_w p_hyd rdz rdzw zx zy "
    "defor11 defor22 defor33 defor12 defor13 defor23 div BN2 xkmh xkmv xkhh xkhv tke "
    "rthblten rublten rvblten rqvblten rqcblten rqiblten rfield "
    "u_1 v_1 w_1 t_1 ph_1 tke_1 c2a ww_1 pm1 a alpha gamma ru_m rv_m ww_m t_2save "
    "th_phy p_phy pi_phy t_phy RQVFTEN refl_10cm h_diabatic qv_diabatic qc_diabatic").split()
# fields of shape (ny, nx)
FIELDS2 = (
    "mu mub msfu msfv msfvx_inv msft msftx msfty msfux msfuy msfvx msfvy f e sina cosa clat ht "
    "mut muu muv mu_tendf mu_tend mu_save mu_1 muus muvs muts mudf muave "
    "rainnc rainncv snownc snowncv graupelnc graupelncv sr").split()
# level arrays (nz,)
FIELDS1 = "rdnw rdn c1h c2h c1f c2f c3h c4h c3f c4f fnm fnp znu znw dn dnw u_base v_base qv_base".split()
# 0-d scalars
FIELDS0 = "cf1 cf2 cf3 cfn cfn1".split()
# (n, nz, ny, nx)
FIELDS4 = {"moist": 7, "moist_tend": 7, "advect_tend": 7, "moist_old": 7,
           "scalar": 3, "scalar_tend": 3, "scalar_old": 3, "advect_tend_sca": 3}
BDY_VARS = ("u", "v", "ph", "t", "mu", "qv")


def allocate_state(nz, ny, nx, device="cpu", dtype=torch.float32):
    """Zero-initialised state with the reference's shapes."""
    z = lambda *shape: torch.zeros(shape, dtype=dtype, device=device)
    s = SimpleNamespace()
    s.grid = Grid(nz, ny, nx)
    for n in FIELDS3:
        setattr(s, n, z(nz, ny, nx))
    for n in FIELDS2:
        setattr(s, n, z(ny, nx))
    for n in FIELDS1:
        setattr(s, n, z(nz))
    _pack_scalars(s, z(len(FIELDS0)))
    for n, c in FIELDS4.items():
        setattr(s, n, z(c, nz, ny, nx))
    s.bdy = {}
    for v in BDY_VARS:
        kz = 1 if v == "mu" else nz
        s.bdy[v] = tuple(boundary_array(kz, ny, nx, n % 4 < 2, dtype, device) for n in range(8))
    s.gcx = torch.tensor([0., 1.1111111e-3, 7.4074074e-4, 3.7037037e-4, 0.], dtype=dtype, device=device)
    s.fcx = torch.tensor([0., 5.5555557e-3, 3.7037039e-3, 1.8518519e-3, 0.], dtype=dtype, device=device)
    s.step, s.it_bdy = 0, 0
    return s


def boundary_array(kz, ny, nx, x_side, dtype=torch.float32, device="cpu"):
    """A zero lateral-boundary array in the reference driver's layout: for an
    x-side, the (kz, ny, 5) view of a (5, kz, ny) array; for a y-side, the
    (kz, 5, nx) view of a (5, kz, nx) array (solver.py:468-470, 478-480)."""
    if x_side:
        return torch.zeros((5, kz, ny), dtype=dtype, device=device).permute(1, 2, 0)
    return torch.zeros((5, kz, nx), dtype=dtype, device=device).permute(1, 0, 2)


def state_tensors(s):
    """(name, tensor) pairs of every array in the state, boundary arrays included."""
    out = []
    for n, v in vars(s).items():
        if isinstance(v, torch.Tensor):
            out.append((n, v))
    for var, arrs in s.bdy.items():
        for tag, a in zip(("bxs", "bxe", "bys", "bye", "btxs", "btxe", "btys", "btye"), arrs):
            out.append(("%s_%s" % (var, tag), a))
    return out


def clone_state(s, device=None, dtype=None):
    """Deep copy of a state, optionally moved to another device / dtype."""
    c = SimpleNamespace(**{k: v for k, v in vars(s).items()})
    conv = lambda t: t.to(device=device or t.device, dtype=dtype or t.dtype, copy=True)
    for n, v in vars(s).items():
        if isinstance(v, torch.Tensor):
            setattr(c, n, conv(v))
    c.bdy = {k: tuple(conv(a) for a in arrs) for k, arrs in s.bdy.items()}
    c.grid = s.grid
    _pack_scalars(c, torch.stack([getattr(c, n) for n in FIELDS0]))
    return c


def _pack_scalars(s, pack):
    """Hold cf1, cf2, cf3, cfn, cfn1 as views of one contiguous vector, so that
    kernels can read (cf1, cf2, cf3) as three consecutive floats."""
    s.scalars0 = pack
    for n, name in enumerate(FIELDS0):
        setattr(s, name, pack[n])


PROGNOSTIC = ("u", "v", "w", "t", "ph", "p", "al", "mu", "moist")
