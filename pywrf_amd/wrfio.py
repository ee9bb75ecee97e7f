"""WRF input and lateral-boundary files into the step's state.

The reference's driver reads wrfinput / wrfbdy / wrfout with xarray
(solver.py:40-134), copies each field into its zero-initialised,
halo-padded state array (solver.py:213-281), forms alb and al
(solver.py:284-300), fixes the edges of mub (solver.py:303-306), and every
bdy_interval steps loads boundary record it_bdy into zero (5, kz, n) arrays
that it hands to the boundary functions as permuted views
(solver.py:457-703).

Here the files are read by pywrf_amd.netcdf (no xarray, no netCDF library),
and the placement follows one rule that covers every copy of the reference:
a field's file array lands at levels 0.. and at the 5-point halo offset in j
and i -- 3-D (k, j, i) arrays at [0:nk, 5:5+nj, 5:5+ni], planes at
[5:5+nj, 5:5+ni], level arrays at [0:nk], scalars as they are; a boundary
record (5, nk, n) at [:, 0:nk, 5:5+n] of the (5, kz, n) storage (mu: (5, n)
at [:, 0, 5:5+n]).  The boundary storage is written in place, so the
pointers the kernels (and a captured HIP graph) hold stay valid across a
refresh.  An array that does not fit its place raises ValueError.
"""
import ctypes

import torch

from . import _lib
from .netcdf import Dataset

HALO = 5
# file variable -> state attribute (solver.py:60-134, 213-281); ("moist", n) is species n
INPUT_FIELDS = {
    "U": "u", "V": "v", "MU": "mu", "MUB": "mub",
    "MAPFAC_U": "msfu", "MAPFAC_V": "msfv", "MF_VX_INV": "msfvx_inv", "MAPFAC_M": "msft", "MAPFAC_MX": "msftx",
    "MAPFAC_MY": "msfty", "MAPFAC_UX": "msfux", "MAPFAC_UY": "msfuy", "MAPFAC_VX": "msfvx", "MAPFAC_VY": "msfvy",
    "QVAPOR": ("moist", 1), "QCLOUD": ("moist", 2), "QRAIN": ("moist", 3), "QICE": ("moist", 4),
    "QSNOW": ("moist", 5), "QGRAUP": ("moist", 6),
    "PB": "pb", "P": "p", "T_INIT": "t_init", "PHB": "phb", "PH": "ph",
    "RDNW": "rdnw", "RDN": "rdn", "C1H": "c1h", "C2H": "c2h", "C3H": "c3h", "C4H": "c4h",
    "C1F": "c1f", "C2F": "c2f", "C3F": "c3f", "C4F": "c4f", "T": "t", "FNM": "fnm", "FNP": "fnp",
    "DN": "dn", "DNW": "dnw", "CF1": "cf1", "CF2": "cf2", "CF3": "cf3", "CFN": "cfn", "CFN1": "cfn1",
    "F": "f", "E": "e", "SINALPHA": "sina", "COSALPHA": "cosa", "CLAT": "clat", "ZNU": "znu", "ZNW": "znw",
    "HGT": "ht", "U_BASE": "u_base", "V_BASE": "v_base", "QV_BASE": "qv_base",
}
# the vertical velocity comes from the output file (solver.py:62)
OUTPUT_FIELDS = {"W": "w"}
# boundary variables of wrfbdy -> the state's boundary sets (solver.py:461-703)
BDY_FIELDS = {"U": "u", "V": "v", "PH": "ph", "T": "t", "MU": "mu", "QVAPOR": "qv"}
BDY_SIDES = ("BXS", "BXE", "BYS", "BYE", "BTXS", "BTXE", "BTYS", "BTYE")


def _target(s, name):
    if isinstance(name, tuple):
        return getattr(s, name[0])[name[1]]
    return getattr(s, name)


def place(dst, src):
    """Copy the file array src (numpy, without its time axis) into the state
    array dst at levels 0.. and the halo offset in j and i."""
    src_t = torch.from_numpy(src).to(dtype=dst.dtype)
    if dst.dim() == 0:
        if src_t.numel() != 1:
            raise ValueError("scalar target, array of shape %s" % (tuple(src.shape),))
        dst.copy_(src_t.reshape(()))
        return
    if dst.dim() == 1:
        region = (slice(0, src_t.shape[0]),)
    elif dst.dim() == 2:
        region = (slice(HALO, HALO + src_t.shape[-2]), slice(HALO, HALO + src_t.shape[-1]))
    else:
        region = (slice(0, src_t.shape[0]), slice(HALO, HALO + src_t.shape[1]), slice(HALO, HALO + src_t.shape[2]))
    if src_t.dim() != dst.dim() or any(r.stop > n for r, n in zip(region, dst.shape)):
        raise ValueError("array of shape %s does not fit a state array of shape %s at the halo offset"
                         % (tuple(src.shape), tuple(dst.shape)))
    dst[region] = src_t.to(dst.device)


def _constant(name):
    v = ctypes.c_double()
    if _lib.LIB.pywrf_constant(name.encode(), ctypes.byref(v)) != 0:
        raise KeyError(name)
    return v.value


def init_diagnostics(s):
    """alb and al of the loaded state (solver.py:284-300) and the edge copy
    of mub (solver.py:303-306), in the reference's expressions."""
    g = s.grid
    r_d, p1000, t0, cvpm = (_constant(n) for n in ("r_d", "p1000mb", "t0", "cvpm"))
    s.alb.copy_(r_d / p1000 * (s.t_init + t0) * (s.pb / p1000) ** cvpm)
    k0, k1, j0, j1, i0, i1 = g.kts, g.kte - 1, g.jts, g.jde - 1, g.its, g.ide - 1
    mub = s.mub[j0:j1, i0:i1]
    mu = s.mu[j0:j1, i0:i1]
    rdnw = s.rdnw[k0:k1].view(-1, 1, 1)
    s.al[k0:k1, j0:j1, i0:i1] = -1.0 / (mub + mu) * (s.alb[k0:k1, j0:j1, i0:i1] * mu + rdnw * (
        s.ph[k0 + 1:k1 + 1, j0:j1, i0:i1] - s.ph[k0:k1, j0:j1, i0:i1]))
    s.mub[g.jte - 1, :] = s.mub[g.jte - 2, :]
    s.mub[:, g.ite - 1] = s.mub[:, g.ite - 2]
    s.mub[g.jts - 1, :] = s.mub[g.jts, :]
    s.mub[:, g.its - 1] = s.mub[:, g.its]


def load_input(s, wrfinput, wrfout=None, record=0):
    """Place the fields of wrfinput (and w from wrfout, else from wrfinput)
    into state s, then form the initial diagnostics.  wrfinput / wrfout:
    paths or open Datasets."""
    inp = wrfinput if isinstance(wrfinput, Dataset) else Dataset(wrfinput)
    out = wrfout if (wrfout is None or isinstance(wrfout, Dataset)) else Dataset(wrfout)
    for var, name in INPUT_FIELDS.items():
        if var not in inp:
            raise KeyError("%s: no variable %s" % (inp.path, var))
        place(_target(s, name), inp[var][record])
    src = out if (out is not None and "W" in out) else inp
    place(s.w, src["W"][record])
    init_diagnostics(s)
    return s


def _storage(view, x_side):
    """The (5, kz, n) storage of a boundary array given as the driver's view."""
    return view.permute(2, 0, 1) if x_side else view.permute(1, 0, 2)


def load_boundary(s, wrfbdy, it_bdy):
    """Load boundary record it_bdy of wrfbdy into the state's boundary arrays
    in place (solver.py:457-703): each side's (5, nk, n) slab at levels 0..
    and the halo offset along the side, zero elsewhere."""
    bdy = wrfbdy if isinstance(wrfbdy, Dataset) else Dataset(wrfbdy)
    for var, key in BDY_FIELDS.items():
        for n, side in enumerate(BDY_SIDES):
            name = "%s_%s" % (var, side)
            if name not in bdy:
                raise KeyError("%s: no variable %s" % (bdy.path, name))
            if it_bdy >= bdy[name].shape[0]:
                raise IndexError("%s: boundary record %d, the file holds %d" % (bdy.path, it_bdy, bdy[name].shape[0]))
            a = torch.from_numpy(bdy[name][it_bdy])
            if a.dim() == 2:                      # mu: (5, n)
                a = a.unsqueeze(1)
            st = _storage(s.bdy[key][n], n % 4 < 2)
            if a.shape[0] != st.shape[0] or a.shape[1] > st.shape[1] or HALO + a.shape[2] > st.shape[2]:
                raise ValueError("%s of shape %s does not fit a boundary array of storage %s"
                                 % (name, tuple(a.shape), tuple(st.shape)))
            st.zero_()
            st[:, :a.shape[1], HALO:HALO + a.shape[2]] = a.to(device=st.device, dtype=st.dtype)


def refresh_boundaries(s, wrfbdy, cfg):
    """Before a step: when the coming step starts a boundary interval
    (solver.py:457-458, step % bdy_interval == 1), load its record.  Returns
    the record loaded, or None.  With HIP-graph replays, call it between
    replays for the first step of each interval (the arrays are updated in
    place, so the graph stays valid)."""
    nxt = s.step + 1
    if nxt % cfg.bdy_interval == 1:
        it = nxt // cfg.bdy_interval
        load_boundary(s, wrfbdy, it)
        return it
    return None


def run(s, ops, cfg, nsteps, wrfbdy=None, clock=None):
    """nsteps RK3 steps of state s with the boundary refresh of the
    reference's time loop (solver.py:454-458)."""
    from .step import rk3_step
    for _ in range(nsteps):
        if wrfbdy is not None:
            refresh_boundaries(s, wrfbdy, cfg)
        rk3_step(s, ops, cfg, clock=clock)
    return s
