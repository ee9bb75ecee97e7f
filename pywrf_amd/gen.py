"""Seeded synthetic initial and lateral-boundary state.

The reference initialises from NetCDF files that are not shipped
(solver.py:40-134, 213-281).  This generator builds the same set of arrays
analytically so every workload is reproducible from a seed:

* base state: theta(z) = theta_s exp(N^2 z / g) with theta_s = 300 K,
  N = 0.01 1/s, p = 1000 hPa at z = 0, Exner from hydrostatic balance;
  41 full eta levels with p_top = 50 hPa, placed at heights
  z_k = z_top * s (0.55 + 0.45 s), s = k/40 (about 275 m at the surface,
  725 m aloft); terrain-following sigma coordinate (c1 = 1, c2 = 0,
  c3 = eta, c4 = 0), so pb = eta * mub + p_top and mub = p_s(ht) - p_top;
  phb integrated from g*ht, alb from the same formula as solver.py:290;
* terrain: Gaussian hills h0 exp(-ln2 (r/a)^2) (a is the half width at half
  maximum);
* winds on the staggered points, w = 0; warm bubbles with a cos^2 profile
  inside an ellipsoid (horizontal radius R, vertical radius 1.5 km);
* qv = RH(z) q_sat(T, p), RH = 0.8 at the ground falling linearly to 0.1 at
  8 km; optional seeded cloud / precipitation blobs;
* p, al and ph in moist hydrostatic balance with mu = 0, by the downward
  integration used in WRF's idealised initialisation, so that al and p agree
  with the reference's own equation of state;
* map factors close to 1 with a small x/y anisotropy (so x/y swaps show), a
  beta-plane Coriolis f, e and a small grid rotation sin(alpha);
* lateral boundary values = the mass-coupled initial state in the 5-point
  strips (u*muu/msfuy, v*muv/msfvx, t*mut, ph*mut, mu, qv*mut), zero outside
  the ranges the reference fills (solver.py:468-703); boundary tendencies
  zero, or (key "bdy_tend") the coupled difference between a seeded smooth
  perturbation of the initial state and the initial state, divided by the
  boundary period bdy_interval * dt (a boundary that drifts towards another
  state, as consecutive wrfbdy records do).
"""
import math

import numpy as np
import torch

from .state import allocate_state

R_D, CP, G, P0, T0, P1000 = 287.0, 1004.5, 9.81, 100000.0, 300.0, 100000.0
CVPM, RVOVRD = -0.7142857, 1.608362
KAPPA = R_D / CP
OMEGA = 7.292e-5

CONFIGS = {
    # reference CPU case: 40x40x41, one hill, uniform 10 m/s, one warm bubble
    "config1": dict(n=40, hills=[(0.5, 0.5, 500.0, 45e3)], jet=None, u0=10.0, noise=0.0,
                    bubbles=[(0.5, 0.5, 1500.0, 2.0, 20e3)], clouds=0, seed=0),
    # config1 with seeded cloud / precipitation blobs (exercises every WSM6 branch from step 1)
    "config1c": dict(n=40, hills=[(0.5, 0.5, 500.0, 45e3)], jet=None, u0=10.0, noise=0.5,
                     bubbles=[(0.5, 0.5, 1500.0, 2.0, 20e3)], clouds=6, seed=1),
    # real shape: 240x240x41, 3 seeded hills, 20 m/s jet at 10 km, noise, seeded bubbles, 6 species
    "config2": dict(n=240, hills="random3", jet=(20.0, 10e3, 5e3), u0=0.0, noise=2.0,
                    bubbles="random", clouds=6, seed=0),
    # flat terrain, air at rest, no perturbations: an invariant case (must stay at rest)
    "rest": dict(n=40, hills=[], jet=None, u0=0.0, noise=0.0, bubbles=[], clouds=0, seed=0),
    # adjoint-check shape: 120x120x41 with the config2 recipe
    "config4": dict(n=120, hills="random3", jet=(20.0, 10e3, 5e3), u0=0.0, noise=2.0,
                    bubbles="random", clouds=4, seed=0),
    # time-varying lateral boundaries: 24x24x41 (14 interior mass points), clouds and noise, and
    # boundary tendencies towards a perturbed state (amplitudes of the uncoupled perturbation:
    # u, v in m/s, t in K, qv relative, mu in Pa, ph in m^2/s^2)
    "bdy24": dict(n=24, hills=[(0.35, 0.6, 400.0, 40e3)], jet=None, u0=8.0, noise=0.5,
                  bubbles=[(0.6, 0.4, 1500.0, 2.0, 15e3)], clouds=6, seed=3,
                  bdy_tend=dict(u=6.0, v=4.0, t=2.0, qv=0.15, mu=300.0, ph=300.0)),
}
BDY_PERIOD = 360 * 54.0     # bdy_interval * dt of the step's default run configuration (seconds)


def _exner_of_z(z, ths=300.0, nbv=0.01):
    return 1.0 - G * G / (CP * nbv * nbv * ths) * (1.0 - np.exp(-nbv * nbv * z / G))


def _z_of_p(p, ths=300.0, nbv=0.01):
    pi = (p / P0) ** KAPPA
    return -(G / nbv ** 2) * np.log(1.0 - (1.0 - pi) * CP * nbv * nbv * ths / (G * G))


def _theta_of_z(z, ths=300.0, nbv=0.01):
    return ths * np.exp(nbv * nbv * z / G)


def _qsat(T, p):
    es = 611.2 * np.exp(17.67 * (T - 273.15) / (T - 29.65))
    es = np.minimum(es, 0.5 * p)
    return 0.62175 * es / (p - es)


def eta_levels(nz=41, p_top=5000.0):
    """Full-level eta (1 at the ground, 0 at p_top) from stretched heights."""
    z_top = _z_of_p(np.array(p_top))
    s = np.arange(nz) / (nz - 1)
    zk = z_top * s * (0.55 + 0.45 * s)
    pk = P0 * _exner_of_z(zk) ** (1.0 / KAPPA)
    znw = (pk - p_top) / (P0 - p_top)
    znw[0], znw[-1] = 1.0, 0.0
    return znw


def _vertical_coeffs(znw):
    nz = znw.size
    a = {k: np.zeros(nz) for k in ("dnw", "rdnw", "znu", "dn", "rdn", "fnm", "fnp")}
    for k in range(nz - 1):
        a["dnw"][k] = znw[k + 1] - znw[k]
        a["rdnw"][k] = 1.0 / a["dnw"][k]
        a["znu"][k] = 0.5 * (znw[k + 1] + znw[k])
    for k in range(1, nz - 1):
        a["dn"][k] = 0.5 * (a["dnw"][k] + a["dnw"][k - 1])
        a["rdn"][k] = 1.0 / a["dn"][k]
        a["fnp"][k] = 0.5 * a["dnw"][k] / a["dn"][k]
        a["fnm"][k] = 0.5 * a["dnw"][k - 1] / a["dn"][k]
    dn, dnw = a["dn"], a["dnw"]
    cof1 = (2. * dn[1] + dn[2]) / (dn[1] + dn[2]) * dnw[0] / dn[1]
    cof2 = dn[1] / (dn[1] + dn[2]) * dnw[0] / dn[2]
    a["cf1"] = a["fnp"][1] + cof1
    a["cf2"] = a["fnm"][1] - cof1 - cof2
    a["cf3"] = cof2
    kl = nz - 2
    a["cfn"] = (.5 * dnw[kl] + dn[kl]) / dn[kl]
    a["cfn1"] = -.5 * dnw[kl] / dn[kl]
    return a


def _hills(cfg, rng, L):
    if cfg["hills"] == "random3":
        out = []
        for _ in range(3):
            out.append((rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.0, 2000.0),
                        rng.uniform(20e3, 80e3)))
        return out
    return cfg["hills"]


def _bubbles(cfg, rng):
    if cfg["bubbles"] == "random":
        return [(rng.uniform(0.15, 0.85), rng.uniform(0.15, 0.85), 1500.0, rng.uniform(1.0, 3.0),
                 rng.uniform(10e3, 25e3)) for _ in range(6)]
    return cfg["bubbles"]


def synthetic_state(config="config1", seed=None, device="cpu", dtype=torch.float32, p_top=5000.0,
                    dx=9000.0, dy=9000.0):
    """Build the seeded state for `config` (a key of CONFIGS or a dict)."""
    cfg = dict(CONFIGS[config]) if isinstance(config, str) else dict(config)
    seed = cfg.get("seed", 0) if seed is None else seed
    rng = np.random.default_rng(seed)
    n, nz, halo = cfg["n"], 41, 5
    ny = nx = n
    L = (n - 2 * halo - 1) * dx
    # coordinates of mass (t), u and v points; i = halo is the first interior column
    xt = (np.arange(nx) - halo + 0.5) * dx
    yt = (np.arange(ny) - halo + 0.5) * dy
    xu = (np.arange(nx) - halo) * dx
    yv = (np.arange(ny) - halo) * dy
    Xt, Yt = np.meshgrid(xt, yt)
    Xu, Yu = np.meshgrid(xu, yt)
    Xv, Yv = np.meshgrid(xt, yv)

    ht = np.zeros((ny, nx))
    for (cx, cy, h0, a) in _hills(cfg, rng, L):
        r2 = (Xt - cx * L) ** 2 + (Yt - cy * L) ** 2
        ht += h0 * np.exp(-math.log(2.0) * r2 / a ** 2)

    def msf(X, Y):
        return 1.0 + 0.02 * ((Y - 0.5 * L) / L) ** 2 + 0.01 * ((X - 0.5 * L) / L) ** 2

    aniso = lambda X: 1.0 + 0.003 * np.sin(2 * np.pi * X / L)
    m_t, m_u, m_v = msf(Xt, Yt), msf(Xu, Yu), msf(Xv, Yv)
    lat = np.deg2rad(35.0) + (Yt - 0.5 * L) / 111e3 * np.pi / 180.0
    f = 2 * OMEGA * np.sin(lat)
    e = 2 * OMEGA * np.cos(lat)
    sina = 0.02 * np.sin(2 * np.pi * Xt / L)
    cosa = np.sqrt(1 - sina ** 2)

    znw = eta_levels(nz, p_top)
    vc = _vertical_coeffs(znw)
    znu = vc["znu"][:nz - 1]
    dnw = vc["dnw"]

    # base state on mass columns
    ps = P0 * _exner_of_z(ht) ** (1.0 / KAPPA)
    mub = ps - p_top
    pb = znu[:, None, None] * mub[None] + p_top
    theta_b = _theta_of_z(_z_of_p(pb))
    t_init = theta_b - T0
    alb = (R_D / P1000) * (t_init + T0) * (pb / P1000) ** CVPM
    phb = np.zeros((nz, ny, nx))
    phb[0] = G * ht
    for k in range(nz - 1):
        phb[k + 1] = phb[k] - dnw[k] * mub * alb[k]
    zmass = 0.5 * (phb[:-1] + phb[1:]) / G

    # perturbations
    theta = theta_b.copy()
    for (cx, cy, zc, amp, rad) in _bubbles(cfg, rng):
        r = np.sqrt(((Xt - cx * L) / rad) ** 2 + ((Yt - cy * L) / rad) ** 2 + ((zmass - zc) / 1500.0) ** 2)
        theta += np.where(r < 1.0, amp * np.cos(0.5 * np.pi * r) ** 2, 0.0)
    Tk = theta * (pb / P0) ** KAPPA
    rh = np.where(zmass < 8000.0, 0.8 - 0.7 * zmass / 8000.0, 0.1)
    qv = rh * _qsat(Tk, pb)
    moist = np.zeros((7, nz, ny, nx))
    moist[1, :nz - 1] = qv
    if cfg["clouds"]:
        for sp, (lo, hi, amp) in zip((2, 3, 4, 5, 6), ((1500, 5000, 3e-4), (500, 4000, 2e-4), (6000, 11000, 1e-4),
                                                       (4000, 10000, 2e-4), (3000, 8000, 1.5e-4))):
            if sp - 1 > cfg["clouds"]:
                continue
            for _ in range(3):
                cx, cy, rad = rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(15e3, 40e3)
                r = np.sqrt(((Xt - cx * L) / rad) ** 2 + ((Yt - cy * L) / rad) ** 2)
                zc, zh = 0.5 * (lo + hi), 0.5 * (hi - lo)
                rz = np.abs(zmass - zc) / zh
                blob = amp * np.clip(1.0 - r, 0.0, None) * np.clip(1.0 - rz, 0.0, None)
                moist[sp, :nz - 1] += blob

    # moist hydrostatic balance with mu = 0 (WRF idealised initialisation)
    t = theta - T0
    p = np.zeros((nz, ny, nx))
    al = np.zeros((nz, ny, nx))
    rdnw, rdn = vc["rdnw"], vc["rdn"]
    k = nz - 2
    qvf1 = 0.5 * (qv[k] + qv[k])
    qvf2 = 1.0 / (1.0 + qvf1)
    qvf1 = qvf1 * qvf2
    p[k] = -0.5 * (qvf1 * mub) / rdnw[k] / qvf2
    for k in range(nz - 2, -1, -1):
        if k < nz - 2:
            qvf1 = 0.5 * (qv[k] + qv[k + 1])
            qvf2 = 1.0 / (1.0 + qvf1)
            qvf1 = qvf1 * qvf2
            p[k] = p[k + 1] - (qvf1 * mub) / qvf2 / rdn[k + 1]
        qvf = 1.0 + RVOVRD * qv[k]
        alt = (R_D / P1000) * (t[k] + T0) * qvf * (((p[k] + pb[k]) / P1000) ** CVPM)
        al[k] = alt - alb[k]
    ph = np.zeros((nz, ny, nx))
    for k in range(nz - 1):
        ph[k + 1] = ph[k] - dnw[k] * (mub * al[k])

    # winds on staggered points
    zu = np.concatenate([zmass[:, :, :1], 0.5 * (zmass[:, :, 1:] + zmass[:, :, :-1])], axis=2)
    if cfg["jet"] is not None:
        amp, zj, wj = cfg["jet"]
        u = amp * np.exp(-((zu - zj) / wj) ** 2)
    else:
        u = np.full((nz - 1, ny, nx), cfg["u0"])
    v = np.zeros((nz - 1, ny, nx))
    if cfg["noise"] > 0:
        u = u + rng.normal(0.0, cfg["noise"], u.shape)
        v = v + rng.normal(0.0, cfg["noise"], v.shape)

    s = allocate_state(nz, ny, nx, device="cpu", dtype=torch.float64)
    T = lambda a: torch.as_tensor(a, dtype=torch.float64)
    s.u[:nz - 1] = T(u)
    s.v[:nz - 1] = T(v)
    s.t[:nz - 1] = T(t)
    s.ph[:] = T(ph)
    s.p[:nz - 1] = T(p[:nz - 1])
    s.al[:nz - 1] = T(al[:nz - 1])
    s.pb[:nz - 1] = T(pb)
    s.phb[:] = T(phb)
    s.t_init[:nz - 1] = T(t_init)
    s.alb[:nz - 1] = T(alb)
    s.alb[nz - 1] = s.alb[nz - 2]
    s.moist[:] = T(moist)
    s.mub[:] = T(mub)
    s.ht[:] = T(ht)
    s.msft[:], s.msftx[:], s.msfty[:] = T(m_t), T(m_t * aniso(Xt)), T(m_t)
    s.msfu[:], s.msfux[:], s.msfuy[:] = T(m_u), T(m_u * aniso(Xu)), T(m_u)
    s.msfv[:], s.msfvx[:], s.msfvy[:] = T(m_v), T(m_v * aniso(Xv)), T(m_v)
    s.msfvx_inv[:] = 1.0 / s.msfvx
    s.f[:], s.e[:], s.sina[:], s.cosa[:], s.clat[:] = T(f), T(e), T(sina), T(cosa), T(np.rad2deg(lat))
    for name in ("rdnw", "rdn", "fnm", "fnp", "dn", "dnw"):
        getattr(s, name)[:] = T(vc[name])
    s.znu[:nz - 1] = T(znu)
    s.znw[:] = T(znw)
    s.c1h[:nz - 1] = 1.0
    s.c1f[:] = 1.0
    s.c3h[:nz - 1] = T(znu)
    s.c3f[:] = T(znw)
    for name in ("cf1", "cf2", "cf3", "cfn", "cfn1"):
        setattr(s, name, torch.tensor(vc[name], dtype=torch.float64))
    s.u_base[:nz - 1] = s.u[:nz - 1].mean(dim=(1, 2))
    s.v_base[:nz - 1] = s.v[:nz - 1].mean(dim=(1, 2))
    s.qv_base[:nz - 1] = s.moist[1, :nz - 1].mean(dim=(1, 2))
    # edge fix of mub (solver.py:443-446)
    g = s.grid
    s.mub[g.jte - 1, :] = s.mub[g.jte - 2, :]
    s.mub[:, g.ite - 1] = s.mub[:, g.ite - 2]
    s.mub[g.jts - 1, :] = s.mub[g.jts, :]
    s.mub[:, g.its - 1] = s.mub[:, g.its]
    _boundary_arrays(s, cfg.get("bdy_tend"), seed)
    s = _cast(s, device, dtype)
    return s


def _pattern(rng, nz, ny, nx):
    """Smooth seeded field in [-1, 1]: three horizontal waves of non-integer
    wavenumber (so that no two strips of the domain see equal values) with a
    vertical half-sine envelope plus a level-wise tilt."""
    y, x = np.meshgrid(np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    kk = np.arange(nz)[:, None, None] / max(nz - 1, 1)
    out = np.zeros((nz, ny, nx))
    for _ in range(3):
        ax, ay = rng.uniform(0.6, 2.4, size=2)
        px, py, a = rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), rng.uniform(0.5, 1.0)
        out += a * np.sin(2 * np.pi * ax * x + px)[None] * np.cos(2 * np.pi * ay * y + py)[None] \
            * (0.6 + 0.4 * np.sin(np.pi * kk + rng.uniform(0, 1)))
    return out / max(np.abs(out).max(), 1e-30)


def _boundary_arrays(s, bdy_tend=None, seed=0):
    """Mass-coupled boundary values of the initial state; tendencies zero, or
    towards a perturbed state over BDY_PERIOD when bdy_tend gives the
    perturbation amplitudes."""
    g = s.grid
    nz, ny, nx = g.nz, g.ny, g.nx
    ids, ide, jds, jde = g.ids, g.ide, g.jds, g.jde
    mu_t = s.mub + s.mu
    muu = torch.zeros_like(mu_t)
    muu[:, 1:] = 0.5 * (mu_t[:, 1:] + mu_t[:, :-1])
    muu[:, 0] = mu_t[:, 0]
    muv = torch.zeros_like(mu_t)
    muv[1:, :] = 0.5 * (mu_t[1:, :] + mu_t[:-1, :])
    muv[0, :] = mu_t[0, :]
    coupled = {
        "u": s.u * muu / s.msfuy, "v": s.v * muv / s.msfvx, "t": s.t * mu_t, "ph": s.ph * mu_t,
        "qv": s.moist[1] * mu_t, "mu": s.mu.unsqueeze(0),
    }
    # (k range, j range for x-boundaries, i range for y-boundaries) filled by solver.py:468-703
    fill = {
        "u": ((0, nz - 1), (5, ny - 6), (5, nx - 5)), "v": ((0, nz - 1), (5, ny - 5), (5, nx - 6)),
        "ph": ((0, nz), (5, ny - 6), (5, nx - 6)), "t": ((0, nz - 1), (5, ny - 6), (5, nx - 6)),
        "qv": ((0, nz - 1), (5, ny - 6), (5, nx - 6)), "mu": ((0, 1), (5, ny - 6), (5, nx - 6)),
    }
    tends = {}
    if bdy_tend:
        rng = np.random.default_rng(1000 + seed)
        T = lambda a: torch.as_tensor(a, dtype=torch.float64)
        P = {var: T(_pattern(rng, c.shape[0], ny, nx)) for var, c in coupled.items()}
        scale = {"u": muu / s.msfuy, "v": muv / s.msfvx, "t": mu_t, "ph": mu_t, "qv": s.moist[1] * mu_t,
                 "mu": torch.ones_like(mu_t)}
        for var, c in coupled.items():
            tends[var] = bdy_tend[var] * P[var] * scale[var] / BDY_PERIOD
    for var, c in coupled.items():
        (k0, k1), (j0, j1), (i0, i1) = fill[var]
        ibe = ide if var == "u" else ide - 1
        jbe = jde if var == "v" else jde - 1
        for arrs, src in ((s.bdy[var][:4], c), (s.bdy[var][4:], tends.get(var))):
            if src is None:
                continue
            bxs, bxe, bys, bye = arrs
            for b in range(5):
                bxs[k0:k1, j0:j1, b] = src[k0:k1, j0:j1, ids + b]
                bxe[k0:k1, j0:j1, b] = src[k0:k1, j0:j1, ibe - 1 - b]
                bys[k0:k1, b, i0:i1] = src[k0:k1, jds + b, i0:i1]
                bye[k0:k1, b, i0:i1] = src[k0:k1, jbe - 1 - b, i0:i1]


def _cast(s, device, dtype):
    from .state import clone_state
    return clone_state(s, device=device, dtype=dtype)
