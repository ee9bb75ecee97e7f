"""WRF files into the state (CPU).

* pywrf_amd.netcdf against an independent writer (scipy.io.netcdf_file) in
  both the classic and the 64-bit-offset format: record and fixed variables
  of every numeric type, attributes, several records.
* load_input / load_boundary against files written from a seeded state with
  the reference's own slices (solver.py:213-281 for the input, 461-703 for
  the boundary sides): every field comes back at exactly the place it was
  taken from, the rest stays zero, alb / al / the mub edges follow
  solver.py:284-306, and a boundary record other than the first is loaded
  when asked for.
* the refresh schedule of the time loop (solver.py:457-458).
"""
import os

import numpy as np
import pytest
import torch

scipy_io = pytest.importorskip("scipy.io")


def _write(path, dims, variables, version=2, records=None):
    """variables: name -> (dims, numpy array[, attrs])."""
    f = scipy_io.netcdf_file(path, "w", version=version)
    f.title = "test"
    for d, n in dims.items():
        f.createDimension(d, n)
    for name, spec in variables.items():
        vdims, arr = spec[0], spec[1]
        v = f.createVariable(name, arr.dtype, vdims)
        v[:] = arr
        for k, val in (spec[2] if len(spec) > 2 else {}).items():
            setattr(v, k, val)
    f.close()


@pytest.mark.parametrize("version", [1, 2])
def test_reader_matches_an_independent_writer(tmp_path, version):
    from pywrf_amd.netcdf import Dataset
    rng = np.random.default_rng(0)
    path = str(tmp_path / "t.nc")
    a = rng.standard_normal((3, 4, 5, 6)).astype(np.float32)
    b = rng.standard_normal((3, 7)).astype(np.float64)
    c = rng.integers(-5, 5, (3, 2)).astype(np.int32)
    d = rng.standard_normal((4, 5)).astype(np.float32)
    e = rng.integers(-5, 5, (3, 3)).astype(np.int16)
    _write(path, {"Time": None, "z": 4, "y": 5, "x": 6, "w": 7, "two": 2, "three": 3},
           {"A": (("Time", "z", "y", "x"), a, {"units": "m s-1", "scale": np.float32(2.5)}),
            "B": (("Time", "w"), b), "C": (("Time", "two"), c), "D": (("z", "y"), d), "E": (("Time", "three"), e)},
           version=version)
    with Dataset(path) as ds:
        assert ds.version == version and ds.numrecs == 3
        assert ds.attrs["title"] == "test"
        assert ds["A"].dims == ("Time", "z", "y", "x")
        assert ds["A"].attrs["units"] == "m s-1" and ds["A"].attrs["scale"] == 2.5
        for name, ref in (("A", a), ("B", b), ("C", c), ("D", d), ("E", e)):
            got = ds[name][...]
            assert got.dtype == ref.dtype and np.array_equal(got, ref), name
        assert np.array_equal(ds["A"][1], a[1])


def test_reader_refuses_other_files(tmp_path):
    from pywrf_amd.netcdf import Dataset
    p = tmp_path / "h5.nc"
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(ValueError, match="HDF5"):
        Dataset(str(p))
    q = tmp_path / "x.nc"
    q.write_bytes(b"not a netcdf file")
    with pytest.raises(ValueError):
        Dataset(str(q))


def _reference_slices(nz, ny, nx):
    """The reference's copies into the state (solver.py:213-281): state
    attribute -> (file variable, region of the state array)."""
    K1, K = slice(0, nz - 1), slice(0, nz)
    jm, jv = slice(5, ny - 6), slice(5, ny - 5)
    im, iu = slice(5, nx - 6), slice(5, nx - 5)
    r = {"u": ("U", (K1, jm, iu)), "v": ("V", (K1, jv, im)), "mu": ("MU", (jm, im)), "mub": ("MUB", (jm, im)),
         "msfu": ("MAPFAC_U", (jm, iu)), "msfv": ("MAPFAC_V", (jv, im)), "msfvx_inv": ("MF_VX_INV", (jv, im)),
         "msft": ("MAPFAC_M", (jm, im)), "msftx": ("MAPFAC_MX", (jm, im)), "msfty": ("MAPFAC_MY", (jm, im)),
         "msfux": ("MAPFAC_UX", (jm, iu)), "msfuy": ("MAPFAC_UY", (jm, iu)), "msfvx": ("MAPFAC_VX", (jv, im)),
         "msfvy": ("MAPFAC_VY", (jv, im)), "pb": ("PB", (K1, jm, im)), "p": ("P", (K1, jm, im)),
         "t_init": ("T_INIT", (K1, jm, im)), "phb": ("PHB", (K, jm, im)), "ph": ("PH", (K, jm, im)),
         "t": ("T", (K1, jm, im)), "f": ("F", (jm, im)), "e": ("E", (jm, im)), "sina": ("SINALPHA", (jm, im)),
         "cosa": ("COSALPHA", (jm, im)), "clat": ("CLAT", (jm, im)), "ht": ("HGT", (jm, im)),
         "znu": ("ZNU", (K1,)), "znw": ("ZNW", (K,))}
    for n in "rdnw rdn c1h c2h c3h c4h fnm fnp dn dnw u_base v_base qv_base".split():
        r[n] = ({"rdnw": "RDNW", "rdn": "RDN", "fnm": "FNM", "fnp": "FNP", "dn": "DN", "dnw": "DNW",
                 "u_base": "U_BASE", "v_base": "V_BASE", "qv_base": "QV_BASE"}.get(n, n.upper()), (K1,))
    for n in "c1f c2f c3f c4f".split():
        r[n] = (n.upper(), (K,))
    for sp, name in enumerate(("QVAPOR", "QCLOUD", "QRAIN", "QICE", "QSNOW", "QGRAUP"), start=1):
        r[("moist", sp)] = (name, (K1, jm, im))
    return r


def _get(s, key):
    return getattr(s, key[0])[key[1]] if isinstance(key, tuple) else getattr(s, key)


def _dims_of(region, shape):
    names = []
    for n, r in enumerate(region):
        names.append("d%d_%d" % (n, r.stop - r.start))
    return names


def _write_input(path, s):
    g = s.grid
    sl = _reference_slices(g.nz, g.ny, g.nx)
    dims, variables = {"Time": None}, {}
    for key, (var, region) in sl.items():
        arr = _get(s, key)[region].numpy()[None]
        names = _dims_of(region, arr.shape)
        for nm, n in zip(names, arr.shape[1:]):
            dims[nm] = n
        variables[var] = (("Time",) + tuple(names), arr.astype(np.float32))
    for var in ("CF1", "CF2", "CF3", "CFN", "CFN1"):
        variables[var] = (("Time",), np.array([float(getattr(s, var.lower()))], dtype=np.float32))
    w = s.w[0:g.nz, 5:g.ny - 6, 5:g.nx - 6].numpy()[None]
    wd = _dims_of((slice(0, g.nz), slice(5, g.ny - 6), slice(5, g.nx - 6)), w.shape)
    for nm, n in zip(wd, w.shape[1:]):
        dims[nm] = n
    variables["W"] = (("Time",) + tuple(wd), w.astype(np.float32))
    _write(path, dims, variables)


def test_input_fields_land_where_the_reference_puts_them(tmp_path):
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.state import allocate_state
    from pywrf_amd import wrfio
    src = synthetic_state("config1c")
    path = str(tmp_path / "wrfinput")
    _write_input(path, src)
    g = src.grid
    s = allocate_state(g.nz, g.ny, g.nx)
    mub_before_fix = None
    wrfio.load_input(s, path)
    for key, (var, region) in _reference_slices(g.nz, g.ny, g.nx).items():
        got, ref = _get(s, key), _get(src, key)
        if key == "mub":
            continue
        assert torch.equal(got[region], ref[region]), var
        rest = torch.ones_like(got, dtype=torch.bool)
        rest[region] = False
        assert not bool(got[rest].any()), var
    assert torch.equal(s.w[0:g.nz, 5:g.ny - 6, 5:g.nx - 6], src.w[0:g.nz, 5:g.ny - 6, 5:g.nx - 6])
    for n in ("cf1", "cf2", "cf3", "cfn", "cfn1"):
        assert float(getattr(s, n)) == float(getattr(src, n))
    # alb over the whole array and al on the interior (solver.py:284-300), in FP64 from the loaded fields
    t_init, pb, ph, mu = (getattr(s, n).double() for n in ("t_init", "pb", "ph", "mu"))
    alb = 287.0 / 100000.0 * (t_init + 300.0) * (pb / 100000.0) ** -0.7142857
    assert torch.allclose(s.alb.double(), alb, rtol=2e-6, atol=0)
    k1, j0, j1, i0, i1 = g.kte - 1, g.jts, g.jde - 1, g.its, g.ide - 1
    mub = torch.zeros_like(s.mub, dtype=torch.float64)
    region = _reference_slices(g.nz, g.ny, g.nx)["mub"][1]
    mub[region] = src.mub[region].double()
    al = -1.0 / (mub[j0:j1, i0:i1] + mu[j0:j1, i0:i1]) * (alb[:k1, j0:j1, i0:i1] * mu[j0:j1, i0:i1] +
                                                          s.rdnw[:k1].double().view(-1, 1, 1) *
                                                          (ph[1:k1 + 1, j0:j1, i0:i1] - ph[:k1, j0:j1, i0:i1]))
    err = float((s.al[:k1, j0:j1, i0:i1].double() - al).abs().max() / al.abs().max())
    assert err < 1e-5
    # mub edges copied from their neighbours (solver.py:303-306)
    assert torch.equal(s.mub[g.jte - 1, 5:g.nx - 6], s.mub[g.jte - 2, 5:g.nx - 6])
    assert torch.equal(s.mub[5:g.ny - 6, g.its - 1], s.mub[5:g.ny - 6, g.its])


def _bdy_regions(nz, ny, nx):
    """The reference's boundary copies (solver.py:461-703): variable ->
    (x-side region, y-side region) of the (5, kz, n) storage."""
    K1, K = slice(0, nz - 1), slice(0, nz)
    A = slice(None)
    return {"U": ((A, K1, slice(5, ny - 6)), (A, K1, slice(5, nx - 5))),
            "V": ((A, K1, slice(5, ny - 5)), (A, K1, slice(5, nx - 6))),
            "PH": ((A, K, slice(5, ny - 6)), (A, K, slice(5, nx - 6))),
            "T": ((A, K1, slice(5, ny - 6)), (A, K1, slice(5, nx - 6))),
            "MU": ((A, 0, slice(5, ny - 6)), (A, 0, slice(5, nx - 6))),
            "QVAPOR": ((A, K1, slice(5, ny - 6)), (A, K1, slice(5, nx - 6)))}


def test_boundary_records_land_in_the_drivers_layout(tmp_path):
    from pywrf_amd.gen import synthetic_state
    from pywrf_amd.state import allocate_state
    from pywrf_amd import wrfio
    src = synthetic_state("bdy24")
    g = src.grid
    keys = {"U": "u", "V": "v", "PH": "ph", "T": "t", "MU": "mu", "QVAPOR": "qv"}
    dims, variables, expect = {"Time": None}, {}, {}
    for var, (rx, ry) in _bdy_regions(g.nz, g.ny, g.nx).items():
        for n, side in enumerate(wrfio.BDY_SIDES):
            x_side = n % 4 < 2
            st = wrfio._storage(src.bdy[keys[var]][n], x_side)
            region = rx if x_side else ry
            base = st[region].numpy()
            rec = np.stack([base, 2.0 * base + 1.0]).astype(np.float32)
            names = ["b%d_%d" % (m, k) for m, k in enumerate(rec.shape[1:])]
            for nm, k in zip(names, rec.shape[1:]):
                dims[nm] = k
            variables["%s_%s" % (var, side)] = (("Time",) + tuple(names), rec)
            expect[(keys[var], n)] = (region, torch.from_numpy(rec[1]))
    path = str(tmp_path / "wrfbdy")
    _write(path, dims, variables)
    s = allocate_state(g.nz, g.ny, g.nx)
    for arrs in s.bdy.values():
        for a in arrs:
            a.fill_(7.0)                      # stale values of a previous record
    ptrs = {k: [a.data_ptr() for a in v] for k, v in s.bdy.items()}
    wrfio.load_boundary(s, path, 1)
    for (key, n), (region, rec) in expect.items():
        st = wrfio._storage(s.bdy[key][n], n % 4 < 2)
        assert torch.equal(st[region], rec), (key, n)
        rest = torch.ones_like(st, dtype=torch.bool)
        rest[region] = False
        assert not bool(st[rest].any()), (key, n)
    assert ptrs == {k: [a.data_ptr() for a in v] for k, v in s.bdy.items()}      # in place
    with pytest.raises(IndexError):
        wrfio.load_boundary(s, path, 2)


def test_refresh_follows_the_time_loop(monkeypatch):
    from types import SimpleNamespace
    from pywrf_amd import wrfio
    loaded = []
    monkeypatch.setattr(wrfio, "load_boundary", lambda s, f, it: loaded.append((s.step + 1, it)))
    s = SimpleNamespace(step=0)
    cfg = SimpleNamespace(bdy_interval=360)
    for _ in range(722):
        wrfio.refresh_boundaries(s, None, cfg)
        s.step += 1
    assert loaded == [(1, 0), (361, 1), (721, 2)]
