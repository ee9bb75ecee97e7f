"""ctypes binding of libpywrf_amd.so (the HIP kernels) for the host mirror.

The argument types of every entry point are read from include/pywrf_amd.h,
so the C declarations are the single description of the ABI.  torch is
imported first: the library then binds the HIP runtime torch has loaded, and
kernels launch on torch's current stream.  A missing library raises
ImportError at import time -- there is no fallback path.
"""
import ctypes
import os
import re

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.environ.get("PYWRF_AMD_LIB") or os.path.join(HERE, "libpywrf_amd.so")   # override: another build
HEADER = os.path.join(ROOT, "include", "pywrf_amd.h")

PYWRF_E_ARG = -1
PYWRF_E_RANGE = -2

_GRID_FIELDS = ("nz ny nx ids ide jds jde kds kde ims ime jms jme kms kme "
                "ips ipe jps jpe kps kpe its ite jts jte kts kte").split()


class Grid(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in _GRID_FIELDS]


_GridPtr = ctypes.POINTER(Grid)


class Surrogate(ctypes.Structure):
    """pywrf_surrogate_t: device pointers of the surrogate network and its
    standardisation, the fallback tolerances, and the weights in fragment
    order."""
    _fields_ = [(n, ctypes.c_void_p) for n in
                "w1 w2 w3 b1 b2 b3 in_mean in_rstd out_mean out_std in_lo in_hi out_absmax".split()] + [
               ("qtol", ctypes.c_float), ("wtol", ctypes.c_float), ("envelope", ctypes.c_float),
               ("dtmax", ctypes.c_float)] + [(n, ctypes.c_void_p) for n in "w1f w2f w3f".split()]


_CTYPES = {
    "const pywrf_grid_t*": ctypes.POINTER(Grid),
    "const float*": ctypes.c_void_p,
    "float*": ctypes.c_void_p,
    "double": ctypes.c_double,
    "int": ctypes.c_int,
    "char": ctypes.c_char,
    "pywrf_stream_t": ctypes.c_void_p,
    "const char*": ctypes.c_char_p,
    "double*": ctypes.POINTER(ctypes.c_double),
    "void* const*": ctypes.POINTER(ctypes.c_void_p),
    "const int*": ctypes.POINTER(ctypes.c_int),
    "const long long*": ctypes.POINTER(ctypes.c_longlong),
    "long long*": ctypes.POINTER(ctypes.c_longlong),
    "long long": ctypes.c_longlong,
    "const void*": ctypes.c_void_p,
    "void*": ctypes.c_void_p,
    "int*": ctypes.c_void_p,
    "const pywrf_surrogate_t*": ctypes.c_void_p,
}


def parse_header(path=HEADER, with_names=False):
    """{name: [ctypes argtypes]} for every `int pywrf_*(...)` declaration
    (with_names: {name: [parameter names]} as well)."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    decls, pnames = {}, {}
    for m in re.finditer(r"\bint\s+(pywrf_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        name, args = m.group(1), " ".join(m.group(2).split())
        types, names = [], []
        if args not in ("", "void"):
            for a in args.split(","):
                a = a.strip()
                t = re.sub(r"\s*\b\w+$", "", a).strip()       # drop the parameter name
                t = re.sub(r"\s*\*", "*", t)
                t = re.sub(r"\*\s*const\*", "* const*", t)
                if t not in _CTYPES:
                    raise ValueError("unsupported C type %r in %s" % (t, name))
                types.append(_CTYPES[t])
                names.append(re.search(r"(\w+)$", a).group(1))
        decls[name] = types
        pnames[name] = names
    return (decls, pnames) if with_names else decls


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("%s is missing: build the HIP kernels first (python -m pywrf_amd.build)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    decls, pnames = parse_header(with_names=True)
    for name, types in decls.items():
        f = getattr(lib, name)          # AttributeError here means the library lacks a declared symbol
        f.argtypes = types
        f.restype = ctypes.c_int
    return lib, decls, pnames


LIB, DECLS, PARAMS = _load()


class KernelError(RuntimeError):
    pass


def check(name, rc):
    if rc == 0:
        return
    if rc == PYWRF_E_ARG:
        raise NameError("%s: argument rejected (unknown staggering / variable flag)" % name)
    if rc == PYWRF_E_RANGE:
        raise IndexError("%s: index sets reach outside the allocation" % name)
    raise KernelError("%s: HIP error %d" % (name, rc))


# Operand roles by C parameter name (include/pywrf_amd.h).  A kernel indexes
# a 2-D operand over the grid's (ny, nx) plane, a level array over nz entries
# (or the fixed count given here) and every other array operand over the
# whole (nz, ny, nx) grid (species arrays: over at least that).
PLANE_PARAMS = frozenset((
    "mu mub mut muu muv mu_1 mu_2 mu_in mu_new mu_old mu_out mu_save mu_tend mu_tendf muave muave_in "
    "muave_out mubf mudf mudf_in mudf_out muf muts muts_in muts_out muuf muus muvf muvs msf msft msftx "
    "msfty msfu msfux msfuy msfv msfvx msfvx_inv msfvy ht f e sina cosa rain rainnc rainncv snow snownc "
    "snowncv graupel graupelnc graupelncv sr mu_base lam_mu lam_mudf lam_muave lam_muts lam_mu_ori "
    "lam_muave_ori lam_muts_ori lam_mudf_ori lam_mu_tend lam_mut lam_mu_1 lam_mu_2 lam_mu_2_in lam_muus lam_muvs "
    "lam_muu lam_muv lam_muu_in lam_muv_in lam_mu_save work2d").split())
LEVEL_PARAMS = dict.fromkeys("c1h dn dnw fnm fnp fzm fzp rdn rdnw rdzu u_base v_base base".split())
LEVEL_PARAMS.update(fcx=5, gcx=5, cf=3, cfn=1, cfn1=1, dtbc=1)
# names whose role depends on the function: the advection operators' rdzw is
# the level array rdnw / rdn (wrf_dynamics.py:1112, 1332, 1535, 1841), the
# diffusion metrics' rdzw a 3-D field
LEVEL_IN = {"rdzw": "pywrf_advect_"}


class TensorPtr(ctypes.c_void_p):
    """Device pointer of a tensor argument; `shape` is checked against the
    call's grid before the launch."""


class BoundaryPtr(ctypes.c_void_p):
    """Device pointer of a lateral-boundary array (checked by bptr)."""


def check_operands(name, g, params, args):
    """Raise IndexError when an array operand is smaller than the part of it
    the kernel addresses for grid g (a Grid), so that no mismatched operand
    reaches a kernel.  params: the C parameter names, args: the arguments."""
    nz, ny, nx = g.nz, g.ny, g.nx
    for pname, a in zip(params, args):
        if not isinstance(a, TensorPtr):
            continue
        shape = a.shape
        if pname in LEVEL_PARAMS or (pname in LEVEL_IN and name.startswith(LEVEL_IN[pname])):
            need = LEVEL_PARAMS.get(pname) or nz
            if pname == "cf" and len(shape) == 0:
                continue                      # (cf1, cf2, cf3) checked adjacent by the caller
            n = 1
            for d in shape:
                n *= d
            if n < need:
                raise IndexError("%s: %s has %d entries, the grid needs %d" % (name, pname, n, need))
            continue
        if len(shape) < 2 or tuple(shape[-2:]) != (ny, nx):
            raise IndexError("%s: %s has shape %s, the grid is (%d, %d, %d)" % (name, pname, tuple(shape), nz, ny, nx))
        if pname in PLANE_PARAMS:
            continue
        lev = shape[-3] if len(shape) >= 3 else 1
        if lev < nz:
            raise IndexError("%s: %s has %d levels, the grid needs %d (shape %s)" % (name, pname, lev, nz,
                                                                                    tuple(shape)))


def call(name, *args):
    if args and isinstance(args[0], _GridPtr):
        check_operands(name, args[0].contents, PARAMS[name], args)
    check(name, getattr(LIB, name)(*args))


_grid_cache = {}


def grid(shape, idx):
    """pywrf_grid_t for an array shape and an index tuple of 16, 18 or 24 ints
    (the reference's 2-D, tile-only and patch+tile argument lists)."""
    key = (tuple(shape), tuple(idx))
    g = _grid_cache.get(key)
    if g is not None:
        return g
    nz, ny, nx = (1,) * (3 - len(shape)) + tuple(shape[-3:])
    v = dict(nz=nz, ny=ny, nx=nx)
    plane = len(shape) == 2
    if len(idx) == 16:
        (v["ids"], v["ide"], v["jds"], v["jde"], v["ims"], v["ime"], v["jms"], v["jme"], v["ips"], v["ipe"],
         v["jps"], v["jpe"], v["its"], v["ite"], v["jts"], v["jte"]) = idx
        v.update(kds=0, kde=1, kms=0, kme=1, kps=0, kpe=1, kts=0, kte=1)
    elif len(idx) == 18:
        (v["ids"], v["ide"], v["jds"], v["jde"], v["kds"], v["kde"], v["ims"], v["ime"], v["jms"], v["jme"],
         v["kms"], v["kme"], v["its"], v["ite"], v["jts"], v["jte"], v["kts"], v["kte"]) = idx
        v.update(ips=v["its"], ipe=v["ite"], jps=v["jts"], jpe=v["jte"], kps=v["kts"], kpe=v["kte"])
    elif len(idx) == 24:
        for n, x in zip(_GRID_FIELDS[3:], idx):
            v[n] = x
    else:
        raise TypeError("expected 16, 18 or 24 index arguments, got %d" % len(idx))
    if plane:
        # a 2-D field: its level count is irrelevant, so take the level
        # indices' extent (the kernels only use the (j, i) plane)
        v["nz"] = max(1, v["kte"], v["kde"])
    g = ctypes.pointer(Grid(**{n: int(v[n]) for n in _GRID_FIELDS}))
    _grid_cache[key] = g
    return g


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_device(t):
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a tensor, got %r" % type(t))
    if t.device.type != "cuda" or t.dtype != torch.float32:
        raise TypeError("expected a float32 CUDA tensor, got %s %s" % (t.device, t.dtype))


def ptr(t):
    """Device pointer of a contiguous FP32 CUDA tensor (raises otherwise)."""
    _require_device(t)
    if not t.is_contiguous():
        raise TypeError("expected a contiguous float32 CUDA tensor, got strides %s" % (t.stride(),))
    p = TensorPtr(t.data_ptr())
    p.shape = tuple(t.shape)
    return p


_bdy_copies = {}


def boundary_layout(t, kind, kz, n):
    """The boundary array t of the reference driver, as a tensor whose storage
    has the kernels' layout: an x-boundary (kind 'x', shape (kz, n, 5)) is
    stored as a contiguous (5, kz, n) array, a y-boundary (kind 'y', shape
    (kz, 5, n)) as a contiguous (5, kz, n) array -- the layout of the
    driver's own permuted views (solver.py:468-703), which are used in place.
    Arrays in any other layout are copied into that layout; the copy is kept
    while the array's data and version are unchanged."""
    _require_device(t)
    want = (kz, n, 5) if kind == "x" else (kz, 5, n)
    if tuple(t.shape) != want:
        raise IndexError("boundary array has shape %s, expected %s (spec_bdy_width 5)" % (tuple(t.shape), want))
    strides = (n, 1, kz * n) if kind == "x" else (n, kz * n, 1)
    if tuple(t.stride()) == strides or t.numel() == 0:
        return t
    key = (t.data_ptr(), tuple(t.stride()), kind)
    hit = _bdy_copies.get(key)
    if hit is not None and hit[0] == t._version:
        return hit[1]
    c = (t.permute(2, 0, 1) if kind == "x" else t.permute(1, 0, 2)).contiguous()
    c = c.permute(1, 2, 0) if kind == "x" else c.permute(1, 0, 2)
    if len(_bdy_copies) > 256:
        _bdy_copies.clear()
    _bdy_copies[key] = (t._version, c)
    return c


def bptr(t, kind, kz, n):
    """Device pointer of a boundary array in the kernels' layout (see
    boundary_layout)."""
    c = boundary_layout(t, kind, kz, n)
    p = BoundaryPtr(c.data_ptr())
    p.keep = c
    return p


def bdy_ptrs(arrs, kz, ny, nx):
    """Pointers of (xs, xe, ys, ye) or (xs, xe, ys, ye, txs, txe, tys, tye)."""
    return [bptr(a, "x" if n % 4 < 2 else "y", kz, ny if n % 4 < 2 else nx) for n, a in enumerate(arrs)]


def flag(s):
    """Staggering / variable string of the reference as one C char."""
    if not isinstance(s, str) or len(s) != 1:
        raise NameError("unsupported variable flag %r" % (s,))
    return ctypes.c_char(s.encode())


def fresh(like):
    return torch.empty_like(like, memory_format=torch.contiguous_format)
