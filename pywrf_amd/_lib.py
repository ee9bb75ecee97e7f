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
}


def parse_header(path=HEADER):
    """{name: [ctypes argtypes]} for every `int pywrf_*(...)` declaration."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"\bint\s+(pywrf_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.S):
        name, args = m.group(1), " ".join(m.group(2).split())
        types = []
        if args not in ("", "void"):
            for a in args.split(","):
                a = a.strip()
                t = re.sub(r"\s*\b\w+$", "", a).strip()       # drop the parameter name
                t = re.sub(r"\s*\*", "*", t)
                t = re.sub(r"\*\s*const\*", "* const*", t)
                if t not in _CTYPES:
                    raise ValueError("unsupported C type %r in %s" % (t, name))
                types.append(_CTYPES[t])
        decls[name] = types
    return decls


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("%s is missing: build the HIP kernels first (python -m pywrf_amd.build)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    decls = parse_header()
    for name, types in decls.items():
        f = getattr(lib, name)          # AttributeError here means the library lacks a declared symbol
        f.argtypes = types
        f.restype = ctypes.c_int
    return lib, decls


LIB, DECLS = _load()


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


def call(name, *args):
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


def ptr(t):
    """Device pointer of a contiguous FP32 CUDA tensor (raises otherwise)."""
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a tensor, got %r" % type(t))
    if t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError("expected a contiguous float32 CUDA tensor, got %s %s contiguous=%s"
                        % (t.device, t.dtype, t.is_contiguous()))
    return ctypes.c_void_p(t.data_ptr())


def flag(s):
    """Staggering / variable string of the reference as one C char."""
    if not isinstance(s, str) or len(s) != 1:
        raise NameError("unsupported variable flag %r" % (s,))
    return ctypes.c_char(s.encode())


def fresh(like):
    return torch.empty_like(like, memory_format=torch.contiguous_format)
