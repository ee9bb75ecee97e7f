"""Build the HIP kernels of the ARW step into pywrf_amd/libpywrf_amd.so (gfx950).

    python -m pywrf_amd.build            # compile every csrc/*.hip, link

Each translation unit is compiled with hipcc in parallel and linked into one
shared library exporting the C ABI declared in include/pywrf_amd.h.  Flags:
-ffp-contract=off keeps every multiply and add separately rounded (the
reference's element-wise tensor arithmetic never fuses), packed FP32
instructions are disabled, and no fast-math is used.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpywrf_amd.so")
ARCH = os.environ.get("PYWRF_AMD_ARCH", "gfx950")

CFLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wdouble-promotion", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "--offload-arch=" + ARCH, "-I" + os.path.join(ROOT, "include")]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=False, jobs=None):
    """Compile and link; raises RuntimeError with the compiler output on failure."""
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        cmd = [hipcc] + CFLAGS + ["-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        return src, obj, r.returncode, out

    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    errs = [(s, o) for s, _, rc, o in results if rc != 0]
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join("%s:\n%s" % e for e in errs))
    if verbose:
        for s, _, _, o in results:
            if o.strip():
                print(o)
    tmp = OUT + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + [r[1] for r in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
