"""Build libprimecount_mi355x.so in-tree with hipcc for gfx950.

    python -m primecount_amd.build        (or primecount_amd.build.build())

Sources live in primecount_amd/csrc; objects go to primecount_amd/_build and
the shared library to primecount_amd/libprimecount_mi355x.so.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libprimecount_mi355x.so")

ARCH = os.environ.get("PCG_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]
# Packed FP32 instructions are kept out of every HIP code object.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# host-only translation units use the HIP runtime API without device code
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]

SOURCES = [
    # (file, extra flags, has device code)
    ("hsl_kernels.hip", [], True),
    ("pix_kernels.hip", [], True),
    ("pix.cpp", [], False),
    ("engine.cpp", [], False),
    ("tables.cpp", [], False),
    # exact IEEE double evaluation of the alpha formulas: no fused multiply-add
    ("params.cpp", ["-ffp-contract=off"], False),
    ("capi.cpp", [], False),
    ("serve.cpp", [], False),
    ("verify.cpp", [], False),
    ("tuner.cpp", [], False),
    ("lds_probe.hip", [], True),
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build the MI355X engine)")


def _compile(src, extra, device):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + extra + (DEVICE if device else HOST) + \
        ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), out))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
