"""ctypes access to the CPU oracle (oracle/hsl_oracle.c, oracle/gourdon_oracle.c,
oracle/dr_oracle.c).

Test infrastructure: the oracle is the checker the HIP engine is compared
against; it is built into oracle/liboracle.so by build() here (or
__graft_entry__.build()).
"""
import ctypes
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "oracle", "liboracle.so")
SRCS = [os.path.join(REPO, "oracle", "hsl_oracle.c"),
        os.path.join(REPO, "oracle", "gourdon_oracle.c"),
        os.path.join(REPO, "oracle", "dr_oracle.c")]
SRC = SRCS[0]


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        for name in ("oracle_D", "oracle_S2_hard"):
            fn = getattr(lib, name)
            fn.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                           ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64),
                           ctypes.POINTER(ctypes.c_int64)]
            fn.restype = None
        lib.oracle_phi_direct.argtypes = [ctypes.c_int64, ctypes.c_int64]
        lib.oracle_phi_direct.restype = ctypes.c_int64
        lib.oracle_gourdon_term.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_gourdon_term.restype = None
        lib.oracle_dr_term.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_dr_term.restype = None
        lib.oracle_pi_sieve.argtypes = [ctypes.c_int64]
        lib.oracle_pi_sieve.restype = ctypes.c_int64

    def _call(self, fn, x, y, z, k, seg):
        x = int(x)
        u = x & ((1 << 128) - 1)
        lo, hi = ctypes.c_uint64(), ctypes.c_int64()
        fn(u & ((1 << 64) - 1), ctypes.c_int64(u >> 64).value, int(y), int(z), int(k), int(seg),
           ctypes.byref(lo), ctypes.byref(hi))
        return (hi.value << 64) | lo.value

    def D(self, x, y, z, k, seg=1 << 16):
        return self._call(self.lib.oracle_D, x, y, z, k, seg)

    def S2_hard(self, x, y, z, c, seg=1 << 16):
        return self._call(self.lib.oracle_S2_hard, x, y, z, c, seg)

    def phi_direct(self, n, a):
        return self.lib.oracle_phi_direct(int(n), int(a))

    def gourdon_term(self, which, x, y, z, k, x_star):
        """which: "AC", "B", "Phi0" or "Sigma" (plain sums over the definitions)."""
        idx = {"AC": 0, "B": 1, "Phi0": 2, "Sigma": 3}[which]
        u = int(x) & ((1 << 128) - 1)
        lo, hi = ctypes.c_uint64(), ctypes.c_int64()
        self.lib.oracle_gourdon_term(idx, u & ((1 << 64) - 1), ctypes.c_int64(u >> 64).value,
                                     int(y), int(z), int(k), int(x_star),
                                     ctypes.byref(lo), ctypes.byref(hi))
        return (hi.value << 64) | lo.value

    def dr_term(self, which, x, y, z, c):
        """which: "S1", "S2_trivial", "S2_easy" or "P2" (plain sums over the
        definitions of the Deleglise-Rivat terms)."""
        idx = {"S1": 0, "S2_trivial": 1, "S2_easy": 2, "P2": 3}[which]
        u = int(x) & ((1 << 128) - 1)
        lo, hi = ctypes.c_uint64(), ctypes.c_int64()
        self.lib.oracle_dr_term(idx, u & ((1 << 64) - 1), ctypes.c_int64(u >> 64).value,
                                int(y), int(z), int(c), ctypes.byref(lo), ctypes.byref(hi))
        return (hi.value << 64) | lo.value

    def pi_sieve(self, n):
        return self.lib.oracle_pi_sieve(int(n))


def build(force=False):
    if force or (not os.path.exists(LIB)) or \
            os.path.getmtime(LIB) < max(os.path.getmtime(s) for s in SRCS):
        subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-o", LIB] + SRCS + ["-lm"], check=True)
    return LIB


def load():
    build()
    return Oracle(ctypes.CDLL(LIB))
