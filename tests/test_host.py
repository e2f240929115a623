"""Host-side logic and the C-ABI boundary, on the CPU (no GPU calls).

- the table builders against the arithmetic definitions (a C++ harness
  compiled here from primecount_amd/csrc/tables.cpp);
- the kernel's leaf quotient (FP64 reciprocal estimate plus integer fix-up
  loops) under reciprocal errors far above the hardware's;
- the reference's pi(x) result check (Schoenfeld's bound) and the typed
  status codes;
- the shared library loads and exports every function that
  include/primecount_mi355x.h declares; the C++ overload header compiles;
- without a device the engine refuses to run (no silent CPU fallback).
"""
import ctypes
import math
import os
import re
import subprocess

import pytest

from tests.conftest import golden

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_table_builders_match_definitions(tmp_path):
    exe = tmp_path / "tables_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", str(exe),
                    os.path.join(REPO, "tests", "native", "tables_check.cpp"),
                    os.path.join(REPO, "primecount_amd", "csrc", "tables.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


def test_leaf_quotient_estimate_is_exact(tmp_path):
    # the sieve kernel's FP64-reciprocal quotient with its integer fix-up
    # loops (the kernel's own code, quotient_fixup.hpp), under reciprocals
    # with relative errors up to 2^-18, 2^-10 and 2^-4 (the hardware's is
    # near 2^-24)
    exe = tmp_path / "leaf_stop_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    os.path.join(REPO, "tests", "native", "leaf_stop_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


def test_easy_leaf_quotient_is_exact(tmp_path):
    # the A/C leaf-sum quotient (reciprocal + two Newton steps + one fix-up)
    # and the table word index derived from it, for quotients below 2^49
    exe = tmp_path / "div49_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(REPO, "tests", "native", "div49_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


def declared_functions():
    text = open(os.path.join(REPO, "include", "primecount_mi355x.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pcg_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(pc):
    lib = ctypes.CDLL(pc.library_path())
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_cpp_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "primecount_mi355x.hpp"\n'
                   'namespace pm = primecount_mi355x;\n'
                   'int main() {\n'
                   '  int64_t x = 1000; pm::int128_t X = 1000;\n'
                   '  int64_t s = pm::D(x, 10, 10, 2) + pm::S2_hard(x, 10, 100, 2) + pm::pi(x)\n'
                   '    + pm::nth_prime(x) + pm::AC(x, 10, 10, 2) + pm::B(x, 10)\n'
                   '    + pm::Phi0(x, 10, 10, 2) + pm::Sigma(x, 10) + pm::S1(x, 10, 5)\n'
                   '    + pm::S2_trivial(x, 10, 100, 5) + pm::S2_easy(x, 10, 100, 5)\n'
                   '    + pm::P2(x, 10, 4) + pm::pi_deleglise_rivat(x, 1);\n'
                   '  pm::int128_t t = pm::D(X, 10, 10, 2) + pm::pi(X) + pm::AC(X, 10, 10, 2)\n'
                   '    + pm::B(X, 10) + pm::Phi0(X, 10, 10, 2) + pm::Sigma(X, 10)\n'
                   '    + pm::S1(X, 10, 5) + pm::S2_trivial(X, 10, 100, 5) + pm::S2_easy(X, 10, 100, 5)\n'
                   '    + pm::P2(X, 10, 4) + pm::pi_deleglise_rivat(X) + pm::S2_hard(X, 10, 100, 2);\n'
                   '  return (int) (s + (int64_t) t);\n'
                   '}\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(REPO, "include"),
                    str(src)], check=True)


def test_version_and_no_silent_fallback(pc):
    lib = pc.load()
    assert b"gfx950" in lib.pcg_version()
    if pc.device_available():
        pytest.skip("a GPU is present; the no-device path is covered on CPU hosts")
    with pytest.raises(pc.PrimecountError, match="HIP device|gfx950"):
        pc.D(10 ** 12, 76500, 76500, 8)
    # a serving context cannot be created without the device either
    with pytest.raises(pc.PrimecountError, match="HIP device|gfx950"):
        pc.Server()


def test_argument_errors(pc):
    with pytest.raises(pc.PrimecountError):
        pc.set_segment_words(1000)            # not a multiple of 128
    with pytest.raises(pc.PrimecountError):
        pc.set_host_threads(-1)
    pc.set_segment_words(0)
    # a prime-count table budget below 1 MiB, and n < 1 for the 128-bit
    # nth_prime, are refused before any device work
    with pytest.raises(pc.PrimecountError):
        pc.set_pi_table_bytes(1000)
    pc.set_pi_table_bytes(0)
    for bad in (0, -7):
        with pytest.raises(pc.PrimecountError):
            pc.nth_prime128(bad)


# li(10^n) - pi(10^n) for n = 10 .. 22: the published table of the error of
# the logarithmic integral (e.g. OEIS A057752); with Li(x) = li(x) - li(2)
# the difference Li(x) - pi(x) is one less, rounded down.
LI_MINUS_PI = {10: 3104, 11: 11588, 12: 38263, 13: 108971, 14: 314890, 15: 1052619,
               16: 3214632, 17: 7956589, 18: 21949555, 19: 99877775, 20: 222744644,
               21: 597394254, 22: 1932355208}


def schoenfeld_bound(x):
    return math.sqrt(x) * math.log(x) / (8 * math.pi)


def test_result_check_bound(pc):
    """pcg_check_pi (the reference's verify_pix, src/util.cpp:397-420):
    the README values pass, and values just inside and just outside
    |pix - Li(x)| < sqrt(x) log(x) / (8 pi) fall on the two sides, which pins
    both Li(x) (to a few units, against the published li(x) - pi(x)) and the
    bound."""
    pis = {int(r["x"]): int(r["pi"]) for r in golden("pi_known.json")["rows"]}
    for n, d in LI_MINUS_PI.items():
        x = 10 ** n
        pc.check_pi(x, pis[x])
        Li = pis[x] + d - 1                       # li(2) = 1.045...
        B = schoenfeld_bound(x)
        # Li(x) is summed in long double: within a unit up to 1e20, within
        # x / 2^60 (a few hundred) above
        margin = 4 + (x >> 60)
        for side in (1, -1):
            edge = Li + side * int(B)
            pc.check_pi(x, edge - side * margin)
            with pytest.raises(pc.PrimecountError, match="result check failed") as e:
                pc.check_pi(x, edge + side * margin)
            assert e.value.status == pc.E_INTERNAL
    # below 2657 nothing is checked (as the reference)
    pc.check_pi(2656, 10 ** 6)


def test_result_check_catches_a_wrong_term(pc):
    """pi(x) assembled from the reference's printed Gourdon terms passes; the
    same sum with one term off by more than the bound fails, and a term off by
    less than it is (as in the reference) not detected."""
    for r in golden("gourdon_terms.json")["rows"]:
        x = int(r["x"])
        t = {k: int(r[k]) for k in ("AC", "B", "Phi0", "Sigma", "D")}
        pix = t["AC"] - t["B"] + t["Phi0"] + t["Sigma"] + t["D"]
        assert pix == int(r["pi"])
        pc.check_pi(x, pix)
        B = int(schoenfeld_bound(x))
        for term in ("D", "AC"):
            wrong = dict(t)
            wrong[term] += 3 * B
            bad = wrong["AC"] - wrong["B"] + wrong["Phi0"] + wrong["Sigma"] + wrong["D"]
            with pytest.raises(pc.PrimecountError, match="result check failed"):
                pc.check_pi(x, bad)
        pc.check_pi(x, pix + B // 100)


def test_status_codes_are_typed(pc):
    """Argument errors return PCG_E_INVALID, even when the message speaks of
    the device; without a GPU the engine returns PCG_E_DEVICE."""
    with pytest.raises(pc.PrimecountError) as e:
        pc.set_segment_words(1000)
    assert e.value.status == pc.E_INVALID
    with pytest.raises(pc.PrimecountError) as e:
        pc.nth_prime(0)
    assert e.value.status == pc.E_INVALID
    if not pc.device_available():
        with pytest.raises(pc.PrimecountError) as e:
            pc.D(10 ** 12, 76500, 76500, 8)
        assert e.value.status == pc.E_DEVICE


def test_schedule_cache_file(pc, tmp_path):
    """The autotuner's cache file: one winner per (device, class) line, a
    later line replacing an earlier one of the same class, lines with a
    schedule the engine would reject or another format skipped."""
    dev = "gfx950:256cu:test_board"
    p = tmp_path / "schedules.txt"
    p.write_text("\n".join([
        "pcg-schedule-v1 %s 0 0 73 0 512 0 100 0 120.5 118.0" % dev,
        "pcg-schedule-v1 %s 0 0 73 0 640 0 100 6 120.5 117.0" % dev,   # same class
        "pcg-schedule-v1 %s 1 0 73 0 768 2 75 0 1.0 0.9" % dev,
        "pcg-schedule-v1 other:8cu:x 1 0 73 0 768 2 75 0 1.0 0.9",      # another device
        "pcg-schedule-v1 %s 1 0 74 0 1000 2 75 0 1.0 0.9" % dev,       # W not 128 k
        "pcg-schedule-v1 %s 1 0 75 0 512 2 0 0 1.0 0.9" % dev,         # no grid
        "pcg-schedule-v1 %s 1 0 76 0 512 9 100 0 1.0 0.9" % dev,       # group of 9
        "pcg-schedule-v1 %s 1 0 78 0 512 2 100 7 1.0 0.9" % dev,       # 7 waves: no build
        "pcg-schedule-v1 %s 1 0 77 0" % dev,                           # cut short
        "garbage",
        "pcg-schedule-v0 %s 0 0 1 0 512 0 100 0 1 1" % dev,
    ]) + "\n")
    try:
        assert pc.set_schedule_cache(str(p)) == 3
        assert pc.set_schedule_cache(str(tmp_path / "absent.txt")) == 0
        with pytest.raises(pc.PrimecountError) as e:
            pc.set_autotune(3)
        assert e.value.status == pc.E_INVALID
    finally:
        pc.set_schedule_cache(None)
        pc.set_autotune(pc.TUNE_CACHED)
