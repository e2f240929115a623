"""A C++ program linked against the engine library calls it through
include/primecount_mi355x.hpp -- the overloads that stand in for the
reference's D() / S2_hard() (include/gourdon.hpp:28,38, include/S.hpp:25,32)
and its pi() / nth_prime() (include/primecount.hpp:57,74,84) -- on the
device, and checks the values against the reference's test rows, its README
and the reference build's printed nth_prime values."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "tests", "native", "_bin", "cpp_api_check")


@pytest.mark.gpu
def test_cpp_program_calls_the_engine():
    # __graft_entry__.build() links the program; a missing binary is an error
    assert os.path.exists(EXE), "tests/native/_bin/cpp_api_check is not built (run build())"
    print("running %s" % EXE, flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    sys.stdout.flush()
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cpp api ok" in r.stdout
    assert r.stdout.count("\nok ") + r.stdout.startswith("ok ") >= 9
