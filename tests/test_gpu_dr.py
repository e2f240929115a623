"""pi(x) by the Deleglise-Rivat algorithm on the device, term by term.

pi(x) = S1 + S2_trivial + S2_easy + S2_hard + pi(y) - 1 - P2
(src/deleglise-rivat/pi_deleglise_rivat.cpp:40-53, 86-92).  The reference's
own values pin every term (tests/golden/dr_terms.json: what the reference
build printed for --S1, --S2-trivial, --S2-easy, --S2-hard and --P2 at
10^10 .. 10^19 and 8 seeded x), pi(10^n) from the reference's README table
(tests/golden/pi_known.json), and the CPU oracle (plain sums over the term
definitions, oracle/dr_oracle.c) on seeded random parameters.
"""
import random

import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def _rows():
    return sorted(golden("dr_terms.json")["rows"], key=lambda r: int(r["x"]))


def test_dr_terms_match_reference(pc):
    for r in _rows():
        x, y, z, c = (int(r[t]) for t in ("x", "y", "z", "c"))
        log("Deleglise-Rivat terms at x=%d (y=%d, z=%d, c=%d)" % (x, y, z, c))
        assert pc.dr_params(x) == {"y": y, "z": z, "c": c}, x
        assert pc.S2_trivial(x, y, z, c) == int(r["S2_trivial"]), x
        assert pc.P2(x, y) == int(r["P2"]), x
        assert pc.S1(x, y, c) == int(r["S1"]), x
        assert pc.S2_easy(x, y, z, c) == int(r["S2_easy"]), x
        assert pc.S2_hard(x, y, z, c) == int(r["S2_hard"]), x


def test_pi_deleglise_rivat_matches_reference(pc):
    for r in _rows():
        log("pi_deleglise_rivat(%s)" % r["x"])
        assert pc.pi_deleglise_rivat(int(r["x"])) == int(r["pi"])
    for r in golden("pi_known.json")["rows"]:
        if int(r["x"]) <= 10 ** 19:
            log("pi_deleglise_rivat(%s), README value" % r["x"])
            assert pc.pi_deleglise_rivat(int(r["x"])) == int(r["pi"])


def test_dr_terms_random_vs_oracle(pc, oracle):
    """Seeded x in [1e9, 1e12] with y between x^(1/3) and 4 x^(1/3), z = x / y
    and c in 5 .. 8 (the terms' own parameter ranges), against the oracle."""
    rng = random.Random(0xD1)
    for _ in range(12):
        x = int(10 ** rng.uniform(9, 12))
        x13 = round(x ** (1 / 3))
        y = rng.randrange(x13 + 1, 4 * x13)
        z = x // y
        c = rng.randrange(5, 9)
        log("random terms x=%d y=%d z=%d c=%d" % (x, y, z, c))
        for name, fn in (("S1", lambda: pc.S1(x, y, c)),
                         ("S2_trivial", lambda: pc.S2_trivial(x, y, z, c)),
                         ("S2_easy", lambda: pc.S2_easy(x, y, z, c)),
                         ("P2", lambda: pc.P2(x, y))):
            assert fn() == oracle.dr_term(name, x, y, z, c), (name, x, y, z, c)


def test_pi_deleglise_rivat_small_and_vs_gourdon(pc, oracle):
    # x <= 1e8: the prime-count table; just above it the full formula
    for x in (0, 1, 2, 3, 100, 7919, 10 ** 8):
        assert pc.pi_deleglise_rivat(x) == oracle.pi_sieve(x), x
    for x in (10 ** 8 + 1, 123456789, 2 ** 28 + 3, 999999937):
        log("pi_deleglise_rivat(%d) vs sieve" % x)
        assert pc.pi_deleglise_rivat(x) == oracle.pi_sieve(x), x
    rng = random.Random(0xD2)
    for _ in range(10):
        x = int(10 ** rng.uniform(9, 17))
        log("pi_deleglise_rivat(%d) vs Gourdon" % x)
        assert pc.pi_deleglise_rivat(x) == pc.pi(x), x


def test_dr_term_errors(pc):
    with pytest.raises(pc.PrimecountError, match="5 <= c <= 8"):
        pc.S1(10 ** 12, 20000, 4)
    with pytest.raises(pc.PrimecountError, match="c must be"):
        pc.S2_easy(10 ** 12, 20000, 10 ** 12 // 20000, 9)
    with pytest.raises(pc.PrimecountError, match="too large"):
        pc.S2_trivial(2 ** 80, 1000, 2 ** 60, 8)


def test_pi_128bit_x_both_formulas(pc):
    """x above 2^64: the 128-bit paths of every term of both formulas.
    pi(1e20) is the README value; at seeded x in (2^64, 4e19) the two
    algorithms (different leaf sets, different tables) must agree."""
    want = {int(r["x"]): int(r["pi"]) for r in golden("pi_known.json")["rows"]}[10 ** 20]
    log("pi(1e20), both formulas")
    assert pc.pi_deleglise_rivat(10 ** 20) == want
    assert pc.pi(10 ** 20) == want
    rng = random.Random(0x128)
    for _ in range(2):
        x = rng.randrange(2 ** 64 + 1, 4 * 10 ** 19)
        log("pi(%d): Gourdon vs Deleglise-Rivat" % x)
        assert pc.pi_deleglise_rivat(x) == pc.pi(x), x


def test_pi_deleglise_rivat_any_alpha(pc, oracle):
    """The split of the leaves between S1, S2_trivial, S2_easy and S2_hard
    moves with alpha (y = alpha x^(1/3)); pi(x) does not.  Alpha 1 puts
    y at x^(1/3) (no trivial and no easy leaves), 30 near x^(1/6)."""
    known = {int(r["x"]): int(r["pi"]) for r in golden("pi_known.json")["rows"]}
    try:
        for alpha in (1.0, 2.5, 7.0, 30.0):
            pc.set_alpha(alpha)
            for x in (10 ** 12, 10 ** 15):
                log("pi_deleglise_rivat(%d), alpha %g, params %s" % (x, alpha, pc.dr_params(x)))
                assert pc.pi_deleglise_rivat(x) == known[x], (x, alpha)
            if alpha < 2:
                continue                 # the oracle's tables would take seconds here
            x = 987654321987
            p = pc.dr_params(x)
            for name, fn in (("S1", lambda: pc.S1(x, p["y"], p["c"])),
                             ("S2_trivial", lambda: pc.S2_trivial(x, p["y"], p["z"], p["c"])),
                             ("S2_easy", lambda: pc.S2_easy(x, p["y"], p["z"], p["c"])),
                             ("P2", lambda: pc.P2(x, p["y"]))):
                assert fn() == oracle.dr_term(name, x, p["y"], p["z"], p["c"]), (name, alpha)
    finally:
        pc.set_alpha(0)
