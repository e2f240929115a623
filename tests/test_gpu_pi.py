"""pi(x) by Gourdon's algorithm on the device, term by term.

The reference's own values pin every term: A + C, B, Phi0, Sigma and D as the
reference build printed them (tests/golden/gourdon_terms.json, default and
alpha_z = 2 parameters), pi(10^n) from the reference's README table
(tests/golden/pi_known.json), and the CPU oracle (plain sums over the term
definitions, oracle/gourdon_oracle.c) on seeded random parameters.
"""
import random

import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_pi_small_x_against_sieve(pc, oracle):
    log("pi(x) for x <= 1e8 (prime-count table)")
    for x in list(range(0, 200)) + [1000, 7919, 7920, 10 ** 6, 99999989, 10 ** 8]:
        assert pc.pi(x) == oracle.pi_sieve(x), x
    rng = random.Random(71)
    for _ in range(20):
        x = rng.randrange(2, 10 ** 8)
        assert pc.pi(x) == oracle.pi_sieve(x), x


def test_terms_match_reference(pc):
    fx = golden("gourdon_terms.json")
    for r in sorted(fx["rows"], key=lambda r: int(r["x"])):
        x, y, z, k = (int(r[t]) for t in ("x", "y", "z", "k"))
        if x > 10 ** 18:
            continue
        log("terms at x=%d" % x)
        assert pc.Sigma(x, y) == int(r["Sigma"]), x
        assert pc.Phi0(x, y, z, k) == int(r["Phi0"]), x
        assert pc.B(x, y) == int(r["B"]), x
        assert pc.AC(x, y, z, k) == int(r["AC"]), x
    for r in fx["alpha_z2"]:
        x, y, z, k = (int(r[t]) for t in ("x", "y", "z", "k"))
        log("terms at x=%d, alpha_z = 2" % x)
        assert (pc.AC(x, y, z, k), pc.B(x, y), pc.Phi0(x, y, z, k), pc.Sigma(x, y)) == \
            (int(r["AC"]), int(r["B"]), int(r["Phi0"]), int(r["Sigma"]))


def test_pi_matches_reference(pc):
    fx = golden("gourdon_terms.json")["rows"]
    for r in sorted(fx, key=lambda r: int(r["x"])):
        log("pi(%s)" % r["x"])
        assert pc.pi(int(r["x"])) == int(r["pi"])
    known = golden("pi_known.json")["rows"]
    for r in known:
        if int(r["x"]) <= 10 ** 19:
            log("pi(%s), README value" % r["x"])
            assert pc.pi(int(r["x"])) == int(r["pi"])


def test_terms_match_oracle_on_random_parameters(pc, oracle):
    # seeded x in [1e9, 1e12] with alpha_y, alpha_z drawn around the
    # defaults (y < z included); the oracle sums the definitions directly
    rng = random.Random(20261203)
    for _ in range(12):
        x = int(10 ** rng.uniform(9, 12))
        p = pc.gourdon_params(x)
        y = max(int(p["y"] * rng.uniform(0.6, 1.0)), int(round(x ** (1 / 3))) + 1)
        z = min(int(y * rng.uniform(1.0, 2.5)), int(x ** 0.5) - 1)
        k = 8
        xs = p["x_star"]
        from primecount_amd import PrimecountError
        try:
            got = (pc.AC(x, y, z, k), pc.B(x, y), pc.Phi0(x, y, z, k), pc.Sigma(x, y))
        except PrimecountError:
            continue
        # x_star depends on y only (src/util.cpp:362-386); recompute it like
        # the engine does through the D parameters of this y
        xs = oracle_x_star(x, y)
        want = tuple(oracle.gourdon_term(t, x, y, z, k, xs) for t in ("AC", "B", "Phi0", "Sigma"))
        log("random x=%d y=%d z=%d" % (x, y, z))
        assert got == want, (x, y, z)


def oracle_x_star(x, y):
    # get_x_star_gourdon (src/util.cpp:362-386): max(x^(1/4), x / y^2),
    # capped at y and sqrt(x / y), evaluated in integers
    from math import isqrt

    def iroot4(n):
        r = int(round(n ** 0.25))
        while r ** 4 > n:
            r -= 1
        while (r + 1) ** 4 <= n:
            r += 1
        return r
    xy2 = -(-x // (y * y))
    xs = max(iroot4(x), xy2)
    return max(1, min(xs, y, isqrt(x // y)))


def test_term_errors(pc):
    with pytest.raises(pc.PrimecountError):
        pc.Phi0(10 ** 12, 76500, 76500, 4)       # the device path needs k >= 5
    with pytest.raises(pc.PrimecountError):
        pc.AC(10 ** 12, 76500, 1000, 8)          # z < y
    assert pc.pi(10 ** 12) == 37607912018        # still fine afterwards


def test_nth_prime_matches_reference(pc, oracle):
    # nth_prime(n) printed by the reference (tests/golden/nth_prime.json):
    # the tiny range, the small-table range, powers of ten to 1e16 and seeded
    # n = 10^(8 + 9u); below pi(1e8) also against a plain sieve
    rows = golden("nth_prime.json")["rows"]
    for r in rows:
        n, p = int(r["n"]), int(r["p"])
        log("nth_prime(%d)" % n)
        assert pc.nth_prime(n) == p, n
    # consistency: pi(nth_prime(n)) == n and nth_prime(pi(p)) == p for primes p
    for n in (5761456, 10 ** 9 + 7, 123456789012):
        p = pc.nth_prime(n)
        assert pc.pi(p) == n and pc.pi(p - 1) == n - 1, n
    with pytest.raises(pc.PrimecountError):
        pc.nth_prime(0)


def test_pi_with_alpha_z(pc):
    """pi(x) does not depend on the tuning factor alpha_z: with z = 2y (a
    FactorTableD that depends on y, so the query gets a table set of its
    own) pi(10^n) still equals the README values."""
    known = {int(r["x"]): int(r["pi"]) for r in golden("pi_known.json")["rows"]}
    try:
        pc.set_alpha_z(2.0)
        for x in (10 ** 10, 10 ** 13, 10 ** 15, 10 ** 17):
            p = pc.gourdon_params(x)
            log("pi(%d) with y=%d z=%d" % (x, p["y"], p["z"]))
            assert p["z"] > p["y"]
            assert pc.pi(x) == known[x], x
    finally:
        pc.set_alpha_z(0)
    assert pc.pi(10 ** 13) == known[10 ** 13]


def test_pi_table_windows(pc):
    # a prime-count table over the budget is built in windows (each with
    # absolute counts, the leaf sums of A, C and B once per window): pi(x) by
    # both algorithms, nth_prime and a batch of pi queries give the README
    # and reference values
    known = {int(r["x"]): int(r["pi"]) for r in golden("pi_known.json")["rows"]}
    nth = [(int(r["n"]), int(r["p"])) for r in golden("nth_prime.json")["rows"]
           if 10 ** 10 <= int(r["n"]) <= 10 ** 13][:4]
    try:
        for e, budget in ((13, 1 << 20), (15, 8 << 20), (17, 128 << 20), (19, 2 << 30)):
            x = 10 ** e
            pc.set_pi_table_bytes(budget)
            log("pi(1e%d) with a %d-byte prime-count table budget" % (e, budget))
            assert pc.pi(x) == known[x]
            st = pc.last_pi_stats()
            assert st["table_windows"] >= 2 and st["table_limit"] > budget // 16 * 240
            assert pc.pi_deleglise_rivat(x) == known[x]
        pc.set_pi_table_bytes(1 << 20)
        log("nth_prime with a 1 MiB prime-count table budget")
        for n, p in nth:
            assert pc.nth_prime(n) == p, n
        log("a batch of pi queries with a 1 MiB prime-count table budget")
        xs = [10 ** 12, 10 ** 13, 10 ** 14, 10 ** 15]
        with pc.Server() as sv:
            ids = sv.submit([(pc.KIND_PI, x) for x in xs])
            got = {r["id"]: r for r in sv.poll_all(len(ids), timeout_ms=120000)}
        for i, x in zip(ids, xs):
            assert got[i]["status"] == 0 and got[i]["value"] == known[x], x
        pc.set_pi_table_bytes(0)
        assert pc.pi(10 ** 15) == known[10 ** 15]
        assert pc.last_pi_stats()["table_windows"] == 1
    finally:
        pc.set_pi_table_bytes(0)


def test_nth_prime_128(pc):
    # answers from 2^62 on: the sieve interval anchored at a 128-bit base
    # with 64-bit sieving primes.  Each answer p is pinned by pi(p) = n and
    # pi(p - 1) = n - 1 (pi itself is pinned by the README values); at the
    # 64-bit boundary n = pi(2^63 - 1) the answers are the largest prime
    # below 2^63 and the smallest above it.
    max_n = 216289611853439384
    cases = [2 * 10 ** 17, max_n, max_n + 1, 10 ** 18 + 7, 2 * 10 ** 18 + 123457]
    for n in cases:
        log("nth_prime(%d) through the 128-bit entry point" % n)
        p = pc.nth_prime128(n)
        assert pc.pi(p) == n and pc.pi(p - 1) == n - 1, (n, p)
        if n <= max_n:
            assert p < 2 ** 63 and pc.nth_prime(n) == p
        else:
            assert p > 2 ** 63 and pc.nth_prime(n) == p
    assert pc.nth_prime128(max_n) == 2 ** 63 - 25
    assert pc.nth_prime128(max_n + 1) == 2 ** 63 + 29
    # the 128-bit entry point gives the 64-bit answers below
    rows = [(int(r["n"]), int(r["p"])) for r in golden("nth_prime.json")["rows"]]
    for n, p in [r for r in rows if 10 ** 6 <= r[0] <= 10 ** 15][:4] + [(1, 2), (169, 1009)]:
        assert pc.nth_prime128(n) == p, n
    for bad in (0, -5):
        with pytest.raises(pc.PrimecountError):
            pc.nth_prime128(bad)
