"""The CPU oracle, pinned before it is trusted (no GPU needed).

- the reference's own known-answer rows (test/gourdon/D.cpp:36-89,
  test/deleglise-rivat/S2_hard_alpha.cpp:36-73), both through the oracle and
  through an independent closed-form brute force (tests/closed_form.py);
- fixtures produced by the reference build (oracle/_ref) at x = 1e13;
- invariants: segment-size independence, phi semantics.
"""
import random

import numpy as np
import pytest

from tests import closed_form
from tests.conftest import golden


def kat_rows(name):
    return [[int(v) for v in r] for r in golden(name)["rows"]]


@pytest.mark.parametrize("row", kat_rows("kat_D.json"), ids=lambda r: "x%d_y%d_z%d" % tuple(r[:3]))
def test_oracle_D_kat(oracle, row):
    x, y, z, k, expected = row
    assert oracle.D(x, y, z, k) == expected


@pytest.mark.parametrize("row", kat_rows("kat_S2_hard.json"), ids=lambda r: "x%d_y%d_z%d" % tuple(r[:3]))
def test_oracle_S2_hard_kat(oracle, row):
    x, y, z, c, expected = row
    assert oracle.S2_hard(x, y, z, c) == expected


def test_closed_form_matches_kat_rows():
    for x, y, z, k, expected in kat_rows("kat_D.json"):
        if x <= 10 ** 8:
            assert closed_form.D(x, y, z, k) == expected, (x, y, z, k)
    for x, y, z, c, expected in kat_rows("kat_S2_hard.json"):
        if x <= 10 ** 8:
            assert closed_form.S2_hard(x, y, z, c) == expected, (x, y, z, c)


def test_oracle_vs_closed_form_random(oracle):
    rng = random.Random(20261203)
    for _ in range(40):
        x = rng.randint(10 ** 5, 3 * 10 ** 7)
        x13 = round(x ** (1 / 3))
        y = rng.randint(x13 + 1, 4 * x13)
        z = rng.randint(y, min(4 * y, int(x ** 0.5) - 1))
        assert oracle.D(x, y, z, 8, seg=1 << 12) == closed_form.D(x, y, z, 8), (x, y, z)
        yd = rng.randint(x13, 3 * x13)
        zd = x // yd
        assert oracle.S2_hard(x, yd, zd, 8, seg=1 << 12) == closed_form.S2_hard(x, yd, zd, 8), (x, yd)


def test_oracle_segment_size_invariance(oracle):
    x, y, z, k = 10 ** 11, 13825, 13825, 8
    ref = oracle.D(x, y, z, k, seg=1 << 16)
    for seg in (1 << 10, 3 * (1 << 11), 1 << 18):
        assert oracle.D(x, y, z, k, seg=seg) == ref


def test_oracle_reference_fixture_1e13(oracle):
    d = {r["x"]: r for r in golden("configs.json")["D"]}
    r = d[str(10 ** 13)]
    assert oracle.D(10 ** 13, int(r["y"]), int(r["z"]), int(r["k"])) == int(r["value"])
    s = {r["x"]: r for r in golden("configs.json")["S2_hard"]}
    r = s[str(10 ** 13)]
    assert oracle.S2_hard(10 ** 13, int(r["y"]), int(r["z"]), int(r["c"])) == int(r["value"])


def test_oracle_int128_entry_matches_int64(oracle):
    # the 128-bit overloads must agree with the 64-bit ones (test/gourdon/D.cpp:104-111)
    for x, y, z, k, expected in kat_rows("kat_D.json")[20:40]:
        assert oracle.D(x, y, z, k) == expected


def test_phi_direct_small(oracle):
    # phi(n, a) counts n' <= n free of the first a primes: closed forms
    assert oracle.phi_direct(100, 0) == 100
    assert oracle.phi_direct(100, 1) == 50
    assert oracle.phi_direct(30, 3) == 8          # 1, 7, 11, 13, 17, 19, 23, 29
    assert oracle.phi_direct(210, 4) == 48        # totient(210)
    phi = closed_form.make_phi(tuple(closed_form.primes_upto(200)))
    for n in (1, 2, 29, 30, 31, 97, 1000):
        for a in range(0, 9):
            assert oracle.phi_direct(n, a) == phi(n, a)


# ---- the other Gourdon terms (oracle/gourdon_oracle.c) ----------------------

def test_oracle_gourdon_terms_match_reference(oracle):
    """Plain sums over the definitions of A + C, B, Phi0 and Sigma reproduce
    the values the reference build printed (default and alpha_z = 2
    parameters), and with D they assemble pi(x) = the reference's pi(x)."""
    fx = golden("gourdon_terms.json")
    rows = [r for r in fx["rows"] if int(r["x"]) <= 10 ** 12] + \
           [r for r in fx["alpha_z2"] if int(r["x"]) <= 10 ** 12]
    assert len(rows) >= 8
    for r in rows:
        x, y, z, k, xs = (int(r[t]) for t in ("x", "y", "z", "k", "x_star"))
        got = {t: oracle.gourdon_term(t, x, y, z, k, xs) for t in ("AC", "B", "Phi0", "Sigma")}
        for t in got:
            assert got[t] == int(r[t]), (x, t)
        if "pi_cli" in r:
            d = oracle.D(x, y, z, k)
            assert got["AC"] - got["B"] + d + got["Phi0"] + got["Sigma"] == int(r["pi_cli"])


def test_oracle_pi_known_values(oracle):
    # pi(10^n) of the reference's README table, by sieving (n <= 8 here) and
    # by the assembled terms for 10^10 .. 10^11
    assert [oracle.pi_sieve(10 ** n) for n in range(1, 9)] == \
        [4, 25, 168, 1229, 9592, 78498, 664579, 5761455]
    known = {int(r["x"]): int(r["pi"]) for r in golden("pi_known.json")["rows"]}
    fx = {int(r["x"]): r for r in golden("gourdon_terms.json")["rows"]}
    for x in (10 ** 10, 10 ** 11):
        r = fx[x]
        y, z, k, xs = (int(r[t]) for t in ("y", "z", "k", "x_star"))
        total = sum(s * oracle.gourdon_term(t, x, y, z, k, xs)
                    for s, t in ((1, "AC"), (-1, "B"), (1, "Phi0"), (1, "Sigma")))
        assert total + oracle.D(x, y, z, k) == known[x]


def test_oracle_dr_terms_match_reference(oracle):
    """Plain sums over the definitions of S1, S2_trivial, S2_easy and P2
    reproduce the values the reference build printed, and the printed terms
    assemble pi(x) = S1 + S2_trivial + S2_easy + S2_hard + pi(y) - 1 - P2,
    equal to the README's pi(10^n)."""
    rows = golden("dr_terms.json")["rows"]
    small = [r for r in rows if int(r["x"]) <= 10 ** 12]
    assert len(small) >= 6
    for r in small:
        x, y, z, c = (int(r[t]) for t in ("x", "y", "z", "c"))
        for t in ("S1", "S2_trivial", "S2_easy", "P2"):
            assert oracle.dr_term(t, x, y, z, c) == int(r[t]), (x, t)
    known = {int(r["x"]): int(r["pi"]) for r in golden("pi_known.json")["rows"]}
    for r in rows:
        x, y = int(r["x"]), int(r["y"])
        if y > 3 * 10 ** 7:
            continue
        total = sum(int(r[t]) for t in ("S1", "S2_trivial", "S2_easy", "S2_hard")) \
            + oracle.pi_sieve(y) - 1 - int(r["P2"])
        assert total == int(r["pi"]), x
        if x in known:
            assert total == known[x], x


def _phi_numpy(x, primes):
    """phi(x, a) by marking the multiples of the first a primes."""
    keep = np.ones(x + 1, dtype=bool)
    keep[0] = False
    for p in primes:
        keep[p::p] = False
    return int(keep.sum())


def test_oracle_dr_identities_brute_force(oracle):
    """At x = 2e7, for c = 5 .. 8 and two y: S1 + S2_trivial + S2_easy +
    S2_hard = phi(x, pi(y)) by brute force (each c moves the boundaries of
    S1, S2_trivial and S2_easy differently), and P2 = the number of
    n = p q <= x with y < p <= q, counted directly."""
    x = 2 * 10 ** 7
    sieve = np.ones(x + 1, dtype=bool)
    sieve[:2] = False
    for i in range(2, int(x ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    primes = np.nonzero(sieve)[0]
    pi = np.cumsum(sieve)
    for y in (272, 1000):
        z = x // y
        a = int(pi[y])
        phi = _phi_numpy(x, primes[:a])
        for c in (5, 6, 7, 8):
            t = {n: oracle.dr_term(n, x, y, z, c) for n in ("S1", "S2_trivial", "S2_easy")}
            assert sum(t.values()) + oracle.S2_hard(x, y, z, c) == phi, (y, c)
        big = primes[primes > y]
        p2 = sum(int(pi[x // p]) - int(pi[p]) + 1 for p in big[big * big <= x])
        semi = 0
        for p in big[big * big <= x]:
            semi += int(np.count_nonzero((big >= p) & (big <= x // p)))
        assert p2 == semi
        assert oracle.dr_term("P2", x, y, z, 8) == semi, y
        assert phi + a - 1 - semi == int(pi[x])


def _is_prime64(n):
    """Deterministic Miller-Rabin for n < 3.3e24 (bases 2 .. 41)."""
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)
    for p in small:
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_serving_mix_fixture():
    """The serving-mix answers (printed by the reference build) belong to the
    seeded query list, and they agree with what the mathematics fixes: every
    nth_prime answer is prime, pi(x) never decreases with x and lies between
    x / ln x and 1.26 x / ln x (Rosser-Schoenfeld), the n-th prime lies
    between n ln n and n (ln n + ln ln n) (Rosser, Dusart, n >= 6), and
    pi(nth_prime(n)) = n wherever the list holds both."""
    import importlib.util
    import math
    import os
    if not os.path.exists(os.path.join(os.path.dirname(__file__), "golden", "serve_c5.json")):
        pytest.skip("serving-mix fixture not generated (tests/golden/gen_fixtures.py --only serve)")
    fx = golden("serve_c5.json")
    spec = importlib.util.spec_from_file_location(
        "gen_fixtures", os.path.join(os.path.dirname(__file__), "golden", "gen_fixtures.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    qs = gen.serve_queries(fx["seed"], len(fx["rows"]))
    assert len(qs) == 4096
    assert [(r["kind"], int(r["arg"])) for r in fx["rows"]] == qs
    pis = sorted((int(r["arg"]), int(r["value"])) for r in fx["rows"] if r["kind"] == "pi")
    for (x0, v0), (x1, v1) in zip(pis, pis[1:]):
        assert v0 <= v1 and v1 - v0 <= x1 - x0
    for x, v in pis:
        lx = math.log(x)
        assert x / lx < v < 1.26 * x / lx, x
    pi_of = dict(pis)
    for r in fx["rows"]:
        if r["kind"] != "nth":
            continue
        n, p = int(r["arg"]), int(r["value"])
        assert _is_prime64(p), n
        ln = math.log(n)
        assert n * ln < p < n * (ln + math.log(ln)), n
        if p in pi_of:
            assert pi_of[p] == n
