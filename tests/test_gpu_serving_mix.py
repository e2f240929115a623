"""The serving mix on the device: a seeded slice of the 4,096-query mix of
tests/golden/serve_c5.json (80% pi(x) with x = 10^(9+10u), 20% nth_prime(n)
with n = 10^(8+9u), answers printed by the reference build:
`primecount X` -> src/api.cpp:75-91, `primecount N --nth-prime` ->
src/nth_prime.cpp:99-139), submitted through one serving context (pcg_ctx_*)
with Poisson arrivals from the committed seed.  Every answer is checked, and a
few of the same queries are then answered alone, one per submit, and must
give the same values.
"""
import time

import pytest

from primecount_amd import serving_bench
from tests.conftest import golden

pytestmark = pytest.mark.gpu

N_QUERIES = 512                # includes 44 pi(x) with x >= 1e18 and 12 nth_prime(n >= 1e16)
RATE = 60.0                    # mean arrivals per second
SEED = 0xC5C5                  # arrival seed of the serving benchmark


def log(msg):
    print(msg, flush=True)


def mix_rows(n):
    kinds = {"pi": 2, "nth": 3}
    return [(kinds[r["kind"]], int(r["arg"]), int(r["value"]))
            for r in golden("serve_c5.json")["rows"][:n]]


def test_serving_mix_poisson_slice(pc):
    qs = mix_rows(N_QUERIES)
    assert pc.KIND_PI == 2 and pc.KIND_NTH_PRIME == 3
    heavy_pi = [q for q in qs if q[0] == pc.KIND_PI and q[1] >= 10 ** 18]
    heavy_nth = [q for q in qs if q[0] == pc.KIND_NTH_PRIME and q[1] >= 10 ** 16]
    assert len(heavy_pi) >= 40 and len(heavy_nth) >= 10
    times = serving_bench.arrivals(len(qs), RATE, SEED)
    log("serving mix: %d queries (%d pi >= 1e18, %d nth_prime n >= 1e16), Poisson %.0f/s, "
        "seed %#x" % (len(qs), len(heavy_pi), len(heavy_nth), RATE, SEED))
    with pc.Server() as sv:
        t0 = time.perf_counter()
        res, elapsed, bad, s0, s1 = serving_bench.run(sv, qs, times)
        log("served %d queries in %.2f s (%.1f/s), %d launch sets, %d mismatches"
            % (len(res), elapsed, len(res) / elapsed, s1["batches"] - s0["batches"], bad))
        assert len(res) == len(qs)
        assert all(r["status"] == 0 for r in res)
        assert bad == 0
        # queries shared launch sets
        assert s1["batches"] - s0["batches"] < len(qs)
        # The same values alone: heavy pi, heavy nth_prime and light queries,
        # each submitted by itself and waited for before the next.
        alone = [max(heavy_pi, key=lambda q: q[1]), min(heavy_pi, key=lambda q: q[1]),
                 max(heavy_nth, key=lambda q: q[1])]
        alone += [q for q in qs if q[1] < 10 ** 12][:3]
        for kind, arg, want in alone:
            log("alone: kind %d arg %d" % (kind, arg))
            i = sv.submit([(kind, arg)])[0]
            r = sv.poll_all(1, timeout_ms=60000)
            assert len(r) == 1 and r[0]["id"] == i
            assert r[0]["status"] == 0 and r[0]["value"] == want, (kind, arg)
        log("mix test done in %.1f s" % (time.perf_counter() - t0))
