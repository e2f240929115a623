"""Serving context (pcg_ctx_*, continuous batching) on the device.

Every answer is checked against the reference's own value for that x (the
256-query fixture printed by the reference build, tests/golden/batch_D.json,
and the default-parameter configs), whether the query ran alone or inside a
mixed launch, submitted at once, in waves or from several threads.
"""
import threading
import time

import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def fixture_rows(limit):
    return [(int(r["x"]), int(r["D"])) for r in golden("batch_D.json")["rows"]
            if int(r["x"]) <= limit]


def test_serving_batch_matches_reference(pc):
    rows = fixture_rows(10 ** 16)[:64]
    log("serving: %d queries in one submit" % len(rows))
    with pc.Server(z_max=10 ** 7) as sv:
        ids = sv.submit([(pc.KIND_D, x) for x, _ in rows])
        res = {r["id"]: r for r in sv.poll_all(len(ids), timeout_ms=120000)}
        assert len(res) == len(ids)
        for i, (x, want) in zip(ids, rows):
            assert res[i]["status"] == 0, x
            assert res[i]["value"] == want, x
            assert res[i]["seconds"] > 0
        st = sv.stats()
        assert st["completed"] == len(ids)
        assert st["batches"] < len(ids)          # queries shared launches


def test_serving_waves_alone_and_mixed(pc):
    rows = fixture_rows(10 ** 15)[:40]
    with pc.Server() as sv:
        # one query alone
        x0, want0 = rows[0]
        log("serving: one query alone")
        i0 = sv.submit([(pc.KIND_D, x0)])[0]
        r0 = sv.poll_all(1, timeout_ms=60000)
        assert r0[0]["id"] == i0 and r0[0]["value"] == want0
        # the same query again inside waves of others, arriving while
        # launches run
        log("serving: 4 waves")
        ids = []
        for w in range(4):
            part = rows[w * 10:(w + 1) * 10]
            ids += sv.submit([(pc.KIND_D, x) for x, _ in part])
            time.sleep(0.005)
        got = {r["id"]: r for r in sv.poll_all(len(ids), timeout_ms=120000)}
        for i, (x, want) in zip(ids, rows):
            assert got[i]["status"] == 0 and got[i]["value"] == want, x


def test_serving_threads_and_errors(pc):
    rows = fixture_rows(10 ** 15)[:48]
    results = {}
    lock = threading.Lock()
    with pc.Server() as sv:
        def client(part):
            ids = sv.submit([(pc.KIND_D, x) for x, _ in part])
            with lock:
                for i, (x, want) in zip(ids, part):
                    results[i] = want
        log("serving: 4 client threads")
        ts = [threading.Thread(target=client, args=(rows[t::4],)) for t in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        # a query outside the domain (z < y) is answered with an error and
        # does not disturb the others
        bad = sv.submit([(pc.KIND_D, 10 ** 12, 100, 50, 8)])[0]
        got = {r["id"]: r for r in sv.poll_all(len(rows) + 1, timeout_ms=120000)}
        assert got[bad]["status"] < 0
        for i, want in results.items():
            assert got[i]["status"] == 0 and got[i]["value"] == want


def test_serving_S2_hard_and_explicit_parameters(pc):
    cfg = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}
    kat = [[int(v) for v in r] for r in golden("kat_D.json")["rows"]][:10]
    with pc.Server() as sv:
        log("serving: S2_hard defaults and explicit D parameters")
        ids = sv.submit([(pc.KIND_S2_HARD, 10 ** 13), (pc.KIND_S2_HARD, 10 ** 15)])
        ids += sv.submit([(pc.KIND_D, x, y, z, k) for x, y, z, k, _ in kat])
        got = {r["id"]: r for r in sv.poll_all(len(ids), timeout_ms=120000)}
        assert got[ids[0]]["value"] == int(cfg[10 ** 13]["value"])
        assert got[ids[1]]["value"] == int(cfg[10 ** 15]["value"])
        for i, row in zip(ids[2:], kat):
            assert got[i]["status"] == 0 and got[i]["value"] == row[4], row


def test_serving_pi_and_nth_prime(pc, oracle):
    """pi(x) and nth_prime queries in a serving context, mixed with D queries
    in the same submit: the README pi(10^n) values, the reference's
    nth_prime values, and the prime-count sieve for small x."""
    known = [(int(r["x"]), int(r["pi"])) for r in golden("pi_known.json")["rows"]
             if int(r["x"]) <= 10 ** 17]
    nth = [(int(r["n"]), int(r["p"])) for r in golden("nth_prime.json")["rows"]
           if int(r["n"]) <= 10 ** 15]
    small = [0, 1, 2, 100, 7919, 10 ** 6, 10 ** 8 + 7]
    drow = fixture_rows(10 ** 14)[:4]
    with pc.Server() as sv:
        log("serving: %d pi, %d nth_prime, %d D queries in one submit"
            % (len(known) + len(small), len(nth), len(drow)))
        qs = [(pc.KIND_PI, x) for x, _ in known] + [(pc.KIND_PI, x) for x in small]
        qs += [(pc.KIND_NTH_PRIME, n) for n, _ in nth] + [(pc.KIND_D, x) for x, _ in drow]
        want = [v for _, v in known] + [oracle.pi_sieve(x) for x in small]
        want += [p for _, p in nth] + [v for _, v in drow]
        ids = sv.submit(qs)
        # parameters on a pi query, and n = 0, are answered with an error
        bad = sv.submit([(pc.KIND_PI, 10 ** 12, 1, 2, 3), (pc.KIND_NTH_PRIME, 0)])
        got = {r["id"]: r for r in sv.poll_all(len(ids) + 2, timeout_ms=180000)}
        assert len(got) == len(ids) + 2
        for i, q, w in zip(ids, qs, want):
            assert got[i]["status"] == 0 and got[i]["value"] == w, q
        assert all(got[i]["status"] < 0 for i in bad)
        # the same queries again, one per submit, arriving while launches run
        log("serving: the same queries one by one")
        ids2 = [sv.submit([q])[0] for q in qs]
        got = {r["id"]: r for r in sv.poll_all(len(ids2), timeout_ms=180000)}
        for i, q, w in zip(ids2, qs, want):
            assert got[i]["status"] == 0 and got[i]["value"] == w, q


def test_serving_small_pass_slots(pc):
    """Small passes (pcg_set_pass_slots) cut the D launches of pi(x) and D
    queries into many passes by segment range, with pass points between them;
    the answers do not change."""
    known = [(int(r["x"]), int(r["pi"])) for r in golden("pi_known.json")["rows"]
             if int(r["x"]) <= 10 ** 15]
    drow = fixture_rows(10 ** 15)[:8]
    try:
        pc.set_pass_slots(1 << 16)
        with pc.Server() as sv:
            log("serving with 2^16-slot passes: %d pi, %d D" % (len(known), len(drow)))
            qs = [(pc.KIND_PI, x) for x, _ in known] + [(pc.KIND_D, x) for x, _ in drow]
            want = [v for _, v in known] + [v for _, v in drow]
            ids = sv.submit(qs)
            got = {r["id"]: r for r in sv.poll_all(len(ids), timeout_ms=180000)}
            for i, q, w in zip(ids, qs, want):
                assert got[i]["status"] == 0 and got[i]["value"] == w, q
    finally:
        pc.set_pass_slots(0)


def test_serving_memory_admission(pc):
    """A context with a 2 GiB device-memory budget answers the queries whose
    tables and scratch fit it and refuses pi(1e19) (7.5 GB of prime-count
    table alone) at once with PCG_E_NOMEM; an out-of-domain pi query is
    refused with PCG_E_INVALID (pi(1e30): x / y beyond the 2^49 limit of the
    prime-count table) and the others still share launches."""
    small = [(int(r["x"]), int(r["pi"])) for r in golden("pi_known.json")["rows"]
             if int(r["x"]) <= 10 ** 14]
    with pc.Server(hbm_budget=2 << 30) as sv:
        log("serving with a 2 GiB budget")
        big = sv.submit([(pc.KIND_PI, 10 ** 19)])[0]
        bad = sv.submit([(pc.KIND_PI, 10 ** 30)])[0]       # x / y beyond 2^49
        ids = sv.submit([(pc.KIND_PI, x) for x, _ in small])
        got = {r["id"]: r for r in sv.poll_all(len(ids) + 2, timeout_ms=120000)}
        assert got[big]["status"] == pc.E_NOMEM
        assert got[bad]["status"] == pc.E_INVALID
        for i, (x, v) in zip(ids, small):
            assert got[i]["status"] == 0 and got[i]["value"] == v, x
        st = sv.stats()
        assert st["rejected_nomem"] == 1
        assert st["batches"] < len(ids)


def test_serving_light_queries_interleave(pc):
    """Light queries submitted while heavy batches run are answered before
    the heavy queries finish, with the same values: at a pass point of the
    running heavy batch when the host reaches one after they arrive, else
    right after that batch (which of the two depends on host timing, so the
    number answered at pass points is only bounded here)."""
    heavy = [(int(r["x"]), int(r["pi"])) for r in golden("pi_known.json")["rows"]
             if 10 ** 18 <= int(r["x"]) <= 10 ** 19]
    light = [(int(r["x"]), int(r["pi"])) for r in golden("pi_known.json")["rows"]
             if int(r["x"]) <= 10 ** 13]
    with pc.Server() as sv:
        # tables for the largest query first
        w = sv.submit([(pc.KIND_PI, heavy[-1][0])])
        sv.poll_all(len(w), timeout_ms=120000)
        log("serving: %d heavy pi, then %d light pi during them" % (len(heavy), len(light)))
        hid = sv.submit([(pc.KIND_PI, x) for x, _ in heavy] * 3)
        time.sleep(0.03)
        lid = sv.submit([(pc.KIND_PI, x) for x, _ in light])
        order = []
        got = {}
        while len(got) < len(hid) + len(lid):
            for r in sv.poll(64, 120000):
                got[r["id"]] = r
                order.append(r["id"])
        for i, (x, v) in zip(hid, heavy * 3):
            assert got[i]["status"] == 0 and got[i]["value"] == v, x
        for i, (x, v) in zip(lid, light):
            assert got[i]["status"] == 0 and got[i]["value"] == v, x
        last_light = max(order.index(i) for i in lid)
        last_heavy = max(order.index(i) for i in hid)
        assert last_light < last_heavy
        st = sv.stats()
        log("light queries answered at pass points: %d of %d" % (st["interleaved"], len(lid)))
        assert st["interleaved"] <= len(lid)
