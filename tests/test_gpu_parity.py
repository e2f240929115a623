"""Parity of the MI355X engine (HIP, through the C ABI) with the reference.

Mirrors the reference's own tests: test/gourdon/D.cpp (every row through the
64-bit and the 128-bit overload) and test/deleglise-rivat/S2_hard_alpha.cpp,
plus fixtures made by the reference build at the workload sizes, seeded
random parameters checked against the CPU oracle, and schedule invariance.
"""
import random
import sys

import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def kat_rows(name):
    return [[int(v) for v in r] for r in golden(name)["rows"]]


def test_device_present(pc):
    assert pc.device_available(), "no gfx950 device: the engine must not run without one"


def test_D_kat_rows(pc):
    log("D KAT rows, 64-bit and 128-bit entry points")
    for x, y, z, k, expected in kat_rows("kat_D.json"):
        assert pc.D(x, y, z, k) == expected, ("D", x, y, z, k)
        assert pc.D128(x, y, z, k) == expected, ("D128", x, y, z, k)


def test_S2_hard_kat_rows(pc):
    log("S2_hard KAT rows, 64-bit and 128-bit entry points")
    for x, y, z, c, expected in kat_rows("kat_S2_hard.json"):
        assert pc.S2_hard(x, y, z, c) == expected, ("S2_hard", x, y, z, c)
        assert pc.S2_hard128(x, y, z, c) == expected, ("S2_hard128", x, y, z, c)


def test_alpha_rows(pc):
    log("non-default alpha rows")
    for r in golden("alpha_D.json")["rows"]:
        x, y, z = int(r["x"]), int(r["y"]), int(r["z"])
        if "D" in r:
            assert pc.D(x, y, z, int(r["k"])) == int(r["D"]), r
        else:
            assert pc.S2_hard(x, y, z, int(r["c"])) == int(r["S2_hard"]), r


def test_random_vs_oracle(pc, oracle):
    rng = random.Random(7)
    for i in range(24):
        e = rng.uniform(6, 12)
        x = int(10 ** e)
        x13 = round(x ** (1 / 3))
        y = rng.randint(x13 + 1, max(x13 + 2, 12 * x13))
        z = rng.randint(y, max(y, min(3 * y, int(x ** 0.5) - 1)))
        log("random D x=%d y=%d z=%d" % (x, y, z))
        assert pc.D(x, y, z, 8) == oracle.D(x, y, z, 8), (x, y, z)
        yd = rng.randint(x13, 3 * x13)
        log("random S2_hard x=%d y=%d z=%d" % (x, yd, x // yd))
        assert pc.S2_hard(x, yd, x // yd, 8) == oracle.S2_hard(x, yd, x // yd, 8), (x, yd)


@pytest.mark.parametrize("words", [128, 256, 384, 512, 640, 1024, 2048, 4096])
def test_segment_size_invariance(pc, words):
    log("segment words %d" % words)
    r = {int(v["x"]): v for v in golden("configs.json")["D"]}[10 ** 15]
    s2 = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}[10 ** 15]
    try:
        pc.set_segment_words(words)
        assert pc.D(10 ** 15, int(r["y"]), int(r["z"]), int(r["k"])) == int(r["value"])
        assert pc.D128(10 ** 13, 241615, 241615, 8) == 304739017794
        assert pc.S2_hard_default(10 ** 15) == int(s2["value"])
    finally:
        pc.set_segment_words(0)


@pytest.mark.parametrize("e", [13, 15, 17])
def test_D_default_configs(pc, e):
    log("D default 1e%d" % e)
    r = {int(v["x"]): v for v in golden("configs.json")["D"]}[10 ** e]
    assert pc.D_default(10 ** e) == int(r["value"])
    st = pc.last_stats()
    assert st["leaves"] > 0


@pytest.mark.parametrize("e", [13, 15, 17])
def test_S2_hard_default_configs(pc, e):
    log("S2_hard default 1e%d" % e)
    r = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}[10 ** e]
    assert pc.S2_hard_default(10 ** e) == int(r["value"])


def test_D_1e18_1e19(pc):
    rows = {int(v["x"]): v for v in golden("configs.json")["D"]}
    for e in (18, 19):
        log("D default 1e%d" % e)
        assert pc.D_default(10 ** e) == int(rows[10 ** e]["value"])


def test_errors_are_reported(pc):
    with pytest.raises(pc.PrimecountError):
        pc.D(10 ** 12, 100, 50, 8)            # z < y is outside the domain
    # the engine still works after an error
    assert pc.D(10 ** 6, 207, 207, 8) == 2465


def test_S2_hard_1e19(pc):
    log("S2_hard default 1e19")
    r = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}[10 ** 19]
    assert pc.S2_hard_default(10 ** 19) == int(r["value"])


def test_int128_path_1e21(pc):
    # x > 2^64: x / p_b needs more than 64 bits (128-bit path of D.cpp:456-505)
    big = golden("configs_big.json")
    r = big["D"][0]
    log("D default 1e21 (128-bit)")
    x = int(r["x"])
    assert x > 2 ** 64
    assert pc.D_default(x) == int(r["value"])
    assert pc.D128(x, int(r["y"]), int(r["z"]), int(r["k"])) == int(r["value"])
    r = big["S2_hard"][0]
    log("S2_hard default 1e21 (128-bit)")
    assert pc.S2_hard_default(x) == int(r["value"])


def test_int128_entry_near_2_63(pc):
    # x just below and above 2^63: the 64-bit and the 128-bit entry points agree
    for x in (2 ** 63 - 1, 2 ** 63 + 12345):
        p = pc.gourdon_params(x)
        log("D near 2^63: x=%d" % x)
        got = pc.D128(x, p["y"], p["z"], p["k"])
        if x < 2 ** 63:
            assert pc.D(x, p["y"], p["z"], p["k"]) == got
        assert got > 0


def test_batch_matches_reference(pc):
    # 256 seeded queries x = 10^(12 + 6u) in one launch, against the values the
    # reference build printed for each x
    fx = golden("batch_D.json")
    xs = [int(r["x"]) for r in fx["rows"]]
    want = [int(r["D"]) for r in fx["rows"]]
    log("D batch of %d queries" % len(xs))
    assert pc.D_batch(xs) == want
    # order and duplicates do not matter
    sub = xs[::-7] + xs[:3]
    assert pc.D_batch(sub) == [want[xs.index(v)] for v in sub]


@pytest.mark.parametrize("words", [128, 512])
def test_shared_segments(pc, words):
    # queries of one launch sieve their common segments once (groups of up
    # to 4); the result of each query is the same with and without sharing
    # and equals the reference's value
    fx = golden("batch_D.json")["rows"]
    xs = [int(r["x"]) for r in fx][:96]
    want = [int(r["D"]) for r in fx][:96]
    try:
        pc.set_segment_words(words)
        for on in (True, False):
            pc.set_shared_segments(on)
            log("D batch of %d, %d-word segments, shared %s" % (len(xs), words, on))
            assert pc.D_batch(xs) == want
            # a few large queries and many small ones
            big = sorted(range(len(xs)), key=lambda i: -xs[i])[:6] + list(range(0, 96, 9))
            assert pc.D_batch([xs[i] for i in big]) == [want[i] for i in big]
    finally:
        pc.set_shared_segments(True)
        pc.set_segment_words(0)


def test_shared_segments_mixed_k(pc, oracle):
    # explicit parameters with y = z and different k in one serving launch:
    # queries with another first b are not grouped together
    rng = random.Random(77)
    qs = []
    for _ in range(12):
        x = rng.randrange(10 ** 9, 3 * 10 ** 10)
        x13 = round(x ** (1 / 3))
        y = rng.randint(x13 + 1, 4 * x13)
        qs.append((x, y, y, rng.choice([5, 6, 7, 8])))
    log("serving launch of %d D queries with mixed k" % len(qs))
    with pc.Server() as sv:
        ids = sv.submit([(pc.KIND_D, x, y, z, k) for x, y, z, k in qs])
        got = {r["id"]: r for r in sv.poll_all(len(ids), timeout_ms=120000)}
    for i, (x, y, z, k) in zip(ids, qs):
        assert got[i]["status"] == 0
        assert got[i]["value"] == oracle.D(x, y, z, k), (x, y, z, k)


@pytest.mark.parametrize("slots", [1 << 14, 1 << 20])
def test_pass_size_invariance(pc, slots):
    # many kernel passes (carry kept on the device between them) give the
    # same sums as one pass, for single queries and for a batch
    rows = {int(v["x"]): v for v in golden("configs.json")["D"]}
    fx = [r for r in golden("batch_D.json")["rows"] if int(r["x"]) <= 10 ** 15][:24]
    xs = [int(r["x"]) for r in fx]
    want = [int(r["D"]) for r in fx]
    try:
        pc.set_pass_slots(slots)
        log("pass slots %d: D default 1e13 and 1e15" % slots)
        for e in (13, 15):
            assert pc.D_default(10 ** e) == int(rows[10 ** e]["value"])
        s2 = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}[10 ** 15]
        assert pc.S2_hard_default(10 ** 15) == int(s2["value"])
        log("pass slots %d: batch of %d" % (slots, len(xs)))
        assert pc.D_batch(xs) == want
    finally:
        pc.set_pass_slots(0)


@pytest.mark.parametrize("kind,y,limit", [
    ("D", 241615, 241615), ("D", 11914935, 11914935), ("D", 88560164, 88560164),
    ("D", 1000, 100000), ("D", 5000, 2000000), ("D", 13, 10 ** 6),
    ("S2", 494134, 494134), ("S2", 10620399, 10620399), ("S2", 1, 50000)])
def test_device_factor_table(pc, kind, y, limit):
    # the factor tables are built on the device; every entry equals the host
    # builder's, which tests/native/tables_check.cpp checks against the
    # definitions (lpf, mpf, mu)
    log("factor table %s y=%d limit=%d" % (kind, y, limit))
    k = pc.KIND_D if kind == "D" else pc.KIND_S2_HARD
    assert pc.factor_table_check(k, y, limit) == 0


@pytest.mark.parametrize("bits", [20, 24, 28])
def test_slot_escape_invariance(pc, bits):
    # C and M share one 32-bit slot word; a slot whose M does not fit the
    # field above C goes to the pass's escape table.  More bits for C leave
    # fewer for M (28: |M| < 8), so that many slots escape: D, S2_hard,
    # a batch and a many-pass run give the same sums (an escape table too
    # small for them would fail the call, never return a wrong sum).
    rows = {int(v["x"]): v for v in golden("configs.json")["D"]}
    s2 = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}
    kat = [[int(v) for v in r] for r in golden("kat_D.json")["rows"]][::5]
    fx = [r for r in golden("batch_D.json")["rows"] if int(r["x"]) <= 10 ** 15][:16]
    try:
        pc.set_slot_bits(bits)
        log("slot bits %d: D default 1e13, 1e15, 1e17, S2_hard 1e15" % bits)
        for e in (13, 15, 17):
            assert pc.D_default(10 ** e) == int(rows[10 ** e]["value"])
            st = pc.last_stats()
            # the host's bound holds, and with 28 bits for C slots do escape
            assert st["escaped_slots"] <= st["escape_bound"]
            if bits == 28:
                assert st["escaped_slots"] > 0
        assert pc.S2_hard_default(10 ** 15) == int(s2[10 ** 15]["value"])
        for x, y, z, k, want in kat:
            assert pc.D(x, y, z, k) == want, (x, y, z, k)
        log("slot bits %d: batch of %d" % (bits, len(fx)))
        assert pc.D_batch([int(r["x"]) for r in fx]) == [int(r["D"]) for r in fx]
        pc.set_pass_slots(1 << 16)
        assert pc.D_default(10 ** 15) == int(rows[10 ** 15]["value"])
    finally:
        pc.set_pass_slots(0)
        pc.set_slot_bits(0)


@pytest.mark.parametrize("words", [256, 640, 1024, 2048])
def test_leaf_batching_invariance(pc, oracle, words):
    # The leaves of several primes counted on one prefix build (with the
    # recorded first-time removals as corrections) give the same sums as the
    # one-prime-at-a-time path; segments above 1024 words never batch.
    rows = {int(v["x"]): v for v in golden("configs.json")["D"]}
    s2 = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}
    kat = [[int(v) for v in r] for r in golden("kat_D.json")["rows"]][::4]
    fx = [r for r in golden("batch_D.json")["rows"] if int(r["x"]) <= 10 ** 15][:16]
    rng = random.Random(4242 + words)
    rnd = []
    for _ in range(4):
        x = rng.randrange(10 ** 10, 10 ** 11)
        x13 = round(x ** (1 / 3))
        y = rng.randint(x13 + 1, 3 * x13)
        rnd.append((x, y, rng.randint(y, 2 * y), rng.choice([6, 7, 8])))
    try:
        pc.set_segment_words(words)
        for on in (True, False):
            pc.set_leaf_batching(on)
            log("leaf batching %s, %d-word segments: D 1e15, 1e17, S2_hard 1e15" % (on, words))
            for e in (15, 17):
                assert pc.D_default(10 ** e) == int(rows[10 ** e]["value"])
                st = pc.last_stats()
                assert (st["leaf_batches"] > 0) == (on and words <= 1024)
            assert pc.S2_hard_default(10 ** 15) == int(s2[10 ** 15]["value"])
            for x, y, z, k, want in kat:
                assert pc.D(x, y, z, k) == want, (x, y, z, k)
            for x, y, z, k in rnd:
                assert pc.D(x, y, z, k) == oracle.D(x, y, z, k), (x, y, z, k)
            assert pc.D_batch([int(r["x"]) for r in fx]) == [int(r["D"]) for r in fx]
    finally:
        pc.set_leaf_batching(True)
        pc.set_segment_words(0)


def test_phase_timer_build(pc):
    # the kernel build with phase timers gives the same sums and reports the
    # cycles of every phase that ran
    rows = {int(v["x"]): v for v in golden("configs.json")["D"]}
    try:
        pc.set_phase_timers(True)
        for e in (15, 17):
            log("phase timers: D default 1e%d" % e)
            assert pc.D_default(10 ** e) == int(rows[10 ** e]["value"])
            cyc = pc.last_stats()["phase_cycles"]
            assert sum(cyc.values()) > 0
            for k in ("segment", "window", "dense_crossoff", "group_crossoff", "prefix_build",
                      "entry_steps", "counting_steps"):
                assert cyc[k] > 0, k
        pc.set_phase_timers(False)
        assert pc.D_default(10 ** 15) == int(rows[10 ** 15]["value"])
        assert sum(pc.last_stats()["phase_cycles"].values()) == 0
    finally:
        pc.set_phase_timers(False)


@pytest.mark.parametrize("words", [128, 256, 512, 640])
def test_kernel_builds_agree(pc, oracle, words):
    # the plain sieve kernel and the build for 6 waves per SIMD (fewer
    # registers, a few spills) give the reference's sums at each segment
    # size, single queries and a batch
    rows = {int(v["x"]): v for v in golden("configs.json")["D"]}
    s2 = {int(v["x"]): v for v in golden("configs.json")["S2_hard"]}
    kat = [[int(v) for v in r] for r in golden("kat_D.json")["rows"]][1::5]
    rng = random.Random(777 + words)
    rnd = []
    for _ in range(3):
        x = rng.randrange(10 ** 10, 10 ** 11)
        x13 = round(x ** (1 / 3))
        y = rng.randint(x13 + 1, 3 * x13)
        rnd.append((x, y, rng.randint(y, 2 * y), rng.choice([6, 7, 8])))
    try:
        for waves in (0, 6):
            pc.set_schedule(words=words, waves=waves)
            log("kernel build %s, %d-word segments: D 1e17, 1e18, S2_hard 1e17"
                % ("6 waves" if waves else "plain", words))
            for e in (17, 18):
                assert pc.D_default(10 ** e) == int(rows[10 ** e]["value"])
                st = pc.last_stats()
                assert st["segment_words"] == words and st["kernel_waves"] == waves
            assert pc.S2_hard_default(10 ** 17) == int(s2[10 ** 17]["value"])
            for x, y, z, k, want in kat:
                assert pc.D(x, y, z, k) == want, (x, y, z, k)
            for x, y, z, k in rnd:
                assert pc.D(x, y, z, k) == oracle.D(x, y, z, k), (x, y, z, k)
    finally:
        pc.set_schedule()


@pytest.mark.parametrize("words", [256, 640])
def test_shared_segment_batching(pc, words):
    # the shared-segment kernels count each member's leaves of several
    # primes on one prefix build; the sums equal the reference's with
    # batched leaf counting on and off, for groups of 4 and of 2
    fx = golden("batch_D.json")
    xs = [int(r["x"]) for r in fx["rows"]][:96]
    want = [int(r["D"]) for r in fx["rows"]][:96]
    try:
        for group, on in ((4, True), (4, False), (2, True), (2, False)):
            pc.set_leaf_batching(on)
            pc.set_schedule(words=words, group=group)
            log("D batch of %d, groups of %d, leaf batching %s, %d-word segments"
                % (len(xs), group, on, words))
            assert pc.D_batch(xs) == want
            st = pc.last_stats()
            assert st["group_size"] == group
            assert (st["leaf_batches"] > 0) == on
    finally:
        pc.set_leaf_batching(True)
        pc.set_schedule()
