"""Serving benchmark: queries with Poisson arrivals through a serving context
(continuous batching), every answer checked against the reference.

    python -m primecount_amd.serving_bench [--workload mix|d] [--queries N]
                                           [--rate QPS] [--seed N]

Workloads:
  mix  the 4096-query serving mix of tests/golden/serve_c5.json: 80% pi(x)
       with x = 10^(9 + 10u), 20% nth_prime(n) with n = 10^(8 + 9u), u from
       SplitMix64, with the answers the reference build printed for them.
       Queries arrive in the fixture's order.
  d    D(x) queries, each picking one of the 256 x values of
       tests/golden/batch_D.json (x = 10^(12 + 6u)) with its reference value.

Inter-arrival times are exponential with mean 1/rate, from SplitMix64 with the
given seed.  --rate 0 submits every query at once (closed loop), which
measures the capacity.  A client thread submits each query at its arrival
time; the main thread polls.  --plain instead answers the queries one call
at a time (pi(), nth_prime(), D_default(): one query per launch), the
baseline the serving context is compared with.  Prints one JSON line:
throughput, latency percentiles (submit to result), batching, and whether
every answer matched.
"""
import argparse
import json
import math
import os
import sys
import threading
import time

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import primecount_amd as pc  # noqa: E402

MASK = (1 << 64) - 1


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def load_workload(name, n, seed, fixture=None):
    """[(kind, arg, expected)] for the workload, n queries."""
    gold = os.path.join(REPO, "tests", "golden")
    if name == "mix":
        path = fixture or os.path.join(gold, "serve_c5.json")
        if path.endswith(".json"):
            rows = json.load(open(path))["rows"]
        else:                                    # JSON lines
            rows = [json.loads(line) for line in open(path)]
        kinds = {"pi": pc.KIND_PI, "nth": pc.KIND_NTH_PRIME}
        qs = [(kinds[r["kind"]], int(r["arg"]), int(r["value"])) for r in rows]
        return qs[:n]
    rows = [(int(r["x"]), int(r["D"])) for r in
            json.load(open(os.path.join(gold, "batch_D.json")))["rows"]]
    st = (seed ^ 0xD) & MASK
    qs = []
    for _ in range(n):
        st, r = splitmix64(st)
        x, v = rows[r % len(rows)]
        qs.append((pc.KIND_D, x, v))
    return qs


def arrivals(n, rate, seed):
    """Arrival times (s) of n queries: exponential gaps with mean 1/rate."""
    if rate <= 0:
        return [0.0] * n
    st = seed & MASK
    t, out = 0.0, []
    for _ in range(n):
        st, r = splitmix64(st)
        u = ((r >> 11) + 1) * 2.0 ** -53            # (0, 1]
        t += -math.log(u) / rate
        out.append(t)
    return out


def run(sv, qs, times):
    want = {}
    lock = threading.Lock()

    def client():
        t0 = time.perf_counter()
        if times[-1] == 0.0:
            ids = sv.submit([(k, a) for k, a, _ in qs])
            with lock:
                for i, (_, _, v) in zip(ids, qs):
                    want[i] = v
            return
        for at, (k, a, v) in zip(times, qs):
            dt = at - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            i = sv.submit([(k, a)])[0]
            with lock:
                want[i] = v

    s0 = sv.stats()
    t_start = time.perf_counter()
    th = threading.Thread(target=client)
    th.start()
    res = []
    last = time.perf_counter()
    while len(res) < len(qs):
        res += sv.poll(1024, 1000)
        if time.perf_counter() - last > 30:
            print("served %d of %d" % (len(res), len(qs)), file=sys.stderr, flush=True)
            last = time.perf_counter()
    elapsed = time.perf_counter() - t_start
    th.join()
    s1 = sv.stats()
    with lock:
        bad = sum(1 for r in res if r["status"] != 0 or r["value"] != want[r["id"]])
    return res, elapsed, bad, s0, s1


TERMS = ("table", "sigma", "phi0", "b", "ac", "d")


def split_plain(qs, lat, terms):
    """Where the plain path's time goes: per kind and decade of the argument,
    the queries, their wall time and the device-stream time of each pi(x)
    term (stream events between the term launches; A/C includes the host's
    descriptor lists, D its planning; for nth_prime the terms are those of
    pi at the starting point, the rest is the interval sieve)."""
    rows = {}
    for (k, a, _), w, t in zip(qs, lat, terms):
        key = ("pi" if k == pc.KIND_PI else "nth" if k == pc.KIND_NTH_PRIME else "D",
               int(math.log10(max(a, 1))))
        r = rows.setdefault(key, {"queries": 0, "wall_s": 0.0, **{n: 0.0 for n in TERMS}})
        r["queries"] += 1
        r["wall_s"] += w
        for n in TERMS:
            r[n] += t.get(n, 0.0)
    out = []
    for (kind, dec), r in sorted(rows.items()):
        other = r["wall_s"] - sum(r[n] for n in TERMS)
        out.append({"kind": kind, "decade": dec, "queries": r["queries"],
                    "wall_s": round(r["wall_s"], 4),
                    **{n + "_s": round(r[n], 4) for n in TERMS},
                    "other_s": round(other, 4)})
    tot = {"wall_s": round(sum(x["wall_s"] for x in out), 3)}
    for n in TERMS:
        tot[n + "_s"] = round(sum(x[n + "_s"] for x in out), 3)
    tot["other_s"] = round(sum(x["other_s"] for x in out), 3)
    return {"rows": out, "total": tot}


def run_plain(qs, terms=None):
    """One call per query, back to back: (latencies, elapsed, mismatches).
    With a list `terms`, each query's pi(x) term times are appended to it."""
    call = {pc.KIND_PI: pc.pi, pc.KIND_NTH_PRIME: pc.nth_prime, pc.KIND_D: pc.D_default}
    lat, bad = [], 0
    t_start = time.perf_counter()
    last = t_start
    for i, (k, a, v) in enumerate(qs):
        t = time.perf_counter()
        got = call[k](a)
        lat.append(time.perf_counter() - t)
        if terms is not None:
            st = pc.last_pi_stats() if k != pc.KIND_D else {}
            terms.append({"table": st.get("seconds_table", 0.0),
                          **{n: st.get("seconds_" + n, 0.0) for n in TERMS[1:]}})
        bad += got != v
        if time.perf_counter() - last > 30:
            print("plain: %d of %d" % (i + 1, len(qs)), file=sys.stderr, flush=True)
            last = time.perf_counter()
    return lat, time.perf_counter() - t_start, bad


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("mix", "d"), default="mix")
    ap.add_argument("--queries", type=int, default=4096)
    ap.add_argument("--rate", type=float, default=None,
                    help="mean arrivals per second (0: all at once); default 20 for "
                         "mix, 100 for d")
    ap.add_argument("--seed", type=int, default=0xC5C5)
    ap.add_argument("--max-batch", type=int, default=0)
    ap.add_argument("--fixture", default=None, help="serving-mix answers (JSON or JSON lines)")
    ap.add_argument("--plain", action="store_true", help="one call per query, no serving context")
    ap.add_argument("--split", action="store_true",
                    help="with --plain: report the time per pi(x) term by kind and decade")
    a = ap.parse_args()
    rate = a.rate if a.rate is not None else (20.0 if a.workload == "mix" else 100.0)

    qs = load_workload(a.workload, a.queries, a.seed, a.fixture)
    times = arrivals(len(qs), rate, a.seed)

    if a.workload == "d":
        z_max = max(pc.gourdon_params(x)["z"] for _, x, _ in qs)
        warm = [max(qs, key=lambda q: q[1])]
    else:
        z_max = 0
        warm = [max((q for q in qs if q[0] == pc.KIND_PI), key=lambda q: q[1]),
                max((q for q in qs if q[0] == pc.KIND_NTH_PRIME), key=lambda q: q[1])]
    if a.workload == "mix":
        npi = sum(1 for q in qs if q[0] == pc.KIND_PI)
        desc = ("serving mix: %d pi(x), x = 10^(9+10u); %d nth_prime(n), n = 10^(8+9u)"
                % (npi, len(qs) - npi))
    else:
        desc = "D(x), x from the 256 reference x values (1e12..1e18)"

    if a.plain:
        call = {pc.KIND_PI: pc.pi, pc.KIND_NTH_PRIME: pc.nth_prime, pc.KIND_D: pc.D_default}
        t0 = time.perf_counter()
        for k, x, _ in warm:
            call[k](x)
        warm_s = time.perf_counter() - t0
        terms = [] if a.split else None
        lat, elapsed, bad = run_plain(qs, terms)
        split = split_plain(qs, lat, terms) if a.split else None
        lat.sort()

        def pctp(p):
            return round(lat[min(len(lat) - 1, int(p * len(lat)))] * 1000, 3)

        print(json.dumps({
            "workload": desc, "arrivals": "one call per query, back to back (plain path)",
            "queries": len(qs), "bit_exact": bad == 0, "mismatches": bad,
            "throughput_qps": round(len(qs) / elapsed, 2), "elapsed_s": round(elapsed, 3),
            "latency_ms": {"p50": pctp(0.50), "p90": pctp(0.90), "p99": pctp(0.99),
                           "max": round(lat[-1] * 1000, 3)},
            "warmup_s": round(warm_s, 3),
            **({"split": split} if split else {}),
        }), flush=True)
        return 0 if bad == 0 else 1
    sv = pc.Server(z_max=z_max, max_batch=a.max_batch)
    # warm up: the largest queries build the resident tables
    t0 = time.perf_counter()
    sv.submit([(k, x) for k, x, _ in warm])
    sv.poll_all(len(warm))
    warm_s = time.perf_counter() - t0

    res, elapsed, bad, s0, s1 = run(sv, qs, times)
    sv.close()
    lat = sorted(r["seconds"] for r in res)

    def pct(p):
        return round(lat[min(len(lat) - 1, int(p * len(lat)))] * 1000, 3)

    batches = s1["batches"] - s0["batches"]
    print(json.dumps({
        "workload": desc,
        "arrivals": "all at once" if rate <= 0 else "Poisson, %.3g/s" % rate,
        "queries": len(qs), "rate_offered_qps": rate, "seed": a.seed,
        "bit_exact": bad == 0, "mismatches": bad,
        "throughput_qps": round(len(qs) / elapsed, 2), "elapsed_s": round(elapsed, 3),
        "latency_ms": {"p50": pct(0.50), "p90": pct(0.90), "p99": pct(0.99),
                       "max": round(lat[-1] * 1000, 3)},
        "batches": batches, "mean_batch": round(len(qs) / max(batches, 1), 2),
        "max_batch": s1["max_batch"],
        "device_busy_frac": round((s1["busy_seconds"] - s0["busy_seconds"]) / elapsed, 3),
        "warmup_s": round(warm_s, 3),
    }), flush=True)
    return 0 if bad == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
