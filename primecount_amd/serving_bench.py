"""Serving benchmark: D(x) queries with Poisson arrivals through a serving
context (continuous batching), every answer checked against the reference.

    python -m primecount_amd.serving_bench [--queries 4096] [--rate 100] [--seed N]

Workload: the 256 x values of tests/golden/batch_D.json (x = 10^(12 + 6u),
u from SplitMix64, with the values the reference build printed for them);
each query picks one of them and the inter-arrival times are exponential with
mean 1/rate, both from SplitMix64 with the given seed.  A client thread
submits each query at its arrival time; the main thread polls.  Prints one
JSON line: throughput, latency percentiles (submit to result), batching.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--queries", type=int, default=4096)
    ap.add_argument("--rate", type=float, default=100.0, help="mean arrivals per second")
    ap.add_argument("--seed", type=int, default=0xC5C5)
    a = ap.parse_args()

    with open(os.path.join(REPO, "tests", "golden", "batch_D.json")) as f:
        rows = [(int(r["x"]), int(r["D"])) for r in json.load(f)["rows"]]
    st = a.seed & MASK
    plan = []
    t = 0.0
    for _ in range(a.queries):
        st, r1 = splitmix64(st)
        st, r2 = splitmix64(st)
        u = ((r2 >> 11) + 1) * 2.0 ** -53            # (0, 1]
        t += -math.log(u) / a.rate
        plan.append((t, rows[r1 % len(rows)]))

    sv = pc.Server(z_max=max(pc.gourdon_params(x)["z"] for x, _ in rows))
    # warm up: one launch with the largest query builds the resident tables
    big = max(rows)
    sv.submit([(pc.KIND_D, big[0])])
    sv.poll_all(1)

    want = {}
    lock = threading.Lock()

    def client():
        t0 = time.perf_counter()
        for at, (x, value) in plan:
            dt = at - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            i = sv.submit([(pc.KIND_D, x)])[0]
            with lock:
                want[i] = value

    s0 = sv.stats()
    t_start = time.perf_counter()
    th = threading.Thread(target=client)
    th.start()
    res = []
    while len(res) < a.queries:
        res += sv.poll(1024, 1000)
    elapsed = time.perf_counter() - t_start
    th.join()
    s1 = sv.stats()
    sv.close()

    bad = sum(1 for r in res if r["status"] != 0 or r["value"] != want[r["id"]])
    lat = sorted(r["seconds"] for r in res)

    def pct(p):
        return round(lat[min(len(lat) - 1, int(p * len(lat)))] * 1000, 3)

    batches = s1["batches"] - s0["batches"]
    print(json.dumps({
        "workload": "D(x) serving, x from the 256 reference x values (1e12..1e18), "
                    "Poisson arrivals",
        "queries": a.queries, "rate_offered_qps": a.rate, "seed": a.seed,
        "bit_exact": bad == 0, "mismatches": bad,
        "throughput_qps": round(a.queries / elapsed, 2), "elapsed_s": round(elapsed, 3),
        "latency_ms": {"p50": pct(0.50), "p90": pct(0.90), "p99": pct(0.99),
                       "max": round(lat[-1] * 1000, 3)},
        "batches": batches, "mean_batch": round(a.queries / max(batches, 1), 2),
        "max_batch": s1["max_batch"],
        "device_busy_frac": round((s1["busy_seconds"] - s0["busy_seconds"]) / elapsed, 3),
    }), flush=True)
    return 0 if bad == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
