#!/usr/bin/env python3
"""Benchmark: Gourdon's D(x) at x = 1e19 on one MI355X.

One step = one D(1e19, y, z, k) hard-special-leaf sum with primecount's
default parameters (y = z = 88560164, k = 8), through the 128-bit entry point
(x > 2^63, as the reference does), checked bit-exact against the reference
value on every step.  The prime, pi and factor tables are built on the first
(warmup) call and stay resident in HBM, so the timed steps measure the hot
path: planning, the sieve kernel, the carry combine and the read-back.

Prints ONE JSON line.  `roofline` is measured on the sieve kernel with HIP
events on the engine's stream (HBM); `roofline_lds` and `roofline_valu` put
the same kernel time against LDS bandwidth and against VALU issue, the bound
that holds (instructions per launch from the committed SQ-counter profile);
`cpu_baseline` times the reference build
(oracle/_ref/primecount, when present) on this host; `pi` reports the full
pi(x) (all Gourdon terms on the device, and beside it the Deleglise-Rivat
formula, every term on the device) outside the timed D steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--x 1e19] [--no-cpu]
"""
import argparse
import json
import math
import os
import re
import subprocess
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz = 78,643 GB/s (nominal)


def expected_D(x):
    with open(os.path.join(REPO, "tests", "golden", "configs.json")) as f:
        for r in json.load(f)["D"]:
            if int(r["x"]) == x:
                return int(r["value"]), r
    big = os.path.join(REPO, "tests", "golden", "configs_big.json")
    if os.path.exists(big):
        with open(big) as f:
            for r in json.load(f)["D"]:
                if int(r["x"]) == x:
                    return int(r["value"]), r
    return None, None


def measured_lds_peak(pc):
    """LDS read bandwidth of this device in GB/s, from the engine library's
    probe kernel (bank-conflict-free 128-bit LDS loads on every CU)."""
    try:
        return pc.lds_read_gbs()
    except pc.PrimecountError:
        return None


def measure_pi(pc, x):
    """pi(x) by Gourdon's algorithm on the device (all six terms), outside
    the timed D steps: one untimed call builds the prime-count table, then
    the median of 3 calls; checked against the reference's value."""
    known = {}
    with open(os.path.join(REPO, "tests", "golden", "pi_known.json")) as f:
        for r in json.load(f)["rows"]:
            known[int(r["x"])] = int(r["pi"])
    got = pc.pi(x)
    times = []
    for _ in range(3):
        t = time.perf_counter()
        got = pc.pi(x)
        times.append(time.perf_counter() - t)
    st = pc.last_pi_stats()
    times.sort()
    # the Deleglise-Rivat formula on the same device (S1 + S2 + pi(y) - 1 - P2)
    dr = pc.pi_deleglise_rivat(x)
    dr_times = []
    for _ in range(3):
        t = time.perf_counter()
        dr = pc.pi_deleglise_rivat(x)
        dr_times.append(time.perf_counter() - t)
    dr_times.sort()
    return {"x": str(x), "value": str(got), "bit_exact": known.get(x) == got,
            "seconds": round(times[1], 6),
            "terms_s": {k: round(st["seconds_" + k], 6) for k in ("sigma", "phi0", "ac", "b", "d")},
            "ac_leaves": st["ac_leaves"], "table_limit": st["table_limit"],
            "deleglise_rivat": {"value": str(dr), "bit_exact": known.get(x) == dr,
                                "seconds": round(dr_times[1], 6)}}


def measure_serving(pc, n, n_poisson, rate):
    """Config 5, the serving mix (tests/golden/serve_c5.json: 80% pi(x),
    x = 10^(9+10u); 20% nth_prime(n), n = 10^(8+9u); with the reference's
    answers): its first n queries one call per query back to back (the plain
    path), the same queries submitted at once to a serving context (its
    capacity), and the first n_poisson queries with Poisson arrivals at `rate`
    per second (latency under load).  None when the fixture is absent."""
    path = os.path.join(REPO, "tests", "golden", "serve_c5.json")
    if n <= 0 or not os.path.exists(path):
        return None
    from primecount_amd import serving_bench as sb
    qs = sb.load_workload("mix", n, 0)
    for k, a, _ in sorted(qs, key=lambda q: -q[1])[:2]:       # tables for the largest
        (pc.pi if k == pc.KIND_PI else pc.nth_prime)(a)

    def pct(v, q):
        return round(v[min(len(v) - 1, int(q * len(v)))] * 1000, 3)

    lat, el, bad = sb.run_plain(qs)
    lat.sort()
    plain = {"queries_per_s": round(len(qs) / el, 2), "seconds": round(el, 2),
             "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99), "bit_exact": bad == 0}

    def served(qq, times):
        sv = pc.Server()
        try:
            res, el2, bad2, s0, s1 = sb.run(sv, qq, times)
        finally:
            sv.close()
        lat2 = sorted(r["seconds"] for r in res)
        return {"queries_per_s": round(len(qq) / el2, 2), "seconds": round(el2, 2),
                "p50_ms": pct(lat2, 0.5), "p90_ms": pct(lat2, 0.9), "p99_ms": pct(lat2, 0.99),
                "bit_exact": bad2 == 0, "launch_sets": s1["batches"] - s0["batches"],
                "interleaved": s1["interleaved"] - s0["interleaved"]}

    cap = served(qs, [0.0] * len(qs))
    qp = qs[:n_poisson]
    poisson = served(qp, sb.arrivals(len(qp), rate, 0xC5C5)) if qp else None
    return {"workload": "serving mix, first %d queries" % len(qs),
            "plain": plain,
            "served_capacity": dict(cap, arrivals="all %d submitted at once" % len(qs)),
            "served_capacity_vs_plain": round(cap["queries_per_s"] / plain["queries_per_s"], 3),
            "served_poisson": None if poisson is None else dict(
                poisson, arrivals="first %d queries, Poisson %.3g/s (seed 0xC5C5)"
                % (len(qp), rate)),
            "p99_target_ms": round(2 * plain["p99_ms"], 3)}


def measure_batch(pc):
    """Config 3: the 256 seeded D(x) queries of tests/golden/batch_D.json
    (x = 10^(12 + 6u), values printed by the reference build) in one
    persistent launch (pcg_D_batch), median of 3 after one warmup; beside it
    the same queries one per launch."""
    path = os.path.join(REPO, "tests", "golden", "batch_D.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        fx = json.load(f)
    xs = [int(r["x"]) for r in fx["rows"]]
    want = [int(r["D"]) for r in fx["rows"]]
    ok = pc.D_batch(xs) == want
    times, kern = [], []
    for _ in range(3):
        t = time.perf_counter()
        got = pc.D_batch(xs)
        times.append(time.perf_counter() - t)
        kern.append(pc.last_stats()["seconds_kernel"])
        ok = ok and got == want
    times.sort()
    lat = []
    t0 = time.perf_counter()
    for x, v in zip(xs, want):
        t = time.perf_counter()
        ok = ok and pc.D_default(x) == v
        lat.append(time.perf_counter() - t)
    plain_s = time.perf_counter() - t0
    lat.sort()
    return {"workload": "256 D(x) queries, x = 10^(12+6u) from SplitMix64 seed %s, one "
                        "persistent launch" % fx.get("seed", "?"),
            "queries": len(xs), "seconds": round(times[1], 4),
            "queries_per_s": round(len(xs) / times[1], 1),
            "kernel_s": round(sorted(kern)[1], 4), "bit_exact": ok,
            "plain_one_per_launch": {"seconds": round(plain_s, 4),
                                     "queries_per_s": round(len(xs) / plain_s, 1),
                                     "p50_ms": round(lat[len(lat) // 2] * 1000, 3),
                                     "p99_ms": round(lat[int(0.99 * len(lat))] * 1000, 3)}}


def _sched(st):
    return {"segment_words": st["segment_words"], "kernel_waves": st["kernel_waves"],
            "grid_waves": st["grid_waves"], "group_size": st["group_size"]}


def measure_int128(pc, tune=True):
    """Config 4: D(x) and S2_hard(x) at x = 1e21 (x > 2^64, the 128-bit path)
    with the reference's default parameters, against the values the
    reference build printed: the first call (which builds the tables) and
    the second (tables resident) on the plain schedule (the segment-size
    rule), then the autotuner on a fresh cache file: the tuning call (the
    rule's schedule and the candidates around it, each checked to give the
    same value) and a call on the tuned schedule."""
    path = os.path.join(REPO, "tests", "golden", "configs_big.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        big = json.load(f)
    out = {}
    import tempfile
    fd, cache = tempfile.mkstemp(prefix="pcg_schedules_", suffix=".txt")
    os.close(fd)
    os.unlink(cache)                 # a fresh cache: every class is tuned here
    for name, fn in (("D", pc.D_default), ("S2_hard", pc.S2_hard_default)):
        r = big[name][0]
        x, want = int(r["x"]), int(r["value"])
        pc.set_autotune(pc.TUNE_OFF)
        t = time.perf_counter()
        v1 = fn(x)
        first = time.perf_counter() - t
        t = time.perf_counter()
        v2 = fn(x)
        second = time.perf_counter() - t
        st = pc.last_stats()
        out[name] = {"x": str(x), "value": str(v2), "bit_exact": v1 == want and v2 == want,
                     "seconds_first_call": round(first, 3), "seconds": round(second, 3),
                     "kernel_s": round(st["seconds_kernel"], 3),
                     "segment_words": st["segment_words"]}
        if not tune:
            continue
        try:
            pc.set_schedule_cache(cache)
            pc.set_autotune(pc.TUNE_ON)
            t = time.perf_counter()
            v3 = fn(x)
            tune_s = time.perf_counter() - t
            assert pc.last_stats()["schedule_source"] == 2
            t = time.perf_counter()
            v4 = fn(x)
            tuned_s = time.perf_counter() - t
            st4 = pc.last_stats()
            out[name]["plain_schedule"] = _sched(st)
            out[name]["tuned"] = {"seconds": round(tuned_s, 3),
                                  "kernel_s": round(st4["seconds_kernel"], 3),
                                  "bit_exact": v3 == want and v4 == want,
                                  "tuning_call_s": round(tune_s, 3),
                                  "schedule": _sched(st4),
                                  "vs_plain": round(second / tuned_s, 3)}
        finally:
            pc.set_autotune(pc.TUNE_CACHED)
            pc.set_schedule_cache(None)
    if os.path.exists(cache):
        os.unlink(cache)
    return out


def traffic_from_profile(x):
    """HBM bytes per sieve-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel)."""
    path = os.path.join(REPO, "profiles", "hsl_sieve_kernel_D_1e%d_pmc.md" % round(math.log10(x)))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            m = re.match(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", line.strip())
            if m:
                return float(m.group(1))
    return None


def valu_from_profile(x):
    """Wave64 VALU instructions per sieve-kernel launch from the committed
    rocprofv3 --pmc summary of the SQ counters (SQ_INSTS_VALU)."""
    path = os.path.join(REPO, "profiles", "hsl_sieve_kernel_D_1e%d_sq.md" % round(math.log10(x)))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            m = re.match(r"^valu_instructions_per_launch:\s*([0-9.eE+]+)", line.strip())
            if m:
                return float(m.group(1))
    return None


def pi_small(pc, n):
    """Wall time of pcg_pi for small x against the device time of its terms
    (stream events), n calls per size after one warm-up call.  The gap is
    host work plus launch overhead; a kernel trace of the same run counts the
    dispatches per call."""
    out = {}
    for e in (9, 10, 11, 12, 13):
        x = 10 ** e
        pc.pi(x)
        time.sleep(0.05)
        walls, dev = [], []
        for _ in range(n):
            t0 = time.perf_counter()
            pc.pi(x)
            walls.append(time.perf_counter() - t0)
            st = pc.last_pi_stats()
            dev.append(st["seconds_sigma"] + st["seconds_phi0"] + st["seconds_ac"] +
                       st["seconds_b"] + st["seconds_d"])
        walls.sort()
        dev.sort()
        out["1e%d" % e] = {"calls": n, "wall_ms_median": round(1e3 * walls[n // 2], 4),
                           "terms_device_ms_median": round(1e3 * dev[n // 2], 4)}
        time.sleep(0.05)
    return out


def phase_split(pc, x):
    """Share of the sieve kernel's wave time per phase at D(x) (default
    parameters), from the phase-timer build (clock read at every phase
    change, cycles summed over the waves), beside the kernel time of the
    plain build."""
    pc.D_default(x)
    pc.D_default(x)
    plain = pc.last_stats()
    pc.set_phase_timers(True)
    try:
        pc.D_default(x)
        pc.D_default(x)
        st = pc.last_stats()
    finally:
        pc.set_phase_timers(False)
    cyc = st["phase_cycles"]
    tot = sum(cyc.values()) or 1
    return {"x": str(x), "kernel_ms_plain": round(1e3 * plain["seconds_kernel"], 3),
            "kernel_ms_timed": round(1e3 * st["seconds_kernel"], 3),
            "share": {k: round(v / tot, 4) for k, v in cyc.items()},
            "wave_cycles": tot}


SCHEDULES = ((512, 6), (512, 0), (640, 0))


def schedules(pc, x, calls):
    """Kernel time of D(x) (default parameters) under fixed schedules: segment
    words W and the kernel build (6 = the 6-waves build, 0 = the plain one),
    median of `calls` calls each after one call under the default schedule
    (which builds the tables), every result checked against that one."""
    want = pc.D_default(x)
    out = []
    try:
        for w, waves in SCHEDULES:
            pc.set_schedule(words=w, waves=waves)
            ts = []
            for _ in range(calls):
                if pc.D_default(x) != want:
                    raise SystemExit("D(%d) differs under W=%d waves=%d" % (x, w, waves))
                ts.append(1e3 * pc.last_stats()["seconds_kernel"])
            st = pc.last_stats()
            out.append({"segment_words": w, "kernel_build": "6 waves" if waves == 6 else "plain",
                        "kernel_ms": round(statistics.median(ts), 3), "calls": calls,
                        "windows": st["windows"], "slots_visited": st["pairs"],
                        "bit_exact": True})
    finally:
        pc.set_schedule()
    return {"x": str(x), "schedules": out}


def device_copy(nbytes, reps=2):
    """A plain device-to-device copy of `nbytes` (HIP's own copy kernel), used
    to calibrate the rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpyDtoD(dst, src, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    return nbytes


def cpu_baseline(x, timeout):
    cli = os.path.join(REPO, "oracle", "_ref", "primecount")
    if not os.path.exists(cli):
        return None
    # this process's CPU share, at most 16 threads (one GPU's share of a host)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    try:
        r = subprocess.run([cli, str(x), "--D", "-s", "-t", str(cores)], capture_output=True,
                           text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"value": None, "unit": "s", "cores": cores, "kind": "reference",
                "sample": "D(%d) timed out after %d s" % (x, timeout)}
    secs = None
    val = None
    for line in r.stdout.splitlines():
        line = line.split("\r")[-1]
        m = re.match(r"^Seconds:\s*([0-9.]+)", line)
        if m:
            secs = float(m.group(1))
        m = re.match(r"^D = (-?\d+)", line)
        if m:
            val = int(m.group(1))
    return {"value": secs, "unit": "s", "cores": cores, "kind": "reference",
            "sample": "full D(%d) by the reference build (primecount --D -t %d), value %s" %
                      (x, cores, "matches" if val == expected_D(x)[0] else "MISMATCH")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--x", type=float, default=1e19)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-timeout", type=int, default=120)
    ap.add_argument("--segment-words", type=int, default=0,
                    help="64-bit words per sieve segment (0 = automatic)")
    ap.add_argument("--no-pi", action="store_true", help="skip the pi(x) measurement")
    ap.add_argument("--serving", type=int, default=4096,
                    help="queries of the serving mix to run plain and all at once through a "
                         "serving context (0 = skip)")
    ap.add_argument("--serving-poisson", type=int, default=1024,
                    help="queries of the serving mix to serve with Poisson arrivals")
    ap.add_argument("--serving-rate", type=float, default=25.0,
                    help="Poisson arrival rate of the served run (queries/s)")
    ap.add_argument("--no-lds-probe", action="store_true",
                    help="use the nominal LDS peak instead of running the LDS probe kernel")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 256-query batch and the 1e21 (128-bit) measurements")
    ap.add_argument("--pi-small", type=int, default=0,
                    help="only time N calls each of pi(10^9) .. pi(10^13) (launch-overhead "
                         "measurement; 50 ms pauses separate the sizes in a kernel trace) and "
                         "print their JSON line")
    ap.add_argument("--phase-split", action="store_true",
                    help="only measure the sieve kernel's time per phase at D(x) with the "
                         "phase-timer build and print its JSON line")
    ap.add_argument("--schedules", action="store_true",
                    help="only time D(x) under fixed schedules (segment words, kernel build), "
                         "--steps calls each, and print their JSON line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy (for rocprofv3 --pmc "
                         "FETCH_SIZE/WRITE_SIZE calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the hard-special-leaf engine runs on a single MI355X "
              "(sharding by segment range across GPUs is a design, not built); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import primecount_amd as pc
    if a.segment_words:
        pc.set_segment_words(a.segment_words)
    if a.pi_small:
        print(json.dumps({"pi_small": pi_small(pc, a.pi_small)}), flush=True)
        return
    if a.phase_split:
        print(json.dumps({"phase_split": phase_split(pc, int(round(a.x)))}), flush=True)
        return
    if a.schedules:
        print(json.dumps({"schedule_table": schedules(pc, int(round(a.x)), a.steps)}), flush=True)
        return
    x = int(round(a.x))
    p = pc.gourdon_params(x)
    want, _ = expected_D(x)
    print("bench: D(%d) y=%d z=%d k=%d" % (x, p["y"], p["z"], p["k"]), file=sys.stderr, flush=True)

    t = time.time()
    first = None
    for i in range(max(a.warmup, 1)):
        got = pc.D128(x, p["y"], p["z"], p["k"])
        if want is not None and got != want:
            raise SystemExit("warmup result %d != reference %d" % (got, want))
        if first is None:
            first = pc.last_stats()       # the first call builds the tables
    print("bench: warmup done in %.2f s (tables %.2f s)" % (time.time() - t, first["seconds_tables"]),
          file=sys.stderr, flush=True)

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None

    kernel_s, combine_s, plan_s, wall = [], [], [], []
    sync()
    t0 = time.perf_counter()
    for i in range(a.steps):
        ts = time.perf_counter()
        got = pc.D128(x, p["y"], p["z"], p["k"])
        wall.append(time.perf_counter() - ts)
        st = pc.last_stats()
        kernel_s.append(st["seconds_kernel"])
        combine_s.append(st["seconds_combine"])
        plan_s.append(st["seconds_plan"])
        if want is not None and got != want:
            raise SystemExit("step %d result %d != reference %d" % (i, got, want))
    sync()
    elapsed = time.perf_counter() - t0
    sec_per_step = elapsed / a.steps

    k_avg = sum(kernel_s) / len(kernel_s)
    lds_bytes = 8.0 * st["hits"] + 12.0 * st["leaves"]
    hbm_bytes = 2.0 * st["factor_entries"]
    traffic = traffic_from_profile(x)
    roofline = {
        "bound": "hbm",
        "achieved": round(hbm_bytes / k_avg / 1e9, 3),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(hbm_bytes / k_avg / 1e9 / HBM_PEAK_GBS, 6),
        "traffic": traffic,
    }
    lds_peak = None if a.no_lds_probe else measured_lds_peak(pc)
    lds_peak_used = lds_peak if lds_peak else LDS_PEAK_GBS
    roofline_lds = {
        "bound": "lds",
        "achieved": round(lds_bytes / k_avg / 1e9, 3),
        "peak": round(lds_peak_used, 1),
        "peak_source": "measured (LDS read probe)" if lds_peak else "nominal",
        "peak_nominal": LDS_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(lds_bytes / k_avg / 1e9 / lds_peak_used, 6),
        "bytes_per_launch": lds_bytes,
    }
    # the bound that holds: VALU issue.  A wave64 VALU instruction takes a
    # SIMD16 4 cycles; MI355X has 1,024 SIMDs at up to 2.4 GHz.
    valu = valu_from_profile(x)
    valu_peak = 1024 * 2.4e9 / 4          # wave instructions per second
    roofline_valu = None if valu is None else {
        "bound": "valu",
        "achieved": round(valu / k_avg / 1e12, 4),
        "peak": round(valu_peak / 1e12, 4),
        "unit": "T wave64 VALU instructions/s",
        "frac": round(valu / k_avg / valu_peak, 4),
        "instructions_per_launch": valu,
        "source": "SQ_INSTS_VALU of the committed rocprofv3 summary, over this run's kernel time",
    }
    out = {
        "metric": "D(x) secs at x=1e19",
        "value": round(sec_per_step, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1000 * sec_per_step, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int64/int128",
        "data": "synthetic: fully determined by x (no dataset)",
        "config": {
            "workload": "D(x) single query, x=%d, y=z=%d, k=%d, one query per launch" % (x, p["y"], p["k"]),
            "x": str(x), "y": p["y"], "z": p["z"], "k": p["k"], "x_star": p["x_star"],
            "segment_words": st["segment_words"], "segments": st["segments"], "chunks": st["chunks"],
        },
        "bit_exact": want is not None,
        "kernel_ms": round(1000 * k_avg, 3),
        "combine_ms": round(1000 * sum(combine_s) / len(combine_s), 3),
        "host_plan_ms": round(1000 * sum(plan_s) / len(plan_s), 3),
        "tables": "resident: the timed steps reuse the prime, prime-count and factor tables "
                  "that the first warmup call built in HBM (the reference's D() builds them "
                  "on every call); that first call is reported separately",
        "table_build_s_first_call": round(first["seconds_tables"], 3),
        "end_to_end_s_first_call": round(first["seconds_total"], 3),
        "work": {k: st[k] for k in ("leaves", "hits", "factor_entries", "pairs", "windows",
                                    "leaf_primes", "leaf_steps", "dense_primes", "dense_steps",
                                    "group_steps", "prefix_builds", "queue_steps")},
        "roofline": roofline,
        "roofline_lds": roofline_lds,
        "roofline_valu": roofline_valu,
        "cpu_baseline": None if a.no_cpu else cpu_baseline(x, a.cpu_timeout),
    }
    if not a.no_pi:
        out["pi"] = measure_pi(pc, x)
    if not a.no_configs:
        out["batch_256"] = measure_batch(pc)
        out["int128_1e21"] = measure_int128(pc)
    if a.serving:
        out["serving"] = measure_serving(pc, a.serving, a.serving_poisson, a.serving_rate)
    if a.calibrate_copy:
        out["calibration_copy_bytes"] = device_copy(1 << 30)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
