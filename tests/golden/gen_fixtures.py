#!/usr/bin/env python3
"""Generate the golden fixtures for the hard-special-leaf path.

Runs HERE only (it needs the reference sources under /root/reference and the
reference CLI built from them by ``make -f oracle/ref.mk``).  The outputs are
plain JSON data files next to this script: inputs and the reference's exact
outputs.  Nothing on the GPU side ever runs this script.

Fixture files
-------------
kat_D.json          the 51 (x, y, z, k) -> D rows of the reference's own test
                    table (test/gourdon/D.cpp:36-89)
kat_S2_hard.json    the 35 (x, y, z, c) -> S2_hard rows of
                    test/deleglise-rivat/S2_hard_alpha.cpp:36-73
params.json         x -> (y, z, k, x_star) chosen by the reference's double
                    precision alpha formulas (Gourdon, via `--Phi0 -s`) and
                    x -> (y, z, c) (Deleglise-Rivat, via `--S2-trivial -s`)
configs.json        D(x) and S2_hard(x) with default parameters at the
                    workload sizes (1e13 ... 1e19, optionally 1e21)
batch_D.json        the 256-query batch: x_i = floor(10^(12 + 6 u_i)),
                    u_i = (splitmix64 >> 11) * 2^-53, and D(x_i)
alpha_D.json        D(x) at non-default alpha_y / alpha_z (y < z cases)
nth_prime.json      nth_prime(n) at small, power-of-ten and seeded n
gourdon_terms.json  the other Gourdon terms (A + C, B, Phi0, Sigma) and D at
                    default parameters, pi(x) assembled from them and (to
                    1e16) printed by the reference's pi(x)

Usage:  python tests/golden/gen_fixtures.py [--big] [--only NAME]
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
CLI = os.path.join(REPO, "oracle", "_ref", "primecount")
M64 = (1 << 64) - 1


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def batch_xs(seed, n=256, lo_exp=12.0, span=6.0):
    """x_i = floor(10^(lo_exp + span * u_i)) with u_i in [0, 1) from SplitMix64."""
    xs, st = [], seed
    for _ in range(n):
        st, r = splitmix64(st)
        u = (r >> 11) * (2.0 ** -53)
        xs.append(int(10.0 ** (lo_exp + span * u)))
    return xs


def run_cli(args, timeout=3600):
    out = subprocess.run([CLI] + [str(a) for a in args], capture_output=True,
                         text=True, timeout=timeout, check=True).stdout
    vals = {}
    for line in out.splitlines():
        line = line.replace("\x1b[K", "").split("\r")[-1]
        m = re.match(r"^\s*([A-Za-z0-9_]+)\s*=\s*(-?[0-9.]+)\s*$", line)
        if m:
            vals[m.group(1)] = m.group(2)
        m = re.match(r"^Seconds:\s*([0-9.]+)", line)
        if m:
            vals["seconds"] = m.group(1)
    return vals


def parse_kat(path, nfields):
    """Extract the `{ a, b, c, d, e }` integer rows of a reference test table."""
    rows = []
    text = open(path).read()
    for m in re.finditer(r"\{\s*(-?\d+)(?:LL)?" + r"\s*,\s*(-?\d+)(?:LL)?" * (nfields - 1) + r"\s*\}", text):
        rows.append([int(g) for g in m.groups()])
    return rows


def gen_kat():
    d_rows = parse_kat(os.path.join(REF, "test/gourdon/D.cpp"), 5)
    assert len(d_rows) == 51, len(d_rows)
    s_rows = parse_kat(os.path.join(REF, "test/deleglise-rivat/S2_hard_alpha.cpp"), 5)
    assert len(s_rows) == 35, len(s_rows)
    json.dump({"source": "test/gourdon/D.cpp:36-89",
               "fields": ["x", "y", "z", "k", "D"],
               "rows": [[str(v) for v in r] for r in d_rows]},
              open(os.path.join(HERE, "kat_D.json"), "w"))
    json.dump({"source": "test/deleglise-rivat/S2_hard_alpha.cpp:36-73",
               "fields": ["x", "y", "z", "c", "S2_hard"],
               "rows": [[str(v) for v in r] for r in s_rows]},
              open(os.path.join(HERE, "kat_S2_hard.json"), "w"))


def gourdon_params(x):
    v = run_cli([x, "--Phi0", "-s"])
    return {"x": str(x), "y": v["y"], "z": v["z"], "k": v["k"], "x_star": v["x_star"]}


def dr_params(x):
    v = run_cli([x, "--S2-trivial", "-s"])
    return {"x": str(x), "y": v["y"], "z": v["z"], "c": v["c"]}


def gen_params():
    xs = set()
    for e in range(1, 23):
        xs.add(10 ** e)
    for e in range(2, 23):
        xs.add(10 ** e - 1)
        xs.add(10 ** e + 1)
    xs.update([2 ** 63 - 1, 2 ** 63, 2 ** 64 - 1, 2 ** 64, 2 ** 64 + 1, 2, 3, 5, 7, 100000,
               130321, 130320, 279841, 279840])
    st = 0xC0FFEE
    for _ in range(300):
        st, r = splitmix64(st)
        u = (r >> 11) * (2.0 ** -53)
        xs.add(int(10.0 ** (1.0 + 21.0 * u)))
    xs = sorted(xs)
    g = [gourdon_params(x) for x in xs]
    d = [dr_params(x) for x in xs]
    json.dump({"source": "oracle/_ref primecount --Phi0 -s / --S2-trivial -s",
               "gourdon": g, "deleglise_rivat": d},
              open(os.path.join(HERE, "params.json"), "w"), indent=0)


def gen_configs(big):
    out = {"source": "oracle/_ref primecount X --D -s / --S2-hard -s (8 host threads)",
           "D": [], "S2_hard": []}
    exps = [13, 15, 17, 18, 19] + ([21] if big else [])
    for e in exps:
        x = 10 ** e
        v = run_cli([x, "--D", "-s"], timeout=7200)
        out["D"].append({"x": str(x), "y": v["y"], "z": v["z"], "k": v["k"],
                         "x_star": v["x_star"], "value": v["D"],
                         "ref_seconds_8t": v.get("seconds")})
        print("D", e, v["D"], v.get("seconds"), flush=True)
    for e in [13, 15, 17, 19] + ([21] if big else []):
        x = 10 ** e
        v = run_cli([x, "--S2-hard", "-s"], timeout=7200)
        out["S2_hard"].append({"x": str(x), "y": v["y"], "z": v["z"], "c": v["c"],
                               "value": v["S2_hard"], "ref_seconds_8t": v.get("seconds")})
        print("S2_hard", e, v["S2_hard"], v.get("seconds"), flush=True)
    # full pi(x) for the README powers of ten we can afford here
    out["pi"] = [{"x": str(10 ** e),
                  "value": subprocess.run([CLI, str(10 ** e)], capture_output=True, text=True,
                                          check=True).stdout.strip()} for e in range(1, 16)]
    json.dump(out, open(os.path.join(HERE, "configs.json"), "w"), indent=1)


def gen_big():
    """x = 1e21 (x > 2^64: 128-bit path), D and S2_hard; tens of minutes."""
    out = {"source": "oracle/_ref primecount X --D -s / --S2-hard -s (8 host threads)",
           "D": [], "S2_hard": []}
    x = 10 ** 21
    v = run_cli([x, "--D", "-s"], timeout=7200)
    out["D"].append({"x": str(x), "y": v["y"], "z": v["z"], "k": v["k"], "x_star": v["x_star"],
                     "value": v["D"], "ref_seconds_8t": v.get("seconds")})
    v = run_cli([x, "--S2-hard", "-s"], timeout=7200)
    out["S2_hard"].append({"x": str(x), "y": v["y"], "z": v["z"], "c": v["c"],
                           "value": v["S2_hard"], "ref_seconds_8t": v.get("seconds")})
    json.dump(out, open(os.path.join(HERE, "configs_big.json"), "w"), indent=1)


def gen_batch():
    seed = 0x5EED2026
    xs = batch_xs(seed)
    rows = []
    for x in xs:
        v = run_cli([x, "--D", "-s"])
        rows.append({"x": str(x), "y": v["y"], "z": v["z"], "k": v["k"],
                     "x_star": v["x_star"], "D": v["D"]})
    json.dump({"source": "oracle/_ref primecount X --D -s",
               "seed": seed, "rule": "x_i = floor(10^(12 + 6 u_i)), u_i = (splitmix64 >> 11) * 2^-53",
               "rows": rows}, open(os.path.join(HERE, "batch_D.json"), "w"), indent=0)


def gen_alpha():
    rows = []
    for e in [8, 9, 10, 11, 12, 13, 14]:
        x = 10 ** e
        for ay, az in [(1.0, 1.0), (1.0, 2.0), (2.0, 3.0), (3.0, 1.5), (1.5, 4.0)]:
            v = run_cli([x, "--D", "-s", "--alpha-y=%g" % ay, "--alpha-z=%g" % az])
            rows.append({"x": str(x), "y": v["y"], "z": v["z"], "k": v["k"], "D": v["D"]})
            v = run_cli([x, "--S2-hard", "-s", "--alpha=%g" % (ay * az)])
            rows.append({"x": str(x), "y": v["y"], "z": v["z"], "c": v["c"], "S2_hard": v["S2_hard"]})
    json.dump({"source": "oracle/_ref primecount X --D/--S2-hard with --alpha-y/--alpha-z/--alpha",
               "rows": rows}, open(os.path.join(HERE, "alpha_D.json"), "w"), indent=0)


def gourdon_terms(x, extra=()):
    """The reference's Gourdon terms and parameters for x (`--AC/--B/--Phi0/
    --Sigma/--D -s`, each printing the parameters it used)."""
    row = {"x": str(x)}
    for opt, key in (("--AC", "A + C"), ("--B", "B"), ("--Phi0", "Phi0"),
                     ("--Sigma", "Sigma"), ("--D", "D")):
        out = subprocess.run([CLI, str(x), opt, "-s"] + list(extra), capture_output=True,
                             text=True, timeout=3600, check=True).stdout
        for line in out.splitlines():
            line = line.replace("\x1b[K", "").split("\r")[-1]
            m = re.match(r"^\s*([A-Za-z_0-9 +]+?)\s*=\s*(-?[0-9.]+)\s*$", line)
            if m and m.group(1) in (key, "y", "z", "k", "x_star"):
                row[{"A + C": "AC"}.get(m.group(1), m.group(1))] = m.group(2)
    return row


def gen_gourdon():
    """pi(x) = A - B + C + D + Phi0 + Sigma (src/gourdon/pi_gourdon.cpp:82-87)
    term by term, at powers of ten, at seeded x and at alpha_z = 2."""
    xs = [10 ** e for e in range(10, 20)] + [3 * 10 ** e for e in range(10, 19, 2)]
    xs += batch_xs(0x6A0DC0DE, n=16, lo_exp=10.0, span=8.0)
    rows = []
    for x in xs:
        r = gourdon_terms(x)
        r["pi"] = str(int(r["AC"]) - int(r["B"]) + int(r["D"]) + int(r["Phi0"]) + int(r["Sigma"]))
        if x <= 10 ** 16:
            r["pi_cli"] = run_cli([x, "-t", "8"]).get("pi", None) or subprocess.run(
                [CLI, str(x), "-t", "8"], capture_output=True, text=True, check=True).stdout.strip()
        rows.append(r)
        print(x, r, flush=True)
    alpha = []
    for x in (10 ** 12, 10 ** 14, 7 * 10 ** 15, 10 ** 17):
        r = gourdon_terms(x, ("--alpha-z=2",))
        alpha.append(r)
        print(x, r, flush=True)
    json.dump({"source": "oracle/_ref primecount X --AC/--B/--Phi0/--Sigma/--D -s",
               "rule": "pi = AC - B + D + Phi0 + Sigma; pi_cli = primecount X (x <= 1e16)",
               "rows": rows, "alpha_z2": alpha},
              open(os.path.join(HERE, "gourdon_terms.json"), "w"), indent=1)


def gen_nth():
    """nth_prime(n) printed by the reference (src/nth_prime.cpp): the tiny
    and small-table ranges, powers of ten and seeded n = 10^(8 + 9 u)."""
    ns = list(range(1, 200, 7)) + [169, 170, 3314, 3315, 10 ** 4, 10 ** 5, 99999, 10 ** 6]
    ns += [10 ** e for e in range(7, 17)] + [3 * 10 ** e for e in range(7, 16, 2)]
    ns += [int(v) for v in batch_xs(0x17A5E, n=24, lo_exp=8.0, span=9.0)]
    rows = []
    for n in ns:
        out = subprocess.run([CLI, str(n), "--nth-prime", "-t", "8"], capture_output=True,
                             text=True, timeout=3600, check=True).stdout.strip()
        rows.append({"n": str(n), "p": out.splitlines()[-1].strip()})
        print(n, rows[-1]["p"], flush=True)
    json.dump({"source": "oracle/_ref primecount N --nth-prime", "rows": rows},
              open(os.path.join(HERE, "nth_prime.json"), "w"), indent=1)


def gen_dr():
    """The Deleglise-Rivat terms printed by the reference build
    (src/deleglise-rivat/pi_deleglise_rivat.cpp: pi(x) = S1 + S2_trivial +
    S2_easy + S2_hard + pi(y) - 1 - P2) at 10^10 .. 10^19 and 8 seeded x."""
    xs = [10 ** e for e in range(10, 20)] + batch_xs(0xD12, n=8, lo_exp=10.0, span=6.0)
    rows = []
    for x in xs:
        r = {"x": str(x)}
        for opt, key in (("--S1", "S1"), ("--S2-trivial", "S2_trivial"), ("--S2-easy", "S2_easy"),
                         ("--S2-hard", "S2_hard"), ("--P2", "P2")):
            v = run_cli([x, opt, "-s", "-t", "8"])
            r[key] = v[key]
            for k in ("y", "z", "c"):
                if k in v:
                    r[k] = v[k]
        out = subprocess.run([CLI, str(x), "--deleglise-rivat", "-t", "8"], capture_output=True,
                             text=True, timeout=3600, check=True).stdout.strip()
        r["pi"] = out.splitlines()[-1].strip()
        rows.append(r)
        print(r, flush=True)
    json.dump({"source": "oracle/_ref primecount X --S1/--S2-trivial/--S2-easy/--S2-hard/--P2 -s "
                         "and --deleglise-rivat", "rows": rows},
              open(os.path.join(HERE, "dr_terms.json"), "w"), indent=1)


SERVE_SEED = 0xC5C5


def serve_queries(seed=SERVE_SEED, n=4096):
    """The serving mix: per query two SplitMix64 draws u1, u2 in [0, 1);
    u1 < 0.8 gives pi(x) with x = floor(10^(9 + 10 u2)), otherwise
    nth_prime(n) with n = floor(10^(8 + 9 u2))."""
    qs, st = [], seed
    for _ in range(n):
        st, r1 = splitmix64(st)
        st, r2 = splitmix64(st)
        u1 = (r1 >> 11) * (2.0 ** -53)
        u2 = (r2 >> 11) * (2.0 ** -53)
        if u1 < 0.8:
            qs.append(("pi", int(10.0 ** (9.0 + 10.0 * u2))))
        else:
            qs.append(("nth", int(10.0 ** (8.0 + 9.0 * u2))))
    return qs


def gen_serve(threads=7):
    """Answers of the 4096-query serving mix from the reference CLI
    (`primecount X` and `primecount N --nth-prime`).  Progress goes to a
    JSON-lines file in the system's temporary directory so that an
    interrupted run resumes."""
    qs = serve_queries()
    part = os.path.join(tempfile.gettempdir(), "primecount_serve_c5.partial")
    os.makedirs(os.path.dirname(part), exist_ok=True)
    done = {}
    if os.path.exists(part):
        for line in open(part):
            r = json.loads(line)
            done[(r["kind"], r["arg"])] = r["value"]
    with open(part, "a") as f:
        for i, (kind, arg) in enumerate(qs):
            if (kind, str(arg)) in done:
                continue
            args = [CLI, str(arg)] + (["--nth-prime"] if kind == "nth" else []) + ["-t", str(threads)]
            out = subprocess.run(args, capture_output=True, text=True, timeout=3600,
                                 check=True).stdout.strip()
            done[(kind, str(arg))] = out.splitlines()[-1].strip()
            f.write(json.dumps({"kind": kind, "arg": str(arg), "value": done[(kind, str(arg))]}) + "\n")
            f.flush()
            if i % 64 == 0:
                print(i, kind, arg, done[(kind, str(arg))], flush=True)
    rows = [{"kind": k, "arg": str(a), "value": done[(k, str(a))]} for k, a in qs]
    json.dump({"source": "oracle/_ref primecount X / primecount N --nth-prime",
               "seed": SERVE_SEED, "rows": rows},
              open(os.path.join(HERE, "serve_c5.json"), "w"), indent=0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true", help="also run x = 1e21 in `configs`")
    ap.add_argument("--only", default=None)
    a = ap.parse_args()
    if not os.path.exists(CLI):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    steps = {"kat": gen_kat, "params": gen_params, "configs": lambda: gen_configs(a.big),
             "batch": gen_batch, "alpha": gen_alpha, "big": gen_big, "gourdon": gen_gourdon,
             "nth": gen_nth, "serve": gen_serve, "dr": gen_dr}
    for name, fn in steps.items():
        if (a.only and a.only != name) or (not a.only and name in ("big", "serve")):
            continue
        print("==", name, flush=True)
        fn()


if __name__ == "__main__":
    main()
