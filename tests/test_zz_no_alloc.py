"""The no-allocation entry point pcg_D_i64_noalloc, called through the C ABI:
its sums equal pcg_D_i64's (and the reference's) on two calls in a row, the
second with a smaller input, and on a few more; it makes no device or pinned
host allocation (the library's allocation count does not move); a query too
large for its fixed scratch array fails with PCG_E_INVALID and leaves the
entry point usable."""
import ctypes

import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def call_noalloc(lib, x, y, z, k):
    out = ctypes.c_int64()
    rc = lib.pcg_D_i64_noalloc(x, y, z, k, ctypes.byref(out))
    return rc, out.value


def call_plain(lib, x, y, z, k):
    out = ctypes.c_int64()
    rc = lib.pcg_D_i64(x, y, z, k, ctypes.byref(out))
    return rc, out.value


def test_noalloc_matches_plain_entry_point(pc):
    lib = pc.load()
    rows = {int(v["x"]): v for v in golden("configs.json")["D"]}
    big = rows[10 ** 15]
    small = rows[10 ** 13]
    q_big = (10 ** 15, int(big["y"]), int(big["z"]), int(big["k"]))
    q_small = (10 ** 13, int(small["y"]), int(small["z"]), int(small["k"]))
    log("plain D(1e15), D(1e13)")
    rc, want_big = call_plain(lib, *q_big)
    assert rc == 0 and want_big == int(big["value"])
    rc, want_small = call_plain(lib, *q_small)
    assert rc == 0 and want_small == int(small["value"])
    allocs = pc.last_stats()["device_allocations"]
    # two calls in a row, the second with a smaller input
    log("no-allocation D(1e15)")
    rc, got = call_noalloc(lib, *q_big)
    assert rc == 0, pc.load().pcg_last_error()
    assert got == want_big
    assert pc.last_stats()["device_allocations"] == allocs
    log("no-allocation D(1e13)")
    rc, got = call_noalloc(lib, *q_small)
    assert rc == 0 and got == want_small
    assert pc.last_stats()["device_allocations"] == allocs
    # the reference's known-answer rows (tables rebuilt in the array when a
    # row needs other ones), then the larger query again
    kat = [[int(v) for v in r] for r in golden("kat_D.json")["rows"]][::4]
    log("no-allocation: %d known-answer rows" % len(kat))
    for x, y, z, k, want in kat:
        rc, got = call_noalloc(lib, x, y, z, k)
        assert rc == 0 and got == want, (x, y, z, k)
    rc, got = call_noalloc(lib, *q_big)
    assert rc == 0 and got == want_big
    assert pc.last_stats()["device_allocations"] == allocs


def test_noalloc_rejects_queries_beyond_the_scratch(pc):
    lib = pc.load()
    r = {int(v["x"]): v for v in golden("configs.json")["D"]}[10 ** 19 // 10]
    x, y, z, k = 10 ** 18, int(r["y"]), int(r["z"]), int(r["k"])
    log("no-allocation D(1e18): too large")
    rc, _ = call_noalloc(lib, x, y, z, k)
    assert rc == pc.E_INVALID
    assert b"no-allocation" in lib.pcg_last_error()
    # still usable, and the plain entry point answers the large query
    rc, got = call_noalloc(lib, 10 ** 6, 207, 207, 8)
    assert rc == 0 and got == 2465
