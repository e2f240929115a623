"""The schedule autotuner on the device: tuned and plain schedules give the
reference's values (D and S2_hard at x = 1e21 through the 128-bit path, and
a 256-query batch with shared segments), the winner is cached per device
and launch class and read back from the cache file."""
import sys

import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tune_cycle(pc, f, x, want, cache):
    pc.set_schedule_cache(str(cache))
    pc.set_autotune(pc.TUNE_OFF)
    assert f(x) == want
    plain = pc.last_stats()
    assert plain["schedule_source"] == 0
    pc.set_autotune(pc.TUNE_ON)
    assert f(x) == want                        # tunes: every trial checked equal
    assert pc.last_stats()["schedule_source"] == 2
    assert f(x) == want                        # the cached winner
    tuned = pc.last_stats()
    assert tuned["schedule_source"] == 1
    return plain, tuned


@pytest.mark.parametrize("term", ["D", "S2_hard"])
def test_autotune_1e21(pc, tmp_path, term):
    r = golden("configs_big.json")[term][0]
    x, want = int(r["x"]), int(r["value"])
    f = pc.D_default if term == "D" else pc.S2_hard_default
    cache = tmp_path / "schedules.txt"
    try:
        log("autotune %s(%d): plain, tune, cached" % (term, x))
        plain, tuned = _tune_cycle(pc, f, x, want, cache)
        log("plain W=%d %.3f s, tuned W=%d grid=%d %.3f s" % (
            plain["segment_words"], plain["seconds_kernel"], tuned["segment_words"],
            tuned["grid_waves"], tuned["seconds_kernel"]))
        # a fresh read of the file gives the same winner
        assert pc.set_schedule_cache(str(cache)) == 1
        pc.set_autotune(pc.TUNE_CACHED)
        assert f(x) == want
        again = pc.last_stats()
        assert again["schedule_source"] == 1
        assert again["segment_words"] == tuned["segment_words"]
        assert again["grid_waves"] == tuned["grid_waves"]
    finally:
        pc.set_autotune(pc.TUNE_CACHED)
        pc.set_schedule_cache(None)


def test_autotune_batch(pc, tmp_path):
    # a 256-query launch: the group size of shared segments is tuned too
    fx = golden("batch_D.json")["rows"]
    xs = [int(r["x"]) for r in fx]
    want = [int(r["D"]) for r in fx]
    try:
        log("autotune D batch of %d" % len(xs))
        plain, tuned = _tune_cycle(pc, pc.D_batch, xs, want, tmp_path / "schedules.txt")
        assert tuned["group_size"] in (2, 4)
    finally:
        pc.set_autotune(pc.TUNE_CACHED)
        pc.set_schedule_cache(None)
