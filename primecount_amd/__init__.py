"""primecount_amd -- MI355X engine for primecount's hard special leaves.

Python mirror of the reference's interface for this path:

    D(x, y, z, k)          primecount::D            include/gourdon.hpp:28,38
    S2_hard(x, y, z, c)    primecount::S2_hard      include/S.hpp:25,32
    D_default(x)           `primecount x --D`       src/app/main.cpp:205-246
    S2_hard_default(x)     `primecount x --S2-hard` src/app/main.cpp:435-461
    gourdon_params(x)      y, z, k, x_star          src/util.cpp:276-386
    dr_params(x)           y, z, c                  src/util.cpp:222-263

Integers of any size up to 127 bits are accepted; results are Python ints.
The 64-bit overloads are used when x < 2^63 (as the reference's CLI does,
src/app/main.cpp:242-245); D128/S2_hard128 force the 128-bit entry points.
Every call goes through the C ABI of libprimecount_mi355x.so (built by
primecount_amd.build); there is no CPU fallback: without the library, or
without a gfx950 device, calls raise PrimecountError.
"""
import ctypes
import os

__all__ = [
    "PrimecountError", "D", "D128", "S2_hard", "S2_hard128", "D_default",
    "S2_hard_default", "D_batch", "gourdon_params", "dr_params", "set_alpha_y",
    "set_alpha_z", "set_alpha", "set_segment_words", "set_host_threads",
    "last_stats", "device_available", "release", "library_path", "load",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libprimecount_mi355x.so")
_lib = None


class PrimecountError(RuntimeError):
    """Raised when the engine reports an error (mirrors primecount_error).
    `status` is the negative PCG_E_* code of the failed call (E_INVALID,
    E_DEVICE, E_NOMEM or E_INTERNAL), or None for errors raised here."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


E_INVALID, E_DEVICE, E_NOMEM, E_INTERNAL = -1, -2, -3, -4


class _I128(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("seconds_total", ctypes.c_double),
        ("seconds_tables", ctypes.c_double),
        ("seconds_plan", ctypes.c_double),
        ("seconds_kernel", ctypes.c_double),
        ("seconds_combine", ctypes.c_double),
        ("segments", ctypes.c_uint64),
        ("chunks", ctypes.c_uint64),
        ("segment_words", ctypes.c_uint64),
        ("sieve_limit", ctypes.c_uint64),
        ("b_first", ctypes.c_uint64),
        ("b_last", ctypes.c_uint64),
        ("leaves", ctypes.c_uint64),
        ("hits", ctypes.c_uint64),
        ("factor_entries", ctypes.c_uint64),
        ("pairs", ctypes.c_uint64),
        ("windows", ctypes.c_uint64),
        ("leaf_primes", ctypes.c_uint64),
        ("leaf_steps", ctypes.c_uint64),
        ("dense_primes", ctypes.c_uint64),
        ("dense_steps", ctypes.c_uint64),
        ("group_steps", ctypes.c_uint64),
        ("prefix_builds", ctypes.c_uint64),
        ("queue_steps", ctypes.c_uint64),
        ("device_allocations", ctypes.c_uint64),
        ("escaped_slots", ctypes.c_uint64),
        ("escape_bound", ctypes.c_uint64),
        ("schedule_source", ctypes.c_uint64),
        ("group_size", ctypes.c_uint64),
        ("grid_waves", ctypes.c_uint64),
        ("kernel_waves", ctypes.c_uint64),
        ("leaf_batches", ctypes.c_uint64),
        ("batch_records", ctypes.c_uint64),
        ("phase_cycles", ctypes.c_uint64 * 11),
    ]

    PHASES = ("segment", "window", "dense_crossoff", "group_crossoff", "batch_crossoff",
              "prefix_build", "prescan", "entry_steps", "counting_steps", "batch_entry_steps",
              "batch_counting_steps")

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["phase_cycles"] = dict(zip(self.PHASES, list(self.phase_cycles)))
        return d


class _Query(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("x", _I128),
                ("y", ctypes.c_int64), ("z", ctypes.c_int64), ("k", ctypes.c_int64)]


class _Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("value", _I128), ("seconds", ctypes.c_double)]


class _CtxStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("max_batch", ctypes.c_uint64),
                ("busy_seconds", ctypes.c_double), ("light_batches", ctypes.c_uint64),
                ("heavy_batches", ctypes.c_uint64), ("interleaved", ctypes.c_uint64),
                ("rejected_nomem", ctypes.c_uint64)]


class PiStats(ctypes.Structure):
    _fields_ = [("seconds_total", ctypes.c_double), ("seconds_table", ctypes.c_double),
                ("seconds_sigma", ctypes.c_double), ("seconds_phi0", ctypes.c_double),
                ("seconds_ac", ctypes.c_double), ("seconds_b", ctypes.c_double),
                ("seconds_d", ctypes.c_double), ("ac_leaves", ctypes.c_uint64),
                ("table_limit", ctypes.c_uint64), ("table_windows", ctypes.c_uint64)]


KIND_D = 0
KIND_S2_HARD = 1
KIND_PI = 2
KIND_NTH_PRIME = 3


def library_path():
    return _LIB_PATH


def load():
    """Load libprimecount_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise PrimecountError(
            "libprimecount_mi355x.so is not built; run `python -m primecount_amd.build`")
    lib = ctypes.CDLL(_LIB_PATH)
    I64, P64 = ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)
    PI128 = ctypes.POINTER(_I128)
    sig = {
        "pcg_last_error": ([], ctypes.c_char_p),
        "pcg_version": ([], ctypes.c_char_p),
        "pcg_device_available": ([], ctypes.c_int),
        "pcg_D_i64": ([I64, I64, I64, I64, P64], ctypes.c_int),
        "pcg_D_i128": ([_I128, I64, I64, I64, PI128], ctypes.c_int),
        "pcg_D_i64_noalloc": ([I64, I64, I64, I64, P64], ctypes.c_int),
        "pcg_S2_hard_i64": ([I64, I64, I64, I64, P64], ctypes.c_int),
        "pcg_S2_hard_i128": ([_I128, I64, I64, I64, PI128], ctypes.c_int),
        "pcg_D": ([_I128, PI128], ctypes.c_int),
        "pcg_S2_hard": ([_I128, PI128], ctypes.c_int),
        "pcg_gourdon_params": ([_I128, P64, P64, P64, P64], ctypes.c_int),
        "pcg_dr_params": ([_I128, P64, P64, P64], ctypes.c_int),
        "pcg_set_alpha_y": ([ctypes.c_double], ctypes.c_int),
        "pcg_set_alpha_z": ([ctypes.c_double], ctypes.c_int),
        "pcg_set_alpha": ([ctypes.c_double], ctypes.c_int),
        "pcg_D_batch": ([PI128, ctypes.c_size_t, PI128], ctypes.c_int),
        "pcg_set_segment_words": ([ctypes.c_uint32], ctypes.c_int),
        "pcg_set_host_threads": ([ctypes.c_int], ctypes.c_int),
        "pcg_set_pass_slots": ([ctypes.c_uint64], ctypes.c_int),
        "pcg_set_pi_table_bytes": ([ctypes.c_uint64], ctypes.c_int),
        "pcg_set_shared_segments": ([ctypes.c_int], ctypes.c_int),
        "pcg_set_leaf_batching": ([ctypes.c_int], ctypes.c_int),
        "pcg_set_phase_timers": ([ctypes.c_int], ctypes.c_int),
        "pcg_set_slot_bits": ([ctypes.c_uint32], ctypes.c_int),
        "pcg_set_autotune": ([ctypes.c_int], ctypes.c_int),
        "pcg_set_schedule": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32],
                             ctypes.c_int),
        "pcg_set_schedule_cache": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "pcg_factor_table_check": ([ctypes.c_int, I64, I64, ctypes.POINTER(ctypes.c_uint64)],
                                   ctypes.c_int),
        "pcg_last_stats": ([ctypes.POINTER(Stats)], ctypes.c_int),
        "pcg_release": ([], ctypes.c_int),
        "pcg_pi": ([_I128, PI128], ctypes.c_int),
        "pcg_last_pi_stats": ([ctypes.POINTER(PiStats)], ctypes.c_int),
        "pcg_check_pi": ([_I128, _I128], ctypes.c_int),
        "pcg_lds_read_gbs": ([ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "pcg_nth_prime": ([I64, P64], ctypes.c_int),
        "pcg_nth_prime_128": ([_I128, PI128], ctypes.c_int),
        "pcg_AC": ([_I128, I64, I64, I64, PI128], ctypes.c_int),
        "pcg_B": ([_I128, I64, PI128], ctypes.c_int),
        "pcg_Phi0": ([_I128, I64, I64, I64, PI128], ctypes.c_int),
        "pcg_Sigma": ([_I128, I64, PI128], ctypes.c_int),
        "pcg_pi_deleglise_rivat": ([_I128, PI128], ctypes.c_int),
        "pcg_S1": ([_I128, I64, I64, PI128], ctypes.c_int),
        "pcg_S2_trivial": ([_I128, I64, I64, I64, PI128], ctypes.c_int),
        "pcg_S2_easy": ([_I128, I64, I64, I64, PI128], ctypes.c_int),
        "pcg_P2": ([_I128, I64, PI128], ctypes.c_int),
        "pcg_ctx_create": ([I64, I64, ctypes.c_uint64, ctypes.c_uint32,
                            ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "pcg_submit": ([ctypes.c_void_p, ctypes.POINTER(_Query), ctypes.c_size_t,
                        ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "pcg_poll": ([ctypes.c_void_p, ctypes.POINTER(_Result), ctypes.c_size_t, ctypes.c_int,
                      ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "pcg_ctx_get_stats": ([ctypes.c_void_p, ctypes.POINTER(_CtxStats)], ctypes.c_int),
        "pcg_ctx_destroy": ([ctypes.c_void_p], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _i128(v):
    v = int(v)
    if not -(1 << 127) <= v < (1 << 127):
        raise PrimecountError("value does not fit in 128 bits")
    u = v & ((1 << 128) - 1)
    return _I128(u & ((1 << 64) - 1), ctypes.c_int64(u >> 64).value)


def _from128(r):
    v = (r.hi << 64) | r.lo
    return v


def _check(rc):
    if rc != 0:
        msg = load().pcg_last_error().decode(errors="replace")
        raise PrimecountError(msg or "error %d" % rc, rc)


def _fits64(x):
    return -(1 << 63) <= x < (1 << 63)


def D(x, y, z, k, threads=None):
    """Gourdon's D(x, y, z, k) (threads is accepted for signature parity)."""
    x = int(x)
    if _fits64(x):
        out = ctypes.c_int64()
        _check(load().pcg_D_i64(x, int(y), int(z), int(k), ctypes.byref(out)))
        return out.value
    return D128(x, y, z, k)


def D_noalloc(x, y, z, k):
    """D(x, y, z, k) for x < 2^63 through the no-allocation entry point
    (pcg_D_i64_noalloc): all device memory from the library's fixed scratch
    array; queries too large for it raise PrimecountError (E_INVALID)."""
    out = ctypes.c_int64()
    _check(load().pcg_D_i64_noalloc(int(x), int(y), int(z), int(k), ctypes.byref(out)))
    return out.value


def D128(x, y, z, k, threads=None):
    out = _I128()
    _check(load().pcg_D_i128(_i128(x), int(y), int(z), int(k), ctypes.byref(out)))
    return _from128(out)


def S2_hard(x, y, z, c, threads=None):
    """Deleglise-Rivat S2_hard(x, y, z, c)."""
    x = int(x)
    if _fits64(x):
        out = ctypes.c_int64()
        _check(load().pcg_S2_hard_i64(x, int(y), int(z), int(c), ctypes.byref(out)))
        return out.value
    return S2_hard128(x, y, z, c)


def S2_hard128(x, y, z, c, threads=None):
    out = _I128()
    _check(load().pcg_S2_hard_i128(_i128(x), int(y), int(z), int(c), ctypes.byref(out)))
    return _from128(out)


def D_default(x):
    out = _I128()
    _check(load().pcg_D(_i128(x), ctypes.byref(out)))
    return _from128(out)


def S2_hard_default(x):
    out = _I128()
    _check(load().pcg_S2_hard(_i128(x), ctypes.byref(out)))
    return _from128(out)


def D_batch(xs):
    """D(x) with default parameters for every x, in one batched launch."""
    xs = [int(v) for v in xs]
    n = len(xs)
    arr = (_I128 * max(n, 1))(*[_i128(v) for v in xs])
    out = (_I128 * max(n, 1))()
    _check(load().pcg_D_batch(arr, n, out))
    return [_from128(out[i]) for i in range(n)]


def pi(x):
    """pi(x): the number of primes <= x (Gourdon's algorithm on the device)."""
    out = _I128()
    _check(load().pcg_pi(_i128(x), ctypes.byref(out)))
    return _from128(out)


def lds_read_gbs():
    """LDS read bandwidth of the device in GB/s (probe kernel)."""
    v = ctypes.c_double()
    _check(load().pcg_lds_read_gbs(ctypes.byref(v)))
    return v.value


def check_pi(x, pix):
    """The reference's result check (verify_pix, src/util.cpp:397-420): raises
    PrimecountError (status E_INTERNAL) when |pix - Li(x)| >= sqrt(x) log(x) /
    (8 pi).  Host arithmetic only."""
    _check(load().pcg_check_pi(_i128(x), _i128(pix)))


def nth_prime(n):
    """The n-th prime, on the device (n beyond 216289611853439384 through the
    128-bit entry point, answers above 2^63)."""
    if int(n) > 216289611853439384:
        return nth_prime128(n)
    out = ctypes.c_int64()
    _check(load().pcg_nth_prime(int(n), ctypes.byref(out)))
    return out.value


def nth_prime128(n):
    """The n-th prime through the 128-bit entry point (any n >= 1)."""
    out = _I128()
    _check(load().pcg_nth_prime_128(_i128(n), ctypes.byref(out)))
    return _from128(out)


def last_pi_stats():
    st = PiStats()
    _check(load().pcg_last_pi_stats(ctypes.byref(st)))
    return {name: getattr(st, name) for name, _ in st._fields_}


def AC(x, y, z, k):
    out = _I128()
    _check(load().pcg_AC(_i128(x), int(y), int(z), int(k), ctypes.byref(out)))
    return _from128(out)


def B(x, y):
    out = _I128()
    _check(load().pcg_B(_i128(x), int(y), ctypes.byref(out)))
    return _from128(out)


def Phi0(x, y, z, k):
    out = _I128()
    _check(load().pcg_Phi0(_i128(x), int(y), int(z), int(k), ctypes.byref(out)))
    return _from128(out)


def Sigma(x, y):
    out = _I128()
    _check(load().pcg_Sigma(_i128(x), int(y), ctypes.byref(out)))
    return _from128(out)


def pi_deleglise_rivat(x):
    """pi(x) by the Deleglise-Rivat algorithm, every term on the device."""
    out = _I128()
    _check(load().pcg_pi_deleglise_rivat(_i128(x), ctypes.byref(out)))
    return _from128(out)


def S1(x, y, c):
    out = _I128()
    _check(load().pcg_S1(_i128(x), int(y), int(c), ctypes.byref(out)))
    return _from128(out)


def S2_trivial(x, y, z, c):
    out = _I128()
    _check(load().pcg_S2_trivial(_i128(x), int(y), int(z), int(c), ctypes.byref(out)))
    return _from128(out)


def S2_easy(x, y, z, c):
    out = _I128()
    _check(load().pcg_S2_easy(_i128(x), int(y), int(z), int(c), ctypes.byref(out)))
    return _from128(out)


def P2(x, y):
    out = _I128()
    _check(load().pcg_P2(_i128(x), int(y), ctypes.byref(out)))
    return _from128(out)


def gourdon_params(x):
    y, z, k, xs = (ctypes.c_int64() for _ in range(4))
    _check(load().pcg_gourdon_params(_i128(x), ctypes.byref(y), ctypes.byref(z),
                                     ctypes.byref(k), ctypes.byref(xs)))
    return {"y": y.value, "z": z.value, "k": k.value, "x_star": xs.value}


def dr_params(x):
    y, z, c = (ctypes.c_int64() for _ in range(3))
    _check(load().pcg_dr_params(_i128(x), ctypes.byref(y), ctypes.byref(z), ctypes.byref(c)))
    return {"y": y.value, "z": z.value, "c": c.value}


def set_alpha_y(a):
    _check(load().pcg_set_alpha_y(float(a)))


def set_alpha_z(a):
    _check(load().pcg_set_alpha_z(float(a)))


def set_alpha(a):
    _check(load().pcg_set_alpha(float(a)))


def set_segment_words(w):
    _check(load().pcg_set_segment_words(int(w)))


def factor_table_check(kind, y, limit):
    """Entries in which the device-built factor table differs from the host's."""
    out = ctypes.c_uint64()
    _check(load().pcg_factor_table_check(int(kind), int(y), int(limit), ctypes.byref(out)))
    return out.value


def set_shared_segments(on):
    _check(load().pcg_set_shared_segments(1 if on else 0))


def set_leaf_batching(on):
    """Count the leaves of several primes on one prefix build (default on);
    results are the same either way."""
    _check(load().pcg_set_leaf_batching(1 if on else 0))


def set_phase_timers(on):
    """Single-query launches report the sieve kernel's cycles per phase in
    last_stats()["phase_cycles"] (a slower kernel build; same results)."""
    _check(load().pcg_set_phase_timers(1 if on else 0))


def set_slot_bits(bits):
    """Bits of C in a slot word (0 = automatic); more bits force escapes."""
    _check(load().pcg_set_slot_bits(int(bits)))


TUNE_OFF, TUNE_CACHED, TUNE_ON = 0, 1, 2


def set_schedule(words=0, group=0, grid_pct=0, waves=-1):
    """Pins the sieve launch schedule (0 / -1: left to the schedule); any
    knob set turns the autotuner off.  set_schedule() restores the default."""
    _check(load().pcg_set_schedule(int(words), int(group), int(grid_pct), int(waves)))


def set_autotune(mode):
    """Schedule autotuner: TUNE_OFF (the segment-size rule), TUNE_CACHED
    (default: a cached winner of the launch class, else the rule) or TUNE_ON
    (also tune classes not cached yet, on the launch itself)."""
    _check(load().pcg_set_autotune(int(mode)))


def set_schedule_cache(path=None):
    """Cache file of the autotuner (None: the default location); returns the
    number of winners it holds."""
    n = ctypes.c_uint64(0)
    _check(load().pcg_set_schedule_cache(None if path is None else os.fsencode(path),
                                         ctypes.byref(n)))
    return n.value


def set_pi_table_bytes(n):
    """Largest prime-count table pi(x) keeps whole (bytes, 0 = 65% of the
    device memory); a larger one is built and read in windows of this size
    (same results, bounded memory)."""
    _check(load().pcg_set_pi_table_bytes(ctypes.c_uint64(int(n))))


def set_pass_slots(n):
    """(segment, prime) slots per kernel pass; 0 restores the default."""
    _check(load().pcg_set_pass_slots(ctypes.c_uint64(int(n))))


def set_host_threads(t):
    _check(load().pcg_set_host_threads(int(t)))


def last_stats():
    st = Stats()
    _check(load().pcg_last_stats(ctypes.byref(st)))
    return st.as_dict()


def device_available():
    return bool(load().pcg_device_available())


def release():
    _check(load().pcg_release())


class Server:
    """Continuous-batching serving context (pcg_ctx_*): submit queries from any
    thread, poll finished results.  Light queries arriving while a heavy batch
    runs are answered at its next pass point; heavy ones join the next heavy
    batch.  hbm_budget bounds the device memory of the context's work
    (0 = 90% of the device); a query that cannot fit gets status E_NOMEM.

        with Server(z_max=10**8) as sv:
            ids = sv.submit([(KIND_D, 10**15), (KIND_D, 10**16)])
            results = sv.poll_all(len(ids))
    """

    def __init__(self, y_max=0, z_max=0, hbm_budget=0, max_batch=0):
        self._ctx = ctypes.c_void_p()
        _check(load().pcg_ctx_create(int(y_max), int(z_max), int(hbm_budget), int(max_batch),
                                     ctypes.byref(self._ctx)))

    def submit(self, queries):
        """queries: iterable of (kind, x) for default parameters or
        (kind, x, y, z, k).  Returns the ids in order."""
        qs = list(queries)
        n = len(qs)
        arr = (_Query * max(n, 1))()
        for i, q in enumerate(qs):
            arr[i].kind = int(q[0])
            arr[i].x = _i128(q[1])
            if len(q) > 2:
                arr[i].y, arr[i].z, arr[i].k = int(q[2]), int(q[3]), int(q[4])
        ids = (ctypes.c_uint64 * max(n, 1))()
        _check(load().pcg_submit(self._ctx, arr, n, ids))
        return [ids[i] for i in range(n)]

    def poll(self, max_results=1024, timeout_ms=-1):
        """Finished results: dicts with id, status (0 or a negative PCG_E_*
        code), value and seconds (submit to result)."""
        out = (_Result * max(max_results, 1))()
        n = ctypes.c_size_t()
        _check(load().pcg_poll(self._ctx, out, max_results, int(timeout_ms), ctypes.byref(n)))
        return [{"id": out[i].id, "status": out[i].status, "value": _from128(out[i].value),
                 "seconds": out[i].seconds} for i in range(n.value)]

    def poll_all(self, count, timeout_ms=-1):
        res = []
        while len(res) < count:
            got = self.poll(count - len(res), timeout_ms)
            if not got and timeout_ms >= 0:
                break
            res.extend(got)
        return res

    def stats(self):
        st = _CtxStats()
        _check(load().pcg_ctx_get_stats(self._ctx, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def close(self):
        if self._ctx:
            _check(load().pcg_ctx_destroy(self._ctx))
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
