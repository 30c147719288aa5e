"""ctypes binding of include/es_sieve.h (lib/libes_sieve.so), for tests and the bench.

Mirrors the reference's scanner interface at the process level (rung_scan3's flags,
records and exit codes) as Python calls. There is no CPU fallback: opening a scanner
without a HIP device raises EsError.
"""
import array
import itertools
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libes_sieve.so")
BIN_PATH = os.path.join(HERE, "bin", "es_sieve")
FILTERS = os.path.join(HERE, "tables", "filters")
F120 = os.path.join(FILTERS, "f_M120120_mmax2000_runtime.txt")
F204 = os.path.join(FILTERS, "f_M2042040_mmax2000_gpu.txt")
F204_MI355X = os.path.join(FILTERS, "f_M2042040_mmax2000_mi355x.txt")   # chosen by bin/es_prune
F204R = os.path.join(FILTERS, "f_M2042040_mmax2000_runtime.txt")

U64 = ctypes.c_uint64


class EsConfig(ctypes.Structure):
    _fields_ = [("wheel_M", U64), ("mmax", U64), ("sieve_P", U64), ("filter_file", ctypes.c_char_p),
                ("threads", ctypes.c_int), ("device", ctypes.c_int)]


class EsCounts(ctypes.Structure):
    _fields_ = [(n, U64) for n in ("positions", "covered", "sieved", "mr", "composite", "rung", "direct",
                                   "survivors")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class EsTiming(ctypes.Structure):
    _fields_ = [("plan_ms", ctypes.c_double), ("batch_ms", ctypes.c_double), ("device_ms", ctypes.c_double),
                ("sieve_ms", ctypes.c_double), ("tail_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("tiles", U64), ("survivors", U64), ("primes", U64), ("host_primes", U64),
                ("replays", ctypes.c_int), ("mr_ms", ctypes.c_double), ("rung_ms", ctypes.c_double),
                ("sieve_event_ms", ctypes.c_double), ("mr_event_ms", ctypes.c_double),
                ("rung_event_ms", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class EsServeConfig(ctypes.Structure):
    _fields_ = [("slice", U64), ("batch_ints", U64), ("batch_windows", ctypes.c_uint32),
                ("slots", ctypes.c_uint32), ("hbm_fraction", ctypes.c_double), ("hbm_budget", U64),
                ("emit_survivors", ctypes.c_int)]


class EsRequestInfo(ctypes.Structure):
    _fields_ = [("arrival_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("batches", U64), ("reserved", U64), ("exit_code", ctypes.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class EsServeStats(ctypes.Structure):
    _fields_ = [(n, U64) for n in ("submitted", "completed", "refused", "batches", "slices", "ints", "budget",
                                   "resident", "peak_reserved")] + \
               [(n, ctypes.c_uint32) for n in ("max_active", "max_batch_requests", "max_queued")] + \
               [("busy_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class EsError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads lib/libes_sieve.so once (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EsError("libes_sieve.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        sig = {
            "es_config_default": (None, [P(EsConfig)]),
            "es_scanner_open": (ctypes.c_int, [P(EsConfig), P(ctypes.c_void_p)]),
            "es_scanner_close": (None, [ctypes.c_void_p]),
            "es_last_error": (ctypes.c_char_p, []),
            "es_check_interval": (ctypes.c_int, [ctypes.c_void_p, U64, U64]),
            "es_scan_records": (ctypes.c_int, [ctypes.c_void_p, P(U64), P(U64), ctypes.c_size_t, ctypes.c_int,
                                               P(ctypes.c_void_p), P(ctypes.c_size_t), P(ctypes.c_size_t),
                                               P(ctypes.c_int)]),
            "es_scan_submit": (ctypes.c_int, [ctypes.c_void_p, P(U64), P(U64), ctypes.c_size_t]),
            "es_scan_collect": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_void_p), P(ctypes.c_size_t),
                                               P(ctypes.c_size_t), P(ctypes.c_int), P(ctypes.c_size_t)]),
            "es_scan_in_flight": (ctypes.c_int, [ctypes.c_void_p]),
            "es_scan_counts": (ctypes.c_int, [ctypes.c_void_p, P(U64), P(U64), ctypes.c_size_t, P(EsCounts)]),
            "es_survivors": (ctypes.c_int, [ctypes.c_void_p, U64, U64, P(U64), P(ctypes.c_uint8), ctypes.c_size_t,
                                            P(ctypes.c_size_t), P(EsCounts)]),
            "es_primality": (ctypes.c_int, [ctypes.c_void_p, P(U64), ctypes.c_size_t, P(ctypes.c_uint8)]),
            "es_primality_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                    ctypes.c_void_p, ctypes.c_void_p]),
            "es_certify": (ctypes.c_int, [U64, ctypes.c_char_p, ctypes.c_size_t]),
            "es_set_layout": (ctypes.c_int, [ctypes.c_void_p, U64]),
            "es_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]),
            "es_set_residency": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
            "es_set_grouping": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
            "es_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "es_set_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
            "es_set_device_tail": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
            "es_set_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, U64]),
            "es_certify_batch": (ctypes.c_int, [ctypes.c_void_p, P(U64), ctypes.c_size_t, P(ctypes.c_void_p),
                                                P(ctypes.c_size_t)]),
            "es_direct_batch": (ctypes.c_int, [ctypes.c_void_p, P(U64), ctypes.c_size_t, P(ctypes.c_void_p),
                                                P(ctypes.c_size_t)]),
            "es_choose_layout": (U64, [ctypes.c_void_p, U64]),
            "es_autotune": (ctypes.c_int, [ctypes.c_void_p, P(U64), P(U64), ctypes.c_size_t, P(ctypes.c_void_p)]),
            "es_last_timing": (ctypes.c_int, [ctypes.c_void_p, P(EsTiming)]),
            "es_wheel_info": (ctypes.c_int, [ctypes.c_void_p, P(U64), P(U64), P(U64), P(U64)]),
            "es_mark_count": (ctypes.c_int, [ctypes.c_void_p, U64, U64, U64, P(U64), P(U64), P(U64), P(U64)]),
            "es_lds_peak": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_double)]),
            "es_lds_scatter_peak": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_double)]),
            "es_wheel_classes": (ctypes.c_int, [ctypes.c_void_p, P(U64), ctypes.c_size_t, P(ctypes.c_size_t)]),
            "es_serve_config_default": (None, [P(EsServeConfig)]),
            "es_server_open": (ctypes.c_int, [ctypes.c_void_p, P(EsServeConfig), P(ctypes.c_void_p)]),
            "es_server_submit": (ctypes.c_int, [ctypes.c_void_p, U64, U64, P(U64)]),
            "es_server_wait": (ctypes.c_int, [ctypes.c_void_p, U64, ctypes.c_double, P(ctypes.c_void_p),
                                              P(ctypes.c_size_t), P(EsRequestInfo)]),
            "es_server_stats": (ctypes.c_int, [ctypes.c_void_p, P(EsServeStats)]),
            "es_server_close": (None, [ctypes.c_void_p]),
            "es_serve_trace": (ctypes.c_int, [ctypes.c_void_p, P(EsServeConfig), P(ctypes.c_double), P(U64),
                                              P(U64), ctypes.c_size_t, ctypes.c_int, P(ctypes.c_void_p),
                                              P(ctypes.c_size_t), P(EsRequestInfo), P(EsServeStats)]),
            "es_free": (None, [ctypes.c_void_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise EsError(lib().es_last_error().decode())
    return rc


def certify(p):
    """Stage D then E for a prime p on the host: (kind, line)."""
    buf = ctypes.create_string_buffer(256)
    kind = _check(lib().es_certify(p, buf, 256))
    return kind, buf.value.decode()


def _split_text(raw, lens):
    """Records as str, cut by their byte lengths (records are ASCII: one decode, then
    slices by the same offsets)."""
    txt = raw.decode("ascii")
    offs = list(itertools.accumulate(lens, initial=0))
    return [txt[a:b] for a, b in zip(offs, offs[1:])]


def _u64s(vals):
    # array('Q') packs a list of Python ints far faster than the ctypes array constructor.
    buf = array.array("Q", vals)
    return (U64 * len(vals)).from_buffer(buf) if vals else (U64 * 0)()


class Batch:
    """Windows [(lo, hi)] packed once into the u64 arrays the library reads, for a batch
    that is submitted again and again (records(), submit() also take plain lists)."""

    def __init__(self, windows):
        self.n = len(windows)
        self.lo = _u64s([w[0] for w in windows])
        self.hi = _u64s([w[1] for w in windows])


def _packed(windows):
    return windows if isinstance(windows, Batch) else Batch(windows)


MAX_IN_FLIGHT = 8   # ES_MAX_IN_FLIGHT of include/es_sieve.h


class Scanner:
    """One scanner configuration resident on a HIP device (rung_scan3's filter flags)."""

    def __init__(self, wheel=2042040, mmax=10000, sieve=0, filters=None, threads=0, device=0):
        L = lib()
        cfg = EsConfig()
        L.es_config_default(ctypes.byref(cfg))
        cfg.wheel_M, cfg.mmax, cfg.sieve_P = wheel, mmax, sieve
        self._ff = filters.encode() if filters else None
        cfg.filter_file = self._ff
        cfg.threads, cfg.device = threads, device
        h = ctypes.c_void_p()
        _check(L.es_scanner_open(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.wheel = wheel
        self._sizes = []          # window counts of the batches submitted and not collected

    def close(self):
        if self.h:
            lib().es_scanner_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check_interval(self, lo, hi):
        return lib().es_check_interval(self.h, lo, hi) == 0

    def records(self, windows, emit_survivors=True):
        """[(lo, hi)] (or a Batch) -> list of record strings (rung_scan3 stdout per window),
        exit codes."""
        raw, lens, codes = self.records_raw(windows, emit_survivors)
        return _split_text(raw, lens), codes

    def records_raw(self, windows, emit_survivors=True):
        """records() as the library returns them: (all records as one bytes object, each
        window's record length in bytes, exit codes)."""
        b = _packed(windows)
        n = b.n
        text = ctypes.c_void_p()
        tl = ctypes.c_size_t()
        rl = (ctypes.c_size_t * max(1, n))()
        ec = (ctypes.c_int * max(1, n))()
        _check(lib().es_scan_records(self.h, b.lo, b.hi, n, 1 if emit_survivors else 0, ctypes.byref(text),
                                     ctypes.byref(tl), rl, ec))
        try:
            raw = ctypes.string_at(text, tl.value)
        finally:
            lib().es_free(text)
        return raw, rl[:n], ec[:n]

    def submit(self, windows):
        """Pipelined records(): plans [(lo, hi)] (or a Batch) and launches the batch without
        waiting. At most MAX_IN_FLIGHT batches are in flight; collect() returns them in
        submission order."""
        b = _packed(windows)
        _check(lib().es_scan_submit(self.h, b.lo, b.hi, b.n))
        self._sizes.append(b.n)

    def collect_raw(self, emit_survivors=True):
        """The oldest submitted batch as the library returns it: (all records as one bytes
        object, each window's record length in bytes, exit codes)."""
        if not self._sizes:
            raise EsError("no batch in flight")
        n = self._sizes[0]
        text = ctypes.c_void_p()
        tl = ctypes.c_size_t()
        rl = (ctypes.c_size_t * max(1, n))()
        ec = (ctypes.c_int * max(1, n))()
        got = ctypes.c_size_t()
        try:
            _check(lib().es_scan_collect(self.h, 1 if emit_survivors else 0, ctypes.byref(text), ctypes.byref(tl),
                                         rl, ec, ctypes.byref(got)))
        finally:
            self._sizes.pop(0)
        try:
            raw = ctypes.string_at(text, tl.value)
        finally:
            lib().es_free(text)
        k = got.value
        return raw, rl[:k], ec[:k]

    def collect(self, emit_survivors=True):
        """The oldest submitted batch's (records, exit codes), exactly as records() returns them."""
        raw, lens, codes = self.collect_raw(emit_survivors)
        return _split_text(raw, lens), codes

    def in_flight(self):
        return lib().es_scan_in_flight(self.h)

    def counts(self, windows):
        n = len(windows)
        out = (EsCounts * max(1, n))()
        _check(lib().es_scan_counts(self.h, _u64s([w[0] for w in windows]), _u64s([w[1] for w in windows]),
                                    n, out))
        return [out[i].as_dict() for i in range(n)]

    def survivors(self, lo, hi, cap=1 << 22):
        """Stage-C survivors of [lo, hi) ascending, their primality flags, and counters."""
        p = (U64 * cap)()
        f = (ctypes.c_uint8 * cap)()
        n = ctypes.c_size_t()
        c = EsCounts()
        _check(lib().es_survivors(self.h, lo, hi, p, f, cap, ctypes.byref(n), ctypes.byref(c)))
        k = min(n.value, cap)
        return [p[i] for i in range(k)], [f[i] for i in range(k)], c.as_dict()

    def primality(self, vals):
        n = len(vals)
        f = (ctypes.c_uint8 * max(1, n))()
        _check(lib().es_primality(self.h, _u64s(vals), n, f))
        return [f[i] for i in range(n)]

    def set_layout(self, Mp):
        _check(lib().es_set_layout(self.h, Mp))

    def set_launch(self, persistent=True, block=256, max_bits=32768):
        _check(lib().es_set_launch(self.h, 1 if persistent else 0, block, max_bits))

    def set_residency(self, waves_per_cu=16):
        _check(lib().es_set_residency(self.h, waves_per_cu))

    def set_grouping(self, max_tiles=4):
        _check(lib().es_set_grouping(self.h, max_tiles))

    def set_graph(self, on=True):
        _check(lib().es_set_graph(self.h, 1 if on else 0))

    def set_device_tail(self, on=True):
        _check(lib().es_set_device_tail(self.h, 1 if on else 0))

    def set_order(self, order="lane", seg=0):
        """Emission order of certificate lines: "lane" (rung_scan3) or "round" (sieve_cuda)."""
        _check(lib().es_set_order(self.h, {"lane": 0, "round": 1}[order], seg))

    def certify_batch(self, primes):
        """Stage D on the device for a list of primes: (lines, how many the host finished)."""
        text = ctypes.c_void_p()
        host = ctypes.c_size_t()
        _check(lib().es_certify_batch(self.h, _u64s(primes), len(primes), ctypes.byref(text), ctypes.byref(host)))
        try:
            out = ctypes.string_at(text).decode()
        finally:
            lib().es_free(text)
        return out.splitlines(keepends=True), host.value

    def direct_batch(self, primes):
        """Stage E alone on the device for a list of primes: (lines, how many the host
        finished); one SOLVED or SURVIVOR line per prime, in input order."""
        text = ctypes.c_void_p()
        host = ctypes.c_size_t()
        _check(lib().es_direct_batch(self.h, _u64s(primes), len(primes), ctypes.byref(text), ctypes.byref(host)))
        try:
            out = ctypes.string_at(text).decode()
        finally:
            lib().es_free(text)
        return out.splitlines(keepends=True), host.value

    def set_tuning(self, c_pair, c_mark, c_bit):
        _check(lib().es_set_tuning(self.h, c_pair, c_mark, c_bit))

    def autotune(self, sample=()):
        """Measures layouts per window size and the launch shape; returns the report."""
        rep = ctypes.c_void_p()
        n = len(sample)
        _check(lib().es_autotune(self.h, _u64s([w[0] for w in sample]) if n else None,
                                 _u64s([w[1] for w in sample]) if n else None, n, ctypes.byref(rep)))
        try:
            return ctypes.string_at(rep).decode()
        finally:
            lib().es_free(rep)

    def choose_layout(self, w):
        return lib().es_choose_layout(self.h, w)

    def timing(self):
        t = EsTiming()
        _check(lib().es_last_timing(self.h, ctypes.byref(t)))
        return t.as_dict()

    def mark_count(self, lo, hi, Mp=0):
        """Exact kernel work of [lo, hi): marks, bits, (tile, filter) pairs, tiles."""
        a, b, c, d = U64(), U64(), U64(), U64()
        _check(lib().es_mark_count(self.h, lo, hi, Mp, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                   ctypes.byref(d)))
        return {"marks": a.value, "bits": b.value, "pairs": c.value, "tiles": d.value}

    def lds_peak(self):
        """Measured LDS atomic-AND operations per second."""
        v = ctypes.c_double()
        _check(lib().es_lds_peak(self.h, ctypes.byref(v)))
        return v.value

    def lds_scatter_peak(self):
        """Measured LDS scattered byte stores per second (the marking primitive)."""
        v = ctypes.c_double()
        _check(lib().es_lds_scatter_peak(self.h, ctypes.byref(v)))
        return v.value

    def wheel_classes(self):
        n = ctypes.c_size_t()
        _check(lib().es_wheel_classes(self.h, None, 0, ctypes.byref(n)))
        buf = (U64 * n.value)()
        _check(lib().es_wheel_classes(self.h, buf, n.value, ctypes.byref(n)))
        return list(buf)

    def wheel_info(self):
        a, b, c, d = U64(), U64(), U64(), U64()
        _check(lib().es_wheel_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
        return {"lanes": a.value, "wheel_pmin": b.value, "nfilters": c.value, "ncov": d.value}


def serve_config(**kw):
    """es_serve_config with the library defaults, overridden by keyword."""
    c = EsServeConfig()
    lib().es_serve_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Server:
    """A continuous-batching request server over a Scanner (es_server_*). While it is
    open it owns the scanner."""

    def __init__(self, scanner, **cfg):
        self.h = ctypes.c_void_p()
        _check(lib().es_server_open(scanner.h, ctypes.byref(serve_config(**cfg)), ctypes.byref(self.h)))

    def submit(self, lo, hi):
        i = U64()
        _check(lib().es_server_submit(self.h, lo, hi, ctypes.byref(i)))
        return i.value

    def wait(self, rid, timeout_ms=-1.0):
        """(record, info dict), or None on timeout."""
        text, n, info = ctypes.c_void_p(), ctypes.c_size_t(), EsRequestInfo()
        rc = _check(lib().es_server_wait(self.h, rid, timeout_ms, ctypes.byref(text), ctypes.byref(n),
                                         ctypes.byref(info)))
        if rc == 1:
            return None
        try:
            return ctypes.string_at(text, n.value).decode(), info.as_dict()
        finally:
            lib().es_free(text)

    def stats(self):
        st = EsServeStats()
        _check(lib().es_server_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h:
            lib().es_server_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def serve_trace(scanner, requests, realtime=True, **cfg):
    """Replays [(arrival_s, lo, hi)] through a server: (records, infos, stats)."""
    n = len(requests)
    arr = (ctypes.c_double * max(1, n))(*[r[0] for r in requests])
    lo = _u64s([r[1] for r in requests])
    hi = _u64s([r[2] for r in requests])
    text = ctypes.c_void_p()
    rl = (ctypes.c_size_t * max(1, n))()
    info = (EsRequestInfo * max(1, n))()
    st = EsServeStats()
    _check(lib().es_serve_trace(scanner.h, ctypes.byref(serve_config(**cfg)), arr, lo, hi, n,
                                1 if realtime else 0, ctypes.byref(text), rl, info, ctypes.byref(st)))
    try:
        raw = ctypes.string_at(text, sum(rl[i] for i in range(n)))
    finally:
        lib().es_free(text)
    out, pos = [], 0
    for i in range(n):
        out.append(raw[pos:pos + rl[i]].decode())
        pos += rl[i]
    return out, [info[i].as_dict() for i in range(n)], st.as_dict()
