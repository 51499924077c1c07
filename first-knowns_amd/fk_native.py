"""ctypes binding of libfk_amd.so (include/fk_engine.h).

This module is the only place Python touches the C ABI.  It refuses to run
without the built library: there is no CPU fallback anywhere in the
product path, and an engine on a machine without a gfx950 device raises
RuntimeError at construction (FK_E_NODEV), as the reference's GPU engines
raise when CuPy is missing (dickson-ladders/ladder_gpu.py:431-432,
primorial-ap/ap_gpu.py:402-403).
"""

import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libfk_amd.so"

FAMILY_EULER, FAMILY_DICKSON, FAMILY_AP = 1, 2, 3
CLASSIFY_NONE, CLASSIFY_FULL, CLASSIFY_TWO_PASS = 0, 1, 2

FK_OK, FK_E_ARG, FK_E_RANGE, FK_E_OVERFLOW = 0, 1, 2, 3
FK_E_DEVICE, FK_E_OOM, FK_E_NODEV, FK_E_BUSY = 4, 5, 6, 7
ABI_VERSION = 3
CLASSIFY_STATIC_MAX = 65536       # values per fk_engine_classify_static call
BATCH_DEDICATED_QUEUE = 1         # fk_batch_create_ex: a hardware queue of its own
BATCH_THROUGHPUT = 2              # fk_batch_create_ex: planned to run beside others

# The exported entry points, in the order include/fk_engine.h declares them.
EXPORTS = ("fk_abi_version", "fk_last_error", "fk_device_count",
           "fk_engine_create", "fk_engine_destroy", "fk_engine_get_info",
           "fk_engine_sweep", "fk_engine_classify", "fk_batch_create",
           "fk_batch_destroy", "fk_batch_create_ex", "fk_batch_submit",
           "fk_batch_query",
           "fk_batch_results", "fk_batch_timing", "fk_batch_set_graph",
           "fk_batch_graph_stats", "fk_batch_counts", "fk_engine_batch",
           "fk_batch_reruns", "fk_batch_set_budget", "fk_batch_cursor_get",
           "fk_batch_cursor_set", "fk_batch_resume",
           "fk_engine_classify_static", "fk_batch_step_times",
           "fk_batch_window_wheel", "fk_probe_peaks")

u32, i32, u64 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64


class EngineDesc(ctypes.Structure):
    _fields_ = [("family", i32), ("n", i32), ("q2", u32), ("device", i32),
                ("nwheel", i32), ("wheel_primes", u32 * 12),
                ("q_pat", u32), ("q_med", u32), ("threads", u32),
                ("words_per_thread", u32), ("round_lane_prime", u32),
                ("round_growth", u32), ("q_mark", u32), ("q_bits", u32),
                ("short_words", u32), ("segment_cap", u32),
                ("alt_wheels", i32)]


class EngineInfo(ctypes.Structure):
    _fields_ = [("family", i32), ("n", i32), ("q2", u32), ("nwheel", u32),
                ("wheel_primes", u32 * 16), ("line_step", u64),
                ("nlines", u64), ("W", u64), ("d", u64 * 4),
                ("q_pat", u32), ("q_mark", u32), ("q_med", u32),
                ("threads", u32), ("words_per_thread", u32),
                ("n_lds_primes", u32),
                ("n_pattern_primes", u32), ("n_round_primes", u32),
                ("n_rounds", u32), ("short_words", u32), ("device_bytes", u64),
                ("survival_lds", ctypes.c_double),
                ("mark_density", ctypes.c_double)]


class Window(ctypes.Structure):
    _fields_ = [("lo", u64 * 2), ("hi", u64 * 2)]


class ClassifySpec(ctypes.Structure):
    _fields_ = [("mode", i32), ("cap", i32), ("exact_from", i32),
                ("first", i32)]


RECORD_DTYPE = np.dtype([("v0", "<u8"), ("v1", "<u8"), ("run", "<i4"),
                         ("window", "<u4")])

_lib = None


class FkError(RuntimeError):
    """A RuntimeError from the library (device, memory, overflow)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """The loaded library; raises if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__"
            f" as g; g.build()'` (hipcc --offload-arch=gfx950) first")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    P = ctypes.POINTER
    L.fk_abi_version.restype = ctypes.c_int
    L.fk_last_error.argtypes = [ctypes.c_char_p, sz]
    L.fk_device_count.argtypes = [P(ctypes.c_int)]
    L.fk_engine_create.argtypes = [P(EngineDesc), P(vp)]
    L.fk_engine_destroy.argtypes = [vp]
    L.fk_engine_get_info.argtypes = [vp, P(EngineInfo)]
    L.fk_engine_sweep.argtypes = [vp, P(Window), P(ClassifySpec), vp, sz,
                                  P(sz)]
    L.fk_engine_classify.argtypes = [vp, vp, sz, P(ClassifySpec), vp]
    L.fk_batch_create.argtypes = [ctypes.c_int, P(vp)]
    L.fk_batch_destroy.argtypes = [vp]
    L.fk_batch_create_ex.argtypes = [ctypes.c_int, u32, P(vp)]
    L.fk_batch_submit.argtypes = [vp, vp, sz, vp, vp, vp, sz]
    L.fk_batch_query.argtypes = [vp, P(ctypes.c_int)]
    L.fk_batch_results.argtypes = [vp, vp, sz, P(sz), vp]
    L.fk_batch_timing.argtypes = [vp, vp, P(u64), P(u64)]
    L.fk_batch_counts.argtypes = [vp, vp, sz, P(sz)]
    L.fk_batch_set_graph.argtypes = [vp, ctypes.c_int]
    L.fk_batch_graph_stats.argtypes = [vp, P(u64), P(u64)]
    L.fk_engine_batch.argtypes = [vp, P(vp)]
    L.fk_batch_reruns.argtypes = [vp, P(u64)]
    L.fk_batch_set_budget.argtypes = [vp, vp]
    L.fk_batch_cursor_get.argtypes = [vp, sz, vp]
    L.fk_batch_cursor_set.argtypes = [vp, sz, vp]
    L.fk_batch_resume.argtypes = [vp]
    L.fk_engine_classify_static.argtypes = [vp, vp, sz, P(ClassifySpec), vp]
    L.fk_batch_step_times.argtypes = [vp, vp, vp, sz, P(sz)]
    L.fk_batch_window_wheel.argtypes = [vp, sz, P(u64)]
    L.fk_probe_peaks.argtypes = [ctypes.c_int, P(ctypes.c_double),
                                 P(ctypes.c_double), P(ctypes.c_double)]
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    if L.fk_abi_version() != ABI_VERSION:
        raise RuntimeError("libfk_amd.so ABI version mismatch")
    _lib = L
    return L


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().fk_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc):
    """Map a status code to the reference's exception taxonomy."""
    if rc == FK_OK:
        return
    msg = last_error()
    if rc in (FK_E_ARG, FK_E_RANGE):
        raise ValueError(msg)
    raise FkError(rc, msg)


def device_count():
    n = ctypes.c_int(0)
    check(lib().fk_device_count(ctypes.byref(n)))
    return n.value


def split128(v):
    v = int(v)
    if v < 0 or v >> 128:
        raise ValueError(f"value {v} outside [0, 2^128)")
    return v & 0xFFFFFFFFFFFFFFFF, v >> 64


def make_window(lo, hi):
    w = Window()
    w.lo[0], w.lo[1] = split128(lo)
    w.hi[0], w.hi[1] = split128(hi)
    return w


def record_values(recs):
    """Record values as Python ints (exact past 2^64)."""
    if recs.size == 0:
        return []
    if not recs["v1"].any():
        return [int(x) for x in recs["v0"].tolist()]
    return [int(a) | (int(b) << 64)
            for a, b in zip(recs["v0"].tolist(), recs["v1"].tolist())]


class Engine:
    """One family's resident device engine (fk_engine)."""

    def __init__(self, family, n, q2, device=0, nwheel=0, wheel_primes=None,
                 q_pat=0, q_med=0, threads=0, words_per_thread=0,
                 round_lane_prime=0, round_growth=0, q_mark=0, q_bits=0,
                 short_words=0, segment_cap=0, alt_wheels=0):
        """alt_wheels: prefix wheels kept for short windows (0: the
        library's default, -1: none); records never depend on it."""
        d = EngineDesc()
        d.family, d.n, d.q2, d.device = int(family), int(n), int(q2), device
        d.nwheel = int(nwheel)
        for i, q in enumerate(wheel_primes or []):
            d.wheel_primes[i] = int(q)
        d.q_pat, d.q_med = int(q_pat), int(q_med)
        d.threads, d.words_per_thread = int(threads), int(words_per_thread)
        d.round_lane_prime = int(round_lane_prime)
        d.round_growth = int(round_growth)
        d.q_mark = int(q_mark)
        d.q_bits = int(q_bits)
        d.short_words = int(short_words)
        d.segment_cap = int(segment_cap)
        d.alt_wheels = int(alt_wheels)
        h = ctypes.c_void_p()
        check(lib().fk_engine_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.desc = d
        info = EngineInfo()
        check(lib().fk_engine_get_info(h, ctypes.byref(info)))
        self.info = info
        self.family, self.n, self.q2 = int(family), int(n), int(q2)
        self.device = int(device)
        self.line_step = int(info.line_step)
        self.nlines = int(info.nlines)
        self.wheel = [int(info.wheel_primes[i]) for i in range(info.nwheel)]

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_pipe", None) is not None:
            self._pipe.close()
            self._pipe = None
        if getattr(self, "_h", None):
            lib().fk_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # expected sieve survivors one device sweep takes at most; wider
    # windows are swept in consecutive pieces.  The library refuses a
    # window whose survivors would not fit one queue segment (2^30 records,
    # refused from 0.9 of that); a sweep of 9e8 holds about 75 GB of queues
    # and flags, a quarter of the device's memory, and one sweep does less
    # work per position than several (the Euler SCORE_WIDE range, 8.8e8
    # survivors: 0.58 s in one sweep against 0.68 s in four)
    PIECE_SURVIVORS = 900_000_000
    # pieces hold whole 16-word chunks of every line (1024 D - D values:
    # 1024 positions per line, whatever the first value's residue)
    PIECE_POSITIONS = 1024

    def max_span(self):
        """Widest window one device sweep takes (line values)."""
        per = self.nlines / self.line_step * float(self.info.survival_lds)
        return max(self.line_step, int(self.PIECE_SURVIVORS / max(per, 1e-30)))

    def sweep(self, lo, hi, mode=CLASSIFY_NONE, cap=0, exact_from=0):
        """Sorted records (numpy RECORD_DTYPE) of the survivors in [lo, hi)."""
        lo, hi = int(lo), int(hi)
        step = self.max_span()
        # pieces short enough for the short-line sieve: lines of at most
        # short_words 64-position words (the line sieve took the SCORE_WIDE
        # range on the 31# wheel, lines of 1,558 words, 11.3 s in one
        # sweep against 0.85 s in pieces of at most 1,024 words)
        sw = int(self.info.short_words)
        if sw:
            step = min(step, 64 * sw * self.line_step)
        unit = self.PIECE_POSITIONS * self.line_step
        if hi - lo > step and step >= 2 * unit:
            step = (step // unit) * unit - self.line_step
        if hi - lo > step:
            parts = [self._sweep_one(a, min(a + step, hi), mode, cap,
                                     exact_from)
                     for a in range(lo, hi, step)]
            return np.concatenate(parts)
        return self._sweep_one(lo, hi, mode, cap, exact_from)

    def _sweep_one(self, lo, hi, mode, cap, exact_from):
        w = make_window(lo, hi)
        c = ClassifySpec(int(mode), int(cap), int(exact_from))
        n = ctypes.c_size_t(0)
        out = getattr(self, "_out", None)        # reused between sweeps
        if out is None:
            out = self._out = np.empty(1 << 12, dtype=RECORD_DTYPE)
        cap_out = out.size
        rc = lib().fk_engine_sweep(self._h, ctypes.byref(w), ctypes.byref(c),
                                   out.ctypes.data, cap_out, ctypes.byref(n))
        if rc == FK_E_OVERFLOW:
            out = np.empty(n.value, dtype=RECORD_DTYPE)
            b = ctypes.c_void_p()
            check(lib().fk_engine_batch(self._h, ctypes.byref(b)))
            rc = lib().fk_batch_results(b, out.ctypes.data, n.value,
                                        ctypes.byref(n), None)
        check(rc)
        return out[:n.value].copy()

    def sweep_many(self, windows, mode=CLASSIFY_NONE, cap=0, exact_from=0):
        """Yield the sorted records of each window [lo, hi) of `windows` in
        order, PIPELINE_DEPTH windows in flight on as many streams
        (Pipeline): the same records as sweep() window by window."""
        ws = [(int(lo), int(hi)) for lo, hi in windows]
        if getattr(self, "_pipe", None) is None:
            self._pipe = Pipeline(self.device)
        # a second generator while one is paused gets a pipeline of its own
        # (the engine's keeps the first one's windows in flight)
        pipe = self._pipe if not self._pipe.active else Pipeline(self.device)
        # windows of one length repeat their launch sequence: capture it
        # at once
        graph = 2 if len({hi - lo for lo, hi in ws}) <= 1 else 1
        spec = (mode, cap, exact_from)
        try:
            for _, recs in pipe.run([(self, lo, hi, spec) for lo, hi in ws],
                                    graph=graph):
                yield recs
        finally:
            if pipe is not self._pipe:
                pipe.close()

    def classify(self, values, mode=CLASSIFY_FULL, cap=64, exact_from=0,
                 first=0):
        """Chain lengths of line values (Python ints), computed on device:
        consecutive chain values from index `first` that pass, up to cap."""
        vals = list(values)
        if not vals:
            return []
        arr = np.empty(2 * len(vals), dtype=np.uint64)
        for i, v in enumerate(vals):
            arr[2 * i], arr[2 * i + 1] = split128(v)
        runs = np.empty(len(vals), dtype=np.int32)
        c = ClassifySpec(int(mode), int(cap), int(exact_from), int(first))
        check(lib().fk_engine_classify(self._h, arr.ctypes.data, len(vals),
                                       ctypes.byref(c), runs.ctypes.data))
        return [int(r) for r in runs.tolist()]

    def classify_static(self, values, mode=CLASSIFY_FULL, cap=64,
                        exact_from=0, first=0):
        """classify() in the library's fixed device array: no device memory
        is allocated (at most CLASSIFY_STATIC_MAX values; calls on one
        device must not overlap)."""
        vals = list(values)
        arr = np.zeros(2 * max(len(vals), 1), dtype=np.uint64)
        for i, v in enumerate(vals):
            arr[2 * i], arr[2 * i + 1] = split128(v)
        runs = np.zeros(max(len(vals), 1), dtype=np.int32)
        c = ClassifySpec(int(mode), int(cap), int(exact_from), int(first))
        check(lib().fk_engine_classify_static(self._h, arr.ctypes.data,
                                              len(vals), ctypes.byref(c),
                                              runs.ctypes.data))
        return [int(r) for r in runs.tolist()[:len(vals)]]

    def extend_runs(self, values, runs, swept_cap, cap, mode=CLASSIFY_FULL,
                    exact_from=0):
        """Runs of a sweep classified with swept_cap (at most 64 values per
        chain), carried on to `cap`: chains that reached swept_cap are
        classified again on the device up to cap."""
        runs = [int(r) for r in runs]
        if cap <= swept_cap:
            return runs
        again = [i for i, r in enumerate(runs) if r >= swept_cap]
        if again:
            more = self.classify([values[i] for i in again], mode, cap,
                                 exact_from)
            for i, r in zip(again, more):
                runs[i] = r
        return runs

    def last_timing(self):
        """(ms[prep, sieve, rounds, classify, whole], tiles, candidates) of
        the engine's last sweep."""
        b = ctypes.c_void_p()
        check(lib().fk_engine_batch(self._h, ctypes.byref(b)))
        return batch_timing(b)

    def stage_timing(self, on=True):
        """Later sweeps run eagerly with HIP events between stages (on) or,
        the default, replay a repeated launch sequence as a HIP graph (off:
        last_timing then reports only the whole)."""
        b = ctypes.c_void_p()
        check(lib().fk_engine_batch(self._h, ctypes.byref(b)))
        check(lib().fk_batch_set_graph(b, 0 if on else 1))

    def last_step_times(self):
        """[(kind, ms)] per kernel launch of the last eager sweep."""
        b = ctypes.c_void_p()
        check(lib().fk_engine_batch(self._h, ctypes.byref(b)))
        return batch_step_times(b)

    def last_counts(self):
        """Queue sizes after the sieve and each round of the last sweep."""
        b = ctypes.c_void_p()
        check(lib().fk_engine_batch(self._h, ctypes.byref(b)))
        return batch_counts(b)

    def last_wheel(self):
        """Wheel period the last sweep's window was sieved on."""
        b = ctypes.c_void_p()
        check(lib().fk_engine_batch(self._h, ctypes.byref(b)))
        d = u64(0)
        check(lib().fk_batch_window_wheel(b, 0, ctypes.byref(d)))
        return int(d.value)


def probe_peaks(device=0):
    """Measured peaks: (VALU lane-ops/s for double FMAs, VALU lane-ops/s
    for 32-bit funnel shifts, LDS bytes/s)."""
    f, i, l = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib().fk_probe_peaks(int(device), ctypes.byref(f), ctypes.byref(i),
                               ctypes.byref(l)))
    return f.value, i.value, l.value


def batch_counts(handle):
    """Queue sizes after the sieve and after each round of the last run."""
    n = ctypes.c_size_t(0)
    check(lib().fk_batch_counts(handle, None, 0, ctypes.byref(n)))
    out = (u32 * max(n.value, 1))()
    check(lib().fk_batch_counts(handle, ctypes.cast(out, ctypes.c_void_p),
                                n.value, ctypes.byref(n)))
    return [int(x) for x in out[:n.value]]


STEP_SIEVE, STEP_ROUND_CANDIDATE, STEP_ROUND_PRIME, STEP_COMPACT = 1, 2, 3, 4
STEP_GATHER, STEP_CLASSIFY, STEP_READBACK = 5, 6, 7


def batch_step_times(handle):
    """[(kind, ms)] of every launch of the last eager submission."""
    n = ctypes.c_size_t(0)
    check(lib().fk_batch_step_times(handle, None, None, 0, ctypes.byref(n)))
    kinds = (i32 * max(n.value, 1))()
    ms = (ctypes.c_float * max(n.value, 1))()
    check(lib().fk_batch_step_times(handle, ctypes.cast(kinds, ctypes.c_void_p),
                                    ctypes.cast(ms, ctypes.c_void_p), n.value,
                                    ctypes.byref(n)))
    return [(int(kinds[i]), float(ms[i])) for i in range(n.value)]


def batch_timing(handle):
    ms = (ctypes.c_float * 5)()
    tiles, cand = ctypes.c_uint64(0), ctypes.c_uint64(0)
    check(lib().fk_batch_timing(handle, ms, ctypes.byref(tiles),
                                ctypes.byref(cand)))
    return list(ms), int(tiles.value), int(cand.value)


class Batch:
    """An fk_batch: many windows of many engines, submitted asynchronously."""

    def __init__(self, device=0, graph=False, dedicated=False,
                 throughput=False):
        """dedicated: the batch's stream gets a hardware queue of its own
        (FK_BATCH_DEDICATED_QUEUE); throughput: its launches are planned to
        run beside other batches' (FK_BATCH_THROUGHPUT).  The slots of a
        Pipeline are both."""
        h = ctypes.c_void_p()
        flags = ((BATCH_DEDICATED_QUEUE if dedicated else 0) |
                 (BATCH_THROUGHPUT if throughput else 0))
        check(lib().fk_batch_create_ex(int(device), flags, ctypes.byref(h)))
        self._h = h
        self._nwin = 0
        if graph:
            # 1 (True): capture a launch sequence when it repeats; 2: at once
            check(lib().fk_batch_set_graph(h, int(graph)))

    def graph_stats(self):
        """(graphs captured, graph launches) so far."""
        c, n = u64(0), u64(0)
        check(lib().fk_batch_graph_stats(self._h, ctypes.byref(c),
                                         ctypes.byref(n)))
        return int(c.value), int(n.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().fk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, engines, engine_of, windows, specs):
        """engines: list[Engine]; engine_of: engine index per window;
        windows: list of (lo, hi); specs: list of (mode, cap, exact_from)."""
        ne = len(engines)
        eh = (ctypes.c_void_p * ne)(*[e.handle.value for e in engines])
        nw = len(windows)
        eo = (u32 * nw)(*[int(i) for i in engine_of])
        ws = (Window * nw)(*[make_window(lo, hi) for lo, hi in windows])
        cs = (ClassifySpec * nw)(*[ClassifySpec(int(m), int(c), int(x))
                                   for m, c, x in specs])
        self._keep = (eh, eo, ws, cs)
        check(lib().fk_batch_submit(self._h, eh, ne, eo, ws, cs, nw))
        self._nwin = nw

    def done(self):
        d = ctypes.c_int(0)
        check(lib().fk_batch_query(self._h, ctypes.byref(d)))
        return bool(d.value)

    def results(self):
        """(records sorted by (window, value), per-window counts)."""
        n = ctypes.c_size_t(0)
        counts = np.zeros(max(self._nwin, 1), dtype=np.uint64)
        out = np.empty(1 << 12, dtype=RECORD_DTYPE)
        rc = lib().fk_batch_results(self._h, out.ctypes.data, out.size,
                                    ctypes.byref(n), counts.ctypes.data)
        if rc == FK_E_OVERFLOW:
            out = np.empty(n.value, dtype=RECORD_DTYPE)
            rc = lib().fk_batch_results(self._h, out.ctypes.data, out.size,
                                        ctypes.byref(n), counts.ctypes.data)
        check(rc)
        return out[:n.value].copy(), counts[:self._nwin]

    def timing(self):
        return batch_timing(self._h)

    def reruns(self):
        """Submissions rerun after a survivor queue segment overflowed."""
        n = u64(0)
        check(lib().fk_batch_reruns(self._h, ctypes.byref(n)))
        return int(n.value)

    def set_budget(self, values):
        """Later submissions sweep at most `values` line values in all,
        window by window (0: no limit)."""
        v = (u64 * 2)(*split128(values))
        check(lib().fk_batch_set_budget(self._h, v))

    def cursor(self, window):
        """First line value of `window` not yet swept (its end when done)."""
        v = (u64 * 2)()
        check(lib().fk_batch_cursor_get(self._h, int(window), v))
        return int(v[0]) | (int(v[1]) << 64)

    def set_cursor(self, window, value):
        v = (u64 * 2)(*split128(value))
        check(lib().fk_batch_cursor_set(self._h, int(window), v))

    def resume(self):
        """Submit the job again, every window from its cursor."""
        check(lib().fk_batch_resume(self._h))

    def counts(self):
        return batch_counts(self._h)

    def window_wheel(self, window):
        """Wheel period the window was sieved on in the last submission
        (the engine's own, or a prefix wheel's for a short window)."""
        d = u64(0)
        check(lib().fk_batch_window_wheel(self._h, int(window), ctypes.byref(d)))
        return int(d.value)


# Windows in flight: HIP gives a process four hardware queues by default,
# so four streams run their kernels side by side; a window's rounds and
# classification are latency-bound launches over few candidates, and the
# other windows' launches fill the device meanwhile.
PIPELINE_DEPTH = 4


class Pipeline:
    """Windows swept one engine call each, `depth` (PIPELINE_DEPTH) in
    flight: window i is enqueued on its own batch (its own stream, queues
    and captured graph) before window i - depth + 1 is read back, so the
    device works on the next windows while the last ones' rounds and
    classification (few candidates, few busy CUs) finish, and never waits
    on the host.  The launch loop of dickson-ladders/ladder_gpu.py:717-747
    (two CUDA streams, chain i + 1 enqueued before chain i is read back),
    deeper."""

    def __init__(self, device=0, depth=PIPELINE_DEPTH):
        self.device, self.depth = int(device), int(depth)
        self._slots = []
        self._graph = None
        self.active = False           # a run() generator is not finished

    def _ready(self, graph):
        if self._graph != graph:
            self.close()
            # one hardware queue per slot: slots that shared one would
            # run their windows in turn
            self._slots = [Batch(self.device, graph=graph, dedicated=True,
                                 throughput=True)
                           for _ in range(self.depth)]
            self._graph = graph

    def run(self, jobs, graph=1):
        """jobs: (engine, lo, hi, (mode, cap, exact_from)) each; yields
        (index, sorted records) in job order.  Windows still in flight when
        the caller stops early are waited for and dropped.  One run at a
        time: a second while the first is paused raises RuntimeError."""
        if self.active:
            raise RuntimeError("this pipeline is already running windows; "
                               "finish or close that generator first")
        self._ready(graph)
        self.active = True
        pend = [None] * self.depth
        try:
            done = 0
            for i, (eng, lo, hi, spec) in enumerate(jobs):
                s = i % self.depth
                if pend[s] is not None:
                    recs = self._slots[s].results()[0]
                    pend[s] = None
                    yield i - self.depth, recs
                    done += 1
                self._slots[s].submit([eng], [0], [(lo, hi)], [spec])
                pend[s] = i
            for i in range(done, done + sum(p is not None for p in pend)):
                s = i % self.depth
                recs = self._slots[s].results()[0]
                pend[s] = None
                yield i, recs
        finally:
            for s, p in enumerate(pend):
                if p is not None:
                    self._slots[s].results()
            self.active = False

    def close(self):
        for b in self._slots:
            b.close()
        self._slots = []
        self._graph = None
