"""Continuous-batching request loop over the MI355X engine.

Range-sweep requests (family, depth n, [lo, hi), classification) arrive
over time.  The server keeps every engine it has built resident (prime
tables, kill sets, wheel), keyed by (family, n, q2).  It runs up to
`slots` batches at once, each an fk_batch with its own HIP stream, so their
work overlaps on the device.  When a slot is free, the waiting requests are
admitted in arrival order while their HBM footprint fits the budget still
free, and they go to the device together in one batch submission.  A
request that arrives while the slots are busy waits, and joins the next
submission with whatever else has arrived by then.  One request larger
than the whole budget is admitted alone, so nothing starves.

Footprint of a request: the queue entries its sieve stage is expected to
leave (span x the wheel's line density x the LDS-stage survival) times
the bytes each entry holds on the device (two queues of 24-byte records,
a flag and two classification words), with the same margin the engine
sizes its queues with.

Each request's result is exactly what the engine returns for that window
alone: the batch shares nothing between windows but the engine tables
(tested alone versus in a mix).  Latency is measured from arrival to the
moment the server has the request's records.

No reference counterpart: the reference's launchers sweep one window at a
time (dickson-ladders/launch.py:1072-1117).
"""

import collections
import math
import time

import fk_families as F
import fk_native as N

FAMILY_ID = {"euler": N.FAMILY_EULER, "dickson": N.FAMILY_DICKSON,
             "ap": N.FAMILY_AP}
DEFAULT_Q2 = {"euler": 1 << 16, "dickson": 1 << 22, "ap": 1 << 24}
# device bytes per expected queue entry: two 24-byte record queues, a
# 4-byte flag and two 8-byte classification words
BYTES_PER_ENTRY = 2 * 24 + 4 + 16


class Request:
    __slots__ = ("rid", "family", "n", "lo", "hi", "classify", "q2",
                 "arrival", "admitted", "done", "records", "footprint")

    def __init__(self, rid, family, n, lo, hi, classify=None, q2=None,
                 arrival=0.0):
        self.rid = rid
        self.family = family
        self.n = int(n)
        self.lo, self.hi = int(lo), int(hi)
        self.classify = classify
        self.q2 = q2 or DEFAULT_Q2[family]
        self.arrival = float(arrival)
        self.admitted = None
        self.done = None
        self.records = None
        self.footprint = 0

    @property
    def span(self):
        return self.hi - self.lo

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def default_classify(family, n):
    if family == "dickson":
        return (N.CLASSIFY_TWO_PASS, n + 8, 5)
    if family == "ap":
        return (N.CLASSIFY_TWO_PASS, n, 6)
    return (N.CLASSIFY_FULL, n + 3, 0)


def footprint(info, span):
    """Device bytes a window of `span` line values needs beyond the
    resident tables (see the module docstring)."""
    density = float(info.nlines) / float(info.line_step)
    x = span * density * float(info.survival_lds)
    entries = 1.5 * x + 8.0 * math.sqrt(x) + 1024.0
    return int(entries * BYTES_PER_ENTRY)


class Server:
    """The request loop.  `engine_factory(family, n, q2)` and
    `batch_factory()` default to the real engine (tests may pass
    stand-ins); `clock` returns seconds."""

    def __init__(self, slots=4, budget=32 << 30, device=0,
                 engine_factory=None, batch_factory=None,
                 clock=time.perf_counter, idle_sleep=50e-6):
        if slots < 1:
            raise ValueError("at least one slot")
        self.slots = int(slots)
        self.budget = int(budget)
        self.device = device
        self.clock = clock
        self.idle_sleep = idle_sleep
        self._engine_factory = engine_factory or self._make_engine
        # slots on hardware queues of their own, so that they run side by
        # side; planned for latency (a request waits on its own launches)
        self._batch_factory = batch_factory or (
            lambda: N.Batch(device, dedicated=True))
        self.engines = {}
        self.free_batches = [self._batch_factory() for _ in range(self.slots)]
        self.waiting = collections.deque()
        self.inflight = []            # (batch, [requests], bytes)
        self.in_use = 0               # bytes admitted and not yet done
        self.finished = []
        self.submissions = []         # requests per submission

    # ------------------------------------------------------------ engines

    def _make_engine(self, family, n, q2):
        t = F.tuning(family)
        kw = {}
        if family == "euler":
            kw["wheel_primes"] = [2, 3, 5, 7, 11, 13]
        return N.Engine(FAMILY_ID[family], n, q2, device=self.device, **kw,
                        **t)

    def engine(self, family, n, q2):
        key = (family, int(n), int(q2))
        if key not in self.engines:
            self.engines[key] = self._engine_factory(family, n, q2)
        return self.engines[key]

    # -------------------------------------------------------------- queue

    def submit(self, req):
        """Queue a request (its arrival time is the caller's)."""
        if req.hi <= req.lo:
            raise ValueError(f"empty window [{req.lo}, {req.hi})")
        if req.family not in FAMILY_ID:
            raise ValueError(f"unknown family {req.family!r}")
        if req.classify is None:
            req.classify = default_classify(req.family, req.n)
        eng = self.engine(req.family, req.n, req.q2)
        req.footprint = footprint(eng.info, req.span)
        self.waiting.append(req)

    def _admit(self, now):
        """Pack waiting requests into one submission on a free slot."""
        if not self.free_batches or not self.waiting:
            return False
        take, need = [], 0
        while self.waiting:
            r = self.waiting[0]
            fits = self.in_use + need + r.footprint <= self.budget
            alone = not take and not self.inflight
            if not (fits or alone):
                break
            take.append(self.waiting.popleft())
            need += r.footprint
        if not take:
            return False
        engines, index, eo = [], {}, []
        for r in take:
            key = (r.family, r.n, r.q2)
            if key not in index:
                index[key] = len(engines)
                engines.append(self.engines[key])
            eo.append(index[key])
            r.admitted = now
        b = self.free_batches.pop()
        b.submit(engines, eo, [(r.lo, r.hi) for r in take],
                 [r.classify for r in take])
        self.inflight.append((b, take, need))
        self.in_use += need
        self.submissions.append(len(take))
        return True

    def _reap(self, now_fn):
        progressed = False
        still = []
        for b, reqs, need in self.inflight:
            if not b.done():
                still.append((b, reqs, need))
                continue
            recs, counts = b.results()
            t = now_fn()
            pos = 0
            for i, r in enumerate(reqs):
                c = int(counts[i])
                r.records = recs[pos:pos + c].copy()
                pos += c
                r.done = t
                self.finished.append(r)
            self.in_use -= need
            self.free_batches.append(b)
            progressed = True
        self.inflight = still
        return progressed

    def step(self, now_fn):
        """One pass: collect finished batches, then fill free slots."""
        progressed = self._reap(now_fn)
        while self._admit(now_fn()):
            progressed = True
        return progressed

    def drain(self, now_fn):
        while self.waiting or self.inflight:
            if not self.step(now_fn) and self.idle_sleep:
                time.sleep(self.idle_sleep)

    # -------------------------------------------------------------- driver

    def run(self, arrivals):
        """Serve `arrivals` (requests with arrival times in seconds from
        the start, ascending): each is submitted when the clock reaches its
        time.  Returns the statistics of the run."""
        t0 = self.clock()

        def now():
            return self.clock() - t0

        pending = sorted(arrivals, key=lambda r: r.arrival)
        i = 0
        while i < len(pending) or self.waiting or self.inflight:
            t = now()
            while i < len(pending) and pending[i].arrival <= t:
                self.submit(pending[i])
                i += 1
            if not self.step(now) and self.idle_sleep:
                time.sleep(self.idle_sleep)
        return stats(pending, now())


def percentile(xs, q):
    if not xs:
        return None
    xs = sorted(xs)
    k = (len(xs) - 1) * q
    f = math.floor(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def stats(reqs, wall):
    lat = [r.latency for r in reqs if r.latency is not None]
    cand = sum(r.span for r in reqs if r.done is not None)
    return {"requests": len(reqs), "completed": len(lat),
            "p50_ms": 1e3 * percentile(lat, 0.5) if lat else None,
            "p99_ms": 1e3 * percentile(lat, 0.99) if lat else None,
            "wall_s": wall, "candidates": cand,
            "candidates_per_s": cand / wall if wall > 0 else None}


def poisson_requests(seed, count, rate, families=("euler", "dickson", "ap"),
                     log2_span=(18, 32), depth=(8, 14)):
    """A seeded stream of requests: exponential inter-arrival times at
    `rate` per second, family uniform, span 2^U(log2_span) line values,
    depth uniform, heights per family such that 64-, 128- and 192-bit
    values occur (Dickson k up to about 1e22 with m up to 14)."""
    import random
    rng = random.Random(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.expovariate(rate)
        fam = rng.choice(list(families))
        n = rng.randint(*depth)
        span = int(2 ** rng.uniform(*log2_span))
        if fam == "euler":
            lo = rng.choice((rng.randrange(10**12, 10**13),
                             rng.randrange(10**20, 10**21)))
        elif fam == "dickson":
            W = F.dickson_wheel_modulus(n)
            k = rng.choice((rng.randrange(10**14, 10**16),
                            rng.randrange(10**21, 10**22)))
            lo = k // W
        else:
            lo = rng.choice((rng.randrange(10**15, 10**16),
                             rng.randrange(10**19, 10**20)))
            lo -= lo % 30
            span -= span % 30
            span = max(span, 30)
        out.append(Request(i, fam, n, lo, lo + span, arrival=t))
    return out
