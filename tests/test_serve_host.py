"""The serving loop's scheduling, on the CPU: stand-in engines and batches
(results from the oracle, completion after a modelled device time on a
simulated clock) exercise admission by footprint, continuous batching,
overlap of slots, no starvation and the latency statistics.  The same loop
on the device is tested in test_gpu_serve.py."""
import math

import numpy as np
import pytest

import fk_native as N
import fk_oracle as O
import fk_serve as S


class Info:
    nlines, line_step, survival_lds = 21, 221, 0.05


class StubEngine:
    def __init__(self, family, n, q2):
        self.family, self.n, self.q2 = family, n, q2
        self.info = Info()


class Clock:
    def __init__(self, dt=1e-4):
        self.t, self.dt = 0.0, dt

    def __call__(self):
        self.t += self.dt
        return self.t


class StubBatch:
    """Finishes `cost` seconds of simulated time after submission; records
    are the oracle's Dickson survivors of each window."""

    def __init__(self, clock, cost_per_window=2e-3):
        self.clock = clock
        self.cost = cost_per_window
        self.ready = None

    def submit(self, engines, engine_of, windows, specs):
        assert self.ready is None, "submitted while busy"
        self.windows = windows
        self.engs = [engines[i] for i in engine_of]
        self.ready = self.clock.t + self.cost * len(windows)

    def done(self):
        return self.clock.t >= self.ready

    def results(self):
        parts, counts = [], []
        for (lo, hi), e in zip(self.windows, self.engs):
            js = O.dickson_survivors(e.n, lo, hi, q2=e.q2)
            r = np.zeros(len(js), dtype=N.RECORD_DTYPE)
            r["v0"] = np.array(js, dtype=np.uint64)
            parts.append(r)
            counts.append(len(js))
        self.ready = None
        return (np.concatenate(parts) if parts else
                np.zeros(0, dtype=N.RECORD_DTYPE)), np.array(counts)


def server(slots, budget, clock, cost=2e-3):
    return S.Server(slots=slots, budget=budget, clock=clock, idle_sleep=0,
                    engine_factory=lambda f, n, q: StubEngine(f, n, q),
                    batch_factory=lambda: StubBatch(clock, cost))


def dickson_requests(count, gap, span=1 << 16, n=10):
    return [S.Request(i, "dickson", n, 10**9 + i * span, 10**9 + (i + 1) * span,
                      q2=4096, arrival=i * gap) for i in range(count)]


def test_every_request_gets_its_alone_result():
    clock = Clock()
    srv = server(2, 1 << 40, clock)
    reqs = dickson_requests(12, 5e-4)
    st = srv.run(reqs)
    assert st["completed"] == 12
    for r in reqs:
        want = O.dickson_survivors(10, r.lo, r.hi, q2=4096)
        assert N.record_values(r.records) == want
        assert r.latency > 0


def test_requests_arriving_while_busy_are_batched_together():
    clock = Clock()
    srv = server(1, 1 << 40, clock, cost=5e-3)
    reqs = dickson_requests(10, 1e-4)
    srv.run(reqs)
    # the first submission takes what has arrived; everything arriving
    # while it runs waits and goes in the next one together
    assert len(srv.submissions) == 2 and sum(srv.submissions) == 10
    assert srv.submissions[1] >= 5
    late = [r for r in reqs if r.admitted > reqs[0].admitted]
    assert all(r.admitted >= reqs[0].done for r in late)


def test_two_slots_overlap():
    clock = Clock()
    srv = server(2, 1 << 40, clock, cost=5e-3)
    reqs = dickson_requests(2, 1e-4)
    srv.run(reqs)
    # the second started before the first finished
    assert reqs[1].admitted < reqs[0].done


def test_admission_respects_the_memory_budget():
    clock = Clock()
    one = S.footprint(Info(), 1 << 16)
    srv = server(4, int(2.5 * one), clock)
    reqs = dickson_requests(9, 0.0)
    srv.run(reqs)
    assert max(srv.submissions) <= 2
    assert sum(srv.submissions) == 9


def test_oversized_request_is_admitted_alone():
    clock = Clock()
    srv = server(2, 1000, clock)             # smaller than any footprint
    reqs = dickson_requests(3, 0.0)
    st = srv.run(reqs)
    assert st["completed"] == 3 and srv.submissions == [1, 1, 1]


def test_rejects_bad_requests():
    srv = server(1, 1 << 30, Clock())
    with pytest.raises(ValueError):
        srv.submit(S.Request(0, "dickson", 10, 5, 5))
    with pytest.raises(ValueError):
        srv.submit(S.Request(0, "nope", 10, 5, 9, q2=4096))


def test_percentiles_and_poisson_stream():
    assert S.percentile([3, 1, 2], 0.5) == 2
    assert math.isclose(S.percentile(list(range(101)), 0.99), 99.0)
    reqs = S.poisson_requests(4, 200, rate=1000.0)
    gaps = np.diff([r.arrival for r in reqs])
    assert (gaps > 0).all() and 0.5e-3 < gaps.mean() < 2e-3
    assert {r.family for r in reqs} == {"euler", "dickson", "ap"}
    assert all(8 <= r.n <= 14 and r.hi > r.lo for r in reqs)
    assert all(r.lo % 30 == 0 for r in reqs if r.family == "ap")
