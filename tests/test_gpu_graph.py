"""Graph mode of the batched engine: a submission whose launch sequence
repeats is captured once into a HIP graph (on its second submission in a
row) and replayed.  Every replay is
checked against the eager batch and the plain engine on the same windows,
with eager device work and synchronisations between the replays, and the
capture and launch counts show that the replays really were replays."""
import sys

import pytest

import fk_families as F
import fk_native as N

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SPEC = (N.CLASSIFY_TWO_PASS, 18, 5)


def records(b):
    recs, counts = b.results()
    return (N.record_values(recs), recs["run"].tolist(),
            [int(c) for c in counts])


def test_graph_replays_equal_eager():
    eng = N.Engine(N.FAMILY_DICKSON, 10, 1 << 16, **F.tuning("dickson"))
    span = 1 << 28
    windows = [(10**12 + k * 7_000_000_011, 10**12 + k * 7_000_000_011 + span)
               for k in range(5)]
    g = N.Batch(graph=True)
    eager = N.Batch()
    total = 0
    for k, (lo, hi) in enumerate(windows):
        say("graph replay", k)
        g.submit([eng], [0], [(lo, hi)], [SPEC])
        got = records(g)
        eager.submit([eng], [0], [(lo, hi)], [SPEC])
        want = records(eager)
        assert got == want, k
        plain = eng.sweep(lo, hi, *SPEC)          # other work in between
        assert got[0] == N.record_values(plain)
        assert got[1] == plain["run"].tolist()
        total += len(got[0])
        ms, _, _ = g.timing()
        assert ms[4] > 0
    assert total >= 20
    captures, launches = g.graph_stats()
    # the first submission runs eagerly, the second is captured and every
    # later one replays
    assert (captures, launches) == (1, len(windows) - 1)


def test_graph_mixed_batch_repeats():
    engs = [N.Engine(N.FAMILY_DICKSON, 10, 1 << 16, **F.tuning("dickson")),
            N.Engine(N.FAMILY_AP, 12, 1 << 16, **F.tuning("ap")),
            N.Engine(N.FAMILY_EULER, 8, 1 << 16,
                     wheel_primes=[2, 3, 5, 7, 11, 13], **F.tuning("euler"))]
    ws = [(10**12, 10**12 + (1 << 26)),
          (4 * 10**13 - 10, 4 * 10**13 - 10 + 30 * (1 << 23)),
          (10**20, 10**20 + 10**9)]
    specs = [SPEC, (N.CLASSIFY_TWO_PASS, 12, 6), (N.CLASSIFY_FULL, 11, 0)]
    eager = N.Batch()
    eager.submit(engs, [0, 1, 2], ws, specs)
    want = records(eager)
    assert len(want[0]) >= 20
    g = N.Batch(graph=True)
    for k in range(4):
        say("mixed graph", k)
        g.submit(engs, [0, 1, 2], ws, specs)
        assert records(g) == want, k
        engs[0].sweep(*ws[0], *SPEC)               # other work in between
    captures, launches = g.graph_stats()
    assert (captures, launches) == (1, 3)


def test_engine_sweeps_replay_equal_eager():
    """The engine's own sweeps replay a repeated launch sequence as a graph
    (captured on the second sweep of one window shape); with stage timing
    on they run eagerly.  Both give the same records on the same windows."""
    span = 1 << 27
    windows = [(10**13 + k * 3_000_000_019, 10**13 + k * 3_000_000_019 + span)
               for k in range(4)]
    graph = N.Engine(N.FAMILY_DICKSON, 10, 1 << 16, **F.tuning("dickson"))
    eager = N.Engine(N.FAMILY_DICKSON, 10, 1 << 16, **F.tuning("dickson"))
    eager.stage_timing(True)
    total = 0
    for k, (lo, hi) in enumerate(windows):
        say("engine replay", k)
        got = graph.sweep(lo, hi, *SPEC)
        want = eager.sweep(lo, hi, *SPEC)
        assert N.record_values(got) == N.record_values(want), k
        assert got["run"].tolist() == want["run"].tolist(), k
        ms, _, _ = eager.last_timing()
        assert ms[1] > 0                           # stage events recorded
        total += len(got)
    assert total >= 10


def test_pipelined_sweeps_equal_plain():
    """Engine.sweep_many keeps four windows in flight on four batches: equal
    windows (each slot captures its launch sequence at once and replays
    it), windows of mixed lengths (eager until a sequence repeats), a
    caller that stops early, and a mixed-engine Pipeline; every window's
    records equal the plain engine's."""
    eng = N.Engine(N.FAMILY_DICKSON, 10, 1 << 16, **F.tuning("dickson"))
    ap = N.Engine(N.FAMILY_AP, 9, 1 << 16, **F.tuning("ap"))
    span = 1 << 27
    equal = [(10**12 + k * 3_000_000_019, 10**12 + k * 3_000_000_019 + span)
             for k in range(7)]
    mixed = [(lo, lo + (span >> (k % 3))) for k, (lo, _) in enumerate(equal)]
    for name, wins in (("equal", equal), ("mixed", mixed)):
        say("pipelined", name)
        got = list(eng.sweep_many(wins, *SPEC))
        assert len(got) == len(wins)
        for (lo, hi), recs in zip(wins, got):
            plain = eng.sweep(lo, hi, *SPEC)
            assert N.record_values(recs) == N.record_values(plain), (name, lo)
            assert recs["run"].tolist() == plain["run"].tolist()
        if name == "equal":
            # slot 0 swept windows 0, d, 2d, ...: captured at once, replayed
            d = eng._pipe.depth
            assert eng._pipe._slots[0].graph_stats() == (1, len(range(0, len(wins), d)))
    say("pipelined early stop")
    it = eng.sweep_many(equal, *SPEC)
    first = next(it)
    it.close()                                    # one window still in flight
    assert N.record_values(first) == N.record_values(eng.sweep(*equal[0], *SPEC))
    again = list(eng.sweep_many(equal[:3], *SPEC))
    assert [N.record_values(r) for r in again] == \
        [N.record_values(eng.sweep(lo, hi, *SPEC)) for lo, hi in equal[:3]]
    say("pipelined mixed engines")
    pipe = N.Pipeline()
    aps = (N.CLASSIFY_FULL, 12, 0)
    jobs = []
    for k in range(6):
        if k % 2:
            lo = 4 * 10**13 + k * 10**9
            jobs.append((ap, lo, lo + (1 << 26), aps))
        else:
            jobs.append((eng, *equal[k], SPEC))
    out = list(pipe.run(jobs))
    assert [i for i, _ in out] == list(range(len(jobs)))
    for (e, lo, hi, spec), (_, recs) in zip(jobs, out):
        plain = e.sweep(lo, hi, *spec)
        assert N.record_values(recs) == N.record_values(plain)
        assert recs["run"].tolist() == plain["run"].tolist()
    pipe.close()
    eng.close()
    ap.close()
