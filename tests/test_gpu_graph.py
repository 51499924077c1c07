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
