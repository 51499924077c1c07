"""Prefix wheels: a window whose lines hold few positions on its engine's
wheel is sieved on that wheel without its last primes (which the LDS stage
then sieves) when that is less work.  The records never depend on the
wheel: every window here gives the same records, classification included,
on an engine with prefix wheels as on one without, and the oracle's
survivors on the smaller windows; each family runs some windows on a
prefix wheel and some on its own."""
import sys

import pytest

import fk_families as F
import fk_native as N
import fk_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


Q2 = 1 << 16
EULER_WHEEL = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]

# family -> (engine family, n, engine keywords, classification, heights,
# spans (log2)); heights put the values below and above 2^64
CASES = {
    "dickson": (N.FAMILY_DICKSON, 10, {}, (N.CLASSIFY_FULL, 18, 0),
                (10**9 + 17, 2 * 10**18 + 5), (8, 12, 16, 20, 23, 27, 33)),
    "ap": (N.FAMILY_AP, 8, {}, (N.CLASSIFY_TWO_PASS, 8, 6),
           (4 * 10**13, 3 * 10**19), (10, 14, 18, 21, 24, 27, 32)),
    "euler": (N.FAMILY_EULER, 11, {"wheel_primes": EULER_WHEEL},
              (N.CLASSIFY_FULL, 14, 0),
              (10**12 + 3, 10**20 + 11), (10, 16, 20, 24, 28, 36, 43)),
}


def oracle(fam, n, lo, hi):
    if fam == "dickson":
        return O.dickson_survivors(n, lo, hi, q2=Q2)
    if fam == "euler":
        return O.euler_survivors(n, lo, hi, wheel_primes=[2, 3, 5, 7, 11, 13],
                                 q2=Q2)
    return [lo + o for o in O.ap_survivors(n, lo, hi - lo, q2=Q2)]


@pytest.mark.parametrize("fam", sorted(CASES))
def test_prefix_wheels_give_the_same_records(fam):
    family, n, kw, cls, heights, logs = CASES[fam]
    t = F.tuning(fam)
    eng = N.Engine(family, n, Q2, **kw, **t)
    one = N.Engine(family, n, Q2, **kw, **t, alt_wheels=-1)
    wins = []
    for h in heights:
        for lg in logs:
            lo = h - h % 30 if fam == "ap" else h
            span = 1 << lg
            if fam == "ap":
                span -= span % 30
            wins.append((lo, lo + span))
    say(fam, "n", n, "wheel", eng.line_step, len(wins), "windows")
    specs = [cls] * len(wins)
    got, want = N.Batch(), N.Batch()
    got.submit([eng], [0] * len(wins), wins, specs)
    recs, counts = got.results()
    want.submit([one], [0] * len(wins), wins, specs)
    ref, rcounts = want.results()
    wheels = [got.window_wheel(i) for i in range(len(wins))]
    say(fam, "wheels", wheels)
    assert [want.window_wheel(i) for i in range(len(wins))] == \
        [eng.line_step] * len(wins)
    assert counts.tolist() == rcounts.tolist()
    assert N.record_values(recs) == N.record_values(ref)
    assert recs["window"].tolist() == ref["window"].tolist()
    assert recs["run"].tolist() == ref["run"].tolist()
    assert recs.size >= 200
    # prefix wheels ran (each a divisor of the engine's period) and so did
    # the engine's own wheel
    alt = [d for d in wheels if d != eng.line_step]
    assert alt and all(eng.line_step % d == 0 for d in alt)
    assert eng.line_step in wheels
    # the oracle pins the smaller windows
    pos, checked = 0, 0
    for i, (lo, hi) in enumerate(wins):
        mine = recs[pos:pos + int(counts[i])]
        pos += int(counts[i])
        if hi - lo <= 1 << 24:
            assert N.record_values(mine) == oracle(fam, n, lo, hi), (lo, hi)
            checked += 1
    assert checked >= 6
    # one window per sweep: the engine's own batch picks the same wheels
    for i, (lo, hi) in enumerate(wins[:4]):
        r = eng.sweep(lo, hi, *cls)
        assert eng.last_wheel() == wheels[i]
        assert N.record_values(r) == N.record_values(
            recs[recs["window"] == i])


def test_windows_of_one_or_no_position_per_line():
    """Spans far below the wheel period: most lines of the full wheel hold
    no position, a prefix wheel's hold a few; edges on both sides of a
    period boundary, down to one value and an empty window."""
    t = F.tuning("dickson")
    q2 = 1500                # one short round; populated windows
    eng = N.Engine(N.FAMILY_DICKSON, 10, q2, **t)
    one = N.Engine(N.FAMILY_DICKSON, 10, q2, **t, alt_wheels=-1)
    D = eng.line_step
    base = 10**9 - 10**9 % D
    wins = [(base - 5000, base + 5000), (base + D - 1, base + D + 1),
            (base + 7, base + 8), (base + 9, base + 9)]
    wins += [(base + i * D - (1 << 18), base + i * D + (1 << 18))
             for i in range(4)]
    wins += [(base + i * D + 1, base + (i + 1) * D - 1) for i in range(3)]
    say("edges", D, wins)
    a, b = N.Batch(), N.Batch()
    a.submit([eng], [0] * len(wins), wins, [(0, 0, 0)] * len(wins))
    b.submit([one], [0] * len(wins), wins, [(0, 0, 0)] * len(wins))
    ra, ca = a.results()
    rb, cb = b.results()
    assert ca.tolist() == cb.tolist()
    assert N.record_values(ra) == N.record_values(rb)
    assert a.window_wheel(3) == 0
    pos = 0
    for i, (lo, hi) in enumerate(wins):
        mine = ra[pos:pos + int(ca[i])]
        pos += int(ca[i])
        if hi > lo:
            assert N.record_values(mine) == O.dickson_survivors(10, lo, hi,
                                                                q2=q2)
    assert ra.size >= 100
