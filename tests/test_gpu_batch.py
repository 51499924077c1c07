"""The batched engine: many windows of many engines (all three families,
several depths, windows from 2^14 to 2^28 positions so that several tile
shapes and per-engine round schedules run in one submission) give exactly
the records of per-window plain sweeps, which the oracle pins on the
smaller windows; submission is asynchronous and a busy batch refuses a
second submission."""
import random
import sys
import time

import numpy as np
import pytest

import fk_families as F
import fk_native as N
import fk_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


FAM = {"euler": N.FAMILY_EULER, "dickson": N.FAMILY_DICKSON,
       "ap": N.FAMILY_AP}
Q2 = {"euler": 1 << 16, "dickson": 1 << 16, "ap": 1 << 16}


def spec(fam, n):
    if fam == "ap":
        return (N.CLASSIFY_TWO_PASS, n, 6)
    if fam == "dickson":
        return (N.CLASSIFY_TWO_PASS, n + 8, 5)
    return (N.CLASSIFY_FULL, n + 3, 0)


def mixed_windows(seed, count):
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        fam = rng.choice(sorted(FAM))
        n = rng.randint(8, 12)
        length = int(2 ** rng.uniform(14, 28))
        wide = rng.random() < 0.4
        if fam == "euler":
            lo = rng.randrange(10**12, 10**13) if not wide else \
                rng.randrange(10**20, 10**21)
        elif fam == "dickson":
            lo = rng.randrange(10**9, 10**11) if not wide else \
                rng.randrange(10**18, 3 * 10**18)
        else:
            lo = rng.randrange(10**13, 10**14) if not wide else \
                rng.randrange(10**19, 10**20)
            lo -= lo % 30
        out.append((fam, n, lo, lo + length))
    return out


@pytest.fixture(scope="module")
def engines():
    cache = {}

    def get(fam, n):
        if (fam, n) not in cache:
            kw = {"wheel_primes": [2, 3, 5, 7, 11, 13]} if fam == "euler" else {}
            cache[(fam, n)] = N.Engine(FAM[fam], n, Q2[fam], **kw,
                                       **F.tuning(fam))
        return cache[(fam, n)]
    return get


def test_mixed_batch_equals_plain_and_oracle(engines):
    wins = mixed_windows(11, 60)
    engs, index = [], {}
    for fam, n, _, _ in wins:
        if (fam, n) not in index:
            index[(fam, n)] = len(engs)
            engs.append(engines(fam, n))
    eo = [index[(f, n)] for f, n, _, _ in wins]
    ws = [(lo, hi) for _, _, lo, hi in wins]
    specs = [spec(f, n) for f, n, _, _ in wins]
    say("mixed batch", len(wins), "windows", len(engs), "engines")
    b = N.Batch()
    b.submit(engs, eo, ws, specs)
    recs, counts = b.results()
    assert int(counts.sum()) == recs.size
    assert recs.size >= 100
    pos = 0
    checked = 0
    for i, (fam, n, lo, hi) in enumerate(wins):
        mine = recs[pos:pos + int(counts[i])]
        pos += int(counts[i])
        assert (mine["window"] == i).all()
        plain = engs[eo[i]].sweep(lo, hi, *specs[i])
        assert N.record_values(mine) == N.record_values(plain), wins[i]
        assert mine["run"].tolist() == plain["run"].tolist(), wins[i]
        if hi - lo <= 1 << 22:
            if fam == "dickson":
                want = O.dickson_survivors(n, lo, hi, q2=Q2[fam])
            elif fam == "euler":
                want = O.euler_survivors(n, lo, hi,
                                         wheel_primes=[2, 3, 5, 7, 11, 13],
                                         q2=Q2[fam])
            else:
                want = [lo + o for o in O.ap_survivors(n, lo, hi - lo,
                                                       q2=Q2[fam])]
            assert N.record_values(mine) == want, wins[i]
            checked += 1
    assert checked >= 10
    # the batch ran several tile shapes and every family
    assert len({f for f, _, _, _ in wins}) == 3


def test_submit_is_asynchronous_and_busy_is_refused(engines):
    eng = engines("dickson", 10)
    lo = 10**12
    b = N.Batch()
    b.submit([eng], [0], [(lo, lo + (1 << 30))], [spec("dickson", 10)])
    with pytest.raises(RuntimeError):
        b.submit([eng], [0], [(lo, lo + 10)], [spec("dickson", 10)])
    t0 = time.time()
    while not b.done():
        assert time.time() - t0 < 60
        time.sleep(0.001)
    recs, counts = b.results()
    plain = eng.sweep(lo, lo + (1 << 30), *spec("dickson", 10))
    assert N.record_values(recs) == N.record_values(plain)
    assert recs.size > 50 and int(counts[0]) == recs.size
