"""Edge paths of the drop-in engines on the GPU, against the oracle:
classification chains counted from any index and past 64 values (the
reference's Euler run cap is 100, euler_search.py:258-262), AP one-candidate
membership at any height (primorial-ap/ap_gpu.py:637-649), and two
sweep_many generators on one engine at once."""
import random
import sys

import pytest

import fk_native as N
import fk_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def chain_value(fam, v, i, n):
    if fam == N.FAMILY_EULER:
        return v + i * i + i
    if fam == N.FAMILY_DICKSON:
        return (i + 1) * v * v + 1
    return v + i * O.primorial(n)


def oracle_run(fam, v, n, first, cap):
    """Consecutive chain values from index `first` that pass the seven-base
    test (huntlib/primes.py:23-47), up to cap."""
    r = 0
    while r < cap and O.mr_is_prime(chain_value(fam, v, first + r, n)):
        r += 1
    return r


CHAIN_CASES = [
    # (family, n, q2, lo, hi): survivors whose chains are long at the start
    (N.FAMILY_EULER, 6, 1 << 16, 10**12, 10**12 + 10**9),
    (N.FAMILY_DICKSON, 10, 1 << 14, 10**12, 10**12 + (1 << 26)),
    (N.FAMILY_AP, 10, 1 << 14, 4 * 10**13 - 10, 4 * 10**13 - 10 + 10**9),
]


@pytest.mark.parametrize("case", CHAIN_CASES, ids=["euler", "dickson", "ap"])
def test_chains_from_any_index_and_past_64(case):
    fam, n, q2, lo, hi = case
    eng = N.Engine(fam, n, q2, **({"wheel_primes": [2, 3, 5, 7, 11, 13]}
                                  if fam == N.FAMILY_EULER else {}))
    try:
        vals = N.record_values(eng.sweep(lo, hi))
        if fam == N.FAMILY_DICKSON:
            vals = [j * O.wheel_modulus(n) for j in vals]
        assert len(vals) >= 40
        rng = random.Random(5)
        sample = rng.sample(vals, 40)
        for first, cap in ((0, 100), (1, 7), (5, 64), (64, 10), (100, 70),
                           (3, 256)):
            say("chains", fam, first, cap)
            got = eng.classify(sample, N.CLASSIFY_FULL, cap, first=first)
            want = [oracle_run(fam, v, n, first, cap) for v in sample]
            assert got == want, (first, cap)
        if fam == N.FAMILY_AP:
            # the chain of p from index 64 is the chain of p + 64 d
            d = O.primorial(n)
            a = eng.classify(sample, N.CLASSIFY_FULL, 20, first=64)
            b = eng.classify([p + 64 * d for p in sample], N.CLASSIFY_FULL, 20)
            assert a == b
        with pytest.raises(ValueError):
            eng.classify(sample[:1], N.CLASSIFY_FULL, 257)
    finally:
        eng.close()


def test_euler_run_cap_of_100():
    """The drop-in's defaults (cap 100): Euler's 41 runs 40 deep, with no
    clamp; hunt carries runs past the 64 a sweep classifies."""
    import euler_gpu
    eng = euler_gpu.GpuEngine(6, wheel_primes=[2, 3, 5, 7, 11, 13])
    assert eng.mr_run_length(41) == 40 == O.euler_run(41, cap=100)
    assert eng.run_lengths([41, 43, 45], cap=100) == \
        [O.euler_run(p, cap=100) for p in (41, 43, 45)] == [40, 1, 0]
    hits = eng.hunt(10**12, 10**12 + 10**9, cap=100)
    assert len(hits) >= 5
    for p, r in hits:
        assert r == O.euler_run(p, cap=100), p
    with pytest.raises(ValueError):
        eng.mr_run_length(41, cap=300)


def test_ap_survives_matches_oracle_at_any_height():
    import ap_gpu
    n, q2 = 8, 4096
    eng = ap_gpu.GpuEngine(n, q2)
    base = 4 * 10**13 - 10
    offs = [int(o) for o in eng.sweep(base, 30 * 40_000)]
    keep = [base + o for o in offs]
    assert len(keep) >= 30
    rng = random.Random(9)
    probes = keep[:40] + [base + rng.randrange(30 * 40_000)
                          for _ in range(200)]
    # below the engine floor (4096 < p < 10^4, and small sieving primes)
    probes += list(range(4000, 4200)) + [7, 11, 13, 4093, 8191]
    say("ap survives", len(probes))
    for p in probes:
        assert eng.survives(p) == O.ap_survives(n, p, q2), p
    assert all(eng.survives(p) for p in keep[:40])


def test_two_sweep_many_generators_on_one_engine():
    import ladder_gpu
    eng = ladder_gpu.GpuEngine(10, q2=1 << 14)._eng
    wins = [(10**12 + i * (1 << 24), 10**12 + (i + 1) * (1 << 24))
            for i in range(6)]
    alone = [eng.sweep(lo, hi) for lo, hi in wins]
    g1 = eng.sweep_many(wins[:3])
    first = next(g1)
    g2 = list(eng.sweep_many(wins[3:]))          # while g1 is paused
    rest = list(g1)
    got = [first] + rest + g2
    for a, b in zip(got, alone):
        assert N.record_values(a) == N.record_values(b)
