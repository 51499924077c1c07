"""Parity at the benchmark configurations' own parameters.

* Euler-form runs, n = 8, the 13# wheel (180 of 30,030 residues), every
  prime below 2^16 sieved, 64 seeded windows of 2^24 consecutive p in
  [1e12, 1e13]: each window's survivors and its hits (run >= 8 over x =
  0..10, cap 11, all seven bases) against the oracle, window by window
  through the drop-in and all 64 in one batch submission.
* Dickson ladders, n = 10, every prime below 2^22 sieved, a full seeded
  window of 2^32 consecutive j near k = 1e15: the survivor stream against
  the oracle's sieve, and every survivor's ladder depth (all seven bases,
  cap 18) against the oracle's Python-integer chain, with the default
  round kernels and each round kernel forced.
* Primorial-difference APs, n = 17, d = 17#, primes 61..2^24 sieved
  (7..59 divide d and kill residue 0):
  - a window around the a(17) evidence value, against the oracle at 2^24,
    with its two-pass depth and its true depth (19);
  - populated windows at n = 5 and 6, primes below 2^24, so that the deep
    rounds have hundreds of candidates to test, both round kernels;
  - full seeded windows of 30 floor(2^32 / 30) integers (near 1e16, and
    one holding a(17)): the oracle at depth 2^16 pins the 2^16 stream on
    the whole window; the 2^24 stream must be exactly the members of it
    that no prime in [2^16, 2^24) divides at any of the 17 values, checked
    by direct division for every prime; cuts and forced round kernels
    must not change it, nor the queue sizes after each stage.
"""
import os
import sys

import numpy as np
import pytest

import ap_direct
import fk_native as N
import fk_oracle as O

pytestmark = pytest.mark.gpu

MASK64 = (1 << 64) - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def threads():
    return max(1, min(16, os.cpu_count() or 1))


LANE_PRIME = {"round_lane_prime": 1}
LANE_CAND = {"round_lane_prime": 1 << 30}

# ----------------------------------------------------------------- Euler

C0_N, C0_CAP, C0_SPAN, C0_WHEEL = 8, 11, 1 << 24, [2, 3, 5, 7, 11, 13]


def c0_window(seed):
    lo = 10**12 + splitmix64(seed) % (9 * 10**12)
    return lo, lo + C0_SPAN


def test_config0_windows_match_oracle():
    import euler_gpu
    eng = euler_gpu.GpuEngine(C0_N, wheel_primes=C0_WHEEL)
    assert (eng.n_offs, eng.M) == (180, 30030)
    wins = [c0_window(s) for s in range(64)]
    say("config 0", len(wins), "windows")
    want_s, want_h = [], []
    for lo, hi in wins:
        surv = O.euler_survivors(C0_N, lo, hi, wheel_primes=C0_WHEEL,
                                 q2=1 << 16, nthreads=threads())
        runs = [O.euler_run(p, cap=C0_CAP) for p in surv]
        want_s.append(surv)
        want_h.append([(p, r) for p, r in zip(surv, runs) if r >= C0_N])
    for (lo, hi), ws, wh in zip(wins, want_s, want_h):
        assert eng.survivors_pre_mr(lo, hi) == ws, lo
        assert eng.hunt(lo, hi, cap=C0_CAP) == wh, lo
    # all 64 windows in one batch: the same streams and runs
    b = N.Batch()
    b.submit([eng._eng], [0] * len(wins), wins,
             [(N.CLASSIFY_FULL, C0_CAP, 0)] * len(wins))
    recs, counts = b.results()
    pos = 0
    for i, ws in enumerate(want_s):
        mine = recs[pos:pos + int(counts[i])]
        pos += int(counts[i])
        assert N.record_values(mine) == ws
        assert mine["run"].tolist() == [O.euler_run(p, cap=C0_CAP)
                                        for p in ws]
    assert sum(len(w) for w in want_s) >= 40
    assert sum(len(w) for w in want_h) >= 1


# --------------------------------------------------------------- Dickson

C1_N, C1_Q2, C1_SPAN, C1_CAP = 10, 1 << 22, 1 << 32, 18


def c1_window(seed):
    j_lo = 432_900_432_900 + splitmix64(seed) % 10**11
    return j_lo, j_lo + C1_SPAN


@pytest.fixture(scope="module")
def c1_oracle():
    """The oracle's survivors and runs of the seed-7 window."""
    lo, hi = c1_window(7)
    say("config 1 oracle", lo, hi)
    js = O.DicksonSieve(C1_N, C1_Q2).survivors(lo, hi, nthreads=threads())
    W = O.wheel_modulus(C1_N)
    runs = [O.dickson_run(j * W, cap=C1_CAP) for j in js]
    return lo, hi, js, runs


@pytest.mark.parametrize("kw", [{}, LANE_PRIME, LANE_CAND],
                         ids=["default", "lane-prime", "lane-candidate"])
def test_config1_full_window_matches_oracle(c1_oracle, kw):
    import fk_families as F
    lo, hi, js, runs = c1_oracle
    assert len(js) >= 25
    t = F.tuning("dickson")
    t.update(kw)
    eng = N.Engine(N.FAMILY_DICKSON, C1_N, C1_Q2, **t)
    try:
        say("config 1 sweep", kw)
        recs = eng.sweep(lo, hi, N.CLASSIFY_FULL, C1_CAP)
        assert N.record_values(recs) == js
        assert [int(r) for r in recs["run"].tolist()] == runs
        # the rounds ran at the configuration's depth
        counts = eng.last_counts()
        assert len(counts) == 1 + eng.info.n_rounds >= 8
        assert counts[-1] == len(js)
    finally:
        eng.close()


def test_config1_split_equals_whole(c1_oracle):
    import ladder_gpu
    lo, hi, js, _ = c1_oracle
    eng = ladder_gpu.GpuEngine(C1_N, q2=C1_Q2)
    cuts = [lo, lo + 1, lo + 987_654_321, lo + (1 << 31) + 12_345,
            hi - 2310, hi]
    got = []
    for a, b in zip(cuts, cuts[1:]):
        say("config 1 piece", a, b)
        got += eng.survivors_j(a, b).tolist()
    assert got == js


# -------------------------------------------------------------------- AP

C2_N, C2_Q2 = 17, 1 << 24
C2_SPAN = 30 * ((1 << 32) // 30)
A17_P = 2097209048106247          # primorial-ap evidence: a(17), 19 deep


def ap_engine(n, q2, **kw):
    import fk_families as F
    t = F.tuning("ap")
    t.update(kw)
    return N.Engine(N.FAMILY_AP, n, q2, **t)


@pytest.mark.parametrize("kw", [{}, LANE_PRIME, LANE_CAND],
                         ids=["default", "lane-prime", "lane-candidate"])
def test_config2_window_around_a17(kw):
    base = A17_P - A17_P % 30 - 30 * (1 << 20)
    span = 30 * (1 << 21)
    say("config 2 around a(17)", base, span, kw)
    want = [base + o for o in O.ap_survivors(C2_N, base, span, q2=C2_Q2,
                                             nthreads=threads())]
    assert A17_P in want
    eng = ap_engine(C2_N, C2_Q2, **kw)
    try:
        recs = eng.sweep(base, base + span, N.CLASSIFY_TWO_PASS, C2_N, 6)
        got = N.record_values(recs)
        assert got == want
        runs = dict(zip(got, recs["run"].tolist()))
        for p in want:
            assert runs[p] == O.ap_chain_depth(p, C2_N), p
        assert runs[A17_P] == C2_N
        # the true depth, past the cap the campaign classifies with
        assert eng.classify([A17_P], N.CLASSIFY_FULL, 40)[0] == 19
        assert O.ap_chain_depth_full(A17_P, C2_N, cap=40) == 19
    finally:
        eng.close()


AP_DEEP_POPULATED = [
    # (n, base, span): hundreds of survivors at 2^24
    (5, 10**16 - 10**16 % 30, 30 * (1 << 20)),
    (6, 10**16 - 10**16 % 30, 30 * (1 << 22)),
    (6, 10**20 - 10**20 % 30, 30 * (1 << 22)),       # past 2^64
]


@pytest.mark.parametrize("case", AP_DEEP_POPULATED,
                         ids=["n5", "n6", "n6-wide"])
def test_ap_deep_rounds_populated(case):
    n, base, span = case
    say("ap deep populated", case)
    want = [base + o for o in O.ap_survivors(n, base, span, q2=C2_Q2,
                                             nthreads=threads())]
    assert len(want) >= 150
    for kw in ({}, LANE_PRIME, LANE_CAND,
               dict(LANE_PRIME, q_med=256, q_pat=64, nwheel=3)):
        eng = ap_engine(n, C2_Q2, **kw)
        try:
            recs = eng.sweep(base, base + span, N.CLASSIFY_TWO_PASS, n, 6)
            assert N.record_values(recs) == want, kw
            if not kw:
                for p, r in list(zip(want, recs["run"].tolist()))[:60]:
                    assert r == O.ap_chain_depth(p, n), p
        finally:
            eng.close()


def c2_window(seed):
    base = 10**16 + splitmix64(seed) % 10**15
    base -= base % 30
    return base


C2_FULL = [("near-1e16", c2_window(0)),
           ("holding-a17", A17_P - A17_P % 30 - 30 * (1 << 27))]


@pytest.mark.parametrize("case", C2_FULL, ids=[c[0] for c in C2_FULL])
def test_config2_full_window(case):
    label, base = case
    hi = base + C2_SPAN
    say("config 2 full window", label, base)
    want16 = [base + o for o in O.ap_survivors(C2_N, base, C2_SPAN,
                                               q2=1 << 16,
                                               nthreads=threads())]
    assert len(want16) >= 10
    shallow = ap_engine(C2_N, 1 << 16)
    try:
        assert N.record_values(shallow.sweep(base, hi)) == want16
    finally:
        shallow.close()
    dead = ap_direct.kills_between(want16, C2_N, 1 << 16, C2_Q2)
    want = [p for p, k in zip(want16, dead) if not k]
    assert sum(dead) >= 10
    if label == "holding-a17":
        assert want == [A17_P]
    ref_counts = None
    for kw in ({}, LANE_PRIME, LANE_CAND):
        eng = ap_engine(C2_N, C2_Q2, **kw)
        try:
            say("config 2 full window", label, kw)
            assert N.record_values(eng.sweep(base, hi)) == want, kw
            counts = eng.last_counts()
            # the table's rounds, the late ones of few candidates joined
            # (fewer stages), and the queue shrinking stage by stage
            assert 2 <= len(counts) <= 1 + eng.info.n_rounds
            assert all(x >= y for x, y in zip(counts, counts[1:]))
            if ref_counts is None:
                ref_counts = counts
                assert counts[0] > 1000 and counts[-1] == len(want)
            assert counts == ref_counts, kw
            if not kw:
                cuts = [base, base + 30 * 12_345_677, base + 30 * 99_999_999,
                        hi - 30, hi]
                got = []
                for a, b in zip(cuts, cuts[1:]):
                    got += N.record_values(eng.sweep(a, b))
                assert got == want
        finally:
            eng.close()
