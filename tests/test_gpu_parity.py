"""GPU parity: the HIP engine's survivor streams and classifications equal
the oracle's, bit for bit, on populated windows of every family, across
heights, launch cuts and tile shapes.  Every comparison asserts a minimum
population (an empty-vs-empty comparison proves nothing)."""
import random
import sys

import numpy as np
import pytest

import fk_oracle as O
import fk_native as N

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------- Dickson

DICKSON_WINDOWS = [
    # (n, j_lo, span, q2, min population): populations measured with the
    # oracle; the deep-sieve windows keep rounds busy, the shallow ones
    # keep the LDS stage populated
    (10, 10**6, 4 * 10**7, 4096, 150),
    (10, 4 * 10**9, 4 * 10**7, 8192, 60),
    (10, 10**12, 1 << 27, 65536, 30),
    (12, 6 * 10**11, 2 * 10**7, 2048, 90),
    (13, 4 * 10**18 - 4 * 10**7, 4 * 10**7, 2048, 40),
    (10, 10**12, 1 << 24, 4096, 60),
    (6, 10**5, 300_000, 300, 1500),
]


@pytest.mark.parametrize("case", DICKSON_WINDOWS)
def test_dickson_stream_matches_oracle(case):
    import ladder_gpu
    n, j_lo, span, q2, need = case
    say("dickson", case)
    eng = ladder_gpu.GpuEngine(n, q2=q2)
    got = eng.survivors_j(j_lo, j_lo + span)
    want = O.dickson_survivors(n, j_lo, j_lo + span, q2=q2)
    assert len(want) >= need
    assert got.dtype == np.uint64
    assert got.tolist() == want


def test_dickson_deep_zone_past_u64():
    import ladder_gpu
    n, q2 = 10, 2048
    eng = ladder_gpu.GpuEngine(n, q2=q2)
    j_lo = eng.J_REACH - 3 * 10**6
    say("dickson deep", j_lo)
    got = eng.survivors_j_deep(j_lo, eng.J_REACH)
    want = O.dickson_survivors(n, j_lo, eng.J_REACH, q2=q2)
    assert j_lo > 1 << 64 and len(want) >= 20
    assert got == want
    with pytest.raises(ValueError):
        eng.survivors_j_deep(j_lo, eng.J_REACH + 1)
    with pytest.raises(ValueError):
        eng.survivors_j(j_lo, eng.J_REACH)


def test_dickson_k_line_view_is_inclusive():
    import ladder_gpu
    eng = ladder_gpu.GpuEngine(10, q2=4096)
    j = int(eng.survivors_j(10**9, 10**9 + 10**6)[3])
    k = j * eng.W
    say("dickson k-line", k)
    assert j in eng.survivors_pre_mr(k, k).tolist()
    assert j in eng.survivors_pre_mr(k - eng.W + 1, k).tolist()
    assert j not in eng.survivors_pre_mr(k + 1, k + eng.W - 1).tolist()


def test_dickson_runs_match_oracle():
    import ladder_gpu
    n, q2 = 10, 2048
    eng = ladder_gpu.GpuEngine(n, q2=q2)
    j_lo, span = 10**12, 1 << 25
    say("dickson runs")
    full = eng.sweep_classified(j_lo, j_lo + span, cap=n + 8,
                                two_pass=False)
    two = eng.sweep_classified(j_lo, j_lo + span, cap=n + 8, two_pass=True)
    assert [j for j, _ in full] == [j for j, _ in two]
    assert len(full) > 300
    rng = random.Random(11)
    sample = rng.sample(range(len(full)), 200)
    for i in sample:
        j, r = full[i]
        k = j * eng.W
        assert r == O.dickson_run(k, cap=n + 8), (j, r)
        assert two[i][1] == O.dickson_sprp_run(k, n + 8), j
    # the deep runs in the window are all exact
    for j, r in full:
        if r >= 4:
            assert r == O.dickson_run(j * eng.W, cap=n + 8)


def test_dickson_evidence_runs_on_device(golden):
    import ladder_gpu
    for rec in golden["evidence"]["dickson"]:
        k, run = int(rec["k"]), rec["run"]
        eng = ladder_gpu.GpuEngine(10, q2=1024)
        say("dickson evidence", k)
        assert eng._eng.classify([k], N.CLASSIFY_FULL, run + 8)[0] == run


# --------------------------------------------------------------- Euler

EULER_WINDOWS = [
    # (n, lo, hi, wheel, min population)
    (5, 10**5, 4 * 10**5, None, 20),
    (5, 10**5, 4 * 10**5, [2, 3, 5, 7, 11, 13], 20),
    (9, 10**6, 6 * 10**7, None, 2),
    (7, 10**15, 10**15 + 10**9, [2, 3, 5, 7, 11, 13, 17], 20),
    (6, 2**64 - 10**8, 2**64 + 10**8, [2, 3, 5, 7, 11, 13], 20),
    (6, 10**24 - 2 * 10**8, 10**24, [2, 3, 5, 7, 11, 13], 20),
]


@pytest.mark.parametrize("case", EULER_WINDOWS)
def test_euler_stream_matches_oracle(case):
    import euler_gpu
    n, lo, hi, wheel, need = case
    say("euler", case)
    eng = euler_gpu.GpuEngine(n, wheel_primes=wheel)
    got = eng.survivors_pre_mr(lo, hi)
    want = O.euler_survivors(n, lo, hi,
                             wheel_primes=[2, 3, 5, 7, 11, 13, 17, 19, 23])
    assert len(want) >= need
    assert got == want


def test_euler_canary_a13_first_occurrence(golden):
    import euler_gpu
    target = int(golden["euler_reference"]["KNOWN"]["value"]["13"])
    say("euler a(13)")
    hits = euler_gpu.GpuEngine(13).hunt(euler_gpu.P_FLOOR, target + 1000)
    firsts = [p for p, r in hits if r == 13]
    assert firsts and firsts[0] == target


def test_euler_runs_match_oracle(golden):
    import euler_gpu
    eng = euler_gpu.GpuEngine(6, wheel_primes=[2, 3, 5, 7, 11, 13])
    lo = 10**12
    say("euler runs")
    recs = eng._eng.sweep(lo, lo + 10**9, N.CLASSIFY_FULL, 11)
    vals = N.record_values(recs)
    assert len(vals) > 100
    for p, r in zip(vals[:300], recs["run"].tolist()[:300]):
        assert r == O.euler_run(p, cap=11), p
    for rec in golden["evidence"]["euler"]:
        p = int(rec["p"])
        assert eng.mr_run_length(p, cap=40) == rec["run"]


# ------------------------------------------------------------------ AP

def test_ap_gate_windows(golden):
    import ap_gpu
    for w in golden["ap_gate_windows"]:
        base, span = int(w["base"]), int(w["span"])
        say("ap", w["name"])
        eng = ap_gpu.GpuEngine(w["n"], w["q2"], launch_u=1 << 16)
        got = np.concatenate([c for c in eng.survivors(base, span)] or
                             [np.zeros(0, np.uint64)])
        want = O.ap_survivors(w["n"], base, span, q2=w["q2"])
        assert len(want) == w["count"]
        assert got.tolist() == want


def test_ap_depths_match_oracle(golden):
    import ap_gpu
    n, q2 = 8, 4096
    eng = ap_gpu.GpuEngine(n, q2)
    base, span = 4 * 10**13 - 10, 3 * 10**6
    say("ap depths")
    hits = eng.hunt(base, span, min_depth=0)
    assert len(hits) > 100
    for p, dep in hits:
        assert dep == O.ap_chain_depth(p, n), p
    for rec in golden["evidence"]["ap"]:
        e = ap_gpu.GpuEngine(rec["n"], 2048)
        p = int(rec["p"])
        assert e.chain_depth(p, cap=40) == O.ap_chain_depth(p, rec["n"],
                                                            cap=40)


# ------------------------------------------------------- slicing

def test_streams_do_not_depend_on_cuts_or_tiles():
    """Lines of 311 positions: the reference stream from the short-line
    sieve, every variant on the line sieve (short_words=1)."""
    import ladder_gpu
    n, q2, lo, hi = 10, 4096, 10**12 + 7, 10**12 + 7 + 30_000_017
    say("slicing")
    ref = ladder_gpu.GpuEngine(n, q2=q2).survivors_j(lo, hi).tolist()
    assert len(ref) > 100
    assert ref == O.dickson_survivors(n, lo, hi, q2=q2)
    for kw in ({},
               {"threads": 512, "words_per_thread": 4},
               {"threads": 1024, "words_per_thread": 1},
               {"threads": 256, "words_per_thread": 16},
               {"q_med": 1024, "q_pat": 64},
               {"q_med": 2048, "q_pat": 256},
               {"nwheel": 1}, {"nwheel": 2},
               {"round_growth": 4, "round_lane_prime": 1},
               {"round_growth": 141}, {"round_growth": 168},
               {"round_lane_prime": 1}, {"round_lane_prime": 1 << 30},
               {"q_bits": 1}, {"q_bits": 300, "q_med": 1024},
               {"q_bits": 1, "q_med": 2048, "q_pat": 64},
               {"round_lane_prime": 1, "q_med": 256},
               {"q_mark": 1024}, {"q_mark": 4096},
               {"q_mark": 256, "threads": 1024, "words_per_thread": 1},
               {"q_mark": 512, "q_med": 3072, "q_pat": 64},
               {"q_pat": 16, "q_bits": 1024, "q_med": 1024}):
        eng = ladder_gpu.GpuEngine(n, q2=q2, short_words=1, **kw)
        assert eng.survivors_j(lo, hi).tolist() == ref, kw
    for bad in (1, 100, 10_000):
        with pytest.raises(ValueError):
            ladder_gpu.GpuEngine(n, q2=q2, round_growth=bad)
    eng = ladder_gpu.GpuEngine(n, q2=q2)
    for cut in (lo + 1, lo + 12_345_678, hi - 1):
        a = eng.survivors_j(lo, cut).tolist()
        b = eng.survivors_j(cut, hi).tolist()
        assert a + b == ref, cut


def test_streams_do_not_depend_on_tile_fit():
    """Windows whose lines are about 45k and 90k positions long, so the
    planner fits them with three- and six-word tiles (256 threads) or
    three-word tiles (512 threads) on the line sieve, against the default
    engine."""
    import ladder_gpu
    n, q2 = 10, 1 << 16
    say("tile fit")
    for npos in (45_000, 90_000):
        lo = 10**12 + 7
        hi = lo + npos * 96_577
        ref = ladder_gpu.GpuEngine(n, q2=q2).survivors_j(lo, hi).tolist()
        assert len(ref) > 1000
        for kw in ({"threads": 256, "words_per_thread": 8},
                   {"threads": 256, "words_per_thread": 6},
                   {"threads": 256, "words_per_thread": 3},
                   {"threads": 512, "words_per_thread": 3}):
            eng = ladder_gpu.GpuEngine(n, q2=q2, short_words=1, **kw)
            assert eng.survivors_j(lo, hi).tolist() == ref, (npos, kw)


# --------------------------------------------- both round kernels

ROUND_CASES = [
    # (label, family, n, q2, lo, hi, engine kwargs, min population)
    ("dickson", N.FAMILY_DICKSON, 10, 1 << 16, 10**12, 10**12 + (1 << 30),
     {}, 250),
    ("dickson-wide", N.FAMILY_DICKSON, 13, 1 << 16, 4 * 10**18 - 10**9,
     4 * 10**18, {}, 12),
    ("euler-wide", N.FAMILY_EULER, 8, 1 << 16, 10**20, 10**20 + 10**9,
     {"wheel_primes": [2, 3, 5, 7, 11, 13]}, 50),
    ("ap", N.FAMILY_AP, 12, 1 << 16, 4 * 10**13 - 10,
     4 * 10**13 - 10 + 3 * 10**8, {}, 80),
    # p past 2^64: the index rows' eight-limb form
    ("ap-wide", N.FAMILY_AP, 12, 1 << 16, 10**20 - 10,
     10**20 - 10 + 3 * 10**8, {}, 60),
]


def round_case_oracle(fam, n, q2, lo, hi, kw):
    if fam == N.FAMILY_DICKSON:
        return O.dickson_survivors(n, lo, hi, q2=q2)
    if fam == N.FAMILY_EULER:
        return O.euler_survivors(n, lo, hi, wheel_primes=kw["wheel_primes"],
                                 q2=q2)
    return [lo + o for o in O.ap_survivors(n, lo, hi - lo, q2=q2)]


@pytest.mark.parametrize("case", ROUND_CASES, ids=[c[0] for c in ROUND_CASES])
def test_round_kernels_match_oracle(case):
    """Lane = prime and lane = candidate rounds (forced by round_lane_prime),
    at small and default LDS cutoffs, against the oracle."""
    label, fam, n, q2, lo, hi, kw, need = case
    say("round kernels", label)
    want = round_case_oracle(fam, n, q2, lo, hi, kw)
    assert len(want) >= need
    for extra in ({"round_lane_prime": 1}, {"round_lane_prime": 1 << 30},
                  {"round_lane_prime": 1, "q_med": 256, "q_pat": 64},
                  {"round_lane_prime": 1 << 30, "q_med": 256, "q_pat": 64},
                  {"q_bits": 1, "q_med": 1024}):
        eng = N.Engine(fam, n, q2, **kw, **extra)
        got = N.record_values(eng.sweep(lo, hi))
        eng.close()
        assert got == want, (label, extra)


# ------------------------------------------------ short lines

SHORT_CASES = [
    # (label, family, n, q2, lo, engine kwargs, words per line): line steps
    # 96577, 2*...*17 = 510510 and 510510; the oracle finds about 6, 40 and
    # 30 survivors per word of line; the Euler engine stride-marks [64, 1024)
    ("dickson", N.FAMILY_DICKSON, 10, 1 << 14, 10**12 + 12_345, {},
     (1, 2, 8, 9, 16, 17, 64, 65)),
    ("ap", N.FAMILY_AP, 12, 1 << 14, 4 * 10**13 - 10, {"nwheel": 7},
     (1, 8, 9, 65)),
    ("euler-marked", N.FAMILY_EULER, 7, 1 << 14, 10**12,
     {"wheel_primes": [2, 3, 5, 7, 11, 13, 17], "q_pat": 64, "q_mark": 1024,
      "q_med": 1024}, (1, 8, 9, 64, 65)),
]


@pytest.mark.parametrize("case", SHORT_CASES, ids=[c[0] for c in SHORT_CASES])
def test_short_lines_match_oracle(case):
    """Windows whose lines hold 1 to 65 words of 64 positions, with ragged
    ends, against the oracle: with short_words=64 lines of up to 64 words
    are sieved one lane per eight words, with short_words=1 every line of
    more than one word goes to the line sieve."""
    label, fam, n, q2, lo0, kw, words_list = case
    short = N.Engine(fam, n, q2, short_words=64, **kw)
    line = N.Engine(fam, n, q2, short_words=1, **kw)
    try:
        assert short.info.short_words == 64 and line.info.short_words == 1
        D = short.line_step
        for words in words_list:
            lo = lo0 + 3 * words
            hi = lo + words * 64 * D - 37 * words
            say("short lines", label, words, lo, hi)
            want = round_case_oracle(fam, n, q2, lo, hi, kw)
            assert len(want) >= 3 * words, (label, words, len(want))
            assert N.record_values(short.sweep(lo, hi)) == want, (label, words)
            assert N.record_values(line.sweep(lo, hi)) == want, (label, words)
    finally:
        short.close()
        line.close()


def test_short_lines_wheel_period_past_32_bits():
    """A wheel of the ten primes 2..29 (period 6469693230 > 2^32, 7.6M
    lines), eight positions per line: the short-line sieve's 64-bit start
    path and the line sieve against the oracle."""
    wp = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]
    n, q2, lo = 7, 1 << 14, 10**12 + 5
    short = N.Engine(N.FAMILY_EULER, n, q2, wheel_primes=wp)
    line = N.Engine(N.FAMILY_EULER, n, q2, wheel_primes=wp, short_words=1)
    try:
        D = short.line_step
        assert D > 1 << 32
        hi = lo + 8 * D - 37
        say("short lines, wide wheel", lo, hi)
        want = O.euler_survivors(n, lo, hi, wheel_primes=wp, q2=q2)
        assert len(want) > 10_000
        assert N.record_values(short.sweep(lo, hi)) == want
        assert N.record_values(line.sweep(lo, hi)) == want
    finally:
        short.close()
        line.close()


# ------------------------------------------- values either side of 2^53

@pytest.mark.parametrize("fam", ["dickson", "ap"])
def test_rounds_either_side_of_2_53(fam):
    """Windows whose values cross 2^53 (Dickson k = 2310 j, AP p): below it
    the round kernels feed the value itself to the residue, above it the
    16-bit limbs; waves hold either or both.  Both round kernels (forced
    by round_lane_prime), and round primes below 32 (a one-prime wheel
    with a small q_med), which must not take the one-value form."""
    if fam == "dickson":
        n, q2 = 10, 1 << 14
        j0 = (1 << 53) // 2310
        lo, hi = j0 - 3 * 10**7, j0 + 3 * 10**7
        want = O.dickson_survivors(n, lo, hi, q2=q2)
        family = N.FAMILY_DICKSON
    else:
        n, q2 = 12, 1 << 14
        lo = (1 << 53) - 6 * 10**8
        lo -= lo % 30
        hi = lo + 12 * 10**8
        want = [lo + o for o in O.ap_survivors(n, lo, hi - lo, q2=q2)]
        family = N.FAMILY_AP
    say("2^53", fam, lo, hi, len(want))
    assert len(want) >= 40
    assert min(want) * (2310 if fam == "dickson" else 1) < 1 << 53
    assert max(want) * (2310 if fam == "dickson" else 1) >= 1 << 53
    small = {"nwheel": 1, "q_med": 24 if fam == "dickson" else 16,
             "q_pat": 16}
    for kw in ({"round_lane_prime": 1 << 30}, {"round_lane_prime": 1},
               small, dict(small, round_lane_prime=1)):
        eng = N.Engine(family, n, q2, **kw)
        try:
            assert N.record_values(eng.sweep(lo, hi)) == want, (fam, kw)
        finally:
            eng.close()
