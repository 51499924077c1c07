"""The oracle, pinned against what the reference itself holds.

Every oracle function is checked against something other than itself:
the reference's frozen fingerprints and gate windows, its literature
tables and discovery ledgers, sympy's independent BPSW test, closed forms
from the reference's own comments, and definition-level brute force on tiny
windows.  A planted mistake in the oracle (the last tests) must trip those
checks.
"""
import os
import random

import pytest

import fk_oracle as O
from conftest import big

sympy = pytest.importorskip("sympy")


# --------------------------------------------------------------- MR stage

SPSP_2_3_5_7 = 3215031751          # 151 * 751 * 28351: strong pseudoprime
#                                    to bases 2, 3, 5 and 7


def test_mr_bases_match_reference(golden):
    assert list(O.MR_BASES) == golden["mr"]["MR_BASES"]["value"]
    assert list(O.SMALL) == golden["mr"]["_SMALL"]["value"]


def test_mr_agrees_with_bpsw_below_2_64():
    rng = random.Random(1)
    vals = [rng.randrange(2, 1 << 64) | 1 for _ in range(3000)]
    vals += [int(sympy.nextprime(rng.randrange(1 << 40, 1 << 64)))
             for _ in range(300)]
    vals += list(range(0, 200)) + [SPSP_2_3_5_7, 2047, 561, 1 << 61]
    for v in vals:
        assert O.mr_is_prime(v) == bool(sympy.isprime(v)), v


def test_mr_agrees_with_bpsw_to_192_bits():
    rng = random.Random(2)
    vals = []
    for bits in (65, 80, 100, 128, 150, 180, 192):
        vals += [rng.getrandbits(bits) | 1 for _ in range(60)]
        vals += [int(sympy.nextprime(rng.getrandbits(bits)))
                 for _ in range(20)]
        a = int(sympy.nextprime(rng.getrandbits(bits // 2)))
        b = int(sympy.nextprime(rng.getrandbits(bits // 2)))
        vals.append(a * b)
    for v in vals:
        assert O.mr_is_prime(v) == bool(sympy.isprime(v)), v


def test_sprp_base2_is_half_a_round():
    # a strong pseudoprime to base 2 passes the base-2 test but not the
    # full set; a failure of the base-2 test is always a proof
    assert O.sprp_base2(SPSP_2_3_5_7)
    assert not O.mr_is_prime(SPSP_2_3_5_7)
    rng = random.Random(3)
    for _ in range(2000):
        v = rng.randrange(3, 1 << 70)
        if not O.sprp_base2(v):
            assert not sympy.isprime(v)
        if sympy.isprime(v):
            assert O.sprp_base2(v)
    # the small-prime pre-check answers m == sp, so 2 and 37 are prime
    assert O.sprp_base2(2) and O.sprp_base2(37) and not O.sprp_base2(1)


# ------------------------------------------------------------ run rules

def test_euler_runs_reproduce_literature(golden):
    ref = golden["euler_reference"]
    for n, p in ref["KNOWN"]["value"].items():
        assert O.euler_run(big(p), cap=100) == int(n), (n, p)
    assert O.euler_run(big(ref["A21_UPPER"]["value"])) == 21
    for rec in golden["evidence"]["euler"]:
        p = big(rec["p"])
        assert O.euler_run(p) == rec["run"]
        x = rec["breaker_x"]
        assert (p + x * x + x) % rec["breaker_factor"] == 0


def test_dickson_runs_reproduce_literature(golden):
    for n, k in golden["dickson_reference"]["KNOWN"]["value"].items():
        n, k = int(n), big(k)
        assert O.dickson_run(k, cap=n + 4) >= n, (n, k)
    for rec in golden["evidence"]["dickson"]:
        k = big(rec["k"])
        assert O.dickson_run(k, cap=rec["run"] + 8) == rec["run"]
        assert O.dickson_sprp_run(k, rec["run"] + 8) == rec["run"]
        m = rec["breaker_m"]
        assert (m * k * k + 1) % rec["breaker_factor"] == 0


def test_ap_depths_reproduce_literature(golden):
    for n, p in golden["ap_reference"]["KNOWN"]["value"].items():
        n, p = int(n), big(p)
        assert O.ap_chain_depth_full(p, n) == n, (n, p)
    for rec in golden["evidence"]["ap"]:
        p, n = big(rec["p"]), rec["n"]
        assert O.primorial(n) == big(rec["difference"])
        assert O.ap_chain_depth_full(p, n, cap=64) == rec["depth"]
        assert O.ap_chain_depth(p, n) == min(rec["depth"], n)


def test_two_pass_rules_equal_full_rules_where_recorded():
    # the two-pass classifiers may differ from the full chain only below
    # their exact_from threshold, and never exceed it there
    rng = random.Random(5)
    for _ in range(300):
        k = 2310 * rng.randrange(10**6, 10**9)
        full, two = O.dickson_run(k, 18), O.dickson_sprp_run(k, 18)
        if two >= O.SPRP_EXACT_FROM:
            assert two == full
        else:
            assert full <= two
    for _ in range(300):
        p = rng.randrange(10**6, 10**12) | 1
        full, two = O.ap_chain_depth_full(p, 12), O.ap_chain_depth(p, 12)
        if two >= O.DEPTH_EXACT_FROM:
            assert two == full
        else:
            assert full <= two


# -------------------------------------------------------------- tables

def test_dickson_table_is_direct_divisibility():
    # ladder_search.py G3: Tonelli-Shanks construction == definition
    for n in (6, 10, 13):
        W = O.wheel_modulus(n)
        for q in [q for q in O.primes_below(400) if q >= n + 2]:
            direct = sorted(j for j in range(q)
                            if any((m * (W * j) ** 2 + 1) % q == 0
                                   for m in range(1, n + 1)))
            assert O.dickson_forbidden(q, n) == direct, (n, q)


def test_dickson_table_sizes_closed_form():
    # w_q = 2 * #{m <= n : (-m | q) = +1}; n = 10 values from the table in
    # the survey of ladder_reference.py's SIEVE argument
    want = {13: 10, 17: 10, 19: 8, 23: 6, 29: 12, 31: 4, 37: 12, 41: 14,
            43: 10, 47: 4, 53: 12, 59: 8, 61: 10}
    for q, w in want.items():
        assert len(O.dickson_forbidden(q, 10)) == w
        legendre = sum(1 for m in range(1, 11)
                       if pow((-m) % q, (q - 1) // 2, q) == 1)
        assert w == 2 * legendre


def test_euler_wheel_counts():
    # 13# at n = 8: 180 of 30030; 29# at n = 17: 1,995,840 (31#: 2.99e7,
    # euler_search.py:68); the count is the product of q - |F_q(n)|
    offs, M = O.euler_wheel(8, [2, 3, 5, 7, 11, 13])
    assert (len(offs), M) == (180, 30030)
    direct = [r for r in range(30030)
              if all(r % q not in O.euler_forbidden(q, 8)
                     for q in (2, 3, 5, 7, 11, 13))]
    assert offs == direct
    offs, M = O.euler_wheel(17, O.EULER_WHEEL_29)
    assert len(offs) == 1_995_840 and M == 6469693230
    count31 = len(offs) * (31 - len(O.euler_forbidden(31, 17)))
    assert count31 == 29_937_600
    # nj(37) = 24 / 20 / 18 at n = 13 / 17 / 21 (euler_search.py:467-471)
    for n, nj in ((13, 24), (17, 20), (21, 18)):
        assert 37 - len(O.euler_forbidden(37, n)) == nj


def test_ap_table_is_direct_divisibility():
    for n in (5, 13, 17):
        d = O.primorial(n)
        for q in O.primes_below(200)[3:]:
            direct = sorted(r for r in range(q)
                            if any((r + j * d) % q == 0 for j in range(n)))
            assert O.ap_forbidden(q, n) == direct


# ------------------------------------------------------------ sieve stage

def test_euler_sieve_equals_brute_force():
    for n, lo, span, q2 in ((3, 100_000, 20_000, 150),
                            (2, 123_457, 20_000, 400),
                            (4, 10**12, 20_000, 100)):
        want = O.euler_direct(n, lo, lo + span, q2=q2)
        assert len(want) >= 20
        for wheel in ([2, 3, 5, 7], [2, 3, 5, 7, 11, 13],
                      [2, 3, 5, 7, 11, 13, 17, 19]):
            got = O.euler_survivors(n, lo, lo + span, wheel_primes=wheel,
                                    q1=min(q2, 1024), q2=q2)
            assert got == want, (n, wheel)


def test_dickson_sieve_equals_brute_force():
    for n, j_lo, span, q2 in ((6, 10**6, 20_000, 500),
                              (10, 4 * 10**9, 20_000, 60)):
        want = O.dickson_direct(n, j_lo, j_lo + span, q2=q2)
        assert len(want) >= 20
        assert O.dickson_survivors(n, j_lo, j_lo + span, q2=q2) == want


def test_ap_sieve_equals_brute_force():
    for n, base, span, q2 in ((5, 30_000, 30_000, 300),
                              (8, 10**7 + 10, 60_000, 200)):
        want = O.ap_direct(n, base, span, q2=q2)
        assert len(want) >= 20
        assert O.ap_survivors(n, base, span, q2=q2) == want


def test_ap_gate_windows(golden):
    # primorial-ap/ap_gpu.py:697-701: measured survivor counts
    for w in golden["ap_gate_windows"]:
        got = O.ap_survivors(w["n"], big(w["base"]), big(w["span"]),
                             q2=w["q2"])
        assert len(got) == w["count"], w["name"]


def test_oracle_streams_do_not_depend_on_cuts():
    n, lo, hi = 10, 10**12, 10**12 + 3_000_017
    whole = O.dickson_survivors(n, lo, hi, q2=4096)
    for cut in (lo + 1, lo + 1_234_567, hi - 3):
        assert (O.dickson_survivors(n, lo, cut, q2=4096)
                + O.dickson_survivors(n, cut, hi, q2=4096)) == whole
    assert O.dickson_survivors(n, lo, hi, q2=4096, block=77_777) == whole
    base, span = 10**9 + 20, 6_000_000
    whole = O.ap_survivors(8, base, span, q2=2048)
    assert O.ap_survivors(8, base, span, q2=2048, block=99_991) == whole
    lo, hi = 10**15, 10**15 + 50_000_000
    whole = O.euler_survivors(5, lo, hi, wheel_primes=O.EULER_WHEEL_29[:8])
    mid = lo + 12_345_678
    assert (O.euler_survivors(5, lo, mid, wheel_primes=O.EULER_WHEEL_29[:8])
            + O.euler_survivors(5, mid, hi,
                                wheel_primes=O.EULER_WHEEL_29[:8])) == whole


@pytest.mark.parametrize("label", ["SCORE", "SCORE12"])
def test_dickson_fingerprints(golden, label):
    # dickson-ladders/score.py:54-59: count and XOR of j over 2^32 j
    s = next(x for x in golden["dickson_fingerprints"] if x["label"] == label)
    j0, span = big(s["j0"]), big(s["span"])
    got = O.dickson_survivors(s["n"], j0, j0 + span, q2=s["q2"])
    x = 0
    for j in got:
        x ^= j
    assert (len(got), x) == (s["count"], big(s["xor"]))


def test_euler_first_occurrences(golden):
    # a(n) is the least p >= the floor whose run is exactly n: the sieve at
    # depth n must hold it, and no earlier survivor may have run == n
    known = golden["euler_reference"]["KNOWN"]["value"]
    for n in (9, 11, 12, 13):
        target = big(known[str(n)])
        hits = O.euler_hunt(n, O.EULER_P_FLOOR, target + 1,
                            wheel_primes=[2, 3, 5, 7, 11, 13, 17, 19, 23])
        firsts = [p for p, r in hits if r == n]
        assert firsts and firsts[0] == target, (n, hits[:3])


@pytest.mark.slow
@pytest.mark.skipif(not os.environ.get("FK_SLOW"),
                    reason="1.6e10-integer oracle sweep; set FK_SLOW=1")
def test_ap_fingerprint(golden):
    # primorial-ap/score.py:42-60: count and XOR of the offsets
    s = golden["ap_fingerprints"][0]
    got = O.ap_survivors(s["n"], big(s["base"]), big(s["span"]), q2=s["q2"])
    x = 0
    for o in got:
        x ^= o
    assert (len(got), x) == (s["count"], big(s["xor"]))


# ------------------------------------------------- planted mistakes

def test_planted_mr_mistake_is_caught():
    # an oracle that forgot six of the seven bases passes a strong
    # pseudoprime; the BPSW comparison above would reject it
    assert O.mr_is_prime(SPSP_2_3_5_7, bases=(2,))
    assert bool(sympy.isprime(SPSP_2_3_5_7)) is False


def test_planted_table_mistake_is_caught(monkeypatch):
    # drop one root of the n = 10 Dickson table: brute force disagrees
    n, j_lo, span, q2 = 10, 4 * 10**9, 20_000, 60
    want = O.dickson_direct(n, j_lo, j_lo + span, q2=q2)
    W = O.wheel_modulus(n)
    bad = [j for j in range(j_lo, j_lo + span)
           if all((m * (W * j) ** 2 + 1) % q
                  for q in O.primes_below(q2) if q >= n + 2 and q != 13
                  for m in range(1, n + 1))]
    assert bad != want


@pytest.mark.parametrize("n,base,span", [
    (6, 10**16 - 10**16 % 30, 30 * (1 << 18)),
    (8, 10**15 - 10**15 % 30, 30 * (1 << 21)),
])
def test_ap_direct_division_pins_deeper_streams(n, base, span):
    """The direct-division helper the config-2 GPU test uses: the oracle's
    stream at 2^16, minus the members some prime in [2^16, 2^20) kills,
    is the oracle's stream at 2^20; with P(n - 1) in place of P(n) (a
    plausible slip) it is not."""
    import ap_direct
    s16 = [base + o for o in O.ap_survivors(n, base, span, q2=1 << 16)]
    s20 = [base + o for o in O.ap_survivors(n, base, span, q2=1 << 20)]
    assert len(s20) >= 20 and len(s16) > len(s20)
    dead = ap_direct.kills_between(s16, n, 1 << 16, 1 << 20)
    assert [p for p, k in zip(s16, dead) if not k] == s20
    slip = ap_direct.kills_between(s16, n - 1, 1 << 16, 1 << 20)
    assert [p for p, k in zip(s16, slip) if not k] != s20


def test_ap_survives_is_the_sieve_membership():
    """The one-candidate membership test agrees with the windowed sieve on
    every p of a populated window (survivors and not), and with the
    residue definition below the engines' floor, where p may itself be a
    sieving prime (which the sieve kills: 0 is always a kill residue)."""
    n, q2 = 8, 4096
    base = 4 * 10**13 - 10
    span = 30 * 6_000
    keep = {base + o for o in O.ap_survivors(n, base, span, q2=q2)}
    assert len(keep) >= 10
    for p in range(base, base + span):
        assert O.ap_survives(n, p, q2) == (p in keep), p
    assert not O.ap_survives(n, 4093, q2)          # a sieving prime itself
    assert not O.ap_survives(n, 7, q2)
    small = [p for p in range(5000, 9000) if O.ap_survives(n, p, q2)]
    assert small and all(p % 2 and p % 3 and p % 5 for p in small)
