"""Verification legs and primality certificates (first-knowns_amd/
fk_verify.py, fk_certificate.py), on the CPU.

Pinned three ways:
* against the verification outputs the reference recorded with its own
  discoveries (tests/golden: the evidence ledgers of the three hunts) --
  Dickson's BLS75 witnesses for every m, its factorizations of k, the
  breakers and their factors, AP's three legs, true depths and proof
  routes;
* against the known-answer samples of the reference's certificate gate
  (huntlib/certificate.py:376-503): a Theorem 1 prime, a Theorem 5 prime
  with only the smooth part of N-1 given, the constructed composite whose
  split must give r^2 - 8s = (a - b)^2, doctored proofs, and the
  subproof recursion;
* against the mathematics: PSI_13 is a strong pseudoprime to all 13 prime
  bases (so the deterministic bound is exactly right), the deterministic
  test agrees with sympy below it, rho splits known semiprimes.
"""
import json
import math
import pathlib
import random

import pytest

import fk_certificate as C
import fk_verify as V

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "reference_vectors.json"
EV = json.loads(GOLDEN.read_text())["evidence"]

# huntlib/certificate.py:376-380: F smooth and even, a prime and a
# composite control in the Theorem 5 range
F5_FAC = {2: 5, 3: 3, 5: 2, 7: 1, 11: 1, 13: 1, 17: 1, 19: 1, 23: 1}
F5 = math.prod(p ** e for p, e in F5_FAC.items())
N5_PRIME = 12131403568136753874573601
RECUR = [10**25 + d for d in (1237, 1291, 1417, 1447, 1483, 1531, 1567,
                              1621, 1747, 1801, 1879, 1951)]


# ------------------------------------------------- the recorded discoveries

@pytest.mark.parametrize("rec", EV["dickson_full"], ids=lambda r: f"run{r['run']}")
def test_dickson_evidence_matches_the_reference(rec):
    ev, why = V.dickson_verify(int(rec["k"]), rec["run"], rec["run"])
    assert why == "ok"
    assert ev["values_prime_m"] == rec["values_prime_m"]
    assert ev["k_factorization"] == rec["k_factorization"]
    assert (ev["breaker_m"], ev["breaker_factor"]) == \
        (rec["breaker_m"], rec["breaker_factor"])
    assert ev["certificates_bls75"] == rec["certificates_bls75"]


@pytest.mark.parametrize("rec", EV["euler_full"], ids=lambda r: f"run{r['run']}")
def test_euler_evidence_matches_the_reference(rec):
    ev, why = V.euler_verify(int(rec["p"]), rec["run"], 8)
    assert why == "ok"
    assert ev["values_prime_x"] == rec["values_prime_x"]
    assert (ev["breaker_x"], str(ev["breaker"]), ev["breaker_factor"]) == \
        (rec["breaker_x"], str(rec["breaker"]), rec["breaker_factor"])


@pytest.mark.parametrize("rec", EV["ap_full"], ids=lambda r: f"n{r['n']}")
def test_ap_evidence_matches_the_reference(rec):
    ev = V.ap_verify(int(rec["p"]), rec["n"], rec["n"], q2=8192)
    assert ev["agree"] and rec["agree"]
    assert ev["legs"] == rec["legs"]
    assert ev["depth"] == rec["depth"]
    assert str(ev["difference"]) == rec["difference"]
    assert [str(v) for v in ev["values"]] == [str(v) for v in rec["values"]]
    assert [(p["proof"], str(p["N"])) for p in ev["proofs"]] == \
        [(p["proof"], str(p["N"])) for p in rec["proofs"]]
    assert all(C.verify(p)[0] for p in ev["proofs"])


def test_wrong_claims_are_refused():
    d = EV["dickson_full"][0]
    k, run = int(d["k"]), d["run"]
    assert V.dickson_verify(k, run + 1, run)[0] is None
    assert V.dickson_verify(k + 2310, run, run)[0] is None   # another k
    e = EV["euler_full"][0]
    assert V.euler_verify(int(e["p"]), e["run"] - 1, 8)[0] is None
    a = EV["ap_full"][0]
    bad = V.ap_verify(int(a["p"]), a["n"], a["n"] - 1, q2=8192)   # exact claim
    assert not bad["agree"]
    W = V.dickson_wheel(10)
    assert k % W == 0
    ok, ev = V.verify_record(2, 10, k // W, run, 1 << 16)
    assert ok and ev["k"] == k
    ok, why = V.verify_record(2, 10, k // W, run + 1, 1 << 16)
    assert not ok and "disagreement" in why


# ------------------------------------------------- the certificate gate

def test_deterministic_route_and_its_bound():
    small = C.prove(122774401)                   # 2^7 3^3 5^2 7^2 29 + 1
    assert small["proof"] == "deterministic-mr" and C.verify(small)[0]
    assert C.prove(122774403) is None
    assert not C.verify(dict(small, N=10**30 + 57))[0]
    # PSI_13 passes all 13 bases and is composite: the bound is exact
    assert all(C.strong_probable_prime(C.PSI_13, a) for a in C.MR_PROOF_BASES)
    from sympy import isprime
    assert not isprime(C.PSI_13)
    assert not C.proved_prime(C.PSI_13)
    rng = random.Random(5)
    for _ in range(400):
        n = rng.randrange(3, C.PSI_13) | 1
        assert C.proved_prime(n) == isprime(n), n


def test_theorem1_route():
    n1 = 2 * 3**2 * 5**18 * 7**14 + 1
    assert n1 > C.PSI_13
    p1 = C.prove(n1, fac={2: 1, 3: 2, 5: 18, 7: 14})
    assert p1["proof"] == "bls75-thm1"
    ok, msg = C.verify(p1)
    assert ok and "Theorem 1" in msg


def test_theorem5_route_and_composite_control():
    assert N5_PRIME > C.PSI_13 and F5 ** 3 > N5_PRIME >= F5 * F5
    p5 = C.prove(N5_PRIME, fac=dict(F5_FAC))
    assert p5["proof"] == "bls75-thm5" and C.verify(p5)[0]
    a, b = 4, 33
    nc = (1 + a * F5) * (1 + b * F5)
    assert nc > C.PSI_13 and F5 ** 3 > nc and F5 <= math.isqrt(nc)
    s, r = C.theorem5_split(F5, (nc - 1) // F5)
    assert (r, s) == (a + b, a * b // 2)
    assert math.isqrt(r * r - 8 * s) == b - a and C.is_square(r * r - 8 * s)
    assert C.prove(nc, fac=dict(F5_FAC)) is None
    for bad in (dict(p5, s=int(p5["s"]) + 1),
                dict(p5, factors=dict(list(p5["factors"].items())[:1])),
                dict(p5, proof="bls75-thm1"),
                dict(p5, witnesses=dict(p5["witnesses"], **{"2": 4}))):
        assert not C.verify(bad)[0]


def test_subproof_recursion():
    found = None
    for n in RECUR:
        pr = C.prove(n)
        if pr is not None and pr.get("subproofs"):
            found = pr
            break
    assert found is not None, "no sample needed a subproof"
    assert C.verify(found)[0]
    stripped = {k: v for k, v in found.items() if k != "subproofs"}
    assert not C.verify(stripped)[0]


def test_bounded_factoring():
    m = 1000003 * 1000033 * 97
    fac, R = C.factor_partial(m)
    assert R == 1 and fac == {97: 1, 1000003: 1, 1000033: 1}
    f = C.brent_rho(1000003 * 998244353)
    assert f in (1000003, 998244353)
    assert V.factor_witness(101 * 103) == 101
    assert V.factor_witness(1000003 * 1000033) == 1000003
    assert V.factor_witness(1000003) is None              # a prime


@pytest.mark.parametrize("rec", EV["dickson_full"], ids=lambda r: f"run{r['run']}")
def test_reference_certificates_verify_here(rec):
    """The Theorem 1 certificates the reference recorded for each value of
    its Dickson discoveries, put in the proof JSON shape with the
    reference's own bases, pass this module's verify (a proof written on
    the other side is checked here), and a wrong base does not."""
    k = int(rec["k"])
    for m_s, bases in rec["certificates_bls75"].items():
        m = int(m_s)
        N = m * k * k + 1
        fac = V.dickson_factor_n_minus_1(m, k)
        proof = {"proof": "bls75-thm1", "N": N,
                 "factors": {str(p): e for p, e in sorted(fac.items())},
                 "witnesses": dict(bases)}
        ok, msg = C.verify(json.loads(json.dumps(proof)))
        assert ok, (m, msg)
        p2 = min(int(p) for p in bases)
        bad = dict(proof, witnesses=dict(bases, **{str(p2): N - 1}))
        assert not C.verify(bad)[0]
