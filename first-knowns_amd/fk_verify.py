"""Verification legs for a recorded discovery, run on the host and kept
independent of the device engine that found it: what the reference's
launchers do before writing a first occurrence's evidence
(euler-prime-runs/launch.py:157-186 three_way_verify,
dickson-ladders/launch.py:403-460 full_verify,
primorial-ap/launch.py:380-446 full_verify), with their evidence fields.

Three legs agree on the claim, none of them the GPU's arithmetic:
  1. this module's own Miller-Rabin chain to the seven bases the run
     lengths are defined with (Python integers);
  2. sympy's isprime (BPSW) on the same values;
  3. an alternate-alignment re-sieve: the candidate checked by direct
     divisibility against every prime below a depth the campaign is not
     running, on a different wheel.
`certify` adds a primality proof for every value of the run (fk_certificate)
and `witness` a factor of the value that ends the run.  Everything is
bounded: factoring gives up and says so rather than stall a campaign.
"""

import math

import fk_certificate as C

# the seven Miller-Rabin bases the run lengths are defined with
# (huntlib/primes.py:17); deterministic below 2^64
RUN_BASES = (2, 325, 9375, 28178, 450775, 9780504, 1795265022)
_SMALL = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)

WITNESS_TDIV = 10**6           # factor_witness: trial division bound
VERIFY_OVERSHOOT = 8           # AP: a full chain is walked this far past n
ALT_Q2 = 8192                  # leg 3 depth unless the campaign runs it
DICKSON_ALT_Q2 = 1 << 16       # Dickson's leg 3: the CPU engine's default
                               # depth (dickson-ladders/ladder_search.py:52)


def is_prime7(m):
    """The seven-base strong test (the engines' prime test)."""
    if m < 2:
        return False
    for p in _SMALL:
        if m % p == 0:
            return m == p
    return all(C.strong_probable_prime(m, a) for a in RUN_BASES)


def _sym_prime(m):
    from sympy import isprime
    return isprime(m)


# ------------------------------------------------------------- the families

def difference(n):
    """P(n), the product of the first n primes."""
    out, c, q = 1, 0, 2
    while c < n:
        if all(q % d for d in range(2, math.isqrt(q) + 1)):
            out *= q
            c += 1
        q += 1
    return out


def dickson_wheel(n):
    """W(n): the product of the primes <= n + 1."""
    w = 1
    for q in range(2, n + 2):
        if all(q % d for d in range(2, math.isqrt(q) + 1)):
            w *= q
    return w


def euler_run(p, cap, prime=is_prime7):
    """Consecutive x = 0, 1, ... with x^2 + x + p prime, up to cap."""
    x = 0
    while x < cap and prime(x * x + x + p):
        x += 1
    return x


def dickson_run(k, cap, prime=is_prime7):
    """Consecutive m = 1, 2, ... with m k^2 + 1 prime, up to cap."""
    m = 0
    while m < cap and prime((m + 1) * k * k + 1):
        m += 1
    return m


def ap_depth(p, n, cap, prime=is_prime7):
    """Consecutive j = 0, 1, ... with p + j P(n) prime, up to cap."""
    d = difference(n)
    j = 0
    while j < cap and prime(p + j * d):
        j += 1
    return j


def _survives(values, depth):
    """No prime below `depth` divides any of `values` (values larger than
    the prime, as every sieved value here is)."""
    for q in C.small_primes(depth):
        for v in values:
            if v % q == 0 and v != q:
                return False
    return True


def alt_q2(q2):
    """Leg 3's depth: ALT_Q2 unless that is the campaign's, then half."""
    return ALT_Q2 if int(q2) != ALT_Q2 else ALT_Q2 // 2


def factor_witness(m, tdiv=WITNESS_TDIV):
    """The smallest prime factor a bounded effort finds for composite m
    (trial division to tdiv, then the certificate module's bounded rho and
    ECM, splitting down to a proved prime), or None."""
    for q in C.small_primes(min(tdiv, C.TDIV_TO)):
        if m % q == 0:
            return q if q < m else None
    d = C.TDIV_TO + 1
    while d < tdiv and d * d <= m:
        if m % d == 0:
            return d
        d += 2
    if d * d > m:
        return None                     # m is prime
    fac, rest = C.factor_partial(m)
    return min(fac) if fac else None


# ------------------------------------------------------------- the legs

def euler_verify(p, claimed, n_filter, q2=1 << 16, witness=True):
    """(evidence, message): euler-prime-runs/launch.py:157-186.  The alt
    re-sieve checks x^2 + x + p, x < n_filter, against every prime below
    alt_q2(q2)."""
    p, claimed = int(p), int(claimed)
    r_own = euler_run(p, claimed + 40)
    r_sym = euler_run(p, claimed + 40, prime=_sym_prime)
    if not (r_own == r_sym == claimed):
        return None, (f"run disagreement own={r_own} sympy={r_sym} "
                      f"claimed={claimed}")
    depth = alt_q2(q2)
    vals = [x * x + x + p for x in range(min(n_filter, claimed))]
    if not _survives(vals, depth):
        return None, f"alternate re-sieve (primes < {depth}) kills p"
    breaker = claimed * claimed + claimed + p
    ev = {"p": p, "run": claimed, "values_prime_x": list(range(claimed)),
          "breaker_x": claimed, "breaker": breaker}
    if witness:
        f = factor_witness(breaker)
        if f is None or breaker % f:
            return None, "no composite witness for the run breaker"
        ev["breaker_factor"] = f
    return ev, "ok"


def dickson_factor_n_minus_1(m, k):
    """N - 1 = m k^2 factored: sympy's factorint of m and of k (k is the
    campaign's own number, at most ~1e23)."""
    from sympy import factorint
    fac = {}
    for p, e in factorint(int(m)).items():
        fac[int(p)] = fac.get(int(p), 0) + int(e)
    for p, e in factorint(int(k)).items():
        fac[int(p)] = fac.get(int(p), 0) + 2 * int(e)
    return fac


def dickson_verify(k, claimed, n_filter, certify=True, witness=True):
    """(evidence, message): dickson-ladders/launch.py:403-460.  The alt
    re-sieve puts k on the coarser wheel W(alt_n) and checks m k^2 + 1,
    m <= alt_n, against every prime below DICKSON_ALT_Q2."""
    k, claimed = int(k), int(claimed)
    r_own = dickson_run(k, claimed + 8)
    r_sym = dickson_run(k, claimed + 8, prime=_sym_prime)
    if not (r_own == r_sym == claimed):
        return None, (f"run disagreement own={r_own} sympy={r_sym} "
                      f"claimed={claimed}")
    alt_n = max(3, min(n_filter, claimed) - 1)
    W = dickson_wheel(alt_n)
    if k % W:
        return None, f"alternate re-sieve: k is not on the {W} wheel"
    if not _survives([m * k * k + 1 for m in range(1, alt_n + 1)],
                     DICKSON_ALT_Q2):
        return None, f"alternate re-sieve (wheel {W}) kills k"
    from sympy import factorint
    certs = {}
    if certify:
        for m in range(1, claimed + 1):
            N = m * k * k + 1
            fac = dickson_factor_n_minus_1(m, k)
            w = C.theorem1(N, fac)
            if w is None:
                return None, f"no primality certificate for m={m}"
            ok, why = C.theorem1_verify(N, fac, w)
            if not ok:
                return None, f"certificate for m={m} fails its re-check: {why}"
            certs[str(m)] = {str(p): int(a) for p, a in w.items()}
    breaker_m = claimed + 1
    breaker = breaker_m * k * k + 1
    bf = 0
    if witness:
        bf = factor_witness(breaker)
        if bf is None or breaker % bf:
            return None, "no composite witness for the run breaker"
    ev = {"k": k, "run": claimed, "values_prime_m": list(range(1, claimed + 1)),
          "k_factorization": {str(p): int(e)
                              for p, e in sorted(factorint(k).items())},
          "breaker_m": breaker_m, "breaker_factor": int(bf),
          "certificates_bls75": certs}
    return ev, "ok"


def _leading_primes(values, test):
    """How many of `values`, from the first, pass `test`."""
    return next((i for i, v in enumerate(values) if not test(v)), len(values))


def _prove_all(values):
    """(proofs, None) with a verified proof for every value, or (None, the
    first value without one)."""
    proofs = []
    for v in values:
        pr = C.prove(v)
        if pr is None or not C.verify(pr)[0]:
            return None, v
        proofs.append(pr)
    return proofs, None


def ap_verify(p, n, claimed, q2=8192, certify=True, witness=True):
    """The legs behind an AP record of depth `claimed` at p, as an evidence
    dict with the fields of primorial-ap/launch.py:380-446.

    The classification walks with cap n, so a claim of n only says that
    the first n values are prime (the chain may run on: a(17)'s runs 19
    deep), while a claim below n is the exact place the chain breaks.  The
    legs walk the chain that far, and VERIFY_OVERSHOOT values further for
    a claim of n, so the true depth and the value that ends it are on
    record.  `agree` is False when a leg disagrees or a proof is missing."""
    p, n, claimed = int(p), int(n), int(claimed)
    d = difference(n)
    at_cap = claimed >= n
    horizon = claimed + 1 + (VERIFY_OVERSHOOT if at_cap else 0)
    chain = [p + j * d for j in range(horizon)]
    own = _leading_primes(chain, is_prime7)
    bpsw = _leading_primes(chain, _sym_prime)
    alt_depth = alt_q2(q2)
    sieved = _survives([p + j * d for j in range(n)], alt_depth)
    holds = own >= claimed if at_cap else own == claimed
    ev = {"p": p, "n": n, "depth": own, "difference": d,
          "values": chain[:min(claimed, n)],
          "legs": {"engine_mr": own, "sympy_bpsw": bpsw,
                   "alt_sieve_survives": sieved, "alt_sieve_depth": alt_depth},
          "agree": own == bpsw and holds and sieved}
    if not ev["agree"]:
        return ev
    if own == horizon:
        ev["depth_is_lower_bound"] = True      # no composite within reach
    if certify and at_cap:
        proofs, missing = _prove_all(ev["values"])
        if proofs is None:
            ev["agree"] = False
            ev["error"] = f"could not prove value {missing} prime"
            return ev
        ev["proofs"] = proofs
        ev["proof_kinds"] = sorted({pr["proof"] for pr in proofs})
    if witness and own < horizon:
        # the first composite of the chain, which both legs stopped at
        ev["chain_breaker"] = chain[own]
        ev["chain_breaker_factor"] = factor_witness(chain[own])
    return ev


def verify_record(family, n, value, run, q2, n_filter=None, certify=True,
                  witness=True):
    """(ok, evidence or message) for one record of an engine of `family`
    and depth n, `value` the record's line value (1 Euler p, 2 Dickson j
    with k = W(n) j, 3 AP p)."""
    nf = n if n_filter is None else n_filter
    if family == 1:
        ev, why = euler_verify(value, run, nf, q2, witness=witness)
        return (ev is not None), (ev if ev is not None else why)
    if family == 2:
        ev, why = dickson_verify(int(value) * dickson_wheel(n), run, nf,
                                 certify, witness)
        return (ev is not None), (ev if ev is not None else why)
    if family == 3:
        ev = ap_verify(value, n, run, q2, certify, witness)
        return bool(ev["agree"]), ev
    raise ValueError(f"unknown family {family}")
