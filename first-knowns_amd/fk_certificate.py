"""Primality proofs for recorded discoveries: the N-1 certificates of
Brillhart, Lehmer and Selfridge (Math. Comp. 29 (1975)), with the routes,
bounds and JSON shape of the reference's certificate module
(huntlib/certificate.py:1-421), so that a proof written by either side is
checked by the other.

The mathematics (the reference module states the same two theorems):

* Pocklington in BLS form.  If N - 1 = F R with F completely factored and
  every prime p | F has a witness a with a^(N-1) = 1 (mod N) and
  gcd(a^((N-1)/p) - 1, N) = 1, then every prime factor of N is 1 mod F.
  With F > sqrt(N) that proves N prime (Theorem 1).
* Theorem 5.  With only F > N^(1/3), F even and gcd(F, R) = 1, write
  R = 2 F s + r with 1 <= r < 2F: N is prime iff s = 0 or r^2 - 8 s is not
  a perfect square.  (A composite N is then (1 + aF)(1 + bF), which forces
  r = a + b and s = ab/2, and r^2 - 8s = (a - b)^2.)

Routes, in the order `prove` tries them:
  "deterministic-mr"   N below PSI_13 = 3317044064679887385961981, the
                       smallest strong pseudoprime to all of the first 13
                       prime bases (Sorenson and Webster, Math. Comp. 86
                       (2017)): Miller-Rabin to the bases 2, 3, ..., 41 is a
                       proof there
  "bls75-thm1"         N-1 factored past sqrt(N)
  "bls75-thm5"         N-1 factored past N^(1/3)

Every prime put inside F is itself proved: below PSI_13 by the
deterministic test, above it by a subproof carried in the certificate (at
most PROOF_DEPTH levels).  Factoring is bounded: trial division to TDIV_TO,
a bounded Brent rho, a bounded number of ECM curves (sympy), then it stops
and says what is left.  Witnesses are the primes in ascending order below
CERT_BASE_CAP, the smallest that works for each prime of F, so a proof is
reproducible from N and F alone.
"""

import math
import random

PSI_13 = 3317044064679887385961981
MR_PROOF_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)

CERT_BASE_CAP = 10_000        # witness bases: the primes below this
TDIV_TO = 1 << 17             # trial division bound in factor_partial
RHO_ITERS = 60_000            # Brent rho: iterations per try
RHO_TRIES = 3
ECM_CURVES = 24               # bounded ECM after rho
SPLIT_BUDGET = 8              # splits attempted per factor_partial call
PROOF_DEPTH = 3               # subproof levels for large prime factors

_SMALL_PRIMES = None


def small_primes(limit=TDIV_TO):
    """Primes below `limit` (sieve of Eratosthenes, cached at TDIV_TO)."""
    global _SMALL_PRIMES
    if limit <= TDIV_TO and _SMALL_PRIMES is not None:
        return [p for p in _SMALL_PRIMES if p < limit]
    n = max(limit, TDIV_TO)
    sieve = bytearray([1]) * n
    sieve[0:2] = b"\0\0"
    for i in range(2, math.isqrt(n - 1) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(range(i * i, n, i)))
    ps = [i for i in range(n) if sieve[i]]
    if n == TDIV_TO:
        _SMALL_PRIMES = ps
    return [p for p in ps if p < limit]


def strong_probable_prime(n, a):
    """Miller-Rabin round to base a (n odd > 2): False proves n composite."""
    d, s = n - 1, 0
    while not d & 1:
        d >>= 1
        s += 1
    x = pow(a % n, d, n)
    if x in (0, 1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def proved_prime(n):
    """True iff n is a prime this module may take as proved without a
    certificate: n < PSI_13 and Miller-Rabin to the first 13 prime bases."""
    if n < 2 or n >= PSI_13:
        return False
    for p in MR_PROOF_BASES:
        if n % p == 0:
            return n == p
    return all(strong_probable_prime(n, a) for a in MR_PROOF_BASES)


def _probable_prime(n):
    """BPSW-strength check (sympy) for deciding where to spend a subproof;
    never a proof by itself."""
    from sympy import isprime
    return isprime(n)


# ------------------------------------------------------------- witnesses

def witnesses(N, primes_of_F, base_cap=CERT_BASE_CAP):
    """{p: a} meeting both BLS conditions for every p in primes_of_F, the
    smallest prime a for each; None if some p has none below base_cap or if
    a base fails Fermat (then N is composite)."""
    out = {}
    bases = small_primes(base_cap)
    for p in primes_of_F:
        e = (N - 1) // int(p)
        for a in bases:
            if pow(a, N - 1, N) != 1:
                return None
            if math.gcd(pow(a, e, N) - 1, N) == 1:
                out[int(p)] = a
                break
        else:
            return None
    return out


# ------------------------------------------------------------- factoring

def brent_rho(m, iters=RHO_ITERS, tries=RHO_TRIES):
    """A nontrivial factor of composite m, or None: Brent's cycle search on
    x -> x^2 + c, the differences multiplied in batches of 128 between
    gcds.  Seeded from m, so a verification reproduces exactly."""
    if m % 2 == 0:
        return 2
    rng = random.Random(m & 0xFFFFFFFF)
    for _ in range(tries):
        y, c = rng.randrange(1, m), rng.randrange(1, m)
        g, r, q = 1, 1, 1
        x = ys = y
        steps = 0
        while g == 1 and steps < iters:
            x = y
            for _ in range(r):
                y = (y * y + c) % m
            k = 0
            while k < r and g == 1:
                ys = y
                for _ in range(min(128, r - k)):
                    y = (y * y + c) % m
                    q = q * abs(x - y) % m
                g = math.gcd(q, m)
                k += 128
            steps += r
            r *= 2
        if g == m:                      # the batch overshot: step singly
            g = 1
            while g == 1:
                ys = (ys * ys + c) % m
                g = math.gcd(abs(x - ys), m)
        if 1 < g < m:
            return g
    return None


def _split(v, ecm_curves):
    f = brent_rho(v)
    if f is None and ecm_curves:
        try:
            from sympy.ntheory import ecm
            found = ecm(v, max_curve=ecm_curves)
            f = min(found) if found else None
        except Exception:               # ECM declining is not an error
            f = None
    return f if (f is not None and 1 < f < v) else None


def factor_partial(m, ecm_curves=ECM_CURVES, budget=SPLIT_BUDGET):
    """(fac, R) with m = F R, fac the factorization of F into proved
    primes, gcd(F, R) = 1 and R what the bounded effort could not split
    (R = 1: m factored completely)."""
    fac, rest = {}, int(m)
    for q in small_primes():
        if q * q > rest:
            break
        if rest % q == 0:
            e = 0
            while rest % q == 0:
                rest //= q
                e += 1
            fac[q] = e
    R = 1
    pending = [rest] if rest > 1 else []
    while pending:
        v = pending.pop()
        if v == 1:
            continue
        if proved_prime(v):
            fac[v] = fac.get(v, 0) + 1
            continue
        if budget <= 0:
            R *= v
            continue
        budget -= 1
        f = _split(v, ecm_curves)
        if f is None:
            R *= v
            continue
        pending += [f, v // f]
    # a prime that landed on both sides goes to F, keeping gcd(F, R) = 1
    for p in list(fac):
        while R % p == 0:
            R //= p
            fac[p] += 1
    return fac, R


# ------------------------------------------------------------- proofs

def theorem5_split(F, R):
    """(s, r) with R = 2 F s + r, 0 <= r < 2F."""
    s, r = divmod(int(R), 2 * int(F))
    return s, r


def is_square(x):
    if x < 0:
        return False
    t = math.isqrt(x)
    return t * t == x


def prove(N, fac=None, base_cap=CERT_BASE_CAP, depth=PROOF_DEPTH, **kw):
    """A proof dict that `verify` re-checks from scratch, or None when no
    proof is found within the bounds (not a claim that N is composite).
    `fac`, a known partial factorization of N-1, is checked, not trusted."""
    N = int(N)
    if N < 2:
        return None
    if N < PSI_13:
        if not proved_prime(N):
            return None
        return {"proof": "deterministic-mr", "N": N, "bound": PSI_13,
                "note": "Miller-Rabin to the 13 prime bases 2..41 is "
                        "deterministic below the bound"}
    if N % 2 == 0:
        return None
    if fac is None:
        fac, R = factor_partial(N - 1, **kw)
    else:
        fac = {int(p): int(e) for p, e in fac.items()}
        F = 1
        for p, e in fac.items():
            F *= p ** e
        if F < 1 or (N - 1) % F:
            return None
        R = (N - 1) // F
    subs = {}
    # a large prime cofactor: prove it one level down and admit it
    if R > 1 and depth > 0 and R >= PSI_13 and _probable_prime(R):
        sub = prove(R, base_cap=base_cap, depth=depth - 1, **kw)
        if sub is not None:
            fac[R] = fac.get(R, 0) + 1
            subs[str(R)] = sub
            R = 1
    F = 1
    for p, e in fac.items():
        if not (proved_prime(p) or str(p) in subs):
            return None
        F *= p ** e
    if F % 2 or math.gcd(F, R) != 1:
        return None
    w = witnesses(N, sorted(fac), base_cap)
    if w is None:
        return None
    out = {"N": N, "F": F, "R": R,
           "factors": {str(p): e for p, e in sorted(fac.items())},
           "witnesses": {str(p): a for p, a in sorted(w.items())}}
    if subs:
        out["subproofs"] = subs
    if F > math.isqrt(N):
        out["proof"] = "bls75-thm1"
        return out
    if F ** 3 > N:
        s, r = theorem5_split(F, R)
        if s != 0 and is_square(r * r - 8 * s):
            return None                 # Theorem 5 is an iff: N composite
        out["proof"] = "bls75-thm5"
        out["s"], out["r"] = s, r
        return out
    return None


def verify(proof):
    """(ok, message): every claim in the dict re-derived, nothing trusted
    (the factorization, the primality of each factor, F | N-1, the
    witnesses, the theorem's side conditions)."""
    if not isinstance(proof, dict):
        return False, "not a proof object"
    kind = proof.get("proof")
    N = int(proof.get("N", 0))
    if N < 2:
        return False, "no N"
    if kind == "deterministic-mr":
        if N >= PSI_13:
            return False, (f"N = {N} is not below {PSI_13}: the deterministic "
                           f"test is no proof there")
        if not proved_prime(N):
            return False, "N fails Miller-Rabin to the 13 prime bases"
        return True, "Miller-Rabin to the 13 prime bases below 3.317e24"
    if kind not in ("bls75-thm1", "bls75-thm5"):
        return False, f"unknown proof kind {kind!r}"
    fac = {int(p): int(e) for p, e in (proof.get("factors") or {}).items()}
    if not fac:
        return False, "no factorization of N-1"
    subs = proof.get("subproofs") or {}
    F = 1
    for p, e in sorted(fac.items()):
        if e < 1:
            return False, f"bad exponent {e} for {p}"
        if not proved_prime(p):
            sub = subs.get(str(p))
            if sub is None:
                return False, (f"factor {p} is past the deterministic bound "
                               f"and carries no subproof")
            if int(sub.get("N", 0)) != p:
                return False, f"subproof for {p} is about {sub.get('N')}"
            ok, why = verify(sub)
            if not ok:
                return False, f"subproof for {p} fails: {why}"
        F *= p ** e
    if (N - 1) % F:
        return False, "F does not divide N-1"
    R = (N - 1) // F
    if F % 2:
        return False, "F must be even"
    if math.gcd(F, R) != 1:
        return False, f"gcd(F, R) = {math.gcd(F, R)} != 1"
    wit = proof.get("witnesses") or {}
    for p in sorted(fac):
        a = wit.get(str(p))
        if a is None:
            return False, f"no witness for {p}"
        a = int(a)
        if pow(a, N - 1, N) != 1:
            return False, f"witness {a} fails Fermat at p = {p}"
        if math.gcd(pow(a, (N - 1) // p, N) - 1, N) != 1:
            return False, f"witness {a} fails the gcd condition at p = {p}"
    if kind == "bls75-thm1":
        if F <= math.isqrt(N):
            return False, f"F = {F} does not exceed sqrt(N)"
        return True, "BLS75 Theorem 1: F > sqrt(N)"
    if F ** 3 <= N:
        return False, f"F = {F} does not exceed N^(1/3)"
    s, r = theorem5_split(F, R)
    if s != int(proof.get("s", -1)) or r != int(proof.get("r", -1)):
        return False, (f"stated (s, r) = ({proof.get('s')}, {proof.get('r')}) "
                       f"but R = 2Fs + r gives ({s}, {r})")
    if s != 0 and is_square(r * r - 8 * s):
        return False, f"r^2 - 8s = {r * r - 8 * s} is a square: N is composite"
    return True, "BLS75 Theorem 5: F > N^(1/3) and the square test passes"


def theorem1(N, fac, base_cap=CERT_BASE_CAP):
    """Witnesses for N when N-1 = prod p^e (fac) is completely factored
    and every p is proved (a Dickson value m k^2 + 1), or None."""
    return witnesses(N, sorted(int(p) for p in fac), base_cap)


def theorem1_verify(N, fac, wit):
    """`verify` on the Theorem 1 certificate (N, fac, wit)."""
    return verify({"proof": "bls75-thm1", "N": int(N),
                   "factors": {str(int(p)): int(e) for p, e in fac.items()},
                   "witnesses": {str(int(p)): int(a) for p, a in wit.items()}})
