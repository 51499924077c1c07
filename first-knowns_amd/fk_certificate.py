"""Primality proofs for recorded discoveries: N-1 certificates after
Brillhart, Lehmer and Selfridge (Math. Comp. 29 (1975)).

A proof is a JSON object that `verify` re-checks from nothing.  Its keys
("proof", "N", "F", "R", "factors", "witnesses", "subproofs", "s", "r",
"bound", "note") and its three route names are those of the reference's
certificate module (huntlib/certificate.py), so that either side checks
the other's proofs; the code here is this package's own.

The mathematics:

* Pocklington in BLS form.  Let N - 1 = F R with F completely factored.
  If every prime p | F has a base a with a^(N-1) = 1 (mod N) and
  gcd(a^((N-1)/p) - 1, N) = 1, then every prime factor of N is 1 mod F.
  Two such factors exceed F each, so F > sqrt(N) proves N prime
  ("bls75-thm1").
* BLS75 Theorem 5.  Under the same conditions with only F^3 > N, F even
  and gcd(F, R) = 1, write R = 2 F s + r with 0 <= r < 2F.  A composite N
  is then (1 + aF)(1 + bF) with a + b = r and ab = 2s, so r^2 - 8s =
  (a - b)^2 is a square; N is prime when s = 0 or r^2 - 8s is not a square
  ("bls75-thm5").
* Below PSI_13 = 3,317,044,064,679,887,385,961,981, the least strong
  pseudoprime to all of the first 13 prime bases (Sorenson and Webster,
  Math. Comp. 86 (2017)), Miller-Rabin to the bases 2, 3, ..., 41 decides
  primality ("deterministic-mr").

Searching for a proof:

1. N - 1 is split as far as a fixed effort allows: trial division by the
   primes below TDIV_TO, then Pollard p-1 (stage one to PM1_BOUND), Brent
   rho and a few ECM curves on each composite piece, all charged to one
   budget of modular multiplications.  A prime found anywhere is divided
   out of every remaining piece, so the factored part F and the rest R are
   coprime by construction.
2. F holds only proved primes.  A piece that is prime past PSI_13 (by
   BPSW, which is only the reason to try) joins F with a proof of its own,
   one level down, carried under "subproofs"; the largest such pieces are
   tried first, and only while F is still too small.
3. The route is Theorem 1 when F > sqrt(N), else Theorem 5 when F^3 > N.
   The base for each prime of F is the smallest prime (below
   CERT_BASE_CAP) meeting both conditions, the rule the reference's
   recorded proofs follow, so a proof depends on N and F alone.
"""

import math
import random

PSI_13 = 3317044064679887385961981
MR_PROOF_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)

CERT_BASE_CAP = 10_000        # candidate bases: the primes below this
TDIV_TO = 1 << 17             # trial division bound
PM1_BOUND = 20_000            # Pollard p-1 stage-one bound
RHO_ITERS = 60_000            # Brent rho iterations per seed
RHO_TRIES = 3                 # Brent rho seeds per piece
ECM_CURVES = 24               # ECM curves per piece (sympy), after rho
EFFORT = 2_000_000            # modular multiplications per factor_partial
PROOF_DEPTH = 3               # nesting of subproofs

_PRIMES = []                  # cached primes below _TOP
_TOP = 0


def small_primes(limit=TDIV_TO):
    """The primes below `limit` (an odd-only sieve, extended on demand)."""
    global _PRIMES, _TOP
    if limit > _TOP:
        top = max(limit, TDIV_TO)
        half = bytearray([1]) * (top // 2)      # index i stands for 2i + 1
        half[0] = 0
        for i in range(1, (math.isqrt(top) + 1) // 2 + 1):
            if half[i]:
                step = 2 * i + 1
                start = (step * step) // 2
                half[start::step] = bytes(len(range(start, len(half), step)))
        _PRIMES = [2] + [2 * i + 1 for i, alive in enumerate(half) if alive]
        _TOP = top
    return _PRIMES[:_bisect(_PRIMES, limit)]


def _bisect(xs, v):
    lo, hi = 0, len(xs)
    while lo < hi:
        mid = (lo + hi) // 2
        if xs[mid] < v:
            lo = mid + 1
        else:
            hi = mid
    return lo


def strong_probable_prime(n, a):
    """One Miller-Rabin round to base a for odd n > 2: False proves n
    composite."""
    e = n - 1
    t = (e & -e).bit_length() - 1      # n - 1 = 2^t e'
    y = pow(a % n, e >> t, n)
    if y in (0, 1, n - 1):
        return True
    for _ in range(t - 1):
        y = y * y % n
        if y == n - 1:
            return True
        if y == 1:
            return False
    return False


def proved_prime(n):
    """True iff n < PSI_13 and n is prime by the 13-prime-base test: a
    prime this module may use without a certificate."""
    if not 1 < n < PSI_13:
        return False
    for b in MR_PROOF_BASES:
        if n % b == 0:
            return n == b
    return all(strong_probable_prime(n, b) for b in MR_PROOF_BASES)


def _bpsw(n):
    """sympy's BPSW: the reason to attempt a subproof, never a proof."""
    from sympy import isprime
    return isprime(n)


# ------------------------------------------------------------- witnesses

def witnesses(N, primes_of_F, base_cap=CERT_BASE_CAP):
    """{p: a}, for each p in primes_of_F the smallest prime a < base_cap
    with a^(N-1) = 1 and gcd(a^((N-1)/p) - 1, N) = 1 (mod N); None when
    some p has no such base or a base fails Fermat's condition (N is then
    composite)."""
    todo = sorted({int(p) for p in primes_of_F})
    found = {}
    if not todo:
        return found
    for a in small_primes(base_cap):
        if pow(a, N - 1, N) != 1:
            return None
        for p in todo:
            if p not in found and math.gcd(pow(a, (N - 1) // p, N) - 1, N) == 1:
                found[p] = a
        if len(found) == len(todo):
            return found
    return None


# ------------------------------------------------------------- factoring

class _Effort:
    """A budget of modular multiplications shared by every method."""

    def __init__(self, total):
        self.left = int(total)

    def take(self, k):
        self.left -= k
        return self.left >= 0


def _pollard_pm1(m, effort, bound=PM1_BOUND):
    """A factor of m when some prime factor p has p - 1 smooth to `bound`
    (stage one: the base raised to every prime power below it), or None."""
    x = 2
    for q in small_primes(bound):
        qk = q
        while qk * q < bound:
            qk *= q
        if not effort.take(qk.bit_length()):
            return None
        x = pow(x, qk, m)
    g = math.gcd(x - 1, m)
    return g if 1 < g < m else None


def brent_rho(m, iters=RHO_ITERS, tries=RHO_TRIES, effort=None):
    """A nontrivial factor of composite m, or None: Brent's cycle search
    on y -> y^2 + c, differences multiplied in batches of 128 between
    gcds.  The seeds come from m, so a verification reproduces exactly."""
    if m % 2 == 0:
        return 2
    rng = random.Random(m & 0xFFFFFFFF)
    for _ in range(tries):
        y, c = rng.randrange(1, m), rng.randrange(1, m)
        g, r, acc = 1, 1, 1
        x = ys = y
        spent = 0
        while g == 1 and spent < iters:
            x = y
            for _ in range(r):
                y = (y * y + c) % m
            k = 0
            while k < r and g == 1:
                ys = y
                for _ in range(min(128, r - k)):
                    y = (y * y + c) % m
                    acc = acc * abs(x - y) % m
                g = math.gcd(acc, m)
                k += 128
            spent += r
            if effort is not None and not effort.take(2 * r):
                return None
            r *= 2
        if g == m:                       # a batch overshot: step singly
            g = 1
            while g == 1:
                ys = (ys * ys + c) % m
                g = math.gcd(abs(x - ys), m)
        if 1 < g < m:
            return g
    return None


def _ecm(m, curves):
    try:
        from sympy.ntheory import ecm
        found = ecm(m, max_curve=curves)
    except Exception:                    # ECM declining is not an error
        return None
    f = min(found) if found else None
    return f if f is not None and 1 < f < m else None


def _one_split(m, effort, ecm_curves):
    for method in (lambda: _pollard_pm1(m, effort),
                   lambda: brent_rho(m, effort=effort),
                   lambda: _ecm(m, ecm_curves) if ecm_curves else None):
        if effort.left <= 0:
            return None
        f = method()
        if f:
            return f
    return None


def _factor(m, ecm_curves=ECM_CURVES, effort=EFFORT):
    """(fac, stuck): m = prod p^e (fac, proved primes) times the product of
    the pieces in `stuck`, which the bounded effort could not split; every
    piece is coprime to every prime of fac."""
    fac, pieces = {}, []
    rest = int(m)
    for q in small_primes(TDIV_TO):
        if q * q > rest:
            break
        if rest % q == 0:
            rest //= q
            e = 1
            while rest % q == 0:
                rest //= q
                e += 1
            fac[q] = e
    if rest > 1:
        pieces.append(rest)
    budget = _Effort(effort)
    stuck = []
    while pieces:
        v = pieces.pop()
        for p in list(fac):              # keep every piece coprime to F
            while v % p == 0:
                v //= p
                fac[p] += 1
        if v == 1:
            continue
        if proved_prime(v):
            fac[v] = fac.get(v, 0) + 1
            continue
        f = _one_split(v, budget, ecm_curves) if budget.left > 0 else None
        if f is None:
            stuck.append(v)
        else:
            g = math.gcd(f, v // f)      # a repeated factor: peel the gcd
            pieces += [g, f // g, v // f // g, g] if g > 1 else [f, v // f]
    # a prime found after a piece got stuck still leaves that piece
    out = []
    for v in stuck:
        for p in fac:
            while v % p == 0:
                v //= p
                fac[p] += 1
        if v > 1:
            out.append(v)
    return fac, out


def factor_partial(m, ecm_curves=ECM_CURVES, effort=EFFORT):
    """(fac, R): m = F R with fac the factorization of F into proved
    primes, gcd(F, R) = 1, and R the part the bounded effort left
    unsplit (R = 1 when m factored completely)."""
    fac, stuck = _factor(m, ecm_curves, effort)
    return fac, math.prod(stuck)


# ------------------------------------------------------------- proofs

def theorem5_split(F, R):
    """(s, r) with R = 2 F s + r, 0 <= r < 2F."""
    return divmod(int(R), 2 * int(F))


def is_square(x):
    return x >= 0 and math.isqrt(x) ** 2 == x


def _product(fac):
    out = 1
    for p, e in fac.items():
        out *= p ** e
    return out


def _route(N, F, R):
    """(route, extra) for the factored part F of N - 1, or (None, why)."""
    if F > math.isqrt(N):
        return "bls75-thm1", {}
    if F ** 3 > N:
        s, r = theorem5_split(F, R)
        if s and is_square(r * r - 8 * s):
            return None, "composite"     # the theorem's other direction
        return "bls75-thm5", {"s": s, "r": r}
    return None, "F too small"


def prove(N, fac=None, base_cap=CERT_BASE_CAP, depth=PROOF_DEPTH, **kw):
    """A proof that `verify` re-checks from nothing, or None when none was
    found within the bounds (which says nothing about N).  `fac`, a known
    factorization of part of N - 1, is checked and then used as it is."""
    N = int(N)
    if N < PSI_13:
        if not proved_prime(N):
            return None
        return {"proof": "deterministic-mr", "N": N, "bound": PSI_13,
                "note": "Miller-Rabin to the 13 prime bases 2..41 decides "
                        "primality below the bound"}
    if N % 2 == 0:
        return None
    if fac is None:
        fac, stuck = _factor(N - 1, **kw)
    else:
        fac = {int(p): int(e) for p, e in fac.items()}
        F0 = _product(fac)
        if F0 < 1 or (N - 1) % F0:
            return None
        stuck = [(N - 1) // F0] if N - 1 > F0 else []
    R = math.prod(stuck)
    subs = {}
    # unsplit pieces that are prime past PSI_13 join F with a subproof
    # each, the largest first, while F is short of Theorem 5
    if depth > 0:
        for v in sorted(stuck, reverse=True):
            if _product(fac) ** 3 > N:
                break
            if v < PSI_13 or v in fac or not _bpsw(v):
                continue
            sub = prove(v, base_cap=base_cap, depth=depth - 1, **kw)
            if sub is None:
                continue
            subs[str(v)] = sub
            fac[v] = 1
            R //= v
    for p in fac:
        if not proved_prime(p) and str(p) not in subs:
            return None
    F = _product(fac)
    if F % 2 or math.gcd(F, R) != 1:
        return None
    route, extra = _route(N, F, R)
    if route is None:
        return None
    w = witnesses(N, fac, base_cap)
    if w is None:
        return None
    out = {"proof": route, "N": N, "F": F, "R": R,
           "factors": {str(p): e for p, e in sorted(fac.items())},
           "witnesses": {str(p): a for p, a in sorted(w.items())}}
    out.update(extra)
    if subs:
        out["subproofs"] = subs
    return out


def _check_mr(proof, N):
    if N >= PSI_13:
        return False, (f"{N} is at or past {PSI_13}, where the 13-base test "
                       f"proves nothing")
    if not proved_prime(N):
        return False, f"{N} fails the 13-base test"
    return True, "deterministic: 13 prime bases, below PSI_13"


def _check_nminus1(proof, N):
    """The part both theorems share: F, its primes and their bases."""
    raw = proof.get("factors") or {}
    try:
        fac = {int(p): int(e) for p, e in raw.items()}
    except (TypeError, ValueError):
        return None, "unreadable factors"
    if not fac:
        return None, "the proof factors nothing of N - 1"
    if any(e < 1 for e in fac.values()):
        return None, "an exponent below 1"
    F = _product(fac)
    if (N - 1) % F:
        return None, f"F = {F} is not a divisor of N - 1"
    R = (N - 1) // F
    if F % 2:
        return None, "F is odd"
    if math.gcd(F, R) != 1:
        return None, "F and R share a factor"
    bases = proof.get("witnesses") or {}
    for p in sorted(fac):
        a = bases.get(str(p))
        if a is None:
            return None, f"prime {p} of F has no base"
        a = int(a)
        if pow(a, N - 1, N) != 1:
            return None, f"base {a} (for {p}): a^(N-1) is not 1 mod N"
        if math.gcd(pow(a, (N - 1) // p, N) - 1, N) != 1:
            return None, f"base {a} (for {p}): a^((N-1)/p) - 1 shares a factor with N"
    return (fac, F, R), None


def verify(proof):
    """(ok, message): the proof re-derived with nothing taken on trust --
    F divides N - 1, every prime of F is proved (below PSI_13 directly,
    above it by its own subproof, checked the same way), every base meets
    both conditions, and the route's bound and side condition hold."""
    todo = [(proof, None)]               # (proof, the prime it must be about)
    top = None
    while todo:
        pf, about = todo.pop()
        if not isinstance(pf, dict):
            return False, "a proof must be a JSON object"
        try:
            N = int(pf.get("N", 0))
        except (TypeError, ValueError):
            return False, "unreadable N"
        if about is not None and N != about:
            return False, f"the subproof for {about} proves {N}"
        if N < 2:
            return False, "N below 2"
        kind = pf.get("proof")
        if kind == "deterministic-mr":
            ok, msg = _check_mr(pf, N)
            if not ok:
                return False, msg if about is None else f"subproof {about}: {msg}"
            top = msg if top is None else top
            continue
        if kind not in ("bls75-thm1", "bls75-thm5"):
            return False, f"no route called {kind!r}"
        got, why = _check_nminus1(pf, N)
        if got is None:
            return False, why if about is None else f"subproof {about}: {why}"
        fac, F, R = got
        subs = pf.get("subproofs") or {}
        for p in fac:
            if not proved_prime(p):
                if str(p) not in subs:
                    return False, (f"prime {p} of F is past PSI_13 and has "
                                   f"no subproof")
                todo.append((subs[str(p)], p))
        if kind == "bls75-thm1":
            if F <= math.isqrt(N):
                return False, f"Theorem 1 needs F > sqrt(N); F = {F}"
            msg = "Theorem 1 (BLS75): F exceeds sqrt(N)"
        else:
            if F ** 3 <= N:
                return False, f"Theorem 5 needs F^3 > N; F = {F}"
            s, r = theorem5_split(F, R)
            if (s, r) != (int(pf.get("s", -1)), int(pf.get("r", -1))):
                return False, (f"R = 2Fs + r gives (s, r) = ({s}, {r}), the "
                               f"proof says ({pf.get('s')}, {pf.get('r')})")
            if s and is_square(r * r - 8 * s):
                return False, (f"r^2 - 8s = {r * r - 8 * s} is a square: "
                               f"Theorem 5 makes N composite")
            msg = "Theorem 5 (BLS75): F^3 exceeds N and r^2 - 8s is no square"
        top = msg if top is None else top
    return True, top


def theorem1(N, fac, base_cap=CERT_BASE_CAP):
    """Bases for N when N - 1 = prod p^e (fac) is completely factored into
    proved primes (a Dickson value m k^2 + 1), or None."""
    return witnesses(N, fac, base_cap)


def theorem1_verify(N, fac, wit):
    """`verify` of the Theorem 1 proof made of (N, fac, wit)."""
    return verify({"proof": "bls75-thm1", "N": int(N),
                   "factors": {str(int(p)): int(e) for p, e in fac.items()},
                   "witnesses": {str(int(p)): int(a) for p, a in wit.items()}})
