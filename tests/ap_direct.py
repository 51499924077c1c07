"""Direct-division check for primorial-difference APs (test helper).

`kills_between(ps, n, q_lo, q_hi)` says, for each p, whether some prime
q_lo <= q < q_hi divides one of p + j P(n), j < n, computed by dividing by
every prime of the range (numpy, vectorised over the primes).  It is the
sieve's definition (primorial-ap/ap_reference.py:159-172 sieve_survivor)
for values far above q_hi, and needs nothing of the oracle's sieve: it
pins deep-sieve streams that the oracle's own sieve would take minutes to
produce.  Values must be below 2^64 * 2^40 and P(n) below 2^104.
"""
import numpy as np

import fk_oracle as O


def _mod(v, qs, t40):
    hi, lo = v >> 40, v & ((1 << 40) - 1)
    if hi >> 64:
        raise ValueError("value too wide for the direct check")
    return ((np.uint64(hi) % qs) * t40 + np.uint64(lo) % qs) % qs


def kills_between(ps, n, q_lo, q_hi):
    if q_hi > 1 << 24:
        raise ValueError("residue products must stay below 2^64")
    qs = np.array([q for q in O.primes_below(q_hi) if q >= q_lo],
                  dtype=np.uint64)
    t20 = np.uint64(1 << 20) % qs
    t40 = (t20 * t20) % qs
    dq = _mod(O.primorial(n), qs, t40)
    out = []
    for p in ps:
        r = _mod(int(p), qs, t40)
        dead = np.zeros(qs.size, dtype=bool)
        for _ in range(n):
            dead |= r == 0
            r = (r + dq) % qs
        out.append(bool(dead.any()))
    return out
