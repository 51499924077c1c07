"""The Euler round test for primes above (n - 1) n, restated on the host:
q kills p exactly when u = (-p) mod q is a product x(x + 1) with x < n,
found as u clamped to (n - 1) n + 1 and 4u + 1 an odd square by a float
square root rounded to the nearest integer (fk_kernels.hip pronic_d).
Checked against the kill sets (euler_search.py:85-87, the oracle's
euler_forbidden) for every residue of small primes, and the square-root
step against a root one ulp off either way for every 4u + 1 it can see."""

import numpy as np

import fk_oracle as O


def pronic_d(r, q, n, ulp=0):
    """The kernel's arithmetic; ulp = -1, 0, 1 perturbs the square root by
    one float32 ulp the way a hardware root within an ulp may."""
    pmax = (n - 1) * n
    u = min(q - r, pmax + 1)
    w = 4 * u + 1
    root = np.sqrt(np.float32(w))
    if ulp:
        root = np.nextafter(root, np.float32(np.inf if ulp > 0 else 0))
    t = int(np.float32(root) + np.float32(0.5))
    d = (t * t) ^ w
    return 0 if r == 0 else d


def test_pronic_rule_matches_kill_sets():
    for n in (2, 3, 8, 11, 14, 17, 20):
        pmax = (n - 1) * n
        qs = [q for q in O.primes_below(pmax + 700) if q > pmax][:40]
        for q in qs:
            kills = {(-x) % q for x in O.euler_forbidden(q, n)}   # u values
            for r in range(q):
                u = (-r) % q
                assert (pronic_d(r, q, n) == 0) == (u in kills), (n, q, r)


def test_square_root_step_is_exact_within_an_ulp():
    for n in (8, 17, 40, 64):
        pmax = (n - 1) * n
        for u in range(pmax + 2):
            w = 4 * u + 1
            is_sq = int(np.sqrt(w)) ** 2 == w
            for ulp in (-1, 0, 1):
                root = np.sqrt(np.float32(w))
                if ulp:
                    root = np.nextafter(root, np.float32(np.inf if ulp > 0 else 0))
                t = int(np.float32(root) + np.float32(0.5))
                assert (t * t == w) == is_sq, (n, u, ulp)
        # the clamp value is odd, so never x(x + 1)
        assert (pmax + 1) % 2 == 1
