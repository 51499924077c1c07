"""Family constants and the small host-side mathematics the drop-in engines
expose as attributes (wheel periods, fold offsets, P(n)), plus the default
tuning per family.

Everything heavy (tables, sieve, strong tests) lives in libfk_amd.so; this
module only answers the questions the reference's callers ask of an engine
object, with the reference's own conventions.
"""

import json
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
TUNING_PATH = HERE / "fk_tuning.json"

# huntlib/primes.py:17-18
MR_BASES = (2, 325, 9375, 28178, 450775, 9780504, 1795265022)

# euler-prime-runs/euler_search.py:50-60, :276
EULER_WHEEL_PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23]
EULER_WHEEL_PRIMES_29 = EULER_WHEEL_PRIMES + [29]
EULER_WHEEL_PRIMES_31 = EULER_WHEEL_PRIMES_29 + [31]
EULER_WHEEL_PRIMES_37 = EULER_WHEEL_PRIMES_31 + [37]
EULER_Q2 = 65536
EULER_P_FLOOR = 100_000
EULER_P_CEIL = 10**24
EULER_WHEEL_BUDGET = 50_000_000
EULER_WHEEL_DEVICE = 1 << 30       # residues the engine holds at most

# dickson-ladders/ladder_reference.py:57; ladder_search.py:52-53;
# dickson-ladders/launch.py:307
DICKSON_K_FLOOR = 10**4
DICKSON_Q2_DEFAULT = 65536
DICKSON_J_CEIL = 4 * 10**18
SPRP_EXACT_FROM = 5

# primorial-ap/ap_reference.py:89-99; ap_search.py:63-86; ap_gpu.py:88-97
AP_W0 = 30
AP_W0_RESIDUES = (1, 7, 11, 13, 17, 19, 23, 29)
AP_P_FLOOR = 10**4
AP_P_CEIL = 10**26
AP_Q2_DEFAULT = 1 << 16
AP_LAUNCH_U = 1 << 27
DEPTH_EXACT_FROM = 6

# The engine's per-candidate reduction sums four 32 x q-bit products in 64
# bits, so sieve primes stay below 2^30 (the reference's Barrett bound is
# 2^31, ladder_gpu.py:433-437 and ap_gpu.py:404-406).
Q2_MAX = 1 << 30


def primes_below(limit):
    limit = int(limit)
    if limit < 3:
        return []
    sieve = bytearray([1]) * limit
    sieve[0:2] = b"\x00\x00"
    for i in range(2, int(limit ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(range(i * i, limit, i)))
    return [i for i in range(limit) if sieve[i]]


def product(xs):
    out = 1
    for x in xs:
        out *= int(x)
    return out


def euler_forbidden(q, n):
    """euler_search.py:85-87: residues r with q | x^2+x+r for some x < n."""
    return {(-(x * x + x)) % q for x in range(n)}


def euler_wheel_count(n, wheel_primes):
    """euler_search.py:90-95: exact admissible-offset count."""
    c = 1
    for q in wheel_primes:
        c *= q - len(euler_forbidden(q, n))
    return c


def euler_best_wheel(n, budget=EULER_WHEEL_BUDGET):
    """The widest wheel whose host-built table fits `budget` residues
    (euler-prime-runs/euler_search.py:200-212).  The 37# wheel is built
    factored: its host table is the 31# one, and the library lifts it over
    37 on the device into every residue (598,752,000 at n = 17, 4.8 GB),
    which must stay within the engine's 2^30 residues.  On the SCORE_WIDE
    shape (n = 17) the 37# wheel sweeps 1.24x the p/s of the 31# one."""
    if (euler_wheel_count(n, EULER_WHEEL_PRIMES_31) <= budget and
            euler_wheel_count(n, EULER_WHEEL_PRIMES_37) <= EULER_WHEEL_DEVICE):
        return list(EULER_WHEEL_PRIMES_37)
    for wp in (EULER_WHEEL_PRIMES_31, EULER_WHEEL_PRIMES_29,
               EULER_WHEEL_PRIMES):
        if euler_wheel_count(n, wp) <= budget:
            return list(wp)
    return list(EULER_WHEEL_PRIMES)


def dickson_wheel_modulus(n):
    """ladder_reference.py:60-66: W(n) = product of primes <= n + 1."""
    return product(primes_below(n + 2))


def dickson_fold_offsets(W, P, n):
    """Surviving j residues mod the fold prime P by the m t + 1 walk
    (ladder_gpu.py:597-614 semantics); 0 always survives."""
    wm = W % P
    offs = []
    for s in range(P):
        t = (wm * s) * (wm * s) % P
        r, dead = (t + 1) % P, False
        for _ in range(n):
            if r == 0:
                dead = True
                break
            r = (r + t) % P
        if not dead:
            offs.append(s)
    return offs


def ap_difference(n):
    """ap_reference.py:99-110: P(n), the product of the first n primes."""
    return product(primes_below(200)[:int(n)])


# ------------------------------------------------------------- tuning

# Defaults measured on MI355X (see the autotuner, fk_autotune.py); a
# fk_tuning.json written by the autotuner overrides them per family.
DEFAULT_TUNING = {
    "euler": {"q_pat": 128, "q_mark": 0, "q_med": 1024, "threads": 256,
              "words_per_thread": 8},
    "dickson": {"q_pat": 128, "q_mark": 0, "q_med": 2048, "threads": 256,
                "words_per_thread": 8},
    "ap": {"q_pat": 128, "q_mark": 0, "q_med": 2048, "threads": 256,
           "words_per_thread": 8},
}


def tuning(family):
    t = dict(DEFAULT_TUNING[family])
    if TUNING_PATH.exists():
        try:
            t.update(json.loads(TUNING_PATH.read_text()).get(family, {}))
        except (OSError, ValueError):
            pass
    return t
