"""CPU oracle for the sieve-and-test path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this module, and only as the checker.  The product path (the HIP
library and its Python wrappers) never imports it and never falls back to
it.

Three layers, each a restatement of a reference function (file:line cited
per function):

* the sieve stage, in C (fk_oracle.c, loaded with ctypes): the three
  reference CPU engines' survivor streams;
* the strong-probable-prime test stage, in Python integers: huntlib's
  `mr_is_prime` / `sprp_base2` and the three families' run/depth rules;
* definition-level brute force for tiny windows (trial division of every
  value by every prime), used to pin the C sieve.

How each function is pinned (tests/test_oracle.py): the reference's frozen
fingerprints (dickson-ladders/score.py:54-59; primorial-ap/score.py:42-60),
its measured gate windows (primorial-ap/ap_gpu.py:697-701), its literature
tables and first occurrences (euler_reference.py:22-45,
ladder_reference.py:37-45, ap_reference.py:61-75), its discovery ledgers,
sympy's independent BPSW `isprime`, and brute force on tiny windows.
"""

import ctypes
import math
import os
import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libfk_oracle.so"

# huntlib/primes.py:17-20
MR_BASES = (2, 325, 9375, 28178, 450775, 9780504, 1795265022)
SMALL = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)

# euler-prime-runs/euler_search.py:56-60, :276
EULER_Q1, EULER_Q2, EULER_P_FLOOR, EULER_P_CEIL = 1024, 65536, 100_000, 10**24
# dickson-ladders/ladder_reference.py:57; ladder_search.py:52-53;
# dickson-ladders/launch.py:307
DICKSON_K_FLOOR, DICKSON_J_CEIL, SPRP_EXACT_FROM = 10**4, 4 * 10**18, 5
# primorial-ap/ap_reference.py:89-99; ap_search.py:63-64, :79-86
AP_P_FLOOR, AP_P_CEIL, AP_W0, DEPTH_EXACT_FROM = 10**4, 10**26, 30, 6
AP_W0_RESIDUES = (1, 7, 11, 13, 17, 19, 23, 29)

_lib = None


def build():
    """Compile the C oracle (gcc + OpenMP) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        u64, u32, i64, i32 = (ctypes.c_uint64, ctypes.c_uint32,
                              ctypes.c_int64, ctypes.c_int)
        p64, p32 = ctypes.POINTER(u64), ctypes.POINTER(u32)
        L.fko_primes.argtypes = [u32, p32, i64]
        L.fko_primes.restype = i64
        L.fko_euler_wheel.argtypes = [i32, p32, i32, p64, i64, p64]
        L.fko_euler_wheel.restype = i64
        L.fko_euler_survivors.argtypes = [i32, p32, i32, u32, u32, u64, u64,
                                          u64, u64, p64, i64, i32]
        L.fko_euler_survivors.restype = i64
        L.fko_dickson_forbidden.argtypes = [u32, i32, u64, p32]
        L.fko_dickson_forbidden.restype = i32
        L.fko_dickson_survivors.argtypes = [i32, u32, u64, u64, u64, u64,
                                            p64, i64, i32, u32]
        L.fko_dickson_survivors.restype = i64
        L.fko_ap_survivors.argtypes = [i32, u32, u64, u64, u64, p64, i64,
                                       i32, u32]
        L.fko_ap_survivors.restype = i64
        L.fko_dickson_ctx_create.argtypes = [i32, u32]
        L.fko_dickson_ctx_create.restype = ctypes.c_void_p
        L.fko_dickson_ctx_free.argtypes = [ctypes.c_void_p]
        L.fko_dickson_ctx_free.restype = None
        L.fko_dickson_ctx_survivors.argtypes = [ctypes.c_void_p, u64, u64,
                                                u64, u64, p64, i64, i32, u32]
        L.fko_dickson_ctx_survivors.restype = i64
        _lib = L
    return _lib


def _split(v):
    v = int(v)
    return v & (2**64 - 1), v >> 64


def _threads(nthreads):
    return int(nthreads or min(8, os.cpu_count() or 1))


def _call_growing(fn, pairs, cap=1 << 16):
    """Call an oracle sieve with a doubling output buffer until it fits."""
    while True:
        width = 2 if pairs else 1
        buf = (ctypes.c_uint64 * (cap * width))()
        got = fn(buf, cap)
        if got < 0:
            raise RuntimeError(f"oracle sieve failed with code {got}")
        if got <= cap:
            if pairs:
                return [buf[2 * i] | (buf[2 * i + 1] << 64)
                        for i in range(got)]
            return [buf[i] for i in range(got)]
        cap = int(got)


# ------------------------------------------------------------------ primes

def primes_below(limit):
    """Primes p < limit (sympy.primerange(2, limit) restated)."""
    n = int(limit)
    cap = max(16, int(1.3 * n / max(1.0, math.log(max(n, 3)))) + 16)
    buf = (ctypes.c_uint32 * cap)()
    got = lib().fko_primes(n, buf, cap)
    return [buf[i] for i in range(got)]


# ------------------------------------------------- test stage (Python ints)

def mr_is_prime(m, bases=MR_BASES):
    """huntlib/primes.py:23-47, line for line in meaning."""
    if m < 2:
        return False
    for sp in SMALL:
        if m % sp == 0:
            return m == sp
    d, s = m - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in bases:
        a %= m
        if a == 0:
            continue
        x = pow(a, d, m)
        if x == 1 or x == m - 1:
            continue
        for _ in range(s - 1):
            x = x * x % m
            if x == m - 1:
                break
        else:
            return False
    return True


def sprp_base2(m):
    """huntlib/primes.py:50-80."""
    return mr_is_prime(m, bases=(2,))


def euler_run(p, cap=100):
    """euler_search.py:258-262 mr_run_length: x from 0 while x^2+x+p is a
    strong probable prime to all seven bases, at most cap."""
    x = 0
    while x < cap and mr_is_prime(x * x + x + p):
        x += 1
    return x


def dickson_run(k, cap=64):
    """ladder_gpu.py:813-817 run_length: full-bases chain over m = 1.."""
    r, kk = 0, k * k
    while r < cap and mr_is_prime((r + 1) * kk + 1):
        r += 1
    return r


def dickson_sprp_run(k, cap, exact_from=SPRP_EXACT_FROM):
    """dickson-ladders/launch.py:310-342 sprp_run: base-2 chain, redone with
    the full bases only when it reaches exact_from."""
    kk, r = k * k, 0
    while r < cap and sprp_base2((r + 1) * kk + 1):
        r += 1
    if r < exact_from:
        return r
    r = 0
    while r < cap and mr_is_prime((r + 1) * kk + 1):
        r += 1
    return r


def primorial(n):
    """ap_reference.py:99-110 difference(n) = product of the first n primes."""
    out, c, q = 1, 0, 2
    while c < n:
        if all(q % d for d in range(2, int(q ** 0.5) + 1)):
            out *= q
            c += 1
        q += 1
    return out


def ap_chain_depth(p, n, cap=None, exact_from=DEPTH_EXACT_FROM):
    """ap_search.py:160-194 chain_depth: two-pass, capped at n."""
    d = primorial(n)
    cap = n if cap is None else int(cap)
    j = 0
    while j < cap and sprp_base2(p + j * d):
        j += 1
    if j < exact_from:
        return j
    j = 0
    while j < cap and mr_is_prime(p + j * d):
        j += 1
    return j


def ap_chain_depth_full(p, n, cap=None):
    """ap_reference.py:113-127 chain_depth, with huntlib's strong test."""
    d = primorial(n)
    cap = n if cap is None else int(cap)
    j = 0
    while j < cap and mr_is_prime(p + j * d):
        j += 1
    return j


# ------------------------------------------------------------ sieve stage

def wheel_modulus(n):
    """ladder_reference.py:60-66 W(n) = product of primes <= n+1."""
    w = 1
    for q in range(2, n + 2):
        if all(q % d for d in range(2, int(q ** 0.5) + 1)):
            w *= q
    return w


def euler_forbidden(q, n):
    """euler_search.py:85-87."""
    return {(-(x * x + x)) % q for x in range(n)}


def euler_wheel(n, wheel_primes):
    """euler_search.py:215-249 build_wheel: (sorted offsets, modulus)."""
    wp = (ctypes.c_uint32 * len(wheel_primes))(*wheel_primes)
    M = ctypes.c_uint64(0)
    cap = 1
    for q in wheel_primes:
        cap *= q
    cap = min(cap, 1 << 26)
    buf = (ctypes.c_uint64 * cap)()
    got = lib().fko_euler_wheel(int(n), wp, len(wheel_primes), buf, cap,
                                ctypes.byref(M))
    if got < 0 or got > cap:
        raise RuntimeError("oracle wheel build failed")
    return [buf[i] for i in range(got)], int(M.value)


EULER_WHEEL_29 = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]


def euler_survivors(n, lo, hi, wheel_primes=None, q1=EULER_Q1,
                    q2=EULER_Q2, nthreads=None):
    """euler_search.py:306-348: sorted survivor p in [lo, hi)."""
    wheel_primes = list(wheel_primes or EULER_WHEEL_29)
    lo, hi = int(lo), int(hi)
    if lo < EULER_P_FLOOR:
        raise ValueError("engine floor is %d" % EULER_P_FLOOR)
    if hi > EULER_P_CEIL:
        raise ValueError("engine ceiling is %d" % EULER_P_CEIL)
    wp = (ctypes.c_uint32 * len(wheel_primes))(*wheel_primes)

    def call(buf, cap):
        return lib().fko_euler_survivors(int(n), wp, len(wheel_primes),
                                         int(q1), int(q2), *_split(lo),
                                         *_split(hi), buf, cap,
                                         _threads(nthreads))
    return _call_growing(call, pairs=True)


def euler_hunt(n, lo, hi, cap=100, **kw):
    """euler_search.py:350-357: [(p, exact run)] with run >= n."""
    out = []
    for p in euler_survivors(n, lo, hi, **kw):
        if all(mr_is_prime(p + x * x + x) for x in range(n)):
            out.append((p, euler_run(p, cap)))
    return out


def dickson_forbidden(q, n):
    """ladder_search.py:71-89 table[q] (via Tonelli-Shanks in C)."""
    buf = (ctypes.c_uint32 * (2 * n + 2))()
    got = lib().fko_dickson_forbidden(int(q), int(n), wheel_modulus(n), buf)
    if got < 0:
        raise ValueError("bad prime for the Dickson table")
    return [buf[i] for i in range(got)]


def dickson_survivors(n, j_lo, j_hi, q2=65536, nthreads=None,
                      block=1 << 19):
    """ladder_search.py:92-112 survivors_j: sorted j in [j_lo, j_hi).

    The reference cuts blocks of 2^22 j; 2^19 keeps a block in a core's L2
    and gives the same stream (the cut is tested not to matter)."""
    j_lo, j_hi = int(j_lo), int(j_hi)
    if j_lo * wheel_modulus(n) < DICKSON_K_FLOOR:
        raise ValueError("engines refuse to run below K_FLOOR")

    def call(buf, cap):
        return lib().fko_dickson_survivors(int(n), int(q2), *_split(j_lo),
                                           *_split(j_hi), buf, cap,
                                           _threads(nthreads), int(block))
    return _call_growing(call, pairs=True)


class DicksonSieve:
    """ladder_search.CpuEngine's tables built once (ladder_search.py:56-69),
    then survivors_j (ladder_search.py:92-112) as often as asked."""

    def __init__(self, n, q2=65536):
        self.n, self.q2 = int(n), int(q2)
        self._c = lib().fko_dickson_ctx_create(self.n, self.q2)
        if not self._c:
            raise RuntimeError("oracle table build failed")

    def survivors(self, j_lo, j_hi, nthreads=None, block=1 << 20):
        def call(buf, cap):
            return lib().fko_dickson_ctx_survivors(
                self._c, *_split(j_lo), *_split(j_hi), buf, cap,
                _threads(nthreads), int(block))
        return _call_growing(call, pairs=True)

    def __del__(self):
        try:
            lib().fko_dickson_ctx_free(self._c)
        except Exception:
            pass


def ap_forbidden(q, n):
    """ap_search.py:102-104: {(-j*d) mod q : j < n}."""
    d = primorial(n)
    return sorted({(-j * d) % q for j in range(n)})


def ap_survives(n, p, q2=1 << 16):
    """ap_search.py:143-157 survives: p kept by the sieve (p is prime to
    30 and p mod q avoids the kill residues of every prime 7 <= q < q2),
    at any height, the engine floor included."""
    p = int(p)
    if p % 2 == 0 or p % 3 == 0 or p % 5 == 0:
        return False
    d = primorial(n)
    for q in primes_below(q2):
        if q < 7:
            continue
        # p is in {(-j d) mod q : j < n} iff some p + j d is 0 mod q
        r, dq = p % q, d % q
        if any((r + j * dq) % q == 0 for j in range(n if dq else 1)):
            return False
    return True


def ap_survivors(n, base, span, q2=1 << 16, nthreads=None, block=1 << 19):
    """ap_search.py:109-141: sorted offsets o, base + o surviving."""
    base, span = int(base), int(span)
    floor = max(AP_P_FLOOR, int(q2))
    if base < floor:
        raise ValueError(f"refuses to sieve below {floor}")
    if base + span > AP_P_CEIL:
        raise ValueError("past the enforced ceiling")

    def call(buf, cap):
        return lib().fko_ap_survivors(int(n), int(q2), *_split(base),
                                      int(span), buf, cap,
                                      _threads(nthreads), int(block))
    return _call_growing(call, pairs=False)


# ------------------------------------------ definition-level brute force

def _primes_py(limit):
    return [q for q in range(2, limit)
            if all(q % d for d in range(2, int(q ** 0.5) + 1))]


def euler_direct(n, lo, hi, q2=EULER_Q2):
    """A p survives iff no prime below q2 divides any x^2+x+p, x < n (the
    sieve definition behind euler_search.py:1-18; valid above P_FLOOR,
    where no value can equal the prime that divides it)."""
    ps = _primes_py(q2)
    return [p for p in range(lo, hi)
            if all((x * x + x + p) % q for q in ps for x in range(n))]


def dickson_direct(n, j_lo, j_hi, q2=65536):
    """ladder_reference.py:95-111 forbidden_k_residues, applied to k = W*j
    for every prime in [n+2, q2)."""
    W = wheel_modulus(n)
    ps = [q for q in _primes_py(q2) if q >= n + 2]
    out = []
    for j in range(j_lo, j_hi):
        k = W * j
        if all((m * k * k + 1) % q for q in ps for m in range(1, n + 1)):
            out.append(j)
    return out


def ap_direct(n, base, span, q2=1 << 16):
    """ap_reference.py:159-172 sieve_survivor over [base, base+span),
    returned as offsets from base."""
    d = primorial(n)
    ps = _primes_py(q2)
    out = []
    for o in range(span):
        p = base + o
        ok = True
        for q in ps:
            for j in range(n):
                v = p + j * d
                if v % q == 0 and v != q:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            out.append(o)
    return out
