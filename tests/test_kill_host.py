"""The round-prime arithmetic of the kernels (first-knowns_amd/csrc/fk_kill.h
and the double-precision residue in fk_arith.h), compiled for the host with
g++, against Python integers: the residue of values up to 2^128 - 1 by
primes up to 2^30 through the 16-bit-limb fused-multiply-add path, and the
value-domain kill set of each family by brute force over every residue
(small q) and by checking each member and the count (large q)."""
import pathlib
import random
import subprocess

import pytest

import fk_oracle as O

HERE = pathlib.Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = tmp_path_factory.mktemp("kill") / "kill_harness"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    str(HERE / "kill_harness.cpp")], check=True)
    return exe


def call(exe, lines):
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True)
    return out.stdout.splitlines()


def is_prime(q):
    return q >= 2 and O.mr_is_prime(q)


def random_prime(rng, lo, hi):
    while True:
        q = rng.randrange(lo, hi) | 1
        if is_prime(q):
            return q


def limbs(v):
    return " ".join("%x" % ((v >> (32 * i)) & 0xFFFFFFFF) for i in range(4))


def test_residue_matches_python(harness):
    rng = random.Random(21)
    qs = [3, 5, 65521, 65537, 2**22 - 3, 2**24 - 3, 2**29 - 3, 2**30 - 35]
    qs += [random_prime(rng, 3, 2**30) for _ in range(40)]
    vals = [0, 1, 2**16 - 1, 2**16, 2**32 - 1, 2**63, 2**64 - 1, 2**64,
            2**66, 2**88 - 1, 2**96 + 12345, 2**128 - 1]
    vals += [rng.getrandbits(b) for b in (20, 40, 63, 64, 80, 100, 127, 128)
             for _ in range(6)]
    cases, want, wq = [], [], []
    for q in qs:
        # multiples of q and their neighbours: where a quotient estimate one
        # off shows
        near = [q * rng.getrandbits(b) + d for b in (20, 40, 60)
                for d in (-1, 0, 1)]
        # and around the half-multiples, where a nearest quotient is one off
        near += [q * rng.getrandbits(b) + q // 2 + d for b in (20, 40, 60)
                 for d in (-1, 0, 1, 2)]
        for v in vals + [x for x in near if 0 <= x < 2**128]:
            nl = 4 if v < 2**64 else 8
            for use in {nl, 8}:
                cases.append(f"R {q} {use} {limbs(v)}")
                want.append(v % q)
                wq.append(q)
    got = [tuple(int(x) for x in line.split()) for line in call(harness, cases)]
    assert [g[0] for g in got] == want          # resid_d (two fixes)
    assert [g[1] for g in got] == want          # resid_d1 (one fix)
    fold = [min(w, q - w) for w, q in zip(want, wq)]
    assert [g[2] for g in got] == fold          # resid_fold_d
    # the 22-bit limb form with 2^22, 2^44, 2^66 mod q, where it applies
    # (q < 2^24, values below 2^88)
    n22 = 0
    for g, w, q in zip(got, want, wq):
        if g[3] != 0xFFFFFFFF:
            assert (g[3], g[4]) == (w, min(w, q - w))
            n22 += 1
    assert n22 > 200


def test_residue_of_a_53_bit_value_in_one_double(harness):
    """The round kernels' one-limb form: a value below 2^53 as one double,
    for primes q >= 32 (then s / q < 2^48), at multiples of q, their
    neighbours and the half-multiples, up to 2^53 - 1."""
    rng = random.Random(53)
    qs = [32 + 5, 37, 65521, 2**22 - 3, 2**30 - 35]
    qs += [random_prime(rng, 32, 2**30) for _ in range(30)]
    cases, want, wq = [], [], []
    for q in qs:
        top = (2**53 - 1) // q
        vals = [2**53 - 1, 2**53 - 2, 0, 1, q - 1, q, q + 1]
        for m in [top, top - 1] + [rng.randrange(1, top) for _ in range(6)]:
            vals += [m * q + d for d in (-1, 0, 1, q // 2, q // 2 + 1)]
        for v in vals:
            if 0 <= v < 2**53:
                cases.append(f"N {q} {v & 0xFFFFFFFF:x} {v >> 32:x}")
                want.append(v % q)
                wq.append(q)
    got = [tuple(int(x) for x in line.split()) for line in call(harness, cases)]
    assert [g[0] for g in got] == want
    assert [g[1] for g in got] == [min(w, q - w) for w, q in zip(want, wq)]


def kill_brute(family, n, q):
    d = O.primorial(n) if family == 3 else 0
    out = set()
    for r in range(q):
        if family == 1:
            dead = any((x * x + x + r) % q == 0 for x in range(n))
        elif family == 2:
            dead = any((m * r * r + 1) % q == 0 for m in range(1, n + 1))
        else:
            dead = any((r + j * d) % q == 0 for j in range(n))
        if dead:
            out.add(r)
    return out


def run_kill(exe, family, n, qs):
    d = O.primorial(n) if family == 3 else 0
    lines = [f"K {family} {n} {q} {limbs(d)}" for q in qs]
    res = []
    for q, row in zip(qs, call(exe, lines)):
        w = [int(x) for x in row.split()]
        c = w[0]
        vals = w[1:1 + c]
        pad = w[1 + c:]
        assert len(set(vals)) == c
        assert all(x == 0xFFFFFFFF for x in pad)
        res.append(set(vals))
    return res


@pytest.mark.parametrize("family,n", [(1, 5), (1, 17), (2, 6), (2, 10),
                                      (2, 13), (3, 8), (3, 16)])
def test_kill_sets_small_primes_brute_force(harness, family, n):
    qs = [q for q in range(3, 700) if is_prime(q) and q > n + 1]
    got = run_kill(harness, family, n, qs)
    for q, g in zip(qs, got):
        want = kill_brute(family, n, q)
        if family == 2:
            # stored folded: one min(r, q - r) per root pair
            assert all(q - r in want for r in want) and 0 not in want
            want = {min(r, q - r) for r in want}
        assert g == want, q


@pytest.mark.parametrize("family,n", [(1, 17), (2, 10), (2, 13), (3, 17)])
def test_kill_sets_large_primes(harness, family, n):
    rng = random.Random(family * 100 + n)
    qs = [random_prime(rng, 2**16, 2**30) for _ in range(60)]
    qs += [2**30 - 35, 2**22 - 3, 1048573]
    got = run_kill(harness, family, n, qs)
    d = O.primorial(n)
    for q, g in zip(qs, got):
        if family == 1:
            assert g == {(-(x * x + x)) % q for x in range(n)}
        elif family == 3:
            assert g == {(-j * d) % q for j in range(n)}
        else:
            want = 0
            for m in range(1, n + 1):
                a = (-pow(m, -1, q)) % q
                if pow(a, (q - 1) // 2, q) == 1:
                    want += 2
            for r in g:
                assert 0 < r <= q // 2
                assert any((m * r * r + 1) % q == 0 for m in range(1, n + 1))
            # distinct m give distinct root pairs (-1/m differ for m < q),
            # stored folded: one per pair
            assert 2 * len(g) == want


@pytest.mark.parametrize("family,n", [(1, 17), (2, 10), (2, 13), (3, 8),
                                      (3, 17), (2, 24)])
def test_round_rows(harness, family, n):
    rng = random.Random(7 * family + n)
    qs = [random_prime(rng, 2**11, 2**30) for _ in range(30)] + [2053]
    d = O.primorial(n) if family == 3 else 0
    lines = [f"W {family} {n} {q} {limbs(d)}" for q in qs]
    kmax = n                      # Dickson rows hold folded root pairs
    for q, row in zip(qs, call(harness, lines)):
        w = row.split()
        stride = int(w[0])
        # header + kill set in groups of four; whole 64-byte lines past 32
        assert stride >= 12 + kmax and stride % 4 == 0
        assert stride == 12 + (kmax + 3) // 4 * 4 or (stride > 32 and stride % 16 == 0)
        assert stride < 12 + kmax + 16
        qd, qi, c16, c32, c48 = (float(x) for x in w[1:6])
        assert (qd, qi) == (float(q), 1.0 / q)
        assert (c16, c32, c48) == (2**16 % q, 2**32 % q, 2**48 % q)
        assert int(w[6]) == q
        cnt = int(w[7])
        res = [int(x) for x in w[8:]]
        assert len(res) == stride - 12
        assert len(set(res[:cnt])) == cnt
        assert all(x == 0xFFFFFFFF for x in res[cnt:])


def test_ap_index_rows(harness):
    """AP index rows: for a prime q not dividing d = P(n), the row's linear
    form gives t = -v / d mod q, and t < n exactly when q kills v (brute
    force over every residue for small q; members of the kill set, their
    neighbours and random values up to 2^128 for large q)."""
    rng = random.Random(31)
    cases, want, kill = [], [], []
    for n in (8, 12, 17):
        d = O.primorial(n)
        pn = [q for q in range(2, 200) if is_prime(q)][n - 1]
        small = [q for q in range(pn + 1, 400) if is_prime(q)]
        # round primes are below 2^30 (engines take q2 <= 2^30)
        large = [random_prime(rng, 2**16, 2**30) for _ in range(12)]
        large += [2**30 - 35, 2**29 - 3]
        for q in small + large:
            ks = {(-j * d) % q for j in range(n)}
            vals = list(range(q)) if q < 400 else (
                [k + q * rng.getrandbits(60) for k in ks] +
                [k + 1 + q * rng.getrandbits(100) for k in ks] +
                [rng.getrandbits(b) for b in (40, 64, 66, 87, 88, 90, 128)
                 for _ in range(8)])
            for v in vals:
                if v >= 2**128:
                    continue
                nl = 4 if v < 2**64 else 8
                cases.append(f"I {n} {q} {nl} {limbs(d)} {limbs(v)}")
                want.append((-v * pow(d, -1, q)) % q)
                kill.append(v % q in ks)
    got = [[int(x) for x in line.split()] for line in call(harness, cases)]
    assert [g[0] for g in got] == want
    assert [g[0] < g[2] for g in got] == kill
    # the 22-bit limb form, wherever it applies (q < 2^24, values < 2^88)
    checked22 = 0
    for g, w in zip(got, want):
        if g[3] != 0xFFFFFFFF:
            assert g[3] == w
            checked22 += 1
    assert checked22 > 1000
