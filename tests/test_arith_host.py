"""The device arithmetic header (first-knowns_amd/csrc/fk_arith.h), compiled
for the host with g++, against the oracle's Python-integer strong test and
sympy's BPSW: random odd values of 10..256 bits, primes on both sides of
every 32-bit word boundary and of every radix-2^29 limb-count boundary
(bit lengths 29 N - 2 and 29 N - 1), strong pseudoprimes to each single
base, and the small-prime edge cases."""
import pathlib
import random
import subprocess

import pytest

import fk_oracle as O

sympy = pytest.importorskip("sympy")
HERE = pathlib.Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = tmp_path_factory.mktemp("arith") / "arith_harness"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    str(HERE / "arith_harness.cpp")], check=True)
    return exe


def run(exe, cases):
    """cases: (nbases, value) from base 0, or (first, nbases, value)."""
    lines = []
    for c in cases:
        first, nb, v = c if len(c) == 3 else (0,) + tuple(c)
        limbs = [(v >> (32 * i)) & 0xFFFFFFFF for i in range(8)]
        lines.append(f"{first} {nb} " + " ".join("%x" % x for x in limbs))
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True)
    return [c == "1" for c in out.stdout.split()]


def test_strong_test_matches_python_integers(harness):
    rng = random.Random(9)
    vals = []
    for bits in (10, 20, 31, 32, 33, 40, 63, 64, 65, 80, 100, 104, 127,
                 128, 129, 150, 160, 191, 192, 193, 250, 255, 256):
        vals += [rng.getrandbits(bits) | 1 for _ in range(40)]
    for bits in (33, 64, 100, 128, 150, 192, 200, 256):
        vals += [int(sympy.nextprime(rng.getrandbits(bits)))
                 for _ in range(15)]
    vals += [0, 1, 2, 3, 4, 5, 37, 41, 43, 3215031751, 2**64 - 59,
             2**127 - 1, 2**128 - 159, 2**192 - 237, 2**255 - 19,
             2**256 - 189]
    cases = [(nb, v) for nb in (1, 7) for v in vals]
    got = run(harness, cases)
    for (nb, v), g in zip(cases, got):
        assert g == O.mr_is_prime(v, bases=O.MR_BASES[:nb]), (nb, v)


def test_strong_test_agrees_with_bpsw(harness):
    rng = random.Random(10)
    vals = [rng.getrandbits(rng.choice((40, 64, 90, 128, 170)))
            | 1 for _ in range(600)]
    got = run(harness, [(7, v) for v in vals])
    for v, g in zip(vals, got):
        assert g == bool(sympy.isprime(v)), v


def test_single_bases_and_limb_boundaries(harness):
    """Each base alone (the device runs one base per lane): composites that
    are strong pseudoprimes to that base, primes and random odd values at
    the bit lengths where the radix-2^29 limb count changes."""
    rng = random.Random(11)
    vals = []
    # composites with no factor <= 37 that pass some single base
    small = [p for p in range(3, 38) if sympy.isprime(p)]
    spsp = []
    for n in range(41, 400000, 2):
        if any(n % p == 0 for p in small) or sympy.isprime(n):
            continue
        if any(O.mr_is_prime(n, bases=[b]) for b in O.MR_BASES):
            spsp.append(n)
    assert len(spsp) > 20
    vals += spsp
    for nlimb in range(2, 10):
        for bits in (29 * nlimb - 3, 29 * nlimb - 2, 29 * nlimb - 1):
            if bits > 256:
                continue
            vals += [rng.getrandbits(bits) | 1 | (1 << (bits - 1))
                     for _ in range(6)]
            for _ in range(4):
                p = int(sympy.prevprime(1 << bits))
                vals.append(p)
                vals.append(int(sympy.nextprime(rng.getrandbits(bits - 1)
                                                | (1 << (bits - 2)))))
    cases = [(b, 1, v) for b in range(7) for v in vals]
    got = run(harness, cases)
    for (b, _, v), g in zip(cases, got):
        assert g == O.mr_is_prime(v, bases=[O.MR_BASES[b]]), (b, v)
