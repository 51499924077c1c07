"""The engine's periodic kill bitmaps (first-knowns_amd/csrc/fk_plan.cpp,
built here with g++): the 64-bit window the sieve kernels funnel-shift out
of three consecutive words at x + 64 k, for any start x < q and every k
below tail / 64, holds exactly the kill bits of positions x + 64 k ..
x + 64 k + 63 taken mod q, for primes below and above 64 (below 64 a window
wraps the period more than once), with the line sieve's tail of 64 bits
and the short-line sieve's of 1024; bitmaps appended to one table keep
their own offsets."""
import pathlib
import random
import subprocess

import pytest

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE.parent / "first-knowns_amd" / "csrc"


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = tmp_path_factory.mktemp("plan") / "plan_harness"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    str(HERE / "plan_harness.cpp"), str(CSRC / "fk_plan.cpp")],
                   check=True)
    return exe


def run(harness, parts, tail=64):
    line = f"B {tail} " + " ; ".join(" ".join(str(v) for v in [q] + ks)
                                     for q, ks in parts) + "\n"
    out = subprocess.run([str(harness)], input=line, capture_output=True,
                         text=True, check=True).stdout.splitlines()
    offs = [int(v) for v in out[0].split()[1:]]
    words = [int(v) for v in out[1].split()]
    return offs, words


def window(words, base, x):
    """The kernels' read: words i, i+1, i+2 from base, i = x >> 5, shifted
    right by x & 31 (two 32-bit funnel shifts)."""
    i, o = x >> 5, x & 31
    w0, w1, w2 = (words[base + i + k] for k in range(3))
    lo = ((w1 << 32 | w0) >> o) & 0xffffffff
    hi = ((w2 << 32 | w1) >> o) & 0xffffffff
    return hi << 32 | lo


@pytest.mark.parametrize("tail", [64, 1024])
@pytest.mark.parametrize("q", [2, 3, 5, 17, 31, 37, 61, 67, 127, 131, 1021])
def test_window_is_the_periodic_kill_set(harness, q, tail):
    rng = random.Random(q)
    kills = sorted(rng.sample(range(q), min(q - 1, rng.randint(1, 12))))
    q2 = 29
    kills2 = [0, 5, 28]
    offs, words = run(harness, [(q, kills), (q2, kills2)], tail)
    assert offs == [0, (q - 1 + tail) // 32 + 1]
    assert len(words) == offs[1] + (q2 - 1 + tail) // 32 + 1
    for base, p, ks in ((offs[0], q, kills), (offs[1], q2, kills2)):
        kset = set(ks)
        for x in range(p):
            for k in range(tail // 64):
                want = sum(1 << b for b in range(64)
                           if (x + 64 * k + b) % p in kset)
                assert window(words, base, x + 64 * k) == want, (p, x, k)


def test_window_past_one_repeat_of_the_period(harness):
    # q = 17, kill residue 3: the window at x = 16 must hold the kills at
    # 20, 37, 54 and 71; a tail holding one repeat of the period (bits
    # q + r) would have only the first of them
    offs, words = run(harness, [(17, [3])])
    w = window(words, 0, 16)
    assert (w >> (37 - 16)) & 1 and (w >> (54 - 16)) & 1 and (w >> (71 - 16)) & 1


@pytest.mark.parametrize("family,n,wheel", [
    (1, 8, [2, 3, 5, 7, 11, 13]),
    (1, 17, [2, 3, 5, 7, 11, 13, 17]),
    (2, 10, [13, 17, 19]),
    (3, 12, [2, 3, 5, 37, 41]),
])
def test_factored_wheel_is_the_wheel(harness, family, n, wheel):
    """The factored wheel (the residues of all but the last prime, lifted
    over the last prime as the device expansion does) is the wheel: the
    same residues, and as many as wheel_count says; the wheel itself is
    every r < prod(wheel) that no wheel prime kills (brute force)."""
    import math
    import fk_oracle as O
    line = f"W {family} {n} " + " ".join(map(str, wheel)) + "\n"
    out = subprocess.run([str(harness)], input=line, capture_output=True,
                         text=True, check=True).stdout.splitlines()
    count = int(out[0].split()[1])
    full = [int(v) for v in out[1].split()]
    fact = [int(v) for v in out[2].split()]
    assert count == len(full) == len(fact) > 0
    assert sorted(fact) == full
    M = math.prod(wheel)

    def kills(q):
        if family == 1:
            return {(-(x * x + x)) % q for x in range(n)}
        if family == 2:
            return set(O.dickson_forbidden(q, n))
        return {(-j * O.primorial(n)) % q for j in range(n)}
    ks = {q: kills(q) for q in wheel}
    brute = [r for r in range(M) if all(r % q not in ks[q] for q in wheel)]
    assert brute == full
