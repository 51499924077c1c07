"""The reference's frozen benchmark fingerprints, reproduced by the HIP
engine at their full shapes: survivor count and XOR of the exact survivor
values (dickson-ladders/score.py:54-59 over 2^32 and 2^38 j,
primorial-ap/score.py:42-60 over 1.6e10 of p-line,
euler-prime-runs/score.py:57-77 over 5e14 and 2e16 of p-line at n = 17).
These windows are far too wide for the oracle to finish in a test; the
fingerprints are the reference's own record of its output there."""
import sys

import numpy as np
import pytest

import fk_native as N
from conftest import big

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def xor_all(vals):
    x = 0
    for v in vals:
        x ^= int(v)
    return x


@pytest.mark.parametrize("label", ["SCORE", "SCORE12", "SCORE13"])
def test_dickson_fingerprint(golden, label):
    import ladder_gpu
    s = next(x for x in golden["dickson_fingerprints"] if x["label"] == label)
    j0, span = big(s["j0"]), big(s["span"])
    say("dickson fingerprint", label)
    eng = ladder_gpu.GpuEngine(s["n"], q2=s["q2"])
    got = eng.survivors_j(j0, j0 + span)
    assert got.dtype == np.uint64
    assert (int(got.size), xor_all(got.tolist())) == (s["count"],
                                                      big(s["xor"]))


def test_ap_fingerprint(golden):
    import ap_gpu
    s = golden["ap_fingerprints"][0]
    say("ap fingerprint")
    eng = ap_gpu.GpuEngine(s["n"], s["q2"], launch_u=s["launch_u"])
    parts = [c for c in eng.survivors(big(s["base"]), big(s["span"]))]
    got = np.concatenate(parts) if parts else np.zeros(0, np.uint64)
    assert (int(got.size), xor_all(got.tolist())) == (s["count"],
                                                      big(s["xor"]))


@pytest.mark.parametrize("label", ["SCORE", "SCORE128", "SCORE_WIDE"])
def test_euler_fingerprint(golden, label):
    import euler_gpu
    s = next(x for x in golden["euler_fingerprints"] if x["label"] == label)
    lo, span = big(s["lo"]), big(s["span"])
    say("euler fingerprint", label)
    eng = euler_gpu.GpuEngine(s["n"])
    got = eng.survivors_pre_mr(lo, lo + span)
    assert (len(got), xor_all(got)) == (s["count"], big(s["xor"]))


WHEEL_31 = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31]


@pytest.mark.parametrize("label", ["SCORE", "SCORE_WIDE"])
def test_euler_fingerprint_on_the_31_wheel(golden, label):
    """The n = 17 fingerprints on the 31# wheel (29,937,600 residues built
    on the host) as well as on the engine's default 37# wheel (598,752,000
    residues: the 31# ones lifted over 37 on the device,
    euler-prime-runs/euler_search.py:113-187): the same stream."""
    import euler_gpu
    s = next(x for x in golden["euler_fingerprints"] if x["label"] == label)
    lo, span = big(s["lo"]), big(s["span"])
    say("euler fingerprint, 31# wheel", label)
    eng = euler_gpu.GpuEngine(s["n"], wheel_primes=WHEEL_31)
    assert eng.n_offs == 29_937_600 and eng.M == 200560490130
    got = eng.survivors_pre_mr(lo, lo + span)
    assert (len(got), xor_all(got)) == (s["count"], big(s["xor"]))
    default = euler_gpu.GpuEngine(s["n"])
    assert default.n_offs == 598_752_000 and default.M == 7420738134810
