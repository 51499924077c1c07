"""The classification entry point that allocates no device memory
(fk_engine_classify_static, a fixed device array per device): through the
C ABI it gives exactly fk_engine_classify's runs on two calls in a row,
the second smaller than the first (the array keeps the first call's
contents), for all three families and both modes, and refuses more values
than the array holds."""
import sys

import pytest

import fk_native as N
import fk_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    # (family, n, q2, lo, hi, mode, cap, exact_from, kwargs): shallow
    # sieves, so that the windows hold thousands of values to classify
    (N.FAMILY_DICKSON, 10, 1024, 10**12, 10**12 + (1 << 27),
     N.CLASSIFY_FULL, 18, 0, {}),
    (N.FAMILY_AP, 12, 1024, 4 * 10**13 - 10, 4 * 10**13 - 10 + 10**8,
     N.CLASSIFY_TWO_PASS, 12, 6, {}),
    (N.FAMILY_EULER, 6, 1024, 10**12, 10**12 + 10**10,
     N.CLASSIFY_FULL, 100, 0, {"wheel_primes": [2, 3, 5, 7, 11, 13]}),
]


@pytest.mark.parametrize("case", CASES, ids=["dickson", "ap", "euler"])
def test_static_classification_equals_allocating_one(case):
    fam, n, q2, lo, hi, mode, cap, xf, kw = case
    eng = N.Engine(fam, n, q2, **kw)
    try:
        vals = N.record_values(eng.sweep(lo, hi))
        if fam == N.FAMILY_DICKSON:
            vals = [j * O.wheel_modulus(n) for j in vals]
        assert len(vals) >= 2000
        first, second = vals[:20_000], vals[-300:]
        for part in (first, second):
            say("static classify", fam, len(part))
            want = eng.classify(part, mode, cap, xf)
            got = eng.classify_static(part, mode, cap, xf)
            assert got == want
            assert max(got) >= 2
        with pytest.raises(ValueError):
            eng.classify_static([vals[0]] * (N.CLASSIFY_STATIC_MAX + 1), mode,
                                cap, xf)
        # the allocating path and the sweeps still work after it
        assert eng.classify(second, mode, cap, xf) == \
            eng.classify_static(second, mode, cap, xf)
    finally:
        eng.close()
