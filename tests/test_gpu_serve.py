"""The serving loop on the device: a seeded Poisson mix of requests of all
three families (64-, 128- and 192-bit values), served with continuous
batching over several slots, gives every request exactly the records
(values and classified runs) of the same window swept alone."""
import sys

import pytest

import fk_native as N
import fk_serve as S

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    reqs = S.poisson_requests(5, 40, rate=2000.0, log2_span=(24, 31))
    print("serving", len(reqs), "requests", flush=True)
    sys.stdout.flush()
    srv = S.Server(slots=3, budget=8 << 30)
    st = srv.run(reqs)
    assert st["completed"] == len(reqs)
    assert max(srv.submissions) > 1          # requests shared submissions
    assert st["p99_ms"] >= st["p50_ms"] > 0
    total = 0
    for r in reqs:
        eng = srv.engines[(r.family, r.n, r.q2)]
        alone = eng.sweep(r.lo, r.hi, *r.classify)
        assert N.record_values(r.records) == N.record_values(alone), r.rid
        assert r.records["run"].tolist() == alone["run"].tolist(), r.rid
        total += alone.size
    assert total >= 20


def oracle_records(r):
    """The oracle's survivors and runs of one request (the reference CPU
    engines' sieve and huntlib's strong-test chains, restated)."""
    import fk_oracle as O
    mode, cap, xf = r.classify
    if r.family == "euler":
        vals = O.euler_survivors(r.n, r.lo, r.hi,
                                 wheel_primes=[2, 3, 5, 7, 11, 13], q2=r.q2)
        runs = [O.euler_run(p, cap=cap) for p in vals]
    elif r.family == "dickson":
        vals = O.dickson_survivors(r.n, r.lo, r.hi, q2=r.q2)
        W = O.wheel_modulus(r.n)
        runs = [O.dickson_sprp_run(j * W, cap, exact_from=xf) for j in vals]
    else:
        vals = [r.lo + o for o in O.ap_survivors(r.n, r.lo, r.hi - r.lo,
                                                 q2=r.q2)]
        runs = [O.ap_chain_depth(p, r.n, cap=cap, exact_from=xf)
                for p in vals]
    return vals, runs


def test_served_results_equal_the_oracle():
    """A served mix small and shallow enough for the oracle (spans 2^20 to
    2^25, depths 6 to 9, primes below 2^16): every request's records and
    runs are the oracle's."""
    reqs = S.poisson_requests(7, 36, rate=4000.0, log2_span=(20, 25),
                              depth=(6, 9))
    for r in reqs:
        r.q2 = 1 << 16
    print("serving (oracle)", len(reqs), "requests", flush=True)
    srv = S.Server(slots=3, budget=8 << 30)
    st = srv.run(reqs)
    assert st["completed"] == len(reqs)
    total = 0
    for r in reqs:
        vals, runs = oracle_records(r)
        assert N.record_values(r.records) == vals, r.rid
        assert r.records["run"].tolist() == runs, r.rid
        total += len(vals)
    assert total >= 200
    assert {r.family for r in reqs} == {"euler", "dickson", "ap"}
